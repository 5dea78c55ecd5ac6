"""Meshes, dof numbering and seeded synthetic inputs.

Plays the role of the Gridap set-up in the reference drivers
(CartesianDiscreteModel, FESpace with dirichlet_tags=["boundary"],
get_cell_dof_ids; profiling/drivers.jl:7-26, test/Demo.jl:25-36):

* ``CartesianModel(D, m, perturb, seed)`` -- the unit square/cube split into
  m^D cells; with ``perturb > 0`` every interior vertex coordinate moves by
  U[-perturb*h, perturb*h] (boundary vertices fixed), so Jacobians vary.
* ``FESpace(model, order, dirichlet=True)`` -- Lagrange Q_p dofs on the
  (p m + 1)^D node lattice; boundary nodes are Dirichlet (id -1), free ids
  are lexicographic over interior lattice nodes (x fastest).  Cell dofs are
  listed tensor-lexicographically in the reference's 1D node order
  [0, 1, 1/p, ..., (p-1)/p].

Random numbers come from a counter-based SplitMix64 stream (seed, stream)
so any host (Python here, C/Julia elsewhere) can reproduce the same bits:
u_i = (splitmix64(state_0 + (i + 1) * golden) >> 11) * 2^-53.
"""
import numpy as np

from .fe import node_offsets

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_STREAM = np.uint64(0xD1B54A32D192ED03)


def _mix(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed, n, stream=0):
    """n consecutive SplitMix64 outputs of the stream (seed, stream)."""
    with np.errstate(over="ignore"):
        s0 = _mix(np.uint64(seed) ^ (np.uint64(stream) * _STREAM))
        i = np.arange(1, n + 1, dtype=np.uint64)
        return _mix(s0 + i * _GOLDEN)


def uniform(seed, n, lo=0.0, hi=1.0, stream=0):
    """n doubles uniform in [lo, hi)."""
    u = (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def normal(seed, n, stream=0):
    """n standard normals (Box-Muller on consecutive uniform pairs)."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, stream=stream)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:n]


# fixed stream ids of the synthetic inputs
STREAM_MESH, STREAM_U, STREAM_F, STREAM_G, STREAM_D, STREAM_DIR = 1, 0, 2, 3, 4, 5


class CartesianModel:
    """m^D cells on [0,1]^D, vertices optionally perturbed (interior only)."""

    def __init__(self, D, m, perturb=0.0, seed=1):
        if D not in (2, 3):
            raise ValueError("D must be 2 or 3")
        self.D, self.m, self.perturb, self.seed = D, int(m), float(perturb), seed
        nv1 = self.m + 1
        h = 1.0 / self.m
        g = np.arange(nv1) * h
        if D == 2:
            yy, xx = np.meshgrid(g, g, indexing="ij")
            V = np.stack([xx.ravel(), yy.ravel()], -1)
            idx = np.stack(np.meshgrid(np.arange(nv1), np.arange(nv1), indexing="ij")[::-1], -1).reshape(-1, 2)
        else:
            zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
            V = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], -1)
            I = np.meshgrid(np.arange(nv1), np.arange(nv1), np.arange(nv1), indexing="ij")
            idx = np.stack([I[2].ravel(), I[1].ravel(), I[0].ravel()], -1)
        if self.perturb > 0.0:
            interior = np.all((idx > 0) & (idx < self.m), axis=1)
            d = uniform(seed, V.size, -self.perturb * h, self.perturb * h,
                        stream=STREAM_MESH).reshape(V.shape)
            V = V + d * interior[:, None]
        self.vertices = V  # ((m+1)^D, D), x fastest
        self.n_cells = self.m ** D

    def cell_vertex_ids(self):
        """(n_cells, 2^D) vertex ids, cells and vertices x fastest."""
        m, nv1, D = self.m, self.m + 1, self.D
        c = np.arange(m)
        if D == 2:
            cy, cx = np.meshgrid(c, c, indexing="ij")
            base = (cy * nv1 + cx).reshape(-1)
            offs = np.array([0, 1, nv1, nv1 + 1])
        else:
            cz, cy, cx = np.meshgrid(c, c, c, indexing="ij")
            base = ((cz * nv1 + cy) * nv1 + cx).reshape(-1)
            offs = np.array([0, 1, nv1, nv1 + 1, nv1 * nv1, nv1 * nv1 + 1,
                             nv1 * nv1 + nv1, nv1 * nv1 + nv1 + 1])
        return base[:, None] + offs[None, :]

    def cell_vertices(self):
        """(n_cells, 2^D, D) vertex coordinates (the library's input layout)."""
        return np.ascontiguousarray(self.vertices[self.cell_vertex_ids()])


class FESpace:
    """Q_p Lagrange space on a CartesianModel with (optional) homogeneous
    Dirichlet conditions on the whole boundary."""

    def __init__(self, model, order, dirichlet=True):
        self.model, self.order, self.dirichlet = model, int(order), bool(dirichlet)
        D, m, p = model.D, model.m, self.order
        self.D = D
        self.n = p + 1
        nL = p * m + 1
        self.n_lattice = nL
        F = nL - 2 if dirichlet else nL
        self.n_free = F ** D
        off = node_offsets(p)  # lattice offset of each 1D position
        c = np.arange(m)
        # lattice coordinate along one axis for (cell index, local position)
        L1 = c[:, None] * p + off[None, :]  # (m, n)
        if dirichlet:
            free1 = (L1 > 0) & (L1 < nL - 1)
            id1 = np.where(free1, L1 - 1, -1)
        else:
            id1 = L1
        n = self.n
        if D == 2:
            # cells (cy, cx) x local (b, a): id = idy * F + idx
            iy = id1[:, None, :, None]  # cy, -, b, -
            ix = id1[None, :, None, :]  # -, cx, -, a
            ids = np.where((iy >= 0) & (ix >= 0), iy * F + ix, -1)
            self.cell_dofs = ids.reshape(m * m, n * n).astype(np.int32)
        else:
            iz = id1[:, None, None, :, None, None]
            iy = id1[None, :, None, None, :, None]
            ix = id1[None, None, :, None, None, :]
            ids = np.where((iz >= 0) & (iy >= 0) & (ix >= 0), (iz * F + iy) * F + ix, -1)
            self.cell_dofs = ids.reshape(m ** 3, n ** 3).astype(np.int32)
        self.n_cells = model.n_cells

    def free_dof_coordinates(self):
        """Physical coordinates (n_free, D) of the free nodes, through each
        cell's multilinear map (used for patch tests)."""
        D, p = self.D, self.order
        t = np.array([0.0, 1.0] + [k / p for k in range(1, p)])
        Xc = self.model.cell_vertices()
        n = self.n
        if D == 2:
            b, a = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
            xi = np.stack([t[a.ravel()], t[b.ravel()]], -1)
        else:
            c3, b, a = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
            xi = np.stack([t[a.ravel()], t[b.ravel()], t[c3.ravel()]], -1)
        nv = 2 ** D
        N = np.ones((xi.shape[0], nv))
        for v in range(nv):
            for d in range(D):
                N[:, v] *= xi[:, d] if (v >> d) & 1 else 1.0 - xi[:, d]
        X = np.einsum("tv,cvd->ctd", N, Xc)
        coords = np.zeros((self.n_free, D))
        m = self.cell_dofs >= 0
        coords[self.cell_dofs[m]] = X[m]
        return coords
