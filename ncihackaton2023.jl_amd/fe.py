"""Operator descriptor: 1D tables and the tensor dof map.

Mirrors SumFactorizationMap{D,SB,SQ} (src/SumFactorizationMaps.jl:2-65):
``SumFactorizationMap(D, fe_orders, quad_orders)`` takes Gridap-style
quadrature *degrees* and builds, for the isotropic case the device supports,
the q x n interpolation matrix A0 (``evalMats``), the derivative matrix B0 and
the D x D ``gradMats`` grid (``gradMats[d][k] = B0 if d == k else A0``).
The tables come from the library (sf_tables_1d), so the descriptor and the
device use one construction.
"""
import numpy as np

from . import _lib


def gauss_points_for_degree(degree):
    """Gauss points per direction for a quadrature degree: ceil((deg+1)/2),
    Gridap's rule behind Quadrature(SEGMENT, degree)
    (src/SumFactorizationMaps.jl:51; test/Demo_StepByStep.jl:53)."""
    return int(_lib.load().sf_gauss_points_for_degree(int(degree)))


def tables_1d(order, nquad):
    """(A0, B0, points, weights) for Lagrange order ``order`` on the 1D node
    order [0, 1, 1/p, ..., (p-1)/p] at ``nquad`` Gauss points of [0, 1]."""
    n = order + 1
    A0 = np.zeros((nquad, n))
    B0 = np.zeros((nquad, n))
    pts = np.zeros(nquad)
    w = np.zeros(nquad)
    L = _lib.load()
    _lib.check(L.sf_tables_1d(order, nquad, A0.ctypes.data, B0.ctypes.data,
                              pts.ctypes.data, w.ctypes.data))
    return A0, B0, pts, w


def node_offsets(order):
    """Lattice offset (units of h/p) of each position of the 1D node order."""
    p = order
    return np.array([0, p] + list(range(1, p)), dtype=np.int64)


class SumFactorizationMap:
    """Sum-factorisation descriptor (isotropic orders).

    Attributes mirror the reference: ``mats = [evalMats, gradMats]`` and
    ``dof_map`` (here the identity on tensor-lexicographic local dofs: the
    cell dof ids this package produces are already in tensor order; a Gridap
    caller passes its own local-to-tensor map to ``sf_plan_set_cells``)."""

    def __init__(self, D, fe_orders, quad_orders):
        fe_orders = tuple(fe_orders) if np.ndim(fe_orders) else (fe_orders,) * D
        quad_orders = tuple(quad_orders) if np.ndim(quad_orders) else (quad_orders,) * D
        if D not in (2, 3):
            raise ValueError("D must be 2 or 3")
        if len(set(fe_orders)) != 1 or len(set(quad_orders)) != 1:
            raise ValueError("only isotropic orders are supported on the device "
                             "(as in src/SumFactorizationMaps.jl:25-32)")
        self.D = D
        self.order = int(fe_orders[0])
        self.quad_degree = int(quad_orders[0])
        self.nquad = gauss_points_for_degree(self.quad_degree)
        A0, B0, pts, w = tables_1d(self.order, self.nquad)
        self.A0, self.B0, self.points, self.weights = A0, B0, pts, w
        self.evalMats = [A0] * D
        self.gradMats = [[B0 if d == k else A0 for k in range(D)] for d in range(D)]
        self.mats = [self.evalMats, self.gradMats]
        self.SB = (self.order + 1,) * D
        self.SQ = (self.nquad,) * D
        self.dof_map = np.arange((self.order + 1) ** D, dtype=np.int32)

    @classmethod
    def with_points(cls, D, order, nquad):
        """Descriptor with ``nquad`` Gauss points per direction (degree 2q-1)."""
        return cls(D, (order,) * D, (2 * nquad - 1,) * D)

    @property
    def tables(self):
        return (self.A0, self.B0, self.points, self.weights)


def get_dims(m):
    return m.D


def get_basis_sizes(m):
    return m.SB


def get_quads_sizes(m):
    return m.SQ


def get_dimensional_parameters(m):
    """(D, SB, SQ), as src/SumFactorizationMaps.jl:15-17."""
    return get_dims(m), get_basis_sizes(m), get_quads_sizes(m)


def count_manual_flops_poisson_matrix_free(ncells, D, SQ, SB, count_only_fma_ops=True,
                                           structured_mesh_opt=True):
    """The reference's analytic FLOP model (src/Utils.jl:33-72), evaluated
    by the library (isotropic SQ, SB)."""
    return int(_lib.load().sf_flops_model(int(ncells), int(D), int(SQ[0]), int(SB[0]),
                                          int(bool(count_only_fma_ops)),
                                          int(bool(structured_mesh_opt))))

