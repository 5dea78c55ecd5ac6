"""TEST INFRASTRUCTURE ONLY -- brute-force (non sum-factorised) numpy oracle.

Independent of both the C restatement (oracle/sumfac_oracle.c) and the
product library:

* 1D tables by a Legendre-Vandermonde change of basis (the reference builds
  its Lagrange basis through a change of basis from a polynomial basis too,
  SumFactorizationMaps.jl:50 via Gridap's LagrangianRefFE), evaluated at
  numpy's Gauss-Legendre points (SumFactorizationMaps.jl:51);
* dense element gradient matrices grad(phi_a)(xi_t) = kron products of the 1D
  tables (no factorised contractions);
* geometry from numpy.linalg on J = dx/dxi (not Gridap's transposed layout);
* element stiffness matrices K_e = sum_t Grad_t^T G_t Grad_t, their
  assembly into scipy CSR -- the "assembled SpMV" side of the reference's
  comparison (test/Demo.jl:39-57, profiling/drivers.jl:7-26);
* the Kronecker-sum closed form of the operator on a uniform Cartesian grid.
"""
import numpy as np
import scipy.sparse as sp
from numpy.polynomial import legendre as npleg


def nodes_1d(p):
    """Reference 1D node order: [0, 1, 1/p, ..., (p-1)/p]."""
    if p == 0:
        return np.array([0.5])
    return np.array([0.0, 1.0] + [k / p for k in range(1, p)])


def node_offsets(p):
    """Lattice offset (in units of h/p) of each 1D node position."""
    return np.array([0, p] + list(range(1, p)), dtype=np.int64)


def gauss01(q):
    x, w = npleg.leggauss(q)
    return 0.5 * (x + 1.0), 0.5 * w


def tables(p, q):
    """(B, Dm, points, weights): B[i, j] = phi_j(xi_i), Dm = phi_j'(xi_i)."""
    t = nodes_1d(p)
    n = p + 1
    # phi_j = sum_k C[k, j] P_k(2x - 1); V[i, k] = P_k(2 t_i - 1)
    V = npleg.legvander(2.0 * t - 1.0, n - 1)
    C = np.linalg.solve(V, np.eye(n))
    pts, w = gauss01(q)
    Vq = npleg.legvander(2.0 * pts - 1.0, n - 1)
    # derivative: d/dx P_k(2x-1) = 2 P_k'(2x-1)
    dV = np.zeros((q, n))
    for k in range(n):
        ck = np.zeros(n)
        ck[k] = 1.0
        dV[:, k] = 2.0 * npleg.legval(2.0 * pts - 1.0, npleg.legder(ck))
    return Vq @ C, dV @ C, pts, w


def grad_matrices(dim, B, Dm):
    """Dense reference gradients: list over directions of (q^D, n^D) arrays,
    row t = quadrature point (x fastest), column a = local dof (x fastest)."""
    if dim == 2:
        return [np.kron(B, Dm), np.kron(Dm, B)]
    return [np.kron(B, np.kron(B, Dm)), np.kron(B, np.kron(Dm, B)),
            np.kron(Dm, np.kron(B, B))]


def geometry_numpy(dim, verts, pts, w):
    """(G, detJ): G[cell, t, r, k] = w_t |det J| (J^T J)^{-1}[r, k]."""
    verts = np.asarray(verts, dtype=np.float64)
    q = len(pts)
    nv = 2 ** dim
    if dim == 2:
        xi = np.stack(np.meshgrid(pts, pts, indexing="xy"), -1).reshape(-1, 2)
        wq = np.outer(w, w).reshape(-1)
    else:
        zz, yy, xx = np.meshgrid(pts, pts, pts, indexing="ij")
        xi = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], -1)
        wq = np.einsum("l,k,i->lki", w, w, w).reshape(-1)
    # dN[t, v, b] = d N_v / d xi_b
    dN = np.zeros((q ** dim, nv, dim))
    for v in range(nv):
        bits = [(v >> d) & 1 for d in range(dim)]
        s = np.stack([xi[:, d] if bits[d] else 1.0 - xi[:, d] for d in range(dim)], -1)
        ds = np.array([1.0 if bits[d] else -1.0 for d in range(dim)])
        for b in range(dim):
            col = ds[b] * np.ones(len(xi))
            for d in range(dim):
                if d != b:
                    col = col * s[:, d]
            dN[:, v, b] = col
    # J[cell, t, a, b] = sum_v X[cell, v, a] dN[t, v, b]
    J = np.einsum("cva,tvb->ctab", verts, dN)
    det = np.linalg.det(J)
    Ji = np.linalg.inv(J)
    G = np.einsum("t,ct,ctrl,ctkl->ctrk", wq, np.abs(det), Ji, Ji)
    return G, det


def element_matrices(dim, B, Dm, G):
    """K[cell] = sum_t Grad_t^T G_t Grad_t (dense, no factorisation)."""
    Gr = np.stack(grad_matrices(dim, B, Dm))  # (D, nq, nl)
    G = np.asarray(G)
    C, nl = G.shape[0], Gr.shape[2]
    K = np.zeros((C, nl, nl))
    # K[c] += Grad_r^T diag(G[c, :, r, k]) Grad_k, as batched matrix products
    # in chunks of cells (a plain einsum may pick a slow contraction order)
    for c0 in range(0, C, 512):
        c1 = min(C, c0 + 512)
        for r in range(dim):
            for k in range(dim):
                tmp = G[c0:c1, :, r, k][:, :, None] * Gr[k][None, :, :]
                K[c0:c1] += np.matmul(Gr[r].T[None, :, :], tmp)
    return K


def diagonal(dim, B, Dm, G, cells, nfree):
    """diag(A) on the free dofs: the element-matrix diagonals
    K_e[a, a] = sum_t Grad_t[:, a]^T G_t Grad_t[:, a] (dense reference
    gradients, no factorisation), summed over the cells that hold dof a."""
    Gr = np.stack(grad_matrices(dim, B, Dm))  # (D, nq, nl)
    de = np.einsum("rta,ctrk,kta->ca", Gr, np.asarray(G), Gr)
    cells = np.asarray(cells)
    d = np.zeros(nfree)
    m = cells >= 0
    np.add.at(d, cells[m], de[m])
    return d


def gll_nodes(p):
    """Gauss-Lobatto-Legendre points on [0, 1] in the reference's node order
    [0, 1, interior ascending]: the interior ones are the roots of P_p'
    (numpy's Legendre roots)."""
    if p == 1:
        return np.array([0.0, 1.0])
    c = np.zeros(p + 1)
    c[-1] = 1.0
    inner = np.sort(npleg.legroots(npleg.legder(c)))
    return np.array([0.0, 1.0] + list(0.5 * (inner + 1.0)))


def lagrange_matrix(nodes, x):
    """L[i, j] = l_j(x_i), the Lagrange polynomials of `nodes` at x."""
    nodes = np.asarray(nodes)
    x = np.asarray(x)
    L = np.ones((x.size, nodes.size))
    for j in range(nodes.size):
        for m in range(nodes.size):
            if m != j:
                L[:, j] *= (x - nodes[m]) / (nodes[j] - nodes[m])
    return L


def gll_to_equispaced(p):
    """T1[a, b] = l_b(t_a): Gauss-Lobatto Lagrange polynomials at the
    equispaced nodes (both in the reference's order)."""
    return lagrange_matrix(gll_nodes(p), nodes_1d(p))


def global_interp(dim, cells, T1, nfree):
    """The global (free x free) matrix of the per-cell tensor product of T1:
    the entries every cell holding a pair of dofs agrees on."""
    Te = T1
    for _ in range(dim - 1):
        Te = np.kron(T1, Te)
    cells = np.asarray(cells)
    rows, cols, vals = [], [], []
    for c in cells:
        a_idx = np.nonzero(c >= 0)[0]
        blk = Te[np.ix_(a_idx, a_idx)]
        nz = np.abs(blk) > 0
        ia, ib = np.nonzero(nz)
        rows.append(c[a_idx][ia])
        cols.append(c[a_idx][ib])
        vals.append(blk[ia, ib])
    rows, cols, vals = (np.concatenate(t) for t in (rows, cols, vals))
    T = sp.coo_matrix((vals, (rows, cols)), shape=(nfree, nfree)).tocsr()
    cnt = sp.coo_matrix((np.ones_like(vals), (rows, cols)), shape=(nfree, nfree)).tocsr()
    T.data /= cnt.data  # duplicates were summed; every cell gives the same value
    return T


def gll_preconditioner(dim, p, B, Dm, G, cells, nfree):
    """(T, dg): z = T (T^T r / dg) with dg = diag(T^T A T), A assembled from
    the brute-force element matrices."""
    T = global_interp(dim, cells, gll_to_equispaced(p), nfree)
    A = assemble_csr(cells, element_matrices(dim, B, Dm, G), nfree)
    dg = (T.T @ A @ T).diagonal()
    return T, dg


def apply_dense(cells, K, x):
    """y = sum_e P_e^T K_e P_e x with Dirichlet ids (< 0) dropped."""
    cells = np.asarray(cells)
    xl = np.where(cells >= 0, x[np.maximum(cells, 0)], 0.0)
    yl = np.einsum("cab,cb->ca", K, xl)
    y = np.zeros_like(x)
    m = cells >= 0
    np.add.at(y, cells[m], yl[m])
    return y


def assemble_csr(cells, K, nfree):
    """Assembled sparse operator on the free dofs (scipy CSR)."""
    cells = np.asarray(cells)
    nl = cells.shape[1]
    rows = np.repeat(cells, nl, axis=1).reshape(-1)
    cols = np.tile(cells, (1, nl)).reshape(-1)
    vals = K.reshape(-1)
    m = (rows >= 0) & (cols >= 0)
    A = sp.coo_matrix((vals[m], (rows[m], cols[m])), shape=(nfree, nfree))
    return A.tocsr()


def kron_operator(dim, p, q, m):
    """Closed form on the uniform m^D grid of [0,1]^D with homogeneous
    Dirichlet boundary:  A = sum_d  M (x) .. K (direction d) .. (x) M, with the
    1D stiffness K1 and mass M1 assembled from the same Gauss rule and
    restricted to interior lattice nodes (free id = sum_d (L_d - 1) F^d)."""
    B, Dm, pts, w = tables(p, q)
    h = 1.0 / m
    K1e = (Dm.T * w) @ Dm / h
    M1e = (B.T * w) @ B * h
    nL = p * m + 1
    off = node_offsets(p)
    K1 = np.zeros((nL, nL))
    M1 = np.zeros((nL, nL))
    for e in range(m):
        L = e * p + off
        K1[np.ix_(L, L)] += K1e
        M1[np.ix_(L, L)] += M1e
    K1 = sp.csr_matrix(K1[1:-1, 1:-1])
    M1 = sp.csr_matrix(M1[1:-1, 1:-1])
    if dim == 2:
        return (sp.kron(M1, K1) + sp.kron(K1, M1)).tocsr()
    return (sp.kron(M1, sp.kron(M1, K1)) + sp.kron(M1, sp.kron(K1, M1)) +
            sp.kron(K1, sp.kron(M1, M1))).tocsr()
