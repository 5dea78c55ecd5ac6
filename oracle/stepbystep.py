"""TEST INFRASTRUCTURE: the reference's step-by-step explanation of one
matrix-free product, restated with explicit Python loops (2D only, small
meshes only).

Follows test/Demo_StepByStep.jl of the reference, step by step:
  A)   gather: free ids read x, Dirichlet ids read 0          (:57-69)
  B.1) vector -> tensor, one copy per gradient component r     (:83-85)
  B.2) contractions: Zk[2][r,i1,j2] += mats[1,r][i1,j1] Zk[1][r,j1,j2],
       Zk[3][r,i1,i2] += mats[2,r][i2,j2] Zk[2][r,i1,j2]       (:93-98)
  B.3) per point: c_r = sum_k (jq^T jq)[r,k] Zk[3][k], then
       Zk[3][r] = wq djq c_r                                   (:102-116)
  B.4) transposed contractions                                 (:119-125)
  B.5) tensor -> vector: yi[i] = sum_r Zk[1][r, I(i)]          (:129-131)
  C)   assembly: y[id] += yi[i] for free ids                   (:141-145)
with mats[d,r] = (d == r ? derivative table : value table)
(src/SumFactorizationMaps.jl:58) and jq = inverse transposed Jacobian,
djq = |det J| (src/LazyMatrices.jl:37-41), recomputed from the bilinear
cell map here.

The cell dofs of this build are already in tensor order (x index fastest,
1D nodes in the reference's order [0, 1, 1/p, ...]), so the reference's
dof_map is the identity: local dof i = j1 + n j2.

It shares no code with the C oracle (oracle/sumfac_oracle.c) or the brute
force (oracle/brute.py); the tests check the three against each other.
"""
import math


def _bilinear_jacobian(V, xi, eta):
    """J[a][b] = d x_a / d xi_b of the bilinear map through the cell's four
    vertices V[v] (v = b0 + 2 b1, lexicographic)."""
    dN_dxi = (-(1 - eta), (1 - eta), -eta, eta)
    dN_deta = (-(1 - xi), -xi, (1 - xi), xi)
    J = [[0.0, 0.0], [0.0, 0.0]]
    for v in range(4):
        for a in range(2):
            J[a][0] += dN_dxi[v] * V[v][a]
            J[a][1] += dN_deta[v] * V[v][a]
    return J


def apply_2d(cells, verts, x, tables):
    """y = A x for a 2D mesh, following the reference's step-by-step demo.

    cells: (ncell, n*n) global free ids (negative = Dirichlet), tensor order;
    verts: (ncell, 4, 2); tables = (B, Dm, points, weights), B and Dm of shape
    (q, n) with rows at the quadrature points."""
    B, Dm, pts, w = tables
    q, n = len(B), len(B[0])
    D = 2
    mats = [[Dm if d == r else B for r in range(D)] for d in range(D)]
    y = [0.0] * len(x)
    for cell in range(len(cells)):
        ids = [int(t) for t in cells[cell]]
        # A) gather
        xi = [x[i] if i >= 0 else 0.0 for i in ids]
        # B.1) vector -> tensor
        Z1 = [[[0.0] * n for _ in range(n)] for _ in range(D)]   # [r][j1][j2]
        for r in range(D):
            for i in range(n * n):
                Z1[r][i % n][i // n] = xi[i]
        # B.2) contractions
        Z2 = [[[0.0] * n for _ in range(q)] for _ in range(D)]   # [r][i1][j2]
        Z3 = [[[0.0] * q for _ in range(q)] for _ in range(D)]   # [r][i1][i2]
        for r in range(D):
            for i1 in range(q):
                for j1 in range(n):
                    for j2 in range(n):
                        Z2[r][i1][j2] += mats[0][r][i1][j1] * Z1[r][j1][j2]
        for r in range(D):
            for i1 in range(q):
                for i2 in range(q):
                    for j2 in range(n):
                        Z3[r][i1][i2] += mats[1][r][i2][j2] * Z2[r][i1][j2]
        # B.3) weights and Jacobian
        V = [[float(verts[cell][v][a]) for a in range(D)] for v in range(4)]
        for i1 in range(q):
            for i2 in range(q):
                J = _bilinear_jacobian(V, pts[i1], pts[i2])
                det = J[0][0] * J[1][1] - J[0][1] * J[1][0]
                # jq = inv(J)^T; change[r][k] = sum_l jq[l][r] jq[l][k]
                inv = [[J[1][1] / det, -J[0][1] / det], [-J[1][0] / det, J[0][0] / det]]
                jq = [[inv[b][a] for b in range(D)] for a in range(D)]
                change = [[sum(jq[l][r] * jq[l][k] for l in range(D)) for k in range(D)]
                          for r in range(D)]
                cr = [sum(change[r][k] * Z3[k][i1][i2] for k in range(D)) for r in range(D)]
                wq = w[i1] * w[i2]
                djq = math.fabs(det)
                for r in range(D):
                    Z3[r][i1][i2] = wq * djq * cr[r]
        # B.4) transposed contractions
        Z2 = [[[0.0] * n for _ in range(q)] for _ in range(D)]
        Z1 = [[[0.0] * n for _ in range(n)] for _ in range(D)]
        for r in range(D):
            for i1 in range(q):
                for i2 in range(q):
                    for j2 in range(n):
                        Z2[r][i1][j2] += mats[1][r][i2][j2] * Z3[r][i1][i2]
        for r in range(D):
            for i1 in range(q):
                for j1 in range(n):
                    for j2 in range(n):
                        Z1[r][j1][j2] += mats[0][r][i1][j1] * Z2[r][i1][j2]
        # B.5) tensor -> vector
        yi = [0.0] * (n * n)
        for r in range(D):
            for i in range(n * n):
                yi[i] += Z1[r][i % n][i // n]
        # C) assembly
        for i, gid in enumerate(ids):
            if gid >= 0:
                y[gid] += yi[i]
    return y
