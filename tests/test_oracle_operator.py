"""Pins of the CPU oracle (oracle/sumfac_oracle.c, the restatement of the
reference's per-cell pipeline) against independent computations:

* brute force: dense element matrices, no sum factorisation, numpy geometry;
* the Kronecker-sum closed form on uniform Cartesian grids;
* textbook Q1 element matrices;
* invariants on perturbed meshes: A 1 = 0, patch test, symmetry, volume;
* the reference's own check, test/Demo.jl:57 (assembled vs lazy, x = 1);
* planted mistakes (in the oracle's inputs, on the CPU) that the checks catch.
"""
import numpy as np
import pytest

from conftest import rel_l2


def setup(pkg, D, m, p, q, perturb=0.2, seed=1, dirichlet=True):
    model = pkg.CartesianModel(D, m, perturb=perturb, seed=seed)
    space = pkg.FESpace(model, p, dirichlet=dirichlet)
    mp = pkg.SumFactorizationMap.with_points(D, p, q)
    return model, space, mp


CASES = [(2, 4, 1, 2), (2, 3, 2, 4), (2, 3, 4, 6), (2, 2, 7, 9), (3, 3, 1, 2),
         (3, 3, 2, 4), (3, 2, 4, 6), (3, 2, 3, 3), (3, 2, 5, 4)]


@pytest.mark.parametrize("D,m,p,q", CASES)
def test_oracle_vs_bruteforce_perturbed(pkg, oracle_mod, brute, D, m, p, q):
    model, space, mp = setup(pkg, D, m, p, q)
    X = model.cell_vertices()
    x = pkg.uniform(0, space.n_free, -1, 1)
    y = oracle_mod.apply(D, p + 1, q, space.cell_dofs, X, x, mp.tables)
    B, Dm, pts, w = brute.tables(p, q)
    G, det = brute.geometry_numpy(D, X, pts, w)
    assert np.all(det > 0)
    K = brute.element_matrices(D, B, Dm, G)
    y_ref = brute.apply_dense(space.cell_dofs, K, x)
    assert rel_l2(y, y_ref) < 1e-13
    # the assembled SpMV gives the same product
    A = brute.assemble_csr(space.cell_dofs, K, space.n_free)
    assert rel_l2(A @ x, y_ref) < 1e-13


@pytest.mark.parametrize("D,m,p,q", [(2, 5, 3, 5), (2, 4, 4, 6), (3, 3, 2, 4), (3, 2, 4, 6),
                                     (3, 2, 3, 3)])
def test_oracle_vs_kronecker_closed_form(pkg, oracle_mod, brute, D, m, p, q):
    model, space, mp = setup(pkg, D, m, p, q, perturb=0.0)
    x = pkg.uniform(3, space.n_free, -1, 1)
    y = oracle_mod.apply(D, p + 1, q, space.cell_dofs, model.cell_vertices(), x, mp.tables)
    A = brute.kron_operator(D, p, q, m)
    assert rel_l2(y, A @ x) < 1e-13


def test_q1_textbook_square(pkg, oracle_mod):
    # one cell [0,1]^2, all 4 dofs free, exact quadrature (2 points)
    model, space, mp = setup(pkg, 2, 1, 1, 2, perturb=0.0, dirichlet=False)
    K = np.stack([oracle_mod.apply(2, 2, 2, space.cell_dofs, model.cell_vertices(), e, mp.tables)
                  for e in np.eye(4)], 1)
    want = np.array([[4, -1, -1, -2], [-1, 4, -2, -1], [-1, -2, 4, -1], [-2, -1, -1, 4]]) / 6.0
    assert np.max(np.abs(K - want)) < 1e-15


def test_q1_textbook_cube(pkg, oracle_mod):
    model, space, mp = setup(pkg, 3, 1, 1, 2, perturb=0.0, dirichlet=False)
    K = np.stack([oracle_mod.apply(3, 2, 2, space.cell_dofs, model.cell_vertices(), e, mp.tables)
                  for e in np.eye(8)], 1)
    want = np.zeros((8, 8))
    for a in range(8):
        for b in range(8):
            hd = bin(a ^ b).count("1")  # 0 self, 1 edge, 2 face diagonal, 3 body diagonal
            want[a, b] = {0: 1.0 / 3.0, 1: 0.0, 2: -1.0 / 12.0, 3: -1.0 / 12.0}[hd]
    assert np.max(np.abs(K - want)) < 1e-15


@pytest.mark.parametrize("D,m,p", [(2, 4, 2), (2, 3, 5), (3, 2, 2), (3, 2, 4)])
def test_oracle_invariants_perturbed(pkg, oracle_mod, brute, D, m, p):
    q = p + 2  # q >= (p + 4) / 2: integrals below are exact
    model, space, mp = setup(pkg, D, m, p, q, dirichlet=False)
    X = model.cell_vertices()
    run = lambda v: oracle_mod.apply(D, p + 1, q, space.cell_dofs, X, v, mp.tables)
    # (i) constants are in the kernel of the full operator
    assert np.max(np.abs(run(np.ones(space.n_free)))) < 1e-12
    # (ii) patch test: for u = a.x + b, rows of interior nodes vanish
    xyz = space.free_dof_coordinates()
    a = np.array([0.7, -1.3, 0.4][:D])
    u = xyz @ a + 0.25
    r = run(u)
    nL = space.n_lattice
    idx = np.arange(space.n_free)
    lat = np.stack([(idx // nL ** d) % nL for d in range(D)], -1)
    interior = np.all((lat > 0) & (lat < nL - 1), axis=1)
    assert np.max(np.abs(r[interior])) < 1e-12
    assert np.max(np.abs(r[~interior])) > 1e-3  # boundary rows carry the flux
    # (iii) u = x_1^2: interior rows equal -2 * integral(phi_i)
    u2 = xyz[:, 0] ** 2
    r2 = run(u2)
    B, Dm, pts, w = brute.tables(p, q)
    G, det = brute.geometry_numpy(D, X, pts, w)
    # integral(phi_i) = sum_t w_t |det J_t| phi_i(xi_t): assemble with the
    # interpolation matrix (B (x) B (x) B) on the same cells
    Bk = np.kron(B, B) if D == 2 else np.kron(B, np.kron(B, B))
    if D == 2:
        wt = np.outer(w, w).reshape(-1)
    else:
        wt = np.einsum("l,k,i->lki", w, w, w).reshape(-1)
    cellint = (wt[None, :] * np.abs(det)) @ Bk  # (cells, nl)
    integ = np.zeros(space.n_free)
    np.add.at(integ, space.cell_dofs.reshape(-1), cellint.reshape(-1))
    assert np.max(np.abs(r2[interior] + 2.0 * integ[interior])) < 1e-12
    # (iv) symmetry
    v1 = pkg.uniform(11, space.n_free, -1, 1)
    v2 = pkg.uniform(12, space.n_free, -1, 1)
    assert abs(v1 @ run(v2) - v2 @ run(v1)) < 1e-12 * abs(v1 @ run(v1))
    # (v) volume of the unit square / cube
    assert abs(np.sum(wt[None, :] * np.abs(det)) - 1.0) < 1e-13


@pytest.mark.parametrize("D,p,q", [(2, 3, 5), (3, 2, 4), (3, 4, 6)])
def test_oracle_geometry_vs_numpy(pkg, oracle_mod, brute, D, p, q):
    model, space, mp = setup(pkg, D, 3, p, q)
    X = model.cell_vertices()
    G = oracle_mod.geometry(D, q, X, mp.tables)
    Gn, _ = brute.geometry_numpy(D, X, mp.points, mp.weights)
    pairs = [(0, 0), (0, 1), (1, 1)] if D == 2 else [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for e, (r, k) in enumerate(pairs):
        assert np.max(np.abs(G[:, e, :] - Gn[:, :, r, k])) < 1e-14 * np.max(np.abs(Gn))


def test_demo_jl_assembled_vs_lazy(pkg, oracle_mod, brute):
    """test/Demo.jl:21-57: 3D, Q4, quadrature degree 6 (4 points), 10^3
    Cartesian cells, homogeneous Dirichlet; norm(A*1 - A_lazy*1) < 1e-6."""
    m = pkg.SumFactorizationMap(3, (4, 4, 4), (6, 6, 6))
    assert m.nquad == 4
    model = pkg.CartesianModel(3, 10)
    space = pkg.FESpace(model, 4)
    X = model.cell_vertices()
    y = np.ones(space.n_free)
    x2 = oracle_mod.apply(3, 5, 4, space.cell_dofs, X, y, m.tables)
    B, Dm, pts, w = brute.tables(4, 4)
    G, _ = brute.geometry_numpy(3, X, pts, w)
    K = brute.element_matrices(3, B, Dm, G)
    A = brute.assemble_csr(space.cell_dofs, K, space.n_free)
    x1 = A @ y
    assert np.linalg.norm(x1 - x2) < 1e-6


@pytest.mark.parametrize("D,p,q", [(2, 3, 5), (3, 2, 4)])
def test_oracle_kappa_gradient(pkg, oracle_mod, brute, D, p, q):
    model, space, mp = setup(pkg, D, 3, p, q)
    X = model.cell_vertices()
    nq = q ** D
    kappa = np.exp(pkg.normal(3, space.n_cells * nq)).reshape(space.n_cells, nq)
    u = pkg.uniform(0, space.n_free, -1, 1)
    lam = pkg.uniform(4, space.n_free, -1, 1)
    g = oracle_mod.kappa_grad(D, p + 1, q, space.cell_dofs, X, u, lam, mp.tables)
    # <lam, A(kappa) u> is linear in kappa: its gradient contracted with any
    # direction equals the product with kappa replaced by that direction
    run = lambda k: lam @ oracle_mod.apply(D, p + 1, q, space.cell_dofs, X, u, mp.tables, kappa=k)
    for s in range(3):
        dk = pkg.normal(50 + s, space.n_cells * nq).reshape(space.n_cells, nq)
        lin = run(dk)
        assert abs(np.sum(g * dk) - lin) < 1e-12 * (abs(lin) + np.linalg.norm(g))
    # and J(kappa) = <lam, A(kappa) u> itself is sum(g * kappa)
    assert abs(np.sum(g * kappa) - run(kappa)) < 1e-12 * abs(run(kappa))
    # brute force: per-quadrature-point dense gradients
    B, Dm, pts, w = brute.tables(p, q)
    G, _ = brute.geometry_numpy(D, X, pts, w)
    Gr = np.stack(brute.grad_matrices(D, B, Dm))
    cd = space.cell_dofs
    ul = np.where(cd >= 0, u[np.maximum(cd, 0)], 0.0)
    ll = np.where(cd >= 0, lam[np.maximum(cd, 0)], 0.0)
    gu = np.einsum("rta,ca->ctr", Gr, ul)
    gl = np.einsum("rta,ca->ctr", Gr, ll)
    g_ref = np.einsum("ctr,ctrk,ctk->ct", gl, G, gu)
    assert np.max(np.abs(g - g_ref)) < 1e-13 * np.max(np.abs(g_ref))


def test_planted_mistakes_are_caught(pkg, oracle_mod, brute):
    """A dropped geometric term, a transposed table and a swapped direction
    in the brute force all break agreement with the oracle."""
    D, p, q = 3, 3, 5
    model, space, mp = setup(pkg, D, 2, p, q)
    X = model.cell_vertices()
    x = pkg.uniform(0, space.n_free, -1, 1)
    y = oracle_mod.apply(D, p + 1, q, space.cell_dofs, X, x, mp.tables)
    B, Dm, pts, w = brute.tables(p, q)
    G, _ = brute.geometry_numpy(D, X, pts, w)

    def brute_y(Bx, Dx, Gx):
        return brute.apply_dense(space.cell_dofs, brute.element_matrices(D, Bx, Dx, Gx), x)

    assert rel_l2(brute_y(B, Dm, G), y) < 1e-13
    G_drop = G.copy()
    G_drop[:, :, 0, 1] = 0.0
    G_drop[:, :, 1, 0] = 0.0  # dropped off-diagonal metric term
    assert rel_l2(brute_y(B, Dm, G_drop), y) > 1e-4
    B_rev = B[::-1, :].copy()  # quadrature points in the wrong order for B only
    assert rel_l2(brute_y(B_rev, Dm, G), y) > 1e-4
    G_swap = G[:, :, [1, 0, 2], :][:, :, :, [1, 0, 2]]  # x <-> y metric swap
    assert rel_l2(brute_y(B, Dm, G_swap), y) > 1e-4


def test_flop_model_contraction_count(pkg):
    """The contraction part of count_manual_flops_poisson_matrix_free
    (src/Utils.jl:33-72) against an explicit enumeration of the reference's
    D x D loop nest (src/SumFactorizationKernels.jl:40-65): one FMA per
    (component, output entry, contracted index)."""
    for D, n, q in [(2, 2, 3), (3, 5, 6), (3, 8, 9), (2, 5, 6)]:
        fma = 0
        for k in range(D):
            shape = [n] * D
            for d in range(D):
                out = list(shape)
                out[d] = q
                fma += int(np.prod(out)) * n
                shape = out
        # structured mesh, FMA-only count: weights add nothing (src/Utils.jl:47-60)
        total_fma = pkg.count_manual_flops_poisson_matrix_free(1, D, (q,) * D, (n,) * D)
        assert total_fma == 2 * fma
        all_flops = pkg.count_manual_flops_poisson_matrix_free(
            1, D, (q,) * D, (n,) * D, count_only_fma_ops=False)
        assert all_flops == 4 * fma + D * q ** D
    # CompareFlops.jl's case: D=2, Q1 (SB=2), 3 points, 16 cells
    assert pkg.count_manual_flops_poisson_matrix_free(16, 2, (3, 3), (2, 2)) == 16 * 120


def _q1_demo_known_answer(m):
    """A 1 on the free dofs of Q1, m x m unit-square cells, homogeneous
    Dirichlet: the Q1 stencil is 8/3 at the centre and -1/3 at each of the 8
    neighbours, so a free dof with k free neighbours gets (8 - k) / 3."""
    F = m - 1
    y = np.empty(F * F)
    for j in range(F):
        for i in range(F):
            k = sum(1 for dj in (-1, 0, 1) for di in (-1, 0, 1)
                    if (di or dj) and 0 <= i + di < F and 0 <= j + dj < F)
            y[i + F * j] = (8 - k) / 3.0
    return y


@pytest.mark.parametrize("p,q,m,perturb", [(1, 3, 4, 0.0), (1, 3, 4, 0.2), (3, 5, 3, 0.2)])
def test_step_by_step_demo_restatement(pkg, oracle_mod, brute, p, q, m, perturb):
    """test/Demo_StepByStep.jl's case (2D, Q1, 4 x 4 Cartesian cells,
    degree 4 -> 3 Gauss points, x = ones) and two perturbed variants: the
    explicit-loop restatement of the demo equals the C oracle and the
    brute-force assembled matrix; on the demo's own case it also equals the
    textbook Q1 stencil.  Swapping the value and derivative tables (a
    transposed operand) breaks the agreement."""
    from oracle import stepbystep
    model = pkg.CartesianModel(2, m, perturb=perturb, seed=1)
    space = pkg.FESpace(model, p)
    mp = pkg.SumFactorizationMap.with_points(2, p, q)
    x = np.ones(space.n_free) if perturb == 0.0 else pkg.uniform(0, space.n_free, -1, 1)
    X = model.cell_vertices()
    tab = [np.asarray(t, dtype=np.float64).tolist() for t in mp.tables]
    y = np.array(stepbystep.apply_2d(space.cell_dofs, X, x, tab))
    y_c = oracle_mod.apply(2, p + 1, q, space.cell_dofs, X, x, mp.tables)
    Bb, Db, pb, wb = brute.tables(p, q)
    G, _ = brute.geometry_numpy(2, X, pb, wb)
    A = brute.assemble_csr(space.cell_dofs, brute.element_matrices(2, Bb, Db, G), space.n_free)
    assert rel_l2(y, y_c) < 1e-13
    assert rel_l2(y, A @ x) < 1e-13
    if perturb == 0.0 and p == 1:
        assert np.abs(y - _q1_demo_known_answer(m)).max() < 1e-14
    else:
        # (on the symmetric demo case with x = 1 the swap goes unnoticed)
        swapped = [tab[1], tab[0], tab[2], tab[3]]
        y_bad = np.array(stepbystep.apply_2d(space.cell_dofs, X, x, swapped))
        assert rel_l2(y_bad, y_c) > 1e-3


def test_flop_model_unstructured_hand_values(pkg):
    """The metric-work branch of count_manual_flops_poisson_matrix_free
    (src/Utils.jl:47-56, structured_mesh_opt = false) that the FP64 apply's
    roofline uses, against values worked out by hand from the reference's
    formula: in 3D the weights-and-Jacobian step is (D-1) D^D + D^2 = 63
    FMA-only operations per quadrature point, 2*63 + 27 + 6 = 159 in all;
    in 2D 8 and 2*8 + 4 + 4 = 24.  The contractions of the 64^3 Q4 workload
    (q = 6) are 3 (6*125 + 36*25 + 216*5) = 8190 FMAs each way per cell."""
    f = pkg.count_manual_flops_poisson_matrix_free
    # 64^3 cells, q = 6, n = 5: 16380 + 159*216 + 16380 = 67104 per cell
    assert f(64 ** 3, 3, (6,) * 3, (5,) * 3, count_only_fma_ops=False,
             structured_mesh_opt=False) == 17_590_910_976
    assert f(1, 3, (6,) * 3, (5,) * 3, count_only_fma_ops=True,
             structured_mesh_opt=False) == 8190 + 63 * 216 + 8190
    # structured: only the D multiplications by the weight per point count
    assert f(1, 3, (6,) * 3, (5,) * 3, count_only_fma_ops=False,
             structured_mesh_opt=True) == 16380 + 3 * 216 + 16380
    # 2D Q4, q = 6: contractions 2 (6*25 + 36*5) = 660 FMAs each way
    assert f(1, 2, (6, 6), (5, 5), count_only_fma_ops=False,
             structured_mesh_opt=False) == 2 * 660 + 24 * 36 + 2 * 660
    assert f(1, 2, (6, 6), (5, 5), count_only_fma_ops=True,
             structured_mesh_opt=False) == 660 + 8 * 36 + 660


@pytest.mark.parametrize("D,m,p,q", [(2, 4, 3, 5), (3, 2, 4, 6), (3, 2, 2, 3)])
def test_brute_diagonal_pins(pkg, oracle_mod, brute, D, m, p, q):
    """The brute-force diagonal (element-matrix diagonals summed over cells)
    equals the diagonal of the assembled CSR matrix on a perturbed mesh with
    a coefficient, e_i^T A e_i by the C oracle's product for a few dofs, and
    the Kronecker closed form's diagonal on the uniform grid."""
    model, space, mp = setup(pkg, D, m, p, q)
    X = model.cell_vertices()
    B, Dm, pts, w = brute.tables(p, q)
    G, _ = brute.geometry_numpy(D, X, pts, w)
    kappa = np.exp(pkg.normal(3, space.n_cells * q ** D)).reshape(space.n_cells, q ** D)
    Gk = G * kappa[:, :, None, None]
    d = brute.diagonal(D, B, Dm, Gk, space.cell_dofs, space.n_free)
    A = brute.assemble_csr(space.cell_dofs, brute.element_matrices(D, B, Dm, Gk), space.n_free)
    assert rel_l2(d, A.diagonal()) < 1e-14
    for i in (0, space.n_free // 3, space.n_free - 1):
        e = np.zeros(space.n_free)
        e[i] = 1.0
        Ae = oracle_mod.apply(D, p + 1, q, space.cell_dofs, X, e, mp.tables, kappa=kappa.reshape(-1))
        assert abs(Ae[i] - d[i]) < 1e-13 * abs(d[i])
    model0, space0, _ = setup(pkg, D, m, p, q, perturb=0.0)
    G0, _ = brute.geometry_numpy(D, model0.cell_vertices(), pts, w)
    d0 = brute.diagonal(D, B, Dm, G0, space0.cell_dofs, space0.n_free)
    assert rel_l2(d0, brute.kron_operator(D, p, q, m).diagonal()) < 1e-13
    # a planted mistake: the derivative table in place of the value table
    # in every direction changes the diagonal
    d_bad = brute.diagonal(D, Dm, Dm, Gk, space.cell_dofs, space.n_free)
    assert rel_l2(d_bad, d) > 1e-2


@pytest.mark.parametrize("p", [1, 2, 3, 5, 7, 9])
def test_brute_gll_interpolation_pins(brute, p):
    """The Gauss-Lobatto points and the GLL-to-equispaced table: endpoints
    and symmetry of the points, the Lobatto rule on them integrates x^k
    exactly for k <= 2p - 1 (which fixes the points), T1 reproduces every
    polynomial of degree <= p, its vertex rows are unit rows, and for p <= 2
    the two node sets coincide (T1 = I)."""
    s = brute.gll_nodes(p)
    assert s[0] == 0.0 and s[1] == 1.0
    inner = np.sort(s[2:])
    assert np.allclose(inner, 1.0 - inner[::-1], atol=1e-14)
    # Lobatto weights from the Lagrange basis: w_j = int_0^1 l_j
    xg, wg = brute.gauss01(p + 2)
    wl = brute.lagrange_matrix(s, xg).T @ wg
    for k in range(2 * p):
        assert abs(wl @ s ** k - 1.0 / (k + 1)) < 1e-13, k
    T1 = brute.gll_to_equispaced(p)
    t = brute.nodes_1d(p)
    for k in range(p + 1):
        assert np.abs(T1 @ s ** k - t ** k).max() < 1e-12
    assert np.array_equal(T1[:2, :2], np.eye(2)) and not T1[:2, 2:].any()
    if p <= 2:
        assert np.abs(T1 - np.eye(p + 1)).max() < 1e-15


@pytest.mark.parametrize("m,p,q", [(2, 4, 6), (2, 3, 5)])
def test_brute_gll_preconditioner_pins(pkg, brute, m, p, q):
    """The global interpolation T of the brute-force GLL preconditioner is
    the per-cell one on every cell (conforming), the diagonal of T^T A T is
    the element-matrix diagonal with the tables B T1, D T1, and
    M^{-1} = T diag^{-1} T^T is symmetric positive definite."""
    model, space, mp = setup(pkg, 3, m, p, q)
    B, Dm, pts, w = brute.tables(p, q)
    G, _ = brute.geometry_numpy(3, model.cell_vertices(), pts, w)
    T, dg = brute.gll_preconditioner(3, p, B, Dm, G, space.cell_dofs, space.n_free)
    T1 = brute.gll_to_equispaced(p)
    Te = np.kron(T1, np.kron(T1, T1))
    Td = T.toarray()
    for c in space.cell_dofs:
        a = np.nonzero(c >= 0)[0]
        assert np.abs(Td[np.ix_(c[a], c[a])] - Te[np.ix_(a, a)]).max() < 1e-15
    d2 = brute.diagonal(3, B @ T1, Dm @ T1, G, space.cell_dofs, space.n_free)
    assert rel_l2(d2, dg) < 1e-13
    u = pkg.uniform(1, space.n_free, -1, 1)
    v = pkg.uniform(2, space.n_free, -1, 1)
    Mi = lambda r: T @ ((T.T @ r) / dg)
    assert abs(u @ Mi(v) - v @ Mi(u)) < 1e-13 * abs(u @ Mi(v))
    assert u @ Mi(u) > 0 and v @ Mi(v) > 0
