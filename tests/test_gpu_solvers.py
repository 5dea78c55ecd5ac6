"""GPU tests of the solver-side pieces through the C ABI: the matrix-free
diagonal, the Jacobi-preconditioned CG forms, the high-order solves of the
precision and highest-order workloads (Q7 q = 9, Q9 q = 11 with a
log-normal coefficient) on reduced meshes in every solver form, the
accumulate convention y += A x, and the central finite-difference check of
the adjoint along 64 directions."""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problem(pkg, D, m, p, q, kappa_seed=None, perturb=0.2):
    model = pkg.CartesianModel(D, m, perturb=perturb, seed=1)
    space = pkg.FESpace(model, p)
    mp = pkg.SumFactorizationMap.with_points(D, p, q)
    kappa = None
    if kappa_seed is not None:
        kappa = np.exp(pkg.normal(kappa_seed, space.n_cells * q ** D))
    return model, space, mp, kappa


def lazy(pkg, mp, space, kappa, monkeypatch, fused=True):
    if not fused:
        monkeypatch.setenv("SFHIP_FUSED_GEOMETRY", "0")
    A = pkg.LazyMatrix(mp, space, kappa=kappa)
    if not fused:
        monkeypatch.delenv("SFHIP_FUSED_GEOMETRY")
    return A


def oracle_rel_residual(oracle_mod, D, p, q, model, space, mp, kappa, x, b):
    r = b - oracle_mod.apply(D, p + 1, q, space.cell_dofs, model.cell_vertices(), x, mp.tables,
                             kappa=kappa)
    return np.linalg.norm(r) / np.linalg.norm(b)


DIAG_CASES = [(2, 6, 4, 6, False), (2, 5, 3, 4, True), (3, 3, 1, 2, True), (3, 3, 2, 4, True),
              (3, 2, 4, 6, True), (3, 2, 5, 7, True), (3, 2, 7, 9, False), (3, 2, 9, 11, True),
              (3, 2, 9, 12, True)]


@pytest.mark.parametrize("D,m,p,q,with_kappa", DIAG_CASES)
def test_diagonal_vs_bruteforce(pkg, brute, cuda, monkeypatch, D, m, p, q, with_kappa):
    """sf_diagonal (sum-factorised element-matrix diagonals, block-
    interleaved G in 3D) against the brute-force diagonal from dense
    reference gradients, with and without a log-normal coefficient, on
    plans built from vertices and with the stored-G form."""
    say("diagonal D=%d m=%d p=%d q=%d kappa=%d" % (D, m, p, q, with_kappa))
    model, space, mp, kappa = problem(pkg, D, m, p, q, 3 if with_kappa else None)
    B, Dm, pts, w = brute.tables(p, q)
    G, _ = brute.geometry_numpy(D, model.cell_vertices(), pts, w)
    if kappa is not None:
        G = G * kappa.reshape(space.n_cells, -1)[:, :, None, None]
    d_ref = brute.diagonal(D, B, Dm, G, space.cell_dofs, space.n_free)
    for fused in (True, False):
        A = lazy(pkg, mp, space, kappa, monkeypatch, fused)
        d = A.diagonal()
        cuda.cuda.synchronize()
        err = rel_l2(d.cpu().numpy(), d_ref)
        A.free()
        assert err < 1e-12, (fused, err)


@pytest.mark.parametrize("D,m,p,q", [(3, 5, 3, 5), (3, 3, 5, 7), (2, 10, 4, 6)])
def test_apply_add_accumulates(pkg, oracle_mod, cuda, monkeypatch, D, m, p, q):
    """y += A x (sf_apply_add, the accumulate convention of the reference's
    GPU kernels, src/GPUKernels.jl:77-89): y0 + A x against the oracle, on
    meshes whose blocks have both exclusive and shared dofs, for the fused
    and the stored-G 3D kernels and the 2D kernel; twice in a row adds
    twice."""
    say("apply_add D=%d m=%d p=%d q=%d" % (D, m, p, q))
    model, space, mp, _ = problem(pkg, D, m, p, q)
    x = pkg.uniform(0, space.n_free, -1, 1)
    y0 = pkg.uniform(9, space.n_free, -1, 1)
    Ax = oracle_mod.apply(D, p + 1, q, space.cell_dofs, model.cell_vertices(), x, mp.tables)
    xd = cuda.from_numpy(x).cuda()
    for fused in ((True, False) if D == 3 else (True,)):
        A = lazy(pkg, mp, space, None, monkeypatch, fused)
        if D == 3:
            info = A.info()
            assert info["n_blocks"] > 1
        y = cuda.from_numpy(y0.copy()).cuda()
        A.mul_add(y, xd)
        cuda.cuda.synchronize()
        assert rel_l2(y.cpu().numpy(), y0 + Ax) < 1e-13, fused
        A.mul_add(y, xd)
        cuda.cuda.synchronize()
        assert rel_l2(y.cpu().numpy(), y0 + 2 * Ax) < 1e-13, fused
        A.free()


# (p, q, m, kappa, Pl): the precision workload's order (Q7, q = 9) and the
# highest-order workload's (Q9, q = 11, log-normal kappa) on reduced meshes.
# Unpreconditioned at Q9 with kappa the mixed forms stall above 1e-10
# (measured: 7.6e-10 / 6e-8 / 1e-2 after 400k iterations for the f64p32 /
# f64 / f32 recurrences: the FP32 operator's rounding against a condition
# number near 1e10), so there only FP64 CG and the refinement run; with the
# Gauss-Lobatto preconditioner every form runs at both orders.
HIGH_ORDER = [(7, 9, 3, False, None), (9, 11, 3, True, None), (7, 9, 3, False, "gll"),
              (9, 11, 3, True, "gll")]


@pytest.mark.parametrize("p,q,m,with_kappa,Pl", HIGH_ORDER)
def test_high_order_solvers_every_form(pkg, oracle_mod, cuda, monkeypatch, p, q, m, with_kappa,
                                       Pl):
    """FP64 CG with the fused step (p.Ap from the apply at N = 8, 10) and
    with the unfused stored-G step, every mixed-CG recurrence and the
    refinement, each to 1e-10 in the FP64 residual recomputed by the oracle,
    all agreeing with the FP64 solution."""
    say("high-order solvers p=%d q=%d m=%d kappa=%d Pl=%s" % (p, q, m, with_kappa, Pl))
    model, space, mp, kappa = problem(pkg, 3, m, p, q, 3 if with_kappa else None)
    b = cuda.from_numpy(pkg.normal(2, space.n_free)).cuda()
    bh = b.cpu().numpy()
    maxit = 400000
    mixed = Pl is not None or not with_kappa
    sols, stats = {}, {}
    for fused in (True, False):
        A = lazy(pkg, mp, space, kappa, monkeypatch, fused)
        x = cuda.zeros_like(b)
        st = pkg.cg_(x, A, b, rtol=1e-10, maxiter=maxit, Pl=Pl)
        say("  f64 fused=%d" % fused, st)
        sols["f64", fused], stats["f64", fused] = x.cpu().numpy(), st
        if fused:
            for rec in (("f64p32", "f64", "f32") if mixed else ()):
                xm = cuda.zeros_like(b)
                sm = pkg.cg_mixed_(xm, A, b, rtol=1e-10, inner="f32", delta=0.1, maxiter=maxit,
                                   recurrence=rec, Pl=Pl)
                say("  mixed", rec, sm)
                sols[rec], stats[rec] = xm.cpu().numpy(), sm
            xi = cuda.zeros_like(b)
            si = pkg.cg_ir_(xi, A, b, rtol=1e-10, inner="f32", inner_rtol=1e-2,
                            inner_maxiter=maxit, max_outer=80, Pl=Pl)
            say("  refinement", si)
            sols["ir"], stats["ir"] = xi.cpu().numpy(), si
        A.free()
    x64 = sols["f64", True]
    for key, st in stats.items():
        assert st["converged"] == 1 and st["rel_residual"] <= 1e-10, (key, st)
        rel = oracle_rel_residual(oracle_mod, 3, p, q, model, space, mp, kappa, sols[key], bh)
        assert rel < 2e-10, (key, rel)
        assert rel_l2(sols[key], x64) < 1e-6, key
    # fused and unfused steps take the same iterations up to the spread of the
    # atomics, unless one of them restarted (outer_iterations: the plain CG's
    # restarts, e.g. a Gauss-Lobatto solve whose first target left the true
    # residual just above rtol), which adds the few iterations of the restart
    sf, su = stats["f64", True], stats["f64", False]
    it_f, it_u = sf["iterations"], su["iterations"]
    slack = 0.02 if sf["outer_iterations"] == su["outer_iterations"] else 0.08
    assert abs(it_f - it_u) <= max(5, slack * it_u), (sf, su)


@pytest.mark.parametrize("m,p,q,with_kappa", [(3, 2, 4, True), (2, 4, 6, True), (2, 7, 9, False),
                                              (2, 9, 11, True), (4, 3, 5, False)])
def test_precond_apply_vs_brute(pkg, brute, cuda, monkeypatch, m, p, q, with_kappa):
    """z = M^{-1} r on the device (sf_precond_apply) for the Jacobi and the
    Gauss-Lobatto preconditioner against the brute force: the assembled
    matrix's diagonal, and T (T^T r / diag(T^T A T)) with T the global
    GLL-to-equispaced interpolation built from numpy's Legendre roots.  The
    inverse diagonals are FP32 on the device, so the reference rounds them
    the same way; fused and stored-G plans; without a preconditioner z = r."""
    say("precond apply m=%d p=%d q=%d kappa=%d" % (m, p, q, with_kappa))
    model, space, mp, kappa = problem(pkg, 3, m, p, q, 3 if with_kappa else None)
    B, Dm, pts, w = brute.tables(p, q)
    G, _ = brute.geometry_numpy(3, model.cell_vertices(), pts, w)
    if kappa is not None:
        G = G * kappa.reshape(space.n_cells, -1)[:, :, None, None]
    rh = pkg.uniform(6, space.n_free, -1, 1)
    d = brute.diagonal(3, B, Dm, G, space.cell_dofs, space.n_free)
    zj = rh * (1.0 / d).astype(np.float32).astype(np.float64)
    T, dg = brute.gll_preconditioner(3, p, B, Dm, G, space.cell_dofs, space.n_free)
    zg = T @ ((T.T @ rh) * (1.0 / dg).astype(np.float32).astype(np.float64))
    r = cuda.from_numpy(rh).cuda()
    for fused in (True, False):
        A = lazy(pkg, mp, space, kappa, monkeypatch, fused)
        z = cuda.empty_like(r)
        for Pl, ref in ((None, rh), ("jacobi", zj), ("gll", zg)):
            A.precond(z, r, Pl)
            cuda.cuda.synchronize()
            err = rel_l2(z.cpu().numpy(), ref)
            assert err < 1e-12, (fused, Pl, err)
        A.free()


@pytest.mark.parametrize("D,m,p,q,with_kappa", [(3, 4, 4, 6, True), (3, 3, 7, 9, False),
                                                (2, 12, 4, 6, True)])
def test_jacobi_pcg(pkg, oracle_mod, cuda, monkeypatch, D, m, p, q, with_kappa):
    """Jacobi-preconditioned CG (the matrix-free diagonal, inverted once):
    FP64 CG fused and unfused, mixed CG and refinement reach 1e-10 in the
    oracle's FP64 residual and agree with unpreconditioned FP64 CG."""
    say("jacobi pcg D=%d m=%d p=%d q=%d kappa=%d" % (D, m, p, q, with_kappa))
    model, space, mp, kappa = problem(pkg, D, m, p, q, 3 if with_kappa else None)
    b = cuda.from_numpy(pkg.normal(2, space.n_free)).cuda()
    bh = b.cpu().numpy()
    A = lazy(pkg, mp, space, kappa, monkeypatch, True)
    x0 = cuda.zeros_like(b)
    s0 = pkg.cg_(x0, A, b, rtol=1e-10, maxiter=200000)
    x0h = x0.cpu().numpy()
    runs = {}
    x = cuda.zeros_like(b)
    runs["cg"] = (pkg.cg_(x, A, b, rtol=1e-10, maxiter=200000, Pl="jacobi"), x)
    recs = ("f64p32", "f64", "f32") if D == 3 else ("f64", "f32")
    for rec in recs:
        xm = cuda.zeros_like(b)
        runs[rec] = (pkg.cg_mixed_(xm, A, b, rtol=1e-10, maxiter=200000, recurrence=rec,
                                   Pl="jacobi"), xm)
    xi = cuda.zeros_like(b)
    runs["ir"] = (pkg.cg_ir_(xi, A, b, rtol=1e-10, inner_rtol=1e-2, inner_maxiter=200000,
                             max_outer=80, Pl="jacobi"), xi)
    A.free()
    if D == 3:
        Au = lazy(pkg, mp, space, kappa, monkeypatch, False)
        xu = cuda.zeros_like(b)
        runs["cg_unfused"] = (pkg.cg_(xu, Au, b, rtol=1e-10, maxiter=200000, Pl="jacobi"), xu)
        Au.free()
    say("  plain", s0)
    for key, (st, xx) in runs.items():
        say("  jacobi", key, st)
        xh = xx.cpu().numpy()
        assert st["converged"] == 1, (key, st)
        rel = oracle_rel_residual(oracle_mod, D, p, q, model, space, mp, kappa, xh, bh)
        assert rel < 2e-10, (key, rel)
        assert rel_l2(xh, x0h) < 1e-6, key
    if D == 3:
        it_f, it_u = runs["cg"][0]["iterations"], runs["cg_unfused"][0]["iterations"]
        assert abs(it_f - it_u) <= max(5, 0.02 * it_u), (it_f, it_u)
    # switching the preconditioner off again gives the plain iteration (the
    # Gauss-Lobatto one first: its count is far from the plain one; the plain
    # count itself moves by about 1% with the order of the apply's atomics)
    x2 = cuda.zeros_like(b)
    A = lazy(pkg, mp, space, kappa, monkeypatch, True)
    s2 = pkg.cg_(x2, A, b, rtol=1e-10, maxiter=200000, Pl="gll" if D == 3 else "jacobi")
    s3 = pkg.cg_(cuda.zeros_like(b), A, b, rtol=1e-10, maxiter=200000, Pl=None)
    say("  switched", s2, s3)
    assert abs(s3["iterations"] - s0["iterations"]) <= max(3, 0.03 * s0["iterations"])
    if D == 3:
        assert s2["iterations"] < 0.5 * s0["iterations"]
    with pytest.raises(ValueError):
        pkg.cg_(x2, A, b, Pl="ilu")
    A.free()


def test_adjoint_fd_64_directions(pkg, oracle_mod, cuda):
    """Central finite differences (h = 1e-6, FP64) of
    J = 1/2 |A(kappa) u - d|^2 along 64 random unit directions in kappa
    and 64 in u, against the adjoint's dJ/dkappa and dJ/du, on a reduced
    differentiation workload (8^3 perturbed hexes, Q5, 7 points per
    direction, log-normal kappa).  J is quadratic in each, so the error is
    rounding only: |fd - <grad, delta>| <= 1e-6 (|<grad, delta>| +
    1e-3 |grad|)."""
    say("adjoint fd 64 directions 8^3 Q5")
    model, space, mp, kappa = problem(pkg, 3, 8, 5, 7, 3)
    A = pkg.LazyMatrix(mp, space, kappa=kappa)
    u = cuda.from_numpy(pkg.uniform(0, space.n_free, -1, 1)).cuda()
    d = cuda.from_numpy(pkg.uniform(4, space.n_free, -1, 1)).cuda()
    J, kb, ub = pkg.adjoint.loss_and_grad(A, kappa, u, d)
    kbh = kb.cpu().numpy()
    kb_ref = oracle_mod.kappa_grad(3, 6, 7, space.cell_dofs, model.cell_vertices(),
                                   u.cpu().numpy(),
                                   oracle_mod.apply(3, 6, 7, space.cell_dofs,
                                                    model.cell_vertices(), u.cpu().numpy(),
                                                    mp.tables, kappa=kappa) - d.cpu().numpy(),
                                   mp.tables).reshape(-1)
    assert rel_l2(kbh, kb_ref) < 1e-12
    kd = cuda.from_numpy(kappa).cuda()
    y = cuda.empty_like(u)
    h = 1e-6

    def resid(v):
        A.mul(y, v)
        return y - d

    def central(rp, rm):
        return 0.5 * float(((rp - rm) * (rp + rm)).sum()) / (2 * h)

    nk, nu = float(kb.norm()), float(ub.norm())
    worst_k = worst_u = 0.0
    for s in range(64):
        dk = cuda.from_numpy(pkg.normal(100 + s, kappa.size)).cuda()
        dk /= dk.norm()
        A.set_kappa((kd + h * dk).cpu().numpy())
        rp = resid(u)
        A.set_kappa((kd - h * dk).cpu().numpy())
        rm = resid(u)
        an = float((kb * dk).sum())
        worst_k = max(worst_k, abs(central(rp, rm) - an) / (abs(an) + 1e-3 * nk))
    A.set_kappa(kappa)
    for s in range(64):
        du = cuda.from_numpy(pkg.normal(200 + s, space.n_free)).cuda()
        du /= du.norm()
        rp = resid(u + h * du)
        rm = resid(u - h * du)
        an = float((ub * du).sum())
        worst_u = max(worst_u, abs(central(rp, rm) - an) / (abs(an) + 1e-3 * nu))
    say("  worst fd error: dkappa %.2e, du %.2e" % (worst_k, worst_u))
    assert worst_k < 1e-6 and worst_u < 1e-6
    A.free()


def test_rrule_pullback_holds_forward_kappa(pkg, oracle_mod, cuda):
    """Forward with kappa1, forward with kappa2 on the same operator, then
    the first pullback: ubar = A(kappa1) ybar (the pullback re-sets its
    kappa), kappabar from kappa-free G either way."""
    say("rrule kappa capture")
    model, space, mp, k1 = problem(pkg, 3, 3, 4, 6, 3)
    k2 = np.exp(pkg.normal(7, k1.size))
    A = pkg.LazyMatrix(mp, space)
    u = cuda.from_numpy(pkg.uniform(0, space.n_free, -1, 1)).cuda()
    yb = pkg.uniform(5, space.n_free, -1, 1)
    y1, pb1 = pkg.adjoint.rrule(A, k1, u)
    y2, pb2 = pkg.adjoint.rrule(A, k2, u)
    kb1, ub1 = pb1(cuda.from_numpy(yb).cuda())
    kb2, ub2 = pb2(cuda.from_numpy(yb).cuda())
    cuda.cuda.synchronize()
    X = model.cell_vertices()
    for k, y, ub in ((k1, y1, ub1), (k2, y2, ub2)):
        assert rel_l2(ub.cpu().numpy(), oracle_mod.apply(3, 5, 6, space.cell_dofs, X, yb,
                                                          mp.tables, kappa=k)) < 1e-12
        assert rel_l2(y.cpu().numpy(), oracle_mod.apply(3, 5, 6, space.cell_dofs, X,
                                                         u.cpu().numpy(), mp.tables,
                                                         kappa=k)) < 1e-12
    kref = oracle_mod.kappa_grad(3, 5, 6, space.cell_dofs, X, u.cpu().numpy(), yb,
                                 mp.tables).reshape(-1)
    assert rel_l2(kb1.cpu().numpy(), kref) < 1e-12
    assert rel_l2(kb2.cpu().numpy(), kref) < 1e-12
    A.free()


def test_set_kappa_ordered_against_side_stream(pkg, oracle_mod, cuda):
    """Applies and pullbacks queued on a non-default torch stream while the
    host changes kappa between them: every result equals the oracle's for
    the kappa that was set when it was queued (set_kappa waits for the
    device before it overwrites the plan's coefficient)."""
    say("set_kappa vs side stream")
    model, space, mp, k0 = problem(pkg, 3, 4, 3, 5, 3)
    A = pkg.LazyMatrix(mp, space, kappa=k0)
    X = model.cell_vertices()
    xh = pkg.uniform(0, space.n_free, -1, 1)
    st = cuda.cuda.Stream()
    with cuda.cuda.stream(st):
        x = cuda.from_numpy(xh).cuda()
        outs, ks = [], []
        for s in range(6):
            k = np.exp(pkg.normal(20 + s, k0.size))
            A.set_kappa(k)
            y = cuda.empty_like(x)
            for _ in range(4):  # a few in flight
                A.mul(y, x)
            kb, ub = pkg.adjoint.pullback(A, x, x)
            outs.append((y, ub))
            ks.append(k)
    st.synchronize()
    for (y, ub), k in zip(outs, ks):
        ref = oracle_mod.apply(3, 4, 5, space.cell_dofs, X, xh, mp.tables, kappa=k)
        assert rel_l2(y.cpu().numpy(), ref) < 1e-12
        assert rel_l2(ub.cpu().numpy(), ref) < 1e-12
    A.free()


def test_set_tables_weights_after_geometry(pkg, oracle_mod, cuda, monkeypatch):
    """New quadrature weights after the geometry: a plan built from
    vertices recomputes G (FP64 and FP32 operators and the stored-G form
    all follow the new weights); a plan with a caller-supplied G refuses."""
    say("set_tables after geometry")
    model, space, mp, kappa = problem(pkg, 3, 3, 3, 5, 3)
    B, Dm, pts, w = (np.ascontiguousarray(t, dtype=np.float64) for t in mp.tables)
    w2 = np.ascontiguousarray(w * np.linspace(0.8, 1.2, w.size))
    X = model.cell_vertices()
    xh = pkg.uniform(0, space.n_free, -1, 1)
    ref = oracle_mod.apply(3, 4, 5, space.cell_dofs, X, xh, (B, Dm, pts, w2), kappa=kappa)
    x = cuda.from_numpy(xh).cuda()
    for fused in (True, False):
        A = lazy(pkg, mp, space, kappa, monkeypatch, fused)
        y = cuda.empty_like(x)
        A.mul(y, x)  # the old weights
        A.prepare("f32")
        pkg._lib.check(A._L.sf_plan_set_tables(A.plan, B.ctypes.data, Dm.ctypes.data,
                                               w2.ctypes.data))
        A.mul(y, x)
        y32 = cuda.empty_like(x.float())
        A.mul_reduced(y32, x.float(), "f32")
        cuda.cuda.synchronize()
        assert rel_l2(y.cpu().numpy(), ref) < 1e-12, fused
        assert rel_l2(y32.cpu().numpy().astype(np.float64), ref) < 1e-5, fused
        A.free()
    Ag = pkg.LazyMatrix(mp, space, geometry=oracle_mod.geometry(3, 5, X, mp.tables))
    rc = Ag._L.sf_plan_set_tables(Ag.plan, B.ctypes.data, Dm.ctypes.data, w2.ctypes.data)
    assert rc == pkg._lib.SF_ERR_NOT_READY
    # tables without weights are still accepted
    pkg._lib.check(Ag._L.sf_plan_set_tables(Ag.plan, B.ctypes.data, Dm.ctypes.data, None))
    Ag.free()


class _Reordered:
    """A space / model pair with the cells listed in another order."""

    def __init__(self, space, order):
        self.D, self.order, self.n = space.D, space.order, space.n
        self.n_free, self.n_cells = space.n_free, space.n_cells
        self.cell_dofs = np.ascontiguousarray(space.cell_dofs[order])
        verts = space.model.cell_vertices()[order]

        class _M:
            def cell_vertices(self_inner):
                return verts

        self.model = _M()


def test_shuffled_cell_order(pkg, oracle_mod, cuda, monkeypatch):
    """Cells listed in a random order (no two consecutive cells neighbours):
    the plan orders them by connectivity, so the block count equals the
    row-wise order's; A x, the kappa gradient (in the caller's cell order),
    G read back, and the GLL-preconditioned CG all match the oracle on the
    shuffled arrays; a row-wise order is kept as it is."""
    say("shuffled cell order")
    model, space, mp, _ = problem(pkg, 3, 6, 3, 5)
    rng = np.random.default_rng(11)
    order = rng.permutation(space.n_cells)
    sh = _Reordered(space, order)
    kappa = np.exp(pkg.normal(3, space.n_cells * 125))
    # kappa per cell follows the caller's cell order
    ks = np.ascontiguousarray(kappa.reshape(space.n_cells, -1)[order].reshape(-1))
    Xs = sh.model.cell_vertices()
    for fused in (True, False):
        A0 = lazy(pkg, mp, space, kappa, monkeypatch, fused)
        As = lazy(pkg, mp, sh, ks, monkeypatch, fused)
        i0, i1 = A0.info(), As.info()
        say("  fused=%d blocks: row-wise %d (reordered %d), shuffled %d (reordered %d)"
            % (fused, i0["n_blocks"], i0["cells_reordered"], i1["n_blocks"],
               i1["cells_reordered"]))
        assert i0["cells_reordered"] == 0 and i1["cells_reordered"] == 1
        assert i1["n_blocks"] == i0["n_blocks"]
        xh = pkg.uniform(0, space.n_free, -1, 1)
        lh = pkg.uniform(1, space.n_free, -1, 1)
        y_ref = oracle_mod.apply(3, 4, 5, sh.cell_dofs, Xs, xh, mp.tables, kappa=ks)
        x, lam = cuda.from_numpy(xh).cuda(), cuda.from_numpy(lh).cuda()
        y = cuda.empty_like(x)
        As.mul(y, x)
        kb, ub = pkg.adjoint.pullback(As, x, lam)
        cuda.cuda.synchronize()
        assert rel_l2(y.cpu().numpy(), y_ref) < 1e-12
        kb_ref = oracle_mod.kappa_grad(3, 4, 5, sh.cell_dofs, Xs, xh, lh, mp.tables).reshape(-1)
        assert rel_l2(kb.cpu().numpy(), kb_ref) < 1e-12
        ub_ref = oracle_mod.apply(3, 4, 5, sh.cell_dofs, Xs, lh, mp.tables, kappa=ks)
        assert rel_l2(ub.cpu().numpy(), ub_ref) < 1e-12
        G_ref = oracle_mod.geometry(3, 5, Xs, mp.tables)
        assert rel_l2(As.geometry().reshape(-1), G_ref.reshape(-1)) < 1e-13
        b = cuda.from_numpy(pkg.normal(2, space.n_free)).cuda()
        x0, x1 = cuda.zeros_like(b), cuda.zeros_like(b)
        s0 = pkg.cg_(x0, A0, b, rtol=1e-10, maxiter=20000, Pl="gll")
        s1 = pkg.cg_(x1, As, b, rtol=1e-10, maxiter=20000, Pl="gll")
        assert s1["converged"] == 1
        assert abs(s1["iterations"] - s0["iterations"]) <= max(2, 0.03 * s0["iterations"])
        assert rel_l2(x1.cpu().numpy(), x0.cpu().numpy()) < 1e-8
        A0.free()
        As.free()
    # a caller-supplied G in the shuffled order goes through the same map
    Ag = pkg.LazyMatrix(mp, sh, geometry=oracle_mod.geometry(3, 5, Xs, mp.tables))
    assert Ag.info()["cells_reordered"] == 1
    xh = pkg.uniform(0, space.n_free, -1, 1)
    y = cuda.empty(space.n_free, dtype=cuda.float64, device="cuda")
    Ag.mul(y, cuda.from_numpy(xh).cuda())
    cuda.cuda.synchronize()
    y_ref = oracle_mod.apply(3, 4, 5, sh.cell_dofs, Xs, xh, mp.tables)
    assert rel_l2(y.cpu().numpy(), y_ref) < 1e-12
    Ag.free()
