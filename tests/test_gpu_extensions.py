"""GPU tests of the extensions beyond the reference's apply, through the C ABI:
reduced-precision applies (error against the FP64 oracle), the fused
adjoint kernel (oracle gradient + central finite differences), and the
device-resident CG / mixed-precision iterative refinement."""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(pkg, D, m, p, q, perturb=0.2, kappa=None):
    model = pkg.CartesianModel(D, m, perturb=perturb, seed=1)
    space = pkg.FESpace(model, p)
    mp = pkg.SumFactorizationMap.with_points(D, p, q)
    A = pkg.LazyMatrix(mp, space, kappa=kappa)
    return model, space, mp, A


@pytest.mark.parametrize("D,m,p,q", [(3, 8, 4, 6), (3, 4, 7, 9), (2, 16, 4, 6), (3, 3, 9, 11)])
def test_reduced_precision_apply_error(pkg, oracle_mod, cuda, D, m, p, q):
    """FP32: relative l2 error ~ FP32 rounding (<= 1e-5); G in bf16: the
    error of a 2^-9 relative perturbation of the coefficient (<= 2e-2)."""
    say("reduced precision D=%d m=%d p=%d q=%d" % (D, m, p, q))
    model, space, mp, A = make(pkg, D, m, p, q)
    x = pkg.uniform(0, space.n_free, -1, 1)
    y_ref = oracle_mod.apply(D, p + 1, q, space.cell_dofs, model.cell_vertices(), x, mp.tables)
    x32 = cuda.from_numpy(x.astype(np.float32)).cuda()
    y32 = cuda.empty_like(x32)
    errs = {}
    cases = [("f32", 1e-5), ("f32_gbf16", 2e-2)]
    if D == 3:
        cases += [("mfma_f32", 1e-5), ("mfma_bf16x3", 2e-2), ("mfma_bf16", 5e-2)]
    for prec, tol in cases:
        A.mul_reduced(y32, x32, prec)
        cuda.cuda.synchronize()
        errs[prec] = rel_l2(y32.cpu().numpy().astype(np.float64), y_ref)
        assert errs[prec] < tol, (prec, errs[prec])
    assert errs["f32_gbf16"] > errs["f32"]
    if D == 3:
        # bf16x3 operands carry FP32 accuracy: same error as bf16 G alone
        assert abs(errs["mfma_bf16x3"] - errs["f32_gbf16"]) < 1e-3 * errs["f32_gbf16"] + 1e-6
        assert errs["mfma_bf16"] > errs["mfma_bf16x3"]
    else:
        with pytest.raises(pkg.SFError) as ei:
            A.mul_reduced(y32, x32, "mfma_bf16")
        assert ei.value.code == 2  # SF_ERR_UNSUPPORTED: the MFMA apply is 3D only
    say("errors", errs)
    A.free()


NQ = [(n, q) for n in range(2, 11) for q in range(n - 1, n + 4)]


def test_mfma_every_instantiated_kernel(pkg, cuda):
    """Every (n, q) of the MFMA apply against the FP32 vector-ALU apply that
    reads the same bf16 G: with three-term operands the two differ by FP32
    rounding only, so a wrong lane layout, table transpose or tile bound
    shows at once.  With one-term operands the gap is bf16 rounding.  The
    FP32-operand form (FP32 G) against the FP32 vector-ALU apply: FP32
    rounding."""
    worst3 = worst1 = 0.0
    for n, q in NQ:
        p = n - 1
        m = 3 if n <= 5 else 2
        say("mfma case p=%d q=%d m=%d" % (p, q, m))
        model, space, mp, A = make(pkg, 3, m, p, q)
        x32 = cuda.from_numpy(pkg.uniform(0, space.n_free, -1, 1).astype(np.float32)).cuda()
        ys = {}
        for prec in ("f32", "mfma_f32", "f32_gbf16", "mfma_bf16x3", "mfma_bf16x2", "mfma_bf16"):
            y = cuda.empty_like(x32)
            A.mul_reduced(y, x32, prec)
            cuda.cuda.synchronize()
            ys[prec] = y.cpu().numpy().astype(np.float64)
        # FP32 operands on the matrix cores against the FP32 vector-ALU apply
        ef = rel_l2(ys["mfma_f32"], ys["f32"])
        assert ef < 1e-5, (p, q, ef)
        e3 = rel_l2(ys["mfma_bf16x3"], ys["f32_gbf16"])
        e2 = rel_l2(ys["mfma_bf16x2"], ys["f32_gbf16"])
        e1 = rel_l2(ys["mfma_bf16"], ys["f32_gbf16"])
        worst3, worst1 = max(worst3, e3), max(worst1, e1)
        assert e3 < 1e-5, (p, q, e3)
        assert e3 < e2 < 2e-4, (p, q, e2)
        assert e2 < e1 < 5e-2, (p, q, e1)
        A.free()
    say("worst rel. l2 vs FP32 apply: bf16x3 %.2e, bf16 %.2e" % (worst3, worst1))


@pytest.mark.parametrize("D,m,p,q", [(3, 4, 5, 7), (3, 3, 3, 5), (2, 8, 4, 6)])
def test_adjoint_vs_oracle_and_fd(pkg, oracle_mod, cuda, D, m, p, q):
    say("adjoint D=%d m=%d p=%d q=%d" % (D, m, p, q))
    model = pkg.CartesianModel(D, m, perturb=0.2, seed=1)
    space = pkg.FESpace(model, p)
    mp = pkg.SumFactorizationMap.with_points(D, p, q)
    nq = q ** D
    kappa = np.exp(pkg.normal(3, space.n_cells * nq))
    A = pkg.LazyMatrix(mp, space, kappa=kappa)
    X = model.cell_vertices()
    u = pkg.uniform(0, space.n_free, -1, 1)
    d = pkg.uniform(4, space.n_free, -1, 1)
    ud = cuda.from_numpy(u).cuda()
    dd = cuda.from_numpy(d).cuda()
    J, kb, ub = pkg.adjoint.loss_and_grad(A, kappa, ud, dd)
    cuda.cuda.synchronize()
    kb = kb.cpu().numpy()
    ub = ub.cpu().numpy()
    # oracle: r = A(kappa) u - d, dJ/dkappa = kappa_grad(u, r), dJ/du = A(kappa) r
    Au = oracle_mod.apply(D, p + 1, q, space.cell_dofs, X, u, mp.tables, kappa=kappa)
    r = Au - d
    J_ref = 0.5 * r @ r
    assert abs(J - J_ref) < 1e-12 * abs(J_ref)
    kb_ref = oracle_mod.kappa_grad(D, p + 1, q, space.cell_dofs, X, u, r, mp.tables).reshape(-1)
    ub_ref = oracle_mod.apply(D, p + 1, q, space.cell_dofs, X, r, mp.tables, kappa=kappa)
    assert rel_l2(kb, kb_ref) < 1e-12
    assert rel_l2(ub, ub_ref) < 1e-12
    # central finite differences, h = 1e-6, FP64, random unit directions.
    # J is quadratic in kappa and in u, so the only error is cancellation;
    # J(+) - J(-) is evaluated as 1/2 (r+ - r-) . (r+ + r-), which keeps the
    # subtraction at the level of the residual vectors (two large sums
    # subtracted would leave rounding of order eps |J| / h).
    def rk(k):
        A.set_kappa(k)
        y = cuda.empty_like(ud)
        A.mul(y, ud)
        return (y - dd).cpu().numpy()

    def ru(v):
        A.set_kappa(kappa)
        y = cuda.empty_like(ud)
        A.mul(y, cuda.from_numpy(v).cuda())
        return (y - dd).cpu().numpy()

    def central(rp, rm):
        return 0.5 * float((rp - rm) @ (rp + rm)) / (2 * h)

    h = 1e-6
    for s in range(4):
        dk = pkg.normal(100 + s, kappa.size)
        dk /= np.linalg.norm(dk)
        fd = central(rk(kappa + h * dk), rk(kappa - h * dk))
        an = kb @ dk
        assert abs(fd - an) < 1e-6 * (abs(an) + 1e-3 * np.linalg.norm(kb)), (s, fd, an)
        du = pkg.normal(200 + s, u.size)
        du /= np.linalg.norm(du)
        fd = central(ru(u + h * du), ru(u - h * du))
        an = ub @ du
        assert abs(fd - an) < 1e-6 * (abs(an) + 1e-3 * np.linalg.norm(ub)), (s, fd, an)
    A.free()


def test_adjoint_every_fused_kernel(pkg, oracle_mod, cuda, monkeypatch):
    """Every 3D (n, q) of the block-assembled fused-geometry adjoint (with
    and without kappa) against the oracle's kappa gradient and A(kappa) lam
    (1e-12), and against the per-cell stored-G kernel of a plan built with
    SFHIP_FUSED_GEOMETRY=0."""
    worst = 0.0
    for n, q in NQ:
        p = n - 1
        m = 3 if n <= 5 else 2
        model = pkg.CartesianModel(3, m, perturb=0.2, seed=1)
        space = pkg.FESpace(model, p)
        mp = pkg.SumFactorizationMap.with_points(3, p, q)
        X = model.cell_vertices()
        u = pkg.uniform(0, space.n_free, -1, 1)
        lam = pkg.uniform(1, space.n_free, -1, 1)
        ud, ld = cuda.from_numpy(u).cuda(), cuda.from_numpy(lam).cuda()
        kb_ref = oracle_mod.kappa_grad(3, n, q, space.cell_dofs, X, u, lam, mp.tables).reshape(-1)
        for with_kappa in (True, False):
            say("fused adjoint case p=%d q=%d m=%d kappa=%d" % (p, q, m, with_kappa))
            kappa = np.exp(pkg.normal(3, space.n_cells * q ** 3)) if with_kappa else None
            ub_ref = oracle_mod.apply(3, n, q, space.cell_dofs, X, lam, mp.tables, kappa=kappa)
            A = pkg.LazyMatrix(mp, space, kappa=kappa)
            kb, ub = pkg.adjoint.pullback(A, ud, ld)
            A.free()
            monkeypatch.setenv("SFHIP_FUSED_GEOMETRY", "0")
            As = pkg.LazyMatrix(mp, space, kappa=kappa)
            monkeypatch.delenv("SFHIP_FUSED_GEOMETRY")
            kbs, ubs = pkg.adjoint.pullback(As, ud, ld)
            As.free()
            cuda.cuda.synchronize()
            kb, ub, kbs, ubs = (t.cpu().numpy() for t in (kb, ub, kbs, ubs))
            errs = (rel_l2(kb, kb_ref), rel_l2(ub, ub_ref), rel_l2(kb, kbs), rel_l2(ub, ubs))
            worst = max(worst, *errs)
            assert max(errs) < 1e-12, (p, q, with_kappa, errs)
    say("worst rel. l2 %.2e" % worst)


def test_adjoint_homogeneity_identity_48cube_q5(pkg, cuda):
    """Size-independent properties of the pullback at the differentiation
    workload's full size (48^3 Q5, q = 7, log-normal kappa): A(kappa) is
    linear in kappa, so sum_q kappa_q dJdk_q = lam . A(kappa) u for the
    cotangent lam (Euler's identity for a degree-1 homogeneous function);
    and dJ/du = A(kappa) lam is the kappa apply itself."""
    say("adjoint homogeneity 48^3 Q5")
    model = pkg.CartesianModel(3, 48, perturb=0.2, seed=1)
    space = pkg.FESpace(model, 5)
    mp = pkg.SumFactorizationMap.with_points(3, 5, 7)
    kappa = np.exp(pkg.normal(3, space.n_cells * 343))
    A = pkg.LazyMatrix(mp, space, kappa=kappa)
    u = cuda.from_numpy(pkg.uniform(0, space.n_free, -1, 1)).cuda()
    lam = cuda.from_numpy(pkg.uniform(4, space.n_free, -1, 1)).cuda()
    kb, ub = pkg.adjoint.pullback(A, u, lam)
    Au = cuda.empty_like(u)
    A.mul(Au, u)
    Al = cuda.empty_like(u)
    A.mul(Al, lam)
    cuda.cuda.synchronize()
    lhs = float((kb * cuda.from_numpy(kappa).cuda()).sum())
    rhs = float((lam * Au).sum())
    say("sum kappa dJdk = %.15e, lam.A u = %.15e" % (lhs, rhs))
    assert abs(lhs - rhs) < 1e-11 * float(kb.abs().mul(cuda.from_numpy(kappa).cuda()).sum())
    assert rel_l2(ub.cpu().numpy(), Al.cpu().numpy()) < 1e-14
    A.free()


def test_adjoint_outputs_optional(pkg, cuda):
    say("adjoint optional outputs")
    model, space, mp, A = make(pkg, 3, 3, 3, 5)
    u = cuda.from_numpy(pkg.uniform(0, space.n_free, -1, 1)).cuda()
    lam = cuda.from_numpy(pkg.uniform(1, space.n_free, -1, 1)).cuda()
    kb1, ub1 = pkg.adjoint.pullback(A, u, lam)
    kb2, _ = pkg.adjoint.pullback(A, u, lam, want_u=False)
    _, ub3 = pkg.adjoint.pullback(A, u, lam, want_kappa=False)
    cuda.cuda.synchronize()
    # dJ/dkappa is written per quadrature point: bitwise reproducible;
    # dJ/du is an atomic scatter-add: equal up to summation order
    assert cuda.equal(kb1, kb2)
    assert rel_l2(ub1.cpu().numpy(), ub3.cpu().numpy()) < 1e-15
    # without kappa, dJ/du = A lam
    y = cuda.empty_like(lam)
    A.mul(y, lam)
    assert rel_l2(ub1.cpu().numpy(), y.cpu().numpy()) < 1e-14
    A.free()


def test_cg_recovers_manufactured_solution(pkg, oracle_mod, cuda):
    say("cg manufactured")
    model, space, mp, A = make(pkg, 3, 6, 3, 5)
    xs = pkg.uniform(7, space.n_free, -1, 1)
    b = oracle_mod.apply(3, 4, 5, space.cell_dofs, model.cell_vertices(), xs, mp.tables)
    bd = cuda.from_numpy(b).cuda()
    x = cuda.zeros_like(bd)
    st = pkg.cg_(x, A, bd, rtol=1e-12, maxiter=5000)
    assert st["converged"] == 1 and st["rel_residual"] <= 1e-12
    xh = x.cpu().numpy()
    # independent residual with the oracle operator
    r = b - oracle_mod.apply(3, 4, 5, space.cell_dofs, model.cell_vertices(), xh, mp.tables)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 2e-12
    assert rel_l2(xh, xs) < 1e-8
    A.free()


def test_cg_maxiter_and_zero_rhs(pkg, cuda):
    say("cg limits")
    model, space, mp, A = make(pkg, 3, 4, 2, 4)
    b = cuda.from_numpy(pkg.normal(2, space.n_free)).cuda()
    x = cuda.zeros_like(b)
    st = pkg.cg_(x, A, b, rtol=1e-14, maxiter=3)
    assert st["iterations"] == 3 and st["converged"] == 0
    with pytest.raises(pkg.SFError):
        pkg.cg_(cuda.zeros_like(b), A, b, rtol=1e-14, maxiter=3, raise_on_fail=True)
    z = cuda.zeros_like(b)
    x = cuda.zeros_like(b)
    st = pkg.cg_(x, A, z, rtol=1e-10)
    assert st["iterations"] == 0 and float(x.abs().max()) == 0.0
    A.free()


def test_cg_fused_step_matches_unfused(pkg, oracle_mod, cuda, monkeypatch):
    """3D CG with the fused-geometry apply takes p.Ap from the apply (per-wave
    shares from its write-back) and fuses x += alpha p into the direction
    update; the stored-G
    plan runs the unfused kernels.  Same iterations (to rounding) and the
    same solution; in 2D (unfused) CG also reaches the oracle residual."""
    say("cg fused vs unfused step")
    model, space, mp, A = make(pkg, 3, 6, 4, 6)
    b = cuda.from_numpy(pkg.normal(5, space.n_free)).cuda()
    xf = cuda.zeros_like(b)
    sf = pkg.cg_(xf, A, b, rtol=1e-10, maxiter=20000)
    A.free()
    monkeypatch.setenv("SFHIP_FUSED_GEOMETRY", "0")
    model, space, mp, A = make(pkg, 3, 6, 4, 6)
    xu = cuda.zeros_like(b)
    su = pkg.cg_(xu, A, b, rtol=1e-10, maxiter=20000)
    A.free()
    monkeypatch.delenv("SFHIP_FUSED_GEOMETRY")
    say("fused", sf, "unfused", su)
    assert sf["converged"] == 1 and su["converged"] == 1
    assert abs(sf["iterations"] - su["iterations"]) <= max(3, 0.01 * su["iterations"])
    bh = b.cpu().numpy()
    for xx in (xf, xu):
        r = bh - oracle_mod.apply(3, 5, 6, space.cell_dofs, model.cell_vertices(),
                                  xx.cpu().numpy(), mp.tables)
        assert np.linalg.norm(r) / np.linalg.norm(bh) < 2e-10
    assert rel_l2(xf.cpu().numpy(), xu.cpu().numpy()) < 1e-8
    # a capped fused run still applies the last step's x update
    x3 = cuda.zeros_like(b)
    model, space, mp, A = make(pkg, 3, 6, 4, 6)
    s3 = pkg.cg_(x3, A, b, rtol=1e-14, maxiter=1)
    assert s3["iterations"] == 1 and float(x3.abs().max()) > 0.0
    # one CG step from x = 0: x = (r.r / r.Ar) r with r = b
    Ab = oracle_mod.apply(3, 5, 6, space.cell_dofs, model.cell_vertices(), bh, mp.tables)
    x1 = (bh @ bh) / (bh @ Ab) * bh
    assert rel_l2(x3.cpu().numpy(), x1) < 1e-12
    A.free()
    # 2D: no blocks, unfused kernels
    model, space, mp, A = make(pkg, 2, 12, 4, 6)
    b2 = cuda.from_numpy(pkg.normal(6, space.n_free)).cuda()
    x2 = cuda.zeros_like(b2)
    s2 = pkg.cg_(x2, A, b2, rtol=1e-10, maxiter=20000)
    assert s2["converged"] == 1
    b2h = b2.cpu().numpy()
    r2 = b2h - oracle_mod.apply(2, 5, 6, space.cell_dofs, model.cell_vertices(),
                                x2.cpu().numpy(), mp.tables)
    assert np.linalg.norm(r2) / np.linalg.norm(b2h) < 2e-10
    A.free()


def test_kappa_fused_geometry_every_kernel(pkg, oracle_mod, cuda, monkeypatch):
    """Every 3D (n, q) of the fused-geometry apply with kappa per quadrature
    point (log-normal, as in the differentiation workload): FP64 against the
    oracle (1e-12) and against the stored kappa-scaled G kernel of a plan
    built with SFHIP_FUSED_GEOMETRY=0; FP32 against the oracle (1e-5)."""
    worst64 = worst32 = 0.0
    for n, q in NQ:
        p = n - 1
        m = 3 if n <= 5 else 2
        say("kappa fused case p=%d q=%d m=%d" % (p, q, m))
        model = pkg.CartesianModel(3, m, perturb=0.2, seed=1)
        space = pkg.FESpace(model, p)
        mp = pkg.SumFactorizationMap.with_points(3, p, q)
        kappa = np.exp(pkg.normal(3, space.n_cells * q ** 3))
        A = pkg.LazyMatrix(mp, space, kappa=kappa)
        x = pkg.uniform(0, space.n_free, -1, 1)
        y_ref = oracle_mod.apply(3, n, q, space.cell_dofs, model.cell_vertices(), x, mp.tables,
                                 kappa=kappa)
        xd = cuda.from_numpy(x).cuda()
        yd = cuda.full_like(xd, float("nan"))
        A.mul(yd, xd)
        x32 = xd.float()
        y32 = cuda.full_like(x32, float("nan"))
        A.mul_reduced(y32, x32, "f32")
        cuda.cuda.synchronize()
        e64 = rel_l2(yd.cpu().numpy(), y_ref)
        e32 = rel_l2(y32.cpu().numpy().astype(np.float64), y_ref)
        A.free()
        monkeypatch.setenv("SFHIP_FUSED_GEOMETRY", "0")
        As = pkg.LazyMatrix(mp, space, kappa=kappa)
        monkeypatch.delenv("SFHIP_FUSED_GEOMETRY")
        ys = cuda.full_like(xd, float("nan"))
        As.mul(ys, xd)
        cuda.cuda.synchronize()
        es = rel_l2(yd.cpu().numpy(), ys.cpu().numpy())
        As.free()
        worst64, worst32 = max(worst64, e64, es), max(worst32, e32)
        assert e64 < 1e-12 and es < 1e-12, (p, q, e64, es)
        assert e32 < 1e-5, (p, q, e32)
    say("worst rel. l2: FP64 %.2e, FP32 %.2e" % (worst64, worst32))


def test_kappa_cg_fused_matches_unfused(pkg, oracle_mod, cuda, monkeypatch):
    """CG with kappa: the fused step (fused-geometry apply reading kappa,
    p.Ap from the apply) against the unfused step on the stored
    kappa-scaled G; FP64 and mixed (FP32 operator, FP64 recurrence)."""
    say("kappa cg fused vs unfused")
    model = pkg.CartesianModel(3, 5, perturb=0.2, seed=1)
    space = pkg.FESpace(model, 4)
    mp = pkg.SumFactorizationMap.with_points(3, 4, 6)
    kappa = np.exp(pkg.normal(3, space.n_cells * 216))
    b = cuda.from_numpy(pkg.normal(5, space.n_free)).cuda()
    res = {}
    for fused in (True, False):
        if not fused:
            monkeypatch.setenv("SFHIP_FUSED_GEOMETRY", "0")
        A = pkg.LazyMatrix(mp, space, kappa=kappa)
        if not fused:
            monkeypatch.delenv("SFHIP_FUSED_GEOMETRY")
        x = cuda.zeros_like(b)
        st = pkg.cg_(x, A, b, rtol=1e-10, maxiter=20000)
        xm = cuda.zeros_like(b)
        sm = pkg.cg_mixed_(xm, A, b, rtol=1e-10, inner="f32", delta=0.1, maxiter=20000)
        res[fused] = (st, sm, x.cpu().numpy(), xm.cpu().numpy())
        A.free()
    bh = b.cpu().numpy()
    for fused in (True, False):
        st, sm, x, xm = res[fused]
        say("fused" if fused else "unfused", st, sm)
        assert st["converged"] == 1 and sm["converged"] == 1
        for xx in (x, xm):
            r = bh - oracle_mod.apply(3, 5, 6, space.cell_dofs, model.cell_vertices(), xx,
                                      mp.tables, kappa=kappa)
            assert np.linalg.norm(r) / np.linalg.norm(bh) < 2e-10
    it_f, it_u = res[True][0]["iterations"], res[False][0]["iterations"]
    assert abs(it_f - it_u) <= max(3, 0.01 * it_u)
    assert rel_l2(res[True][2], res[False][2]) < 1e-8


@pytest.mark.parametrize("inner", ["f32", "f32_gbf16", "mfma_bf16x3"])
def test_mixed_precision_refinement(pkg, oracle_mod, cuda, inner):
    say("IR inner=%s" % inner)
    model, space, mp, A = make(pkg, 3, 6, 5, 7)
    b = cuda.from_numpy(pkg.normal(2, space.n_free)).cuda()
    x64 = cuda.zeros_like(b)
    s64 = pkg.cg_(x64, A, b, rtol=1e-10, maxiter=20000)
    assert s64["converged"] == 1
    x = cuda.zeros_like(b)
    st = pkg.cg_ir_(x, A, b, rtol=1e-10, inner=inner, inner_rtol=1e-2, max_outer=60)
    say("f64", s64, "ir", st)
    assert st["converged"] == 1 and st["rel_residual"] <= 1e-10
    assert st["outer_iterations"] >= 2
    # the FP64 residual, recomputed by the oracle
    xh = x.cpu().numpy()
    bh = b.cpu().numpy()
    r = bh - oracle_mod.apply(3, 6, 7, space.cell_dofs, model.cell_vertices(), xh, mp.tables)
    assert np.linalg.norm(r) / np.linalg.norm(bh) < 2e-10
    assert rel_l2(xh, x64.cpu().numpy()) < 1e-7
    A.free()


@pytest.mark.parametrize("inner", ["f32", "f32_gbf16", "mfma_bf16x3"])
def test_mixed_cg_reliable_updates(pkg, oracle_mod, cuda, inner):
    """One FP32 CG recurrence with FP64 residual replacement: reaches the
    FP64 tolerance, matches the FP64 CG solution, and keeps the Krylov
    space (iteration count close to FP64 CG's, far below a restarted
    refinement's)."""
    say("mixed CG inner=%s" % inner)
    model, space, mp, A = make(pkg, 3, 6, 5, 7)
    b = cuda.from_numpy(pkg.normal(2, space.n_free)).cuda()
    x64 = cuda.zeros_like(b)
    s64 = pkg.cg_(x64, A, b, rtol=1e-10, maxiter=20000)
    assert s64["converged"] == 1
    x = cuda.zeros_like(b)
    st = pkg.cg_mixed_(x, A, b, rtol=1e-10, inner=inner, delta=0.1, recurrence="f32")
    say("f64", s64, "mixed", st)
    assert st["converged"] == 1 and st["rel_residual"] <= 1e-10
    assert st["outer_iterations"] >= 5  # 1e-10 in steps of at most 0.1
    assert st["iterations"] < 1.5 * s64["iterations"]
    xh = x.cpu().numpy()
    bh = b.cpu().numpy()
    r = bh - oracle_mod.apply(3, 6, 7, space.cell_dofs, model.cell_vertices(), xh, mp.tables)
    assert np.linalg.norm(r) / np.linalg.norm(bh) < 2e-10
    assert rel_l2(xh, x64.cpu().numpy()) < 1e-7
    # a converged start returns at once
    st2 = pkg.cg_mixed_(x, A, b, rtol=1e-9, inner=inner, recurrence="f32")
    assert st2["iterations"] == 0 and st2["converged"] == 1
    A.free()


@pytest.mark.parametrize("inner,rec", [("f32", "f64"), ("f32_gbf16", "f64"),
                                       ("f32", "f64p32")])
def test_mixed_cg_fp64_recurrence(pkg, oracle_mod, cuda, inner, rec):
    """FP64 Krylov recurrence with only the operator reduced ('f64p32': the
    search direction held in FP32 as well): reaches the FP64 tolerance in
    the FP64 residual, agrees with FP64 CG, and needs about as many
    iterations as FP64 CG (the FP32 recurrence needs noticeably more)."""
    say("mixed CG, %s recurrence, inner=%s" % (rec, inner))
    model, space, mp, A = make(pkg, 3, 6, 5, 7)
    b = cuda.from_numpy(pkg.normal(3, space.n_free)).cuda()
    x64 = cuda.zeros_like(b)
    s64 = pkg.cg_(x64, A, b, rtol=1e-10, maxiter=20000)
    assert s64["converged"] == 1
    x = cuda.zeros_like(b)
    st = pkg.cg_mixed_(x, A, b, rtol=1e-10, inner=inner, delta=0.1, recurrence=rec)
    say("f64", s64, "mixed/%s recurrence" % rec, st)
    assert st["converged"] == 1 and st["rel_residual"] <= 1e-10
    assert st["outer_iterations"] >= 5
    assert st["iterations"] < 1.25 * s64["iterations"]
    xh = x.cpu().numpy()
    bh = b.cpu().numpy()
    r = bh - oracle_mod.apply(3, 6, 7, space.cell_dofs, model.cell_vertices(), xh, mp.tables)
    assert np.linalg.norm(r) / np.linalg.norm(bh) < 2e-10
    assert rel_l2(xh, x64.cpu().numpy()) < 1e-7
    # nonzero initial guess: continues from it
    x2 = 0.5 * x64.clone()
    st3 = pkg.cg_mixed_(x2, A, b, rtol=1e-10, inner=inner, recurrence=rec)
    assert st3["converged"] == 1
    assert rel_l2(x2.cpu().numpy(), x64.cpu().numpy()) < 1e-7
    with pytest.raises(ValueError):
        pkg.cg_mixed_(x, A, b, recurrence="f16")
    import ctypes
    rc_bad = A._L.sf_cg_mixed(A.plan, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                              1e-10, 10, pkg._lib.PRECISIONS["f32"], 0.1, 3, None, None)
    assert rc_bad == pkg._lib.SF_ERR_ARG
    A.free()
