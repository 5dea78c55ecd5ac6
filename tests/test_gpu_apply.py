"""GPU parity of the HIP apply (FP64) against the CPU oracle, through the
C ABI.  Bar: 1e-12 relative l2 (the FP64 parity tolerance)."""
import ctypes
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(pkg, D, m, p, q, perturb=0.2, seed=1, dirichlet=True, kappa=None):
    model = pkg.CartesianModel(D, m, perturb=perturb, seed=seed)
    space = pkg.FESpace(model, p, dirichlet=dirichlet)
    mp = pkg.SumFactorizationMap.with_points(D, p, q)
    A = pkg.LazyMatrix(mp, space, kappa=kappa)
    return model, space, mp, A


def gpu_apply(cuda, A, x):
    xd = cuda.from_numpy(np.ascontiguousarray(x)).cuda()
    yd = cuda.full_like(xd, float("nan"))  # overwrite semantics: garbage in y
    A.mul(yd, xd)
    cuda.cuda.synchronize()
    return yd.cpu().numpy()


NQ = [(n, q) for n in range(2, 11) for q in range(n - 1, n + 4)]


@pytest.mark.parametrize("D", [2, 3])
def test_every_instantiated_kernel_vs_oracle(pkg, oracle_mod, cuda, D):
    worst = 0.0
    for n, q in NQ:
        p = n - 1
        m = 3 if D == 3 and n <= 5 else (2 if D == 3 else 4)
        say("case D=%d p=%d q=%d m=%d" % (D, p, q, m))
        model, space, mp, A = make(pkg, D, m, p, q)
        x = pkg.uniform(0, space.n_free, -1, 1)
        y = gpu_apply(cuda, A, x)
        y_ref = oracle_mod.apply(D, n, q, space.cell_dofs, model.cell_vertices(), x, mp.tables)
        e = rel_l2(y, y_ref)
        worst = max(worst, e)
        assert e < TOL, (D, p, q, e)
        A.free()
    say("worst rel. l2 over %d kernels: %.2e" % (len(NQ), worst))


def test_full_blocks_every_kernel(pkg, oracle_mod, cuda):
    """3D meshes whose rows hold exactly one full cell block (m = cells per
    block, up to 64 at the lowest orders): every (n, q) of the fused-geometry
    FP64 and FP32 applies and of the block adjoint, with kappa, against the
    oracle.  Blocks of more than NT / 24 cells need several vertex loads per
    thread; the small meshes above never fill them."""
    worst = 0.0
    for n, q in NQ:
        p = n - 1
        _, _, _, A0 = make(pkg, 3, 2, p, q)
        eb = A0.info()["cells_per_block"]
        A0.free()
        m = max(eb, 2)
        say("full blocks p=%d q=%d m=%d" % (p, q, m))
        model = pkg.CartesianModel(3, m, perturb=0.2, seed=1)
        space = pkg.FESpace(model, p)
        mp = pkg.SumFactorizationMap.with_points(3, p, q)
        kappa = np.exp(pkg.normal(3, space.n_cells * q ** 3))
        A = pkg.LazyMatrix(mp, space, kappa=kappa)
        if m == eb:  # one full block per mesh row
            assert A.info()["n_blocks"] == m * m
        X = model.cell_vertices()
        x = pkg.uniform(0, space.n_free, -1, 1)
        lam = pkg.uniform(1, space.n_free, -1, 1)
        y_ref = oracle_mod.apply(3, n, q, space.cell_dofs, X, x, mp.tables, kappa=kappa)
        xd = cuda.from_numpy(x).cuda()
        yd = cuda.full_like(xd, float("nan"))
        A.mul(yd, xd)
        y32 = cuda.full_like(xd.float(), float("nan"))
        A.mul_reduced(y32, xd.float(), "f32")
        kb, ub = pkg.adjoint.pullback(A, xd, cuda.from_numpy(lam).cuda())
        cuda.cuda.synchronize()
        kb_ref = oracle_mod.kappa_grad(3, n, q, space.cell_dofs, X, x, lam, mp.tables).reshape(-1)
        ub_ref = oracle_mod.apply(3, n, q, space.cell_dofs, X, lam, mp.tables, kappa=kappa)
        e = (rel_l2(yd.cpu().numpy(), y_ref), rel_l2(kb.cpu().numpy(), kb_ref),
             rel_l2(ub.cpu().numpy(), ub_ref))
        e32 = rel_l2(y32.cpu().numpy().astype(np.float64), y_ref)
        worst = max(worst, *e)
        assert max(e) < TOL and e32 < 1e-5, (p, q, m, e, e32)
        A.free()
    say("worst rel. l2 %.2e" % worst)


def test_2d_blocks_every_kernel(pkg, oracle_mod, cuda, monkeypatch):
    """2D meshes of m = cells per block + 3 cells per row, so blocks span
    row ends and hold up to 64 cells (several vertex loads per thread):
    every (n, q) of the block-assembled 2D apply -- fused geometry and
    stored G in FP64, fused FP32, and y0 + A x (exclusive dofs
    read-modify-written) -- against the oracle."""
    worst = 0.0
    for n, q in NQ:
        p = n - 1
        _, _, _, A0 = make(pkg, 2, 2, p, q)
        eb = A0.info()["cells_per_block"]
        A0.free()
        m = eb + 3
        say("2D blocks p=%d q=%d m=%d (%d cells per block)" % (p, q, m, eb))
        model = pkg.CartesianModel(2, m, perturb=0.2, seed=2)
        space = pkg.FESpace(model, p)
        mp = pkg.SumFactorizationMap.with_points(2, p, q)
        X = model.cell_vertices()
        x = pkg.uniform(0, space.n_free, -1, 1)
        y0 = pkg.uniform(1, space.n_free, -1, 1)
        y_ref = oracle_mod.apply(2, n, q, space.cell_dofs, X, x, mp.tables)
        xd = cuda.from_numpy(x).cuda()
        errs = []
        for fused in (True, False):
            if not fused:
                monkeypatch.setenv("SFHIP_FUSED_GEOMETRY", "0")
            A = pkg.LazyMatrix(mp, space)
            if not fused:
                monkeypatch.delenv("SFHIP_FUSED_GEOMETRY")
            assert A.info()["n_blocks"] == -(-space.n_cells // eb)
            yd = cuda.full_like(xd, float("nan"))
            A.mul(yd, xd)
            ya = cuda.from_numpy(y0.copy()).cuda()
            A.mul_add(ya, xd)
            y32 = cuda.full_like(xd.float(), float("nan"))
            A.mul_reduced(y32, xd.float(), "f32")
            cuda.cuda.synchronize()
            errs += [rel_l2(yd.cpu().numpy(), y_ref), rel_l2(ya.cpu().numpy(), y0 + y_ref)]
            e32 = rel_l2(y32.cpu().numpy().astype(np.float64), y_ref)
            assert e32 < 1e-5, (p, q, fused, e32)
            A.free()
        worst = max(worst, *errs)
        assert max(errs) < TOL, (p, q, errs)
    say("worst rel. l2 %.2e" % worst)


def test_tiny_meshes(pkg, oracle_mod, cuda):
    """One cell, or two per direction, with and without Dirichlet boundaries
    (down to no free dof at all): blocks holding a single partial cell, cells
    whose every dof is Dirichlet, FP64 and FP32, 2D and 3D, against the
    oracle; the solvers return at once for an empty system."""
    for D in (2, 3):
        for m, p, q in [(1, 1, 2), (1, 2, 3), (1, 3, 4), (2, 1, 2)]:
            for dirichlet in (True, False):
                say("tiny D=%d m=%d p=%d q=%d dirichlet=%d" % (D, m, p, q, dirichlet))
                model, space, mp, A = make(pkg, D, m, p, q, perturb=0.2 if m > 1 else 0.0,
                                           seed=3, dirichlet=dirichlet)
                x = pkg.uniform(0, space.n_free, -1, 1)
                y_ref = oracle_mod.apply(D, p + 1, q, space.cell_dofs, model.cell_vertices(), x,
                                         mp.tables)
                y = gpu_apply(cuda, A, x)
                xd = cuda.from_numpy(x).cuda()
                y32 = cuda.full_like(xd.float(), float("nan"))
                A.mul_reduced(y32, xd.float(), "f32")
                cuda.cuda.synchronize()
                if space.n_free == 0:
                    assert y.size == 0 and y32.numel() == 0
                    b = cuda.zeros(0, dtype=cuda.float64, device="cuda")
                    for st in (pkg.cg_(cuda.zeros_like(b), A, b, rtol=1e-10, maxiter=10),
                               pkg.cg_mixed_(cuda.zeros_like(b), A, b, rtol=1e-10, maxiter=10),
                               pkg.cg_ir_(cuda.zeros_like(b), A, b, rtol=1e-10)):
                        assert st["iterations"] == 0 and st["converged"] == 1
                else:
                    assert rel_l2(y, y_ref) < TOL, (D, m, p, q, dirichlet)
                    assert rel_l2(y32.cpu().numpy().astype(np.float64), y_ref) < 1e-5
                A.free()


@pytest.mark.parametrize("p,q,m,perturb", [(1, 3, 4, 0.0), (1, 3, 4, 0.2), (3, 5, 3, 0.2)])
def test_demo_step_by_step_on_gpu(pkg, oracle_mod, cuda, p, q, m, perturb):
    """test/Demo_StepByStep.jl's case (2D, Q1, 4 x 4 cells, 3 Gauss points,
    x = ones) and two perturbed variants: the HIP apply against the
    explicit-loop restatement of the demo; on the demo's own case also the
    textbook Q1 stencil ((8 - free neighbours) / 3 per free dof)."""
    from oracle import stepbystep
    say("step-by-step demo p=%d q=%d m=%d perturb=%g" % (p, q, m, perturb))
    model, space, mp, A = make(pkg, 2, m, p, q, perturb=perturb)
    x = np.ones(space.n_free) if perturb == 0.0 else pkg.uniform(0, space.n_free, -1, 1)
    y = gpu_apply(cuda, A, x)
    tab = [np.asarray(t, dtype=np.float64).tolist() for t in mp.tables]
    y_sbs = np.array(stepbystep.apply_2d(space.cell_dofs, model.cell_vertices(), x, tab))
    assert rel_l2(y, y_sbs) < TOL
    if perturb == 0.0 and p == 1:
        F = m - 1
        k = np.zeros((F, F))
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if di or dj:
                    k[max(0, -dj):F - max(0, dj), max(0, -di):F - max(0, di)] += 1
        assert np.abs(y - (8 - k.reshape(-1)) / 3.0).max() < 1e-14
    A.free()


def test_2d_64sq_q4_vs_oracle_and_assembled(pkg, oracle_mod, brute, cuda):
    """2D unit square, 64x64 Cartesian quads, Q4, 6 Gauss points, u ~ U[-1,1]
    seed 0: against the oracle and the assembled CSR SpMV."""
    say("2D 64^2 Q4 q=6")
    model, space, mp, A = make(pkg, 2, 64, 4, 6, perturb=0.0)
    assert space.n_free == 65025
    x = pkg.uniform(0, space.n_free, -1, 1)
    y = gpu_apply(cuda, A, x)
    X = model.cell_vertices()
    y_or = oracle_mod.apply(2, 5, 6, space.cell_dofs, X, x, mp.tables)
    B, Dm, pts, w = brute.tables(4, 6)
    G, _ = brute.geometry_numpy(2, X, pts, w)
    Acsr = brute.assemble_csr(space.cell_dofs, brute.element_matrices(2, B, Dm, G), space.n_free)
    y_sp = Acsr @ x
    say("rel l2 vs oracle %.2e, vs assembled %.2e" % (rel_l2(y, y_or), rel_l2(y, y_sp)))
    assert rel_l2(y, y_or) < TOL
    assert rel_l2(y, y_sp) < TOL


@pytest.mark.parametrize("D,m,p,q", [(3, 8, 4, 6), (3, 16, 4, 6), (3, 4, 7, 9), (3, 6, 5, 7),
                                     (3, 3, 9, 11), (3, 8, 2, 4), (3, 8, 3, 5), (3, 5, 6, 8)])
def test_small_meshes_vs_oracle(pkg, oracle_mod, cuda, D, m, p, q):
    say("small mesh D=%d m=%d p=%d q=%d" % (D, m, p, q))
    model, space, mp, A = make(pkg, D, m, p, q)
    x = pkg.uniform(0, space.n_free, -1, 1)
    y = gpu_apply(cuda, A, x)
    y_ref = oracle_mod.apply(D, p + 1, q, space.cell_dofs, model.cell_vertices(), x, mp.tables)
    assert rel_l2(y, y_ref) < TOL


def test_3d_64cube_q4_full_size_vs_oracle(pkg, oracle_mod, cuda):
    """3D 64^3 perturbed hexes, Q4, 6 points (16.6M dofs), u ~ U[-1,1]."""
    say("full size: 3D 64^3 Q4 q=6 perturbed")
    model, space, mp, A = make(pkg, 3, 64, 4, 6)
    assert space.n_free == 16581375
    x = pkg.uniform(0, space.n_free, -1, 1)
    y = gpu_apply(cuda, A, x)
    say("GPU done; oracle (CPU) ...")
    y_ref = oracle_mod.apply(3, 5, 6, space.cell_dofs, model.cell_vertices(), x, mp.tables)
    e = rel_l2(y, y_ref)
    say("rel l2 %.2e" % e)
    assert e < TOL
    # size-independent properties at full size: symmetry and linearity
    v = pkg.uniform(9, space.n_free, -1, 1)
    Av = gpu_apply(cuda, A, v)
    s1, s2 = v @ y, x @ Av
    assert abs(s1 - s2) < 1e-12 * abs(x @ y)
    Axv = gpu_apply(cuda, A, 2.0 * x - 3.0 * v)
    assert rel_l2(Axv, 2.0 * y - 3.0 * Av) < 1e-13
    A.free()


@pytest.mark.parametrize("cfg,m,p,q,with_kappa,n_free", [
    ("32cube_q7", 32, 7, 9, False, 11089567),
    ("48cube_q5_kappa", 48, 5, 7, True, 13651919),
    ("24cube_q9_kappa", 24, 9, 11, True, 9938375)])
def test_large_meshes_full_size_vs_oracle(pkg, oracle_mod, cuda, cfg, m, p, q, with_kappa, n_free):
    """The other 3D workloads at their full sizes (perturbed meshes; log-normal
    kappa per quadrature point where the workload has one): the FP64 apply
    against the oracle (1e-12), symmetry x.Av = v.Ax, and the FP32 apply
    within 1e-5 of the oracle."""
    say("full size: %s (m=%d p=%d q=%d kappa=%d)" % (cfg, m, p, q, with_kappa))
    model = pkg.CartesianModel(3, m, perturb=0.2, seed=1)
    space = pkg.FESpace(model, p)
    assert space.n_free == n_free
    mp = pkg.SumFactorizationMap.with_points(3, p, q)
    kappa = np.exp(pkg.normal(3, space.n_cells * q ** 3)) if with_kappa else None
    A = pkg.LazyMatrix(mp, space, kappa=kappa)
    x = pkg.uniform(0, space.n_free, -1, 1)
    v = pkg.uniform(9, space.n_free, -1, 1)
    y = gpu_apply(cuda, A, x)
    Av = gpu_apply(cuda, A, v)
    xd = cuda.from_numpy(x).cuda()
    y32 = cuda.full_like(xd.float(), float("nan"))
    A.mul_reduced(y32, xd.float(), "f32")
    cuda.cuda.synchronize()
    y32 = y32.cpu().numpy().astype(np.float64)
    say("GPU done; oracle (CPU) ...")
    y_ref = oracle_mod.apply(3, p + 1, q, space.cell_dofs, model.cell_vertices(), x, mp.tables,
                             kappa=kappa)
    e, e32 = rel_l2(y, y_ref), rel_l2(y32, y_ref)
    say("rel l2 FP64 %.2e, FP32 %.2e" % (e, e32))
    assert e < TOL and e32 < 1e-5
    assert abs(v @ y - x @ Av) < 1e-12 * abs(x @ y)
    A.free()


def test_patch_test_full_operator(pkg, cuda):
    """All dofs free (no Dirichlet): A 1 = 0 and interior rows of A (a.x + b)
    vanish on a perturbed mesh (exact quadrature, q = p + 2)."""
    for D, m, p in [(3, 12, 4), (2, 32, 5)]:
        say("patch test D=%d m=%d p=%d" % (D, m, p))
        model, space, mp, A = make(pkg, D, m, p, p + 2, dirichlet=False)
        one = gpu_apply(cuda, A, np.ones(space.n_free))
        assert np.max(np.abs(one)) < 1e-12
        xyz = space.free_dof_coordinates()
        u = xyz @ np.array([0.3, -0.8, 1.1][:D]) - 0.5
        r = gpu_apply(cuda, A, u)
        nL = space.n_lattice
        idx = np.arange(space.n_free)
        lat = np.stack([(idx // nL ** d) % nL for d in range(D)], -1)
        interior = np.all((lat > 0) & (lat < nL - 1), axis=1)
        assert np.max(np.abs(r[interior])) < 1e-12
        A.free()


def test_device_geometry_vs_oracle(pkg, oracle_mod, cuda):
    for D, p, q in [(3, 4, 6), (3, 7, 9), (2, 4, 6)]:
        say("geometry D=%d p=%d q=%d" % (D, p, q))
        model, space, mp, A = make(pkg, D, 5, p, q)
        G = A.geometry()
        G_ref = oracle_mod.geometry(D, q, model.cell_vertices(), mp.tables)
        assert np.max(np.abs(G - G_ref)) < 1e-14 * np.max(np.abs(G_ref))
        # the same operator from host-supplied G
        x = pkg.uniform(0, space.n_free, -1, 1)
        A2 = pkg.LazyMatrix(mp, space, geometry=G_ref)
        assert rel_l2(gpu_apply(cuda, A2, x), gpu_apply(cuda, A, x)) < 1e-14
        A.free()
        A2.free()


def test_host_pointer_apply_and_input_untouched(pkg, cuda):
    say("host apply")
    model, space, mp, A = make(pkg, 3, 6, 3, 5)
    x = pkg.uniform(0, space.n_free, -1, 1)
    x0 = x.copy()
    yh = np.full(space.n_free, np.nan)
    pkg.mul_(yh, A, x)
    assert np.array_equal(x, x0)
    yd = gpu_apply(cuda, A, x)
    assert np.array_equal(yh, yd) or rel_l2(yh, yd) < 1e-15
    # repeated applies overwrite (no accumulation across calls)
    yh2 = np.empty_like(yh)
    pkg.mul_(yh2, A, x)
    assert rel_l2(yh2, yh) < 1e-15


def test_demo_jl_on_gpu(pkg, brute, cuda):
    """test/Demo.jl:57: norm(A*1 - A_lazy*1) < 1e-6, 3D Q4, degree 6, 10^3."""
    say("Demo.jl on GPU")
    m = pkg.SumFactorizationMap(3, (4, 4, 4), (6, 6, 6))
    model = pkg.CartesianModel(3, 10)
    space = pkg.FESpace(model, 4)
    A_lazy = pkg.LazyMatrix(m, space)
    y = np.ones(space.n_free)
    x2 = np.zeros(space.n_free)
    pkg.mul_(x2, A_lazy, y)
    B, Dm, pts, w = brute.tables(4, 4)
    G, _ = brute.geometry_numpy(3, model.cell_vertices(), pts, w)
    A = brute.assemble_csr(space.cell_dofs, brute.element_matrices(3, B, Dm, G), space.n_free)
    x1 = A @ y
    assert np.linalg.norm(x1 - x2) < 1e-6


def test_error_paths(pkg, cuda):
    say("error paths")
    L = pkg.load_library()
    model = pkg.CartesianModel(3, 2)
    space = pkg.FESpace(model, 2)
    mp = pkg.SumFactorizationMap.with_points(3, 2, 4)
    h = ctypes.c_void_p()
    assert L.sf_plan_create(ctypes.byref(h), 3, 2, 4, space.n_cells, space.n_free, 0) == 0
    x = cuda.zeros(space.n_free, dtype=cuda.float64, device="cuda")
    y = cuda.zeros_like(x)
    # not ready
    rc = L.sf_apply(h, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()), None)
    assert rc == pkg._lib.SF_ERR_NOT_READY
    # dof id out of range
    bad = space.cell_dofs.copy()
    bad[0, 5] = space.n_free + 3
    assert L.sf_plan_set_cells(h, bad.ctypes.data, 0, None) == pkg._lib.SF_ERR_DOF_RANGE
    assert b"outside" in L.sf_last_error()
    # a local_to_tensor that is not a permutation
    perm = np.zeros(27, dtype=np.int32)
    assert L.sf_plan_set_cells(h, space.cell_dofs.ctypes.data, 0, perm.ctypes.data) == pkg._lib.SF_ERR_ARG
    # 3D geometry needs the cells first (its layout follows the cell blocks)
    X = model.cell_vertices().copy()
    assert L.sf_plan_set_vertices(h, X.ctypes.data) == pkg._lib.SF_ERR_NOT_READY
    assert L.sf_plan_set_cells(h, space.cell_dofs.ctypes.data, 0, None) == 0
    # inverted cell: det J < 0
    X[0] = X[0][[1, 0, 3, 2, 5, 4, 7, 6]]  # mirror in x
    assert L.sf_plan_set_vertices(h, X.ctypes.data) == pkg._lib.SF_ERR_JACOBIAN
    # non-positive kappa
    k = np.ones(space.n_cells * 64)
    k[3] = -1.0
    assert L.sf_plan_set_kappa(h, k.ctypes.data) == pkg._lib.SF_ERR_ARG
    assert L.sf_plan_destroy(h) == 0
    with pytest.raises(pkg.SFError):
        pkg.LazyMatrix(mp, space, kappa=-np.ones(space.n_cells * 64))
    _ = mp


def test_gridap_convention_cells_and_local_permutation(pkg, oracle_mod, cuda):
    """index_base = 1 (Gridap: free ids 1..n, Dirichlet <= 0) and a
    local-to-tensor permutation (the reference's dof_map) give the same
    operator as 0-based tensor-ordered ids."""
    say("gridap convention cells")
    model, space, mp, A = make(pkg, 3, 4, 3, 5)
    L = pkg.load_library()
    rng = np.random.default_rng(0)
    perm = rng.permutation(64).astype(np.int32)  # local k -> tensor slot perm[k]
    cd = space.cell_dofs
    local = np.empty_like(cd)
    local[:, np.arange(64)] = cd[:, perm]  # local dof k holds tensor slot perm[k]
    gridap = np.where(local >= 0, local + 1, -7).astype(np.int32)
    h = ctypes.c_void_p()
    assert L.sf_plan_create(ctypes.byref(h), 3, 3, 5, space.n_cells, space.n_free, 0) == 0
    assert L.sf_plan_set_cells(h, gridap.ctypes.data, 1, perm.ctypes.data) == 0
    X = model.cell_vertices()
    assert L.sf_plan_set_vertices(h, X.ctypes.data) == 0
    x = pkg.uniform(0, space.n_free, -1, 1)
    y = np.zeros(space.n_free)
    assert L.sf_apply_host(h, y.ctypes.data, x.ctypes.data) == 0
    L.sf_plan_destroy(h)
    assert rel_l2(y, gpu_apply(cuda, A, x)) < 1e-15


def test_apply_kernel_timing(pkg, cuda):
    """sf_plan_set_timing / sf_plan_read_timing: one duration per apply
    kernel while enabled (FP64 and reduced precision), none while disabled,
    the record cleared by a read; each duration is positive and no longer
    than the host-timed span of the applies."""
    import time
    say("apply kernel timing")
    _, space, _, A = make(pkg, 3, 6, 4, 6)
    x = cuda.from_numpy(pkg.uniform(0, space.n_free, -1, 1)).cuda()
    y = cuda.empty_like(x)
    x32, y32 = x.float(), cuda.empty_like(x.float())
    A.mul(y, x)  # set-up and first launches outside the record
    A.mul_reduced(y32, x32, "f32")
    cuda.cuda.synchronize()
    A.timing(True)
    t0 = time.perf_counter()
    for _ in range(3):
        A.mul(y, x)
    A.mul_reduced(y32, x32, "f32")
    cuda.cuda.synchronize()
    span_ms = (time.perf_counter() - t0) * 1e3
    A.timing(False)
    A.mul(y, x)  # not recorded
    ms = A.read_timing()
    say("  kernel ms", ms, "span ms", span_ms)
    assert len(ms) == 4
    assert all(0.0 < m <= span_ms for m in ms)
    assert A.read_timing() == []  # the read cleared the record
    A.free()
