"""GPU test of sf_cg_noalloc, the CG entry point that allocates nothing
(its workspace lives in the library's fixed per-GPU scratch array): the
same iterations and solution as sf_cg on two calls in a row, the second on
a smaller problem, with and without the Gauss-Lobatto preconditioner; the
plan's own device memory does not grow; problems too large for the
scratch and preconditioners not yet built are refused."""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def plan(pkg, m, p, q, kappa=True):
    model = pkg.CartesianModel(3, m, perturb=0.2, seed=1)
    space = pkg.FESpace(model, p)
    mp = pkg.SumFactorizationMap.with_points(3, p, q)
    k = np.exp(pkg.normal(3, space.n_cells * q ** 3)) if kappa else None
    return pkg.LazyMatrix(mp, space, kappa=k), space, model, mp, k


@pytest.mark.parametrize("Pl", [None, "gll"])
def test_cg_noalloc_matches_cg(pkg, oracle_mod, cuda, Pl):
    """Both entry points run the same kernels on the same workspace layout:
    after a fixed number of iterations their iterates agree to 1e-12 (the
    apply's atomic adds leave the summation order, and so the last bits,
    free); run to 1e-10 both converge, with the FP64 residual recomputed by
    the oracle, to the same solution."""
    for m, p, q in ((8, 4, 6), (5, 3, 5)):  # the second call on a smaller problem
        say("cg_noalloc m=%d p=%d q=%d Pl=%s" % (m, p, q, Pl))
        A, space, model, mp, kappa = plan(pkg, m, p, q)
        b = cuda.from_numpy(pkg.normal(2, space.n_free)).cuda()
        if Pl is not None:  # build the preconditioner (the no-alloc call cannot)
            z = cuda.empty_like(b)
            A.precond(z, b, Pl)
        before = A.info()["device_bytes"]
        k = 10
        xn = cuda.zeros_like(b)
        sn = pkg.cg_noalloc_(xn, A, b, rtol=1e-14, maxiter=k)
        cuda.cuda.synchronize()
        assert A.info()["device_bytes"] == before  # nothing allocated on the plan
        xr = cuda.zeros_like(b)
        sr = pkg.cg_(xr, A, b, rtol=1e-14, maxiter=k, Pl=Pl)
        cuda.cuda.synchronize()
        assert A.info()["device_bytes"] > before  # sf_cg allocates its workspace
        assert sn["iterations"] == sr["iterations"] == k
        assert rel_l2(xn.cpu().numpy(), xr.cpu().numpy()) < 1e-12
        assert abs(sn["rel_residual"] - sr["rel_residual"]) <= 1e-10 * sr["rel_residual"]
        xn = cuda.zeros_like(b)
        sn = pkg.cg_noalloc_(xn, A, b, rtol=1e-10, maxiter=20000)
        xr = cuda.zeros_like(b)
        sr = pkg.cg_(xr, A, b, rtol=1e-10, maxiter=20000, Pl=Pl)
        say("  noalloc", sn, "cg", sr)
        # to convergence the counts may differ a little: the apply's FP64
        # atomics leave the summation order free, and the stop (a check of
        # the true residual, continued when above rtol) amplifies that
        assert sn["converged"] == 1
        assert abs(sn["iterations"] - sr["iterations"]) <= max(2, 0.03 * sr["iterations"])
        bh = b.cpu().numpy()
        r = bh - oracle_mod.apply(3, p + 1, q, space.cell_dofs, model.cell_vertices(),
                                  xn.cpu().numpy(), mp.tables, kappa=kappa)
        assert np.linalg.norm(r) / np.linalg.norm(bh) < 2e-10
        assert rel_l2(xn.cpu().numpy(), xr.cpu().numpy()) < 1e-8
        A.free()


def test_cg_noalloc_refusals(pkg, cuda):
    say("cg_noalloc refusals")
    A, space, _, _, _ = plan(pkg, 3, 3, 5, kappa=False)
    b = cuda.from_numpy(pkg.normal(2, space.n_free)).cuda()
    x = cuda.zeros_like(b)
    A._L.sf_plan_set_preconditioner(A.plan, pkg._lib.SF_PC_GLL)
    with pytest.raises(pkg.SFError) as ei:
        pkg.cg_noalloc_(x, A, b)
    assert ei.value.code == pkg._lib.SF_ERR_NOT_READY
    A.free()
    # 42^3 Q4: 167^3 = 4.66M free dofs, 149 MB of workspace > 128 MiB
    A, space, _, _, _ = plan(pkg, 42, 4, 6, kappa=False)
    b = cuda.ones(space.n_free, dtype=cuda.float64, device="cuda")
    x = cuda.zeros_like(b)
    with pytest.raises(pkg.SFError) as ei:
        pkg.cg_noalloc_(x, A, b)
    assert ei.value.code == pkg._lib.SF_ERR_ARG and "scratch" in str(ei.value)
    A.free()
