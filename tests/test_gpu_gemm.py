"""GPU tests of the FP16 matrix-core apply (SF_PREC_MFMA_F16: every
contraction an FP16 GEMM over pencils batched across cells), through the C
ABI: error against the FP64 oracle and against the FP32 vector-ALU apply on
the same bf16 G, power-of-two scale invariance (its FP16 conversions scale
per block), the refusal outside n = 8, and a mixed-precision solve
with it as the inner operator."""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(pkg, m, p, q, kappa=False):
    model = pkg.CartesianModel(3, m, perturb=0.2, seed=1)
    space = pkg.FESpace(model, p)
    mp = pkg.SumFactorizationMap.with_points(3, p, q)
    k = np.exp(pkg.normal(5, space.n_cells * q ** 3)) if kappa else None
    return model, space, mp, k, pkg.LazyMatrix(mp, space, kappa=k)


# (m, q, kappa): rows of m cells against blocks of 4 / 3 / 2 cells (q = 8 /
# 9 / 10), so full and partial blocks both occur
CASES = [(2, 8, False), (3, 8, True), (3, 9, False), (4, 9, True), (5, 9, False),
         (2, 10, False), (3, 10, True)]


@pytest.mark.parametrize("m,q,with_kappa", CASES)
def test_gemm_f16_apply(pkg, oracle_mod, cuda, m, q, with_kappa):
    """FP16 operands carry 11 significant bits: against the FP32 apply that
    reads the same bf16 G the difference is FP16 rounding (<= 3e-3), against
    the FP64 oracle the bf16 G dominates (<= 2e-2).  Scaling x by 2^20 or
    2^-20 scales y by that factor: the same FP16 operands, up to the order
    of the FP32 atomic adds."""
    say("gemm f16 m=%d q=%d kappa=%d" % (m, q, with_kappa))
    model, space, mp, kappa, A = make(pkg, m, 7, q, with_kappa)
    xh = pkg.uniform(0, space.n_free, -1, 1)
    y_ref = oracle_mod.apply(3, 8, q, space.cell_dofs, model.cell_vertices(), xh, mp.tables,
                             kappa=kappa)
    x32 = cuda.from_numpy(xh.astype(np.float32)).cuda()
    ys = {}
    for prec in ("f32", "f32_gbf16", "mfma_f16"):
        y = cuda.full_like(x32, float("nan"))  # every free dof must be written
        A.mul_reduced(y, x32, prec)
        cuda.cuda.synchronize()
        ys[prec] = y.cpu().numpy().astype(np.float64)
    e_or = rel_l2(ys["mfma_f16"], y_ref)
    e_bf = rel_l2(ys["mfma_f16"], ys["f32_gbf16"])
    say("  vs oracle %.2e, vs FP32 apply with bf16 G %.2e (FP32 apply vs oracle %.2e)"
        % (e_or, e_bf, rel_l2(ys["f32"], y_ref)))
    assert np.all(np.isfinite(ys["mfma_f16"]))
    assert e_bf < 3e-3, e_bf
    assert e_or < 2e-2, e_or
    for k in (20, -20):
        y = cuda.empty_like(x32)
        A.mul_reduced(y, x32 * 2.0 ** k, "mfma_f16")
        cuda.cuda.synchronize()
        # identical FP16 operands; only the order of the FP32 atomic adds of
        # shared dofs may differ
        assert rel_l2(y.cpu().numpy().astype(np.float64) * 2.0 ** -k, ys["mfma_f16"]) < 1e-6, k
    A.free()


def test_gemm_f16_refused_elsewhere(pkg, cuda):
    """Built for n = 8 only: other orders (and q outside 8..10) are refused
    with SF_ERR_UNSUPPORTED, not run."""
    for p, q in ((4, 6), (7, 11)):
        say("gemm f16 refusal p=%d q=%d" % (p, q))
        _, space, _, _, A = make(pkg, 2, p, q)
        x32 = cuda.zeros(space.n_free, dtype=cuda.float32, device="cuda")
        y = cuda.empty_like(x32)
        with pytest.raises(pkg.SFError) as ei:
            A.mul_reduced(y, x32, "mfma_f16")
        assert ei.value.code == pkg._lib.SF_ERR_UNSUPPORTED
        A.free()


@pytest.mark.parametrize("Pl", ["gll", None])
def test_gemm_f16_inner_solver(pkg, oracle_mod, cuda, Pl):
    """The FP16 GEMM apply as the inner operator of the reliable-update
    mixed CG and of the refinement (Q7, q = 9, kappa): both reach 1e-10 in
    the FP64 residual the oracle recomputes, and agree with FP64 CG."""
    say("gemm f16 inner solver Pl=%s" % Pl)
    model, space, mp, kappa, A = make(pkg, 3, 7, 9, True)
    b = cuda.from_numpy(pkg.normal(2, space.n_free)).cuda()
    bh = b.cpu().numpy()
    x64 = cuda.zeros_like(b)
    s64 = pkg.cg_(x64, A, b, rtol=1e-10, maxiter=400000, Pl=Pl)
    runs = {}
    xm = cuda.zeros_like(b)
    runs["mixed"] = (pkg.cg_mixed_(xm, A, b, rtol=1e-10, inner="mfma_f16", delta=0.1,
                                   maxiter=400000, recurrence="f64", Pl=Pl), xm)
    xi = cuda.zeros_like(b)
    runs["ir"] = (pkg.cg_ir_(xi, A, b, rtol=1e-10, inner="mfma_f16", inner_rtol=1e-2,
                             inner_maxiter=400000, max_outer=80, Pl=Pl), xi)
    say("  f64", s64)
    for key, (st, xx) in runs.items():
        say("  ", key, st)
        xh = xx.cpu().numpy()
        assert st["converged"] == 1, (key, st)
        r = bh - oracle_mod.apply(3, 8, 9, space.cell_dofs, model.cell_vertices(), xh,
                                  mp.tables, kappa=kappa)
        assert np.linalg.norm(r) / np.linalg.norm(bh) < 2e-10, key
        assert rel_l2(xh, x64.cpu().numpy()) < 1e-6, key
    A.free()
