"""Device-resident solvers around the operator.

``cg_(x, A, b)`` mirrors ``IterativeSolvers.cg!(x, A, b)`` as the reference's
driver calls it (compute_solution, profiling/drivers.jl:85-95): unpreconditioned
CG, x is the initial guess and is overwritten.  ``cg_ir_`` is the
mixed-precision iterative refinement: FP64 outer residual, norms and dot
products, reduced-precision (FP32-vector) inner CG, restarted per outer
step.  ``cg_mixed_`` keeps a single reduced-precision CG recurrence and
replaces its residual by the FP64 one at reliable updates.  All vectors are torch
CUDA float64 tensors; the iteration runs in libsfhip.so on torch's current
stream.
"""
import ctypes

from . import _lib
from .operator import _torch_stream


def _set_pc(A, Pl):
    """The Pl keyword of IterativeSolvers.cg! as a plan setting: None /
    'none' (the reference's unpreconditioned call), 'jacobi' (diag(A)) or
    'gll' (3D: Jacobi in the Gauss-Lobatto nodal basis of the space)."""
    if Pl not in _lib.PRECONDITIONERS:
        raise ValueError("Pl must be one of %s" % sorted(k for k in _lib.PRECONDITIONERS if k))
    _lib.check(A._L.sf_plan_set_preconditioner(A.plan, _lib.PRECONDITIONERS[Pl]))


def _vec_ok(A, v):
    if not (type(v).__module__.startswith("torch") and v.is_cuda and str(v.dtype) == "torch.float64"
            and v.is_contiguous() and v.numel() == A.n_free):
        raise TypeError("solver vectors must be contiguous torch CUDA float64 of n_free entries")


def cg_(x, A, b, rtol=1e-10, maxiter=None, raise_on_fail=False, Pl=None):
    """FP64 CG: stop when ||b - A x|| <= rtol ||b||.  Pl selects the
    preconditioner (None: unpreconditioned, as the reference's driver).
    Returns the statistics dict (iterations, rel_residual, converged,
    seconds)."""
    _vec_ok(A, x)
    _vec_ok(A, b)
    _set_pc(A, Pl)
    st = _lib.SolveStats()
    maxiter = A.n_free if maxiter is None else int(maxiter)
    rc = A._L.sf_cg(A.plan, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                    float(rtol), maxiter, ctypes.byref(st), _torch_stream(x))
    if rc != _lib.SF_OK and (raise_on_fail or rc != _lib.SF_ERR_NOT_CONVERGED):
        _lib.check(rc)
    return st.as_dict()


def cg_noalloc_(x, A, b, rtol=1e-10, maxiter=None, raise_on_fail=False):
    """cg_ with the plan's current preconditioner and no device allocation:
    the workspace lives in the library's fixed per-GPU scratch array
    (sf_cg_noalloc; concurrent calls on one GPU race)."""
    _vec_ok(A, x)
    _vec_ok(A, b)
    st = _lib.SolveStats()
    maxiter = A.n_free if maxiter is None else int(maxiter)
    rc = A._L.sf_cg_noalloc(A.plan, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                            float(rtol), maxiter, ctypes.byref(st), _torch_stream(x))
    if rc != _lib.SF_OK and (raise_on_fail or rc != _lib.SF_ERR_NOT_CONVERGED):
        _lib.check(rc)
    return st.as_dict()


def cg_ir_(x, A, b, rtol=1e-10, inner="f32", inner_rtol=1e-3, inner_maxiter=None,
           max_outer=50, raise_on_fail=False, Pl=None):
    """Mixed-precision iterative refinement (FP64 outer, reduced inner,
    preconditioned inner CG when Pl is set)."""
    _vec_ok(A, x)
    _vec_ok(A, b)
    _set_pc(A, Pl)
    st = _lib.SolveStats()
    prec = _lib.PRECISIONS[inner] if isinstance(inner, str) else int(inner)
    inner_maxiter = max(1, A.n_free) if inner_maxiter is None else int(inner_maxiter)
    rc = A._L.sf_cg_ir(A.plan, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                       float(rtol), int(max_outer), prec, float(inner_rtol), inner_maxiter,
                       ctypes.byref(st), _torch_stream(x))
    if rc != _lib.SF_OK and (raise_on_fail or rc != _lib.SF_ERR_NOT_CONVERGED):
        _lib.check(rc)
    return st.as_dict()


def cg_mixed_(x, A, b, rtol=1e-10, inner="f32", delta=0.1, maxiter=None,
              recurrence=None, raise_on_fail=False, Pl=None):
    """Mixed-precision CG with reliable updates: one CG recurrence on the
    reduced-precision operator (recurrence 'f32': FP32 vectors; 'f64': FP64
    vectors, only the operator reduced; 'f64p32': FP64 solution and residual,
    the search direction only in FP32), residual replaced by the FP64
    b - A x whenever it has fallen by `delta` (search direction kept).
    Pl: the preconditioner, as in cg_.  recurrence None: 'f32' with the
    Gauss-Lobatto preconditioner (its well-conditioned system loses no
    iterations to the FP32 recurrence, which moves the fewest bytes: 1.48x
    FP64 CG at 32^3 Q7 against 1.24x for 'f64p32'), else 'f64p32' (on the
    unpreconditioned system the FP32 recurrence needs about 50% more
    iterations)."""
    _vec_ok(A, x)
    _vec_ok(A, b)
    _set_pc(A, Pl)
    if recurrence is None:
        recurrence = "f32" if Pl == "gll" else "f64p32"
    st = _lib.SolveStats()
    prec = _lib.PRECISIONS[inner] if isinstance(inner, str) else int(inner)
    maxiter = 10 * A.n_free if maxiter is None else int(maxiter)
    modes = {"f32": 0, "f64": 1, "f64p32": 2}
    if recurrence not in modes:
        raise ValueError("recurrence must be one of %s" % sorted(modes))
    rc = A._L.sf_cg_mixed(A.plan, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                          float(rtol), maxiter, prec, float(delta), modes[recurrence],
                          ctypes.byref(st), _torch_stream(x))
    if rc != _lib.SF_OK and (raise_on_fail or rc != _lib.SF_ERR_NOT_CONVERGED):
        _lib.check(rc)
    return st.as_dict()


def compute_solution(A, b, **kw):
    """x = A \\ b from a zero initial guess (profiling/drivers.jl:85-89)."""
    import torch
    x = torch.zeros_like(b)
    cg_(x, A, b, **kw)
    return x
