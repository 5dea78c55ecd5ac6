"""Differentiation of y = A(kappa) u with respect to kappa (at quadrature
points) and u.

``rrule(A, kappa, u)`` follows ChainRulesCore's rrule convention (the
Julia-side binding is an rrule for ``apply(A, kappa, u)``): it returns
``y = A(kappa) u`` and a pullback ``ybar -> (kappabar, ubar)`` with

    kappabar[cell][qp] = grad(ybar)_qp . G_qp grad(u)_qp,
    ubar               = A(kappa)^T ybar = A(kappa) ybar,

both from one fused device kernel (sf_adjoint) that recomputes the
quadrature-point gradients instead of storing them.  The pullback holds the
kappa of its forward call: if the plan's kappa has changed since (another
apply, a taped pass over several kappas), it is set back before the adjoint
runs, so ubar is always A(kappa of the forward) ybar.
"""
import ctypes

from . import _lib
from .operator import _torch_stream


def apply(A, kappa, u):
    """y = A(kappa) u (sets kappa on the plan, FP64, torch CUDA tensors)."""
    import torch
    A.set_kappa(kappa)
    y = torch.empty_like(u)
    return A.mul(y, u)


def pullback(A, u, ybar, want_kappa=True, want_u=True):
    """(kappabar, ubar) of y = A(kappa) u for cotangent ybar (plan's kappa)."""
    import torch
    nq = A.m.nquad ** A.D
    kb = torch.empty(A.n_cells * nq, dtype=torch.float64, device=u.device) if want_kappa else None
    ub = torch.empty_like(u) if want_u else None
    _lib.check(A._L.sf_adjoint(A.plan, ctypes.c_void_p(u.data_ptr()),
                               ctypes.c_void_p(ybar.data_ptr()),
                               ctypes.c_void_p(kb.data_ptr()) if kb is not None else None,
                               ctypes.c_void_p(ub.data_ptr()) if ub is not None else None,
                               _torch_stream(u)))
    return kb, ub


def rrule(A, kappa, u):
    """(y, pullback) for y = A(kappa) u."""
    import numpy as np
    k = np.array(kappa, dtype=np.float64, copy=True).reshape(-1)
    y = apply(A, k, u)
    version = A.kappa_version

    def _pullback(ybar):
        if A.kappa_version != version:
            A.set_kappa(k)
        return pullback(A, u, ybar)

    return y, _pullback


def loss_and_grad(A, kappa, u, d):
    """J = 1/2 ||A(kappa) u - d||^2 and (dJ/dkappa, dJ/du)."""
    y, pb = rrule(A, kappa, u)
    r = y - d
    J = 0.5 * float((r * r).sum())
    kb, ub = pb(r)
    return J, kb, ub


