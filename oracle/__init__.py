"""TEST INFRASTRUCTURE ONLY -- the CPU oracle of the matrix-free Laplacian.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker (or the timed CPU baseline); the
product library never calls it.

Two independent CPU paths:

* ``apply`` / ``geometry`` / ``kappa_grad`` -- ctypes wrappers of
  ``sumfac_oracle.c``, a step-by-step restatement of the reference's
  per-cell pipeline (src/SumFactorizationKernels.jl:40-216,
  src/SumFactorizationMaps.jl:78-89, src/LazyMatrices.jl:37-61).
* ``brute`` -- numpy code with no sum factorisation at all: 1D tables from
  a Vandermonde (monomial) change of basis, dense element gradient matrices,
  element stiffness matrices, scipy CSR assembly, and the Kronecker-sum
  closed form of the Cartesian operator.

Parity pins (tests/test_oracle_*.py): mpmath golden tables, textbook Q1
element matrices, the Kronecker closed form, brute-force element assembly,
and mesh invariants (A 1 = 0, patch test, symmetry, unit volume).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def build():
    """Compile liboracle.so with gcc (see oracle/Makefile)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_apply.argtypes = [ctypes.c_int] * 3 + [ctypes.c_int64] * 2 + [
            _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.oracle_apply.restype = ctypes.c_int
        L.oracle_geometry.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                      _dp, _dp, _dp, _dp]
        L.oracle_geometry.restype = ctypes.c_int
        L.oracle_kappa_grad.argtypes = [ctypes.c_int] * 3 + [ctypes.c_int64] + [
            _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.oracle_kappa_grad.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def apply(dim, n, q, cells, verts, x, tables, kappa=None, nthreads=None):
    """y = A(kappa) x by the reference's per-cell algorithm (C, OpenMP).

    tables = (B, Dm, points, weights) with B, Dm of shape (q, n)."""
    B, Dm, pts, w = (_c(t, np.float64) for t in tables)
    cells = _c(cells, np.int32)
    verts = _c(verts, np.float64)
    x = _c(x, np.float64)
    kap = _c(kappa, np.float64) if kappa is not None else None
    ncell = cells.shape[0]
    nfree = x.shape[0]
    y = np.empty(nfree)
    nt = nthreads or os.cpu_count() or 1
    rc = lib().oracle_apply(dim, n, q, ncell, nfree, _d(B), _d(Dm), _d(pts), _d(w),
                            cells.ctypes.data_as(_ip), _d(verts), _d(kap), _d(x), _d(y), nt)
    if rc:
        raise RuntimeError("oracle_apply: singular Jacobian (rc=%d)" % rc)
    return y


def geometry(dim, q, verts, tables):
    """G[cell][entry][qp] from the reference's wq * djq * jq^T jq product."""
    _, _, pts, w = (_c(t, np.float64) for t in tables)
    verts = _c(verts, np.float64)
    ncell = verts.shape[0]
    ng = dim * (dim + 1) // 2
    G = np.empty((ncell, ng, q ** dim))
    rc = lib().oracle_geometry(dim, q, ncell, _d(pts), _d(w), _d(verts), _d(G))
    if rc:
        raise RuntimeError("oracle_geometry: singular Jacobian")
    return G


def kappa_grad(dim, n, q, cells, verts, u, lam, tables, nthreads=None):
    """dJ/dkappa[cell][qp] = grad(lam) . (wq djq jq^T jq) grad(u)."""
    B, Dm, pts, w = (_c(t, np.float64) for t in tables)
    cells = _c(cells, np.int32)
    verts = _c(verts, np.float64)
    u = _c(u, np.float64)
    lam = _c(lam, np.float64)
    ncell = cells.shape[0]
    out = np.empty((ncell, q ** dim))
    nt = nthreads or os.cpu_count() or 1
    rc = lib().oracle_kappa_grad(dim, n, q, ncell, _d(B), _d(Dm), _d(pts), _d(w),
                                 cells.ctypes.data_as(_ip), _d(verts), _d(u), _d(lam),
                                 _d(out), nt)
    if rc:
        raise RuntimeError("oracle_kappa_grad: singular Jacobian")
    return out
