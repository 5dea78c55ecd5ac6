"""The matrix-free operator: the drop-in for the reference's LazyMatrix.

``LazyMatrix(m, U, V, dOmega)`` (src/LazyMatrices.jl:30-50) becomes
``LazyMatrix(m, space)``: it creates a device plan, uploads the tensor-ordered
cell dofs and the cell vertices, and computes the quadrature-point geometric
factor once on the GPU in FP64.  ``mul_(y, A, x)`` is
``LinearAlgebra.mul!(y, A::LazyMatrix, x)`` (src/LazyMatrices.jl:52-61):
it overwrites y with A x on the free dofs and returns y.

Vectors may be numpy arrays (host; staged through device buffers, returns
when y is complete) or torch CUDA tensors (device; launched on torch's
current stream, asynchronous).  Either way the arithmetic runs in the HIP
kernels of libsfhip.so; there is no CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from .fe import SumFactorizationMap


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _torch_stream(a):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(a.device).cuda_stream)


class LazyMatrix:
    """Matrix-free FE Laplacian (optionally with a coefficient kappa at the
    quadrature points) on the free dofs of ``space``."""

    def __init__(self, m, space, kappa=None, geometry=None):
        if not isinstance(m, SumFactorizationMap):
            raise TypeError("m must be a SumFactorizationMap")
        if m.D != space.D or m.order != space.order:
            raise ValueError("descriptor and space disagree on D / order")
        self.m = m
        self.space = space
        self.D = m.D
        self.n_free = int(space.n_free)
        self.n_cells = int(space.n_cells)
        self.shape = (self.n_free, self.n_free)
        self.dtype = np.float64
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.sf_plan_create(ctypes.byref(h), m.D, m.order, m.nquad,
                                    self.n_cells, self.n_free, 0))
        self._plan = h
        self._L = L
        self.kappa_version = 0  # bumped by every set_kappa
        cells = np.ascontiguousarray(space.cell_dofs, dtype=np.int32)
        _lib.check(L.sf_plan_set_cells(h, cells.ctypes.data, 0, None))
        if geometry is None:
            X = np.ascontiguousarray(space.model.cell_vertices(), dtype=np.float64)
            _lib.check(L.sf_plan_set_vertices(h, X.ctypes.data))
        else:
            G = np.ascontiguousarray(geometry, dtype=np.float64)
            _lib.check(L.sf_plan_set_geometry(h, G.ctypes.data))
        if kappa is not None:
            self.set_kappa(kappa)

    # -- plan management --------------------------------------------------
    @property
    def plan(self):
        if self._plan is None:
            raise RuntimeError("LazyMatrix has been freed")
        return self._plan

    def free(self):
        if getattr(self, "_plan", None) is not None:
            self._L.sf_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self):
        inf = _lib.PlanInfo()
        _lib.check(self._L.sf_plan_get_info(self.plan, ctypes.byref(inf)))
        return {k: getattr(inf, k) for k, _ in inf._fields_}

    def set_kappa(self, kappa):
        k = np.ascontiguousarray(kappa, dtype=np.float64).reshape(-1)
        if k.size != self.n_cells * self.m.nquad ** self.D:
            raise ValueError("kappa must have n_cells * q^D entries")
        _lib.check(self._L.sf_plan_set_kappa(self.plan, k.ctypes.data))
        self.kappa_version += 1

    def geometry(self):
        """The FP64 geometric factor G[cell][entry][qp] held on the device."""
        ng = self.D * (self.D + 1) // 2
        G = np.empty((self.n_cells, ng, self.m.nquad ** self.D))
        _lib.check(self._L.sf_plan_get_geometry(self.plan, G.ctypes.data))
        return G

    def diagonal(self, out=None):
        """diag(A) on the free dofs (FP64 torch CUDA tensor), matrix-free."""
        import torch
        if out is None:
            out = torch.empty(self.n_free, dtype=torch.float64, device="cuda")
        self._check_dev(out, out, "float64")
        _lib.check(self._L.sf_diagonal(self.plan, ctypes.c_void_p(out.data_ptr()),
                                       _torch_stream(out)))
        return out

    def precond(self, z, r, Pl):
        """z = M^{-1} r for the preconditioner Pl (None, 'jacobi', 'gll'),
        FP64 torch CUDA tensors: the ldiv!(z, Pl, r) of a CG iteration."""
        self._check_dev(r, z, "float64")
        if Pl not in _lib.PRECONDITIONERS:
            raise ValueError("unknown preconditioner %r" % (Pl,))
        _lib.check(self._L.sf_plan_set_preconditioner(self.plan, _lib.PRECONDITIONERS[Pl]))
        _lib.check(self._L.sf_precond_apply(self.plan, ctypes.c_void_p(z.data_ptr()),
                                            ctypes.c_void_p(r.data_ptr()), _torch_stream(r)))
        return z

    def timing(self, enable=True):
        """Record the duration of every apply kernel (HIP events around the
        kernel alone) from now on; enable=False stops."""
        _lib.check(self._L.sf_plan_set_timing(self.plan, 1 if enable else 0))

    def read_timing(self, max_count=1 << 16):
        """The recorded apply-kernel durations in ms (and clear them)."""
        buf = (ctypes.c_double * max_count)()
        n = ctypes.c_int(0)
        _lib.check(self._L.sf_plan_read_timing(self.plan, buf, max_count, ctypes.byref(n)))
        return [buf[i] for i in range(min(n.value, max_count))]

    def prepare(self, precision):
        _lib.check(self._L.sf_plan_prepare(self.plan, _lib.PRECISIONS.get(precision, precision)))

    # -- products -----------------------------------------------------------
    def mul(self, y, x):
        """y = A x (FP64, y overwritten); returns y."""
        if _is_torch(x):
            self._check_dev(x, y, "float64")
            _lib.check(self._L.sf_apply(self.plan, ctypes.c_void_p(y.data_ptr()),
                                        ctypes.c_void_p(x.data_ptr()), _torch_stream(x)))
            return y
        x = np.ascontiguousarray(x, dtype=np.float64)
        if y.dtype != np.float64 or not y.flags["C_CONTIGUOUS"] or y.shape != (self.n_free,):
            raise ValueError("y must be a contiguous float64 vector of n_free entries")
        if x.shape != (self.n_free,):
            raise ValueError("x must have n_free entries")
        _lib.check(self._L.sf_apply_host(self.plan, y.ctypes.data, x.ctypes.data))
        return y

    def mul_add(self, y, x):
        """y += A x (FP64 torch CUDA tensors; one kernel launch, no fill)."""
        self._check_dev(x, y, "float64")
        _lib.check(self._L.sf_apply_add(self.plan, ctypes.c_void_p(y.data_ptr()),
                                        ctypes.c_void_p(x.data_ptr()), _torch_stream(x)))
        return y

    def mul_reduced(self, y, x, precision="f32"):
        """y = A x with FP32 device vectors in a reduced-precision operator
        ('f32', 'f32_gbf16', 'mfma_bf16' or 'mfma_bf16x3'; the MFMA forms are 3D only)."""
        self._check_dev(x, y, "float32")
        prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        _lib.check(self._L.sf_apply_f32(self.plan, ctypes.c_void_p(y.data_ptr()),
                                        ctypes.c_void_p(x.data_ptr()), prec, _torch_stream(x)))
        return y

    def __matmul__(self, x):
        if _is_torch(x):
            import torch
            y = torch.empty_like(x)
        else:
            y = np.empty(self.n_free)
        return self.mul(y, x)

    def _check_dev(self, x, y, dt):
        for a in (x, y):
            if not _is_torch(a) or not a.is_cuda:
                raise TypeError("device products need torch CUDA tensors")
            if str(a.dtype) != "torch." + dt:
                raise TypeError("expected torch.%s vectors" % dt)
            if a.numel() != self.n_free or not a.is_contiguous():
                raise ValueError("vectors must be contiguous with n_free entries")


def mul_(y, A, x):
    """LinearAlgebra.mul!(y, A, x): y = A x, overwritten, returned."""
    return A.mul(y, x)
