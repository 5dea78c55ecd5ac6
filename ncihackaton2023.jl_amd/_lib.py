"""ctypes binding of libsfhip.so (the C ABI declared in include/sfhip.h).

The library is the only compute path: if it is missing, loading fails with
an error -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsfhip.so")

# Load torch first when it is available so that libsfhip.so binds to the
# HIP runtime torch already mapped (same SONAME) instead of a second copy:
# device pointers and streams of torch tensors are then valid in the library.
try:  # pragma: no cover - depends on the environment
    import torch as _torch  # noqa: F401
except Exception:  # pragma: no cover
    _torch = None

SF_OK = 0
SF_ERR_ARG = 1
SF_ERR_UNSUPPORTED = 2
SF_ERR_DOF_RANGE = 3
SF_ERR_JACOBIAN = 4
SF_ERR_HIP = 5
SF_ERR_NOT_READY = 6
SF_ERR_NOT_CONVERGED = 7
SF_ERR_NONFINITE = 8

SF_PREC_F64 = 0
SF_PREC_F32 = 1
SF_PREC_F32_GBF16 = 2
SF_PREC_MFMA_BF16 = 3
SF_PREC_MFMA_BF16X3 = 4
SF_PREC_MFMA_BF16X2 = 5
SF_PREC_MFMA_F32 = 6
SF_PREC_MFMA_F16 = 7

SF_PC_NONE = 0
SF_PC_JACOBI = 1
SF_PC_GLL = 2
PRECONDITIONERS = {None: SF_PC_NONE, "none": SF_PC_NONE, "jacobi": SF_PC_JACOBI,
                   "gll": SF_PC_GLL}

PRECISIONS = {"f64": SF_PREC_F64, "f32": SF_PREC_F32, "f32_gbf16": SF_PREC_F32_GBF16,
              "mfma_bf16": SF_PREC_MFMA_BF16, "mfma_bf16x2": SF_PREC_MFMA_BF16X2,
              "mfma_bf16x3": SF_PREC_MFMA_BF16X3, "mfma_f32": SF_PREC_MFMA_F32,
              "mfma_f16": SF_PREC_MFMA_F16}


class SFError(RuntimeError):
    """A non-zero status from libsfhip; ``code`` is the sf_status value."""

    def __init__(self, code, msg):
        super().__init__("sfhip error %d: %s" % (code, msg))
        self.code = code


class PlanInfo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("order", ctypes.c_int), ("n1d", ctypes.c_int),
                ("q1d", ctypes.c_int), ("n_cells", ctypes.c_int64), ("n_free", ctypes.c_int64),
                ("n_qp", ctypes.c_int64), ("has_cells", ctypes.c_int),
                ("has_geometry", ctypes.c_int), ("has_kappa", ctypes.c_int),
                ("threads_per_block", ctypes.c_int), ("cells_per_block", ctypes.c_int),
                ("device_bytes", ctypes.c_int64), ("n_blocks", ctypes.c_int),
                ("launches_per_apply", ctypes.c_int), ("cells_reordered", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("outer_iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every entry point in include/sfhip.h
_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double
SIGNATURES = [
    ("sf_last_error", ctypes.c_char_p, []),
    ("sf_version", ctypes.c_char_p, []),
    ("sf_tables_1d", _i, [_i, _i, _vp, _vp, _vp, _vp]),
    ("sf_gauss_points_for_degree", _i, [_i]),
    ("sf_flops_model", _i64, [_i64, _i, _i, _i, _i, _i]),
    ("sf_plan_create", _i, [ctypes.POINTER(_vp), _i, _i, _i, _i64, _i64, _i]),
    ("sf_plan_destroy", _i, [_vp]),
    ("sf_plan_get_info", _i, [_vp, ctypes.POINTER(PlanInfo)]),
    ("sf_plan_set_timing", _i, [_vp, _i]),
    ("sf_plan_read_timing", _i, [_vp, _vp, _i, ctypes.POINTER(_i)]),
    ("sf_plan_set_tables", _i, [_vp, _vp, _vp, _vp]),
    ("sf_plan_set_cells", _i, [_vp, _vp, _i, _vp]),
    ("sf_plan_set_vertices", _i, [_vp, _vp]),
    ("sf_plan_set_geometry", _i, [_vp, _vp]),
    ("sf_plan_get_geometry", _i, [_vp, _vp]),
    ("sf_plan_set_kappa", _i, [_vp, _vp]),
    ("sf_plan_prepare", _i, [_vp, _i]),
    ("sf_apply", _i, [_vp, _vp, _vp, _vp]),
    ("sf_apply_add", _i, [_vp, _vp, _vp, _vp]),
    ("sf_apply_host", _i, [_vp, _vp, _vp]),
    ("sf_apply_f32", _i, [_vp, _vp, _vp, _i, _vp]),
    ("sf_adjoint", _i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("sf_diagonal", _i, [_vp, _vp, _vp]),
    ("sf_plan_set_preconditioner", _i, [_vp, _i]),
    ("sf_precond_apply", _i, [_vp, _vp, _vp, _vp]),
    ("sf_cg", _i, [_vp, _vp, _vp, _d, _i, ctypes.POINTER(SolveStats), _vp]),
    ("sf_cg_noalloc", _i, [_vp, _vp, _vp, _d, _i, ctypes.POINTER(SolveStats), _vp]),
    ("sf_cg_ir", _i, [_vp, _vp, _vp, _d, _i, _i, _d, _i, ctypes.POINTER(SolveStats), _vp]),
    ("sf_cg_mixed", _i, [_vp, _vp, _vp, _d, _i, _i, _d, _i, ctypes.POINTER(SolveStats), _vp]),
]

_lib = None


def load():
    """Load libsfhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libsfhip.so is not built (%s); run __graft_entry__.build() or "
            "make -C ncihackaton2023.jl_amd/csrc" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    return load().sf_last_error().decode()


def check(rc):
    if rc != SF_OK:
        raise SFError(rc, last_error())
    return rc
