"""ctypes binding of libhpcg_amd.so (declarations in include/hpcg_amd.h).

The library must have been built (``python -m hpcg_amd.build``); importing
this module raises if it is missing -- there is no CPU fallback.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libhpcg_amd.so")

ORDER_LEXICOGRAPHIC = 0
ORDER_MULTICOLOR = 1
KEEP_SCHEDULE = 1
CG_TIMERS = 1
CG_GRAPH = 2

ERRORS = {1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "EUNSUP", 5: "ECHECK"}
ECHECK = 5


class HpcgxError(RuntimeError):
    def __init__(self, code, func, msg):
        super().__init__("%s failed: HPCGX_%s: %s" % (func, ERRORS.get(code, code), msg))
        self.code = code


class HostLevel(ctypes.Structure):
    """hpcgx_host_level: one level in the reference's storage."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("nrow", ctypes.c_int64),
                ("nonzerosInRow", ctypes.c_void_p),
                ("mtxIndL", ctypes.c_void_p),
                ("matrixValues", ctypes.c_void_p),
                ("matrixDiagonal", ctypes.c_void_p),
                ("f2cOperator", ctypes.c_void_p)]


class LevelInfo(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("nrow", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("ncolors", ctypes.c_int), ("nlevels_sched", ctypes.c_int),
                ("ncoarse", ctypes.c_int64), ("wave", ctypes.c_int), ("nslabs", ctypes.c_int),
                ("col16", ctypes.c_int), ("wide_tiles", ctypes.c_int64),
                ("byte_tiles", ctypes.c_int64)]


# name -> (restype, argtypes)
P, I, I64, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
PP = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "hpcgx_last_error": (ctypes.c_char_p, []),
    "hpcgx_abi_version": (I, []),
    "hpcgx_optimize_problem": (I, [ctypes.POINTER(HostLevel), I, I, PP]),
    "hpcgx_generate_problem": (I, [I, I, I, I, I, PP]),
    "hpcgx_check_problem": (I, [P, P, P, P, P]),
    "hpcgx_problem_destroy": (I, [P]),
    "hpcgx_problem_nlevels": (I, [P]),
    "hpcgx_problem_info": (I, [P, I, ctypes.POINTER(LevelInfo)]),
    "hpcgx_optimize_problem_memory_use": (D, [P]),
    "hpcgx_problem_vectors": (I, [P, P, P, P, P]),
    "hpcgx_get_schedule": (I, [P, I, I, P]),
    "hpcgx_to_device_order": (I, [P, I, P, P, P]),
    "hpcgx_to_natural_order": (I, [P, I, P, P, P]),
    "hpcgx_replace_diagonal": (I, [P, I, P, P]),
    "hpcgx_copy_diagonal": (I, [P, I, P, P]),
    "hpcgx_spmv": (I, [P, I, P, P, P]),
    "hpcgx_symgs": (I, [P, I, I, P, P, P]),
    "hpcgx_mg": (I, [P, I, I, P, P, P]),
    "hpcgx_restriction": (I, [P, I, P, P, P, P]),
    "hpcgx_residual_restriction": (I, [P, I, P, P, P, P]),
    "hpcgx_prolongation": (I, [P, I, P, P, P]),
    "hpcgx_dot": (I, [P, I64, P, P, P, P]),
    "hpcgx_dot_host": (I, [P, I64, P, P, P, P]),
    "hpcgx_dot_noalloc": (I, [I64, P, P, P, P]),
    "hpcgx_waxpby": (I, [I64, D, P, D, P, P, P]),
    "hpcgx_cg": (I, [P, I, P, P, I, D, I, I, P, P, P, P, P]),
    "hpcgx_prepare_mixed": (I, [P]),
    "hpcgx_spmv32": (I, [P, I, P, P, P]),
    "hpcgx_symgs32": (I, [P, I, I, P, P, P]),
    "hpcgx_mg32": (I, [P, I, I, P, P, P]),
    "hpcgx_cg_mixed": (I, [P, I, P, P, I, D, I, P, P, P, P, P]),
    "hpcgx_spmv_block": (I, [P, I, I, P, P, P]),
    "hpcgx_autotune": (I, [P, I, I, ctypes.c_char_p, I]),
    "hpcgx_mg32_block": (I, [P, I, I, I, P, P, P]),
    "hpcgx_cg_block_mixed": (I, [P, I, I, P, P, I, D, I, P, P, P, P, P]),
    "hpcgx_generate_fleet": (I, [I, P, P, P, I, I, PP]),
    "hpcgx_fleet_info": (I, [P, I, P, P]),
    "hpcgx_cg_fleet": (I, [P, I, P, P, I, D, I, I, P, P, P, P, P]),
    "hpcgx_tune": (I, [P, ctypes.c_char_p, D]),
    "hpcgx_check_status": (I, [P, P]),
    "hpcgx_kernel_timing": (I, [P, I]),
    "hpcgx_kernel_timing_read": (I, [P, P, P]),
}

_lib = None


def exported_symbols():
    return list(SIGNATURES)


def lib():
    """Load libhpcg_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        try:
            # load torch's HIP runtime first so the process holds one runtime
            # (it has the same SONAME as /opt/rocm's and is then shared)
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhpcg_amd.so is not built: run `python -m hpcg_amd.build` "
                               "(the HIP path is the only implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    """Call an int-returning entry point and raise HpcgxError on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise HpcgxError(rc, name, lib().hpcgx_last_error().decode())
    return rc
