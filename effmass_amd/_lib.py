"""ctypes binding of libeffm_hip.so (include/effm_hpsi.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libeffm_hip.so")

EFFM_OK = 0
EFFM_ERR_RUNTIME = -1
EFFM_ERR_OUT_OF_RANGE = -2
EFFM_ERR_DOMAIN = -3
EFFM_ERR_INVALID = -4
EFFM_ERR_NOMEM = -5
EFFM_ERR_HIP = -6

EFFM_PREC_FP64 = 0
EFFM_PREC_MIXED = 1

EFFM_F_OVERWRITE = 0x1
EFFM_F_HOST = 0x2
EFFM_F_GRAPH = 0x4

# every symbol include/effm_hpsi.h declares
EXPORTS = (
    "effm_ctx_create", "effm_ctx_destroy", "effm_set_stream", "effm_last_error",
    "effm_mem_info", "effm_cell_upload", "effm_kpt_upload", "effm_kpt_dim",
    "effm_kpt_release", "effm_hpsi", "effm_dhdk", "effm_hpsi_multi", "effm_dhdk_multi",
    "effm_hpsi_dhdk_multi", "effm_lincomb", "effm_rayleigh",
    "effm_overlap", "effm_overlap_noalloc", "effm_kp_mass", "effm_autotune",
    "effm_get_shapes", "effm_set_shapes", "effm_set_timing", "effm_get_timing", "effm_reset_timing",
)


class Shapes(C.Structure):
    """effm_shapes: the autotuner's launch-shape knobs of one precision."""
    _fields_ = [("pencil_bands", C.c_int), ("pencil_z", C.c_int), ("gemm_wgs", C.c_int), ("apply_tile", C.c_int)]


class Species(C.Structure):
    _fields_ = [
        ("lmax", C.c_int32), ("nbeta", C.c_int32),
        ("beta_l", C.POINTER(C.c_int32)), ("lam", C.POINTER(C.c_double)),
        ("nq", C.c_int32), ("dq", C.c_double), ("qmax", C.c_double),
        ("beta_tab", C.POINTER(C.c_double)), ("dbeta_tab", C.POINTER(C.c_double)),
    ]


class Cell(C.Structure):
    _fields_ = [
        ("n1", C.c_int32), ("n2", C.c_int32), ("n3", C.c_int32),
        ("lattice", C.c_double * 9), ("vloc", C.POINTER(C.c_double)),
        ("natom", C.c_int32), ("species", C.POINTER(C.c_int32)), ("tau", C.POINTER(C.c_double)),
        ("nspecies", C.c_int32), ("sp", C.POINTER(Species)),
    ]


class EffmError(RuntimeError):
    """Base class; subclasses map the status codes onto the reference's
    exception classes."""


class EffmRuntimeError(EffmError):
    pass


class EffmOutOfRange(EffmError, IndexError):
    pass


class EffmDomainError(EffmError, ValueError):
    pass


class EffmInvalidArgument(EffmError, ValueError):
    pass


class EffmOutOfMemory(EffmError, MemoryError):
    pass


_CODE_EXC = {
    EFFM_ERR_RUNTIME: EffmRuntimeError,
    EFFM_ERR_OUT_OF_RANGE: EffmOutOfRange,
    EFFM_ERR_DOMAIN: EffmDomainError,
    EFFM_ERR_INVALID: EffmInvalidArgument,
    EFFM_ERR_NOMEM: EffmOutOfMemory,
    EFFM_ERR_HIP: EffmError,
}

_lib = None


def load() -> C.CDLL:
    """Load libeffm_hip.so; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the H|psi> engine has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, u64, dp = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.POINTER(C.c_double)
    lib.effm_ctx_create.argtypes = [C.POINTER(vp), i32, u64]
    lib.effm_ctx_destroy.argtypes = [vp]
    lib.effm_set_stream.argtypes = [vp, vp]
    lib.effm_last_error.argtypes = [vp]
    lib.effm_last_error.restype = C.c_char_p
    lib.effm_mem_info.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.effm_cell_upload.argtypes = [vp, C.POINTER(Cell)]
    lib.effm_kpt_upload.argtypes = [vp, i32, i32, vp, vp, vp, vp]
    lib.effm_kpt_dim.argtypes = [vp, i32]
    lib.effm_kpt_release.argtypes = [vp, i32]
    lib.effm_hpsi.argtypes = [vp, i32, i32, vp, i64, vp, i64, i32, i32]
    lib.effm_dhdk.argtypes = [vp, i32, i32, vp, i64, vp, vp, vp, i64, i32, i32]
    lib.effm_hpsi_multi.argtypes = [vp, i32, vp, i32, vp, vp, vp, i32, i32]
    lib.effm_dhdk_multi.argtypes = [vp, i32, vp, i32, vp, vp, vp, vp, vp, i32, i32]
    lib.effm_hpsi_dhdk_multi.argtypes = [vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, i32, i32]
    lib.effm_lincomb.argtypes = [vp, i64, C.c_double, vp, C.c_double, vp, C.c_double, vp, i32, i32]
    lib.effm_rayleigh.argtypes = [vp, i32, i32, vp, i64, vp, i64, i32, vp, i32]
    lib.effm_autotune.argtypes = [vp, i32, i32, i32, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.effm_get_shapes.argtypes = [vp, i32, C.POINTER(Shapes)]
    lib.effm_set_shapes.argtypes = [vp, i32, C.POINTER(Shapes)]
    lib.effm_overlap.argtypes = [vp, i32, i32, i32, vp, i64, vp, i64, vp, i64, i32, i32]
    lib.effm_overlap_noalloc.argtypes = [vp, i32, i32, i32, vp, i64, vp, i64, vp, i64, i32, i32]
    lib.effm_kp_mass.argtypes = [vp, i32, vp, vp, i64, i32, vp, vp, C.c_double, vp, i32]
    lib.effm_set_timing.argtypes = [vp, i32]
    lib.effm_get_timing.argtypes = [vp, C.c_char_p, dp, C.POINTER(C.c_int64)]
    lib.effm_reset_timing.argtypes = [vp]
    for name in EXPORTS:
        getattr(lib, name).restype = getattr(lib, name).restype or C.c_int
    lib.effm_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def check(rc: int, ctx=None, what: str = "") -> int:
    if rc >= 0:
        return rc
    msg = ""
    if ctx is not None:
        raw = load().effm_last_error(ctx)
        msg = raw.decode() if raw else ""
    exc = _CODE_EXC.get(rc, EffmError)
    raise exc(f"{what}: {msg} (status {rc})" if what else f"{msg} (status {rc})")
