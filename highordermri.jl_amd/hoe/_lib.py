"""ctypes binding of libhoe (include/hoe.h).

Loads the in-tree ``libhoe.so`` built next to this package; there is no
fallback implementation: a missing library raises ``LibraryNotBuilt``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent          # highordermri.jl_amd/
LIB_PATH = PKG_DIR / "libhoe.so"

HOE_OK, HOE_EINVAL, HOE_EDIM, HOE_EASSERT, HOE_ENOMEM, HOE_EHIP, HOE_ECLOSED, HOE_ENODEV = range(8)
BACKEND_DIRECT, BACKEND_MIXED_FP16, BACKEND_MIXED_BF16, BACKEND_FP64 = 0, 1, 2, 3
HOE_DEVICE_PTRS, HOE_ASYNC = 1, 2
HOE_CURRENT_WEIGHTS = -1

BACKENDS = {
    "direct": BACKEND_DIRECT,
    "mixed": BACKEND_MIXED_FP16,
    "mixed_fp16": BACKEND_MIXED_FP16,
    "mixed_bf16": BACKEND_MIXED_BF16,
    "fp64": BACKEND_FP64,
}

# Every symbol include/hoe.h declares.
EXPORTS = (
    "hoe_version", "hoe_last_error", "hoe_device_count", "hoe_create", "hoe_get_info",
    "hoe_set_backend", "hoe_forward", "hoe_adjoint", "hoe_normal", "hoe_set_weights",
    "hoe_set_coefficients", "hoe_set_kspha", "hoe_set_kspha_dt", "hoe_grad", "hoe_cg", "hoe_close", "hoe_destroy",
    "hoe_sync", "hoe_last_kernel_ms", "hoe_apply_girf", "hoe_apply_girf_f64", "hoe_girf_fft_length",
    "hoe_normal_noalloc", "hoe_create_f64", "hoe_set_weights_f64",
)


class LibraryNotBuilt(RuntimeError):
    """libhoe.so is absent: run __graft_entry__.build() (make -C highordermri.jl_amd)."""


class ArgumentError(ValueError):
    """Counterpart of Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Counterpart of Julia's DimensionMismatch."""


class AssertionFailure(AssertionError):
    """Counterpart of Julia's AssertionError."""


class HIPError(RuntimeError):
    """A HIP runtime failure inside libhoe."""


class NoDeviceError(RuntimeError):
    """No HIP device is visible."""


_ERRORS = {
    HOE_EINVAL: ArgumentError,
    HOE_EDIM: DimensionMismatch,
    HOE_EASSERT: AssertionFailure,
    HOE_ENOMEM: MemoryError,
    HOE_EHIP: HIPError,
    HOE_ECLOSED: ArgumentError,
    HOE_ENODEV: NoDeviceError,
}


class ModelDesc(C.Structure):
    _fields_ = [
        ("n_x", C.c_int32), ("n_y", C.c_int32), ("n_z", C.c_int32),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
        ("mask", C.c_void_p),
        ("n_term", C.c_int32), ("n_sam", C.c_int32),
        ("kspha", C.c_void_p), ("times", C.c_void_p), ("fieldmap", C.c_void_p),
        ("n_cha", C.c_int32),
        ("csm", C.c_void_p), ("concomitant", C.c_void_p),
        ("recon_terms", C.c_char_p), ("k_nominal", C.c_void_p),
        ("dwell", C.c_double),
        ("backend", C.c_int32), ("device", C.c_int32),
    ]


class ModelDescF64(C.Structure):
    """hoe_model_desc_f64: the FP64 backend's model (double data, no dwell or backend)."""
    _fields_ = [
        ("n_x", C.c_int32), ("n_y", C.c_int32), ("n_z", C.c_int32),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
        ("mask", C.c_void_p),
        ("n_term", C.c_int32), ("n_sam", C.c_int32),
        ("kspha", C.c_void_p), ("times", C.c_void_p), ("fieldmap", C.c_void_p),
        ("n_cha", C.c_int32),
        ("csm", C.c_void_p), ("concomitant", C.c_void_p),
        ("recon_terms", C.c_char_p), ("k_nominal", C.c_void_p),
        ("device", C.c_int32),
    ]


class Info(C.Structure):
    _fields_ = [
        ("n_grid", C.c_int64), ("n_vox", C.c_int64), ("n_sam", C.c_int64), ("n_cha", C.c_int64),
        ("n_term", C.c_int32), ("backend", C.c_int32), ("device", C.c_int32), ("closed", C.c_int32),
        ("horner", C.c_int32), ("n_runs", C.c_int32),
    ]


_lib = None


def load(path: str | os.PathLike | None = None):
    """Load libhoe once and declare the prototypes."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise LibraryNotBuilt(f"{p} not found; build it with `make -C {PKG_DIR}`")
    lib = C.CDLL(str(p))
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    proto = {
        "hoe_version": (C.c_char_p, []),
        "hoe_last_error": (C.c_char_p, [vp]),
        "hoe_device_count": (C.c_int, []),
        "hoe_create": (C.c_int, [C.POINTER(ModelDesc), C.POINTER(vp)]),
        "hoe_create_f64": (C.c_int, [C.POINTER(ModelDescF64), C.POINTER(vp)]),
        "hoe_set_weights_f64": (C.c_int, [vp, vp, i64]),
        "hoe_get_info": (C.c_int, [vp, C.POINTER(Info)]),
        "hoe_set_backend": (C.c_int, [vp, i32]),
        "hoe_forward": (C.c_int, [vp, vp, vp, C.c_int]),
        "hoe_adjoint": (C.c_int, [vp, vp, vp, C.c_int]),
        "hoe_normal": (C.c_int, [vp, vp, i64, vp, vp, C.c_int]),
        "hoe_set_weights": (C.c_int, [vp, vp, i64]),
        "hoe_normal_noalloc": (C.c_int, [vp, vp, vp, C.c_int]),
        "hoe_set_coefficients": (C.c_int, [vp, vp, dbl, C.c_int]),
        "hoe_set_kspha": (C.c_int, [vp, vp, vp]),
        "hoe_apply_girf": (C.c_int, [vp, i64, i64, vp, i32, i32, dbl, i32, vp, C.c_int]),
        "hoe_apply_girf_f64": (C.c_int, [vp, i64, i64, vp, i32, i32, dbl, i32, vp, C.c_int]),
        "hoe_girf_fft_length": (i64, [i64]),
        "hoe_set_kspha_dt": (C.c_int, [vp, vp]),
        "hoe_grad": (C.c_int, [vp, vp, vp, vp, C.POINTER(dbl), C.POINTER(dbl), C.c_int]),
        "hoe_cg": (C.c_int, [vp, vp, vp, i64, i32, dbl, dbl, vp, vp, C.c_int]),
        "hoe_close": (C.c_int, [vp]),
        "hoe_destroy": (None, [vp]),
        "hoe_sync": (C.c_int, [vp]),
        "hoe_last_kernel_ms": (dbl, [vp, C.c_int]),
    }
    for name, (res, args) in proto.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int, handle=None) -> None:
    """Raise the exception mapped to a nonzero hoe_status."""
    if rc == HOE_OK:
        return
    lib = load()
    msg = lib.hoe_last_error(handle)
    msg = msg.decode() if msg else ""
    raise _ERRORS.get(rc, RuntimeError)(msg or f"libhoe status {rc}")
