"""ctypes binding of libscir_hip.so (the C ABI declared in include/scir_hip.h).

The library is loaded from this package's directory and nowhere else.  When
it is missing, :func:`lib` raises :class:`LibraryMissing`: there is no
fallback implementation behind this module.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscir_hip.so")

# scir_hip_status
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_NO_DEVICE = 2
ERR_BACKEND = 3
ERR_OUT_OF_MEMORY = 4
ERR_UNSUPPORTED = 5
ERR_CAPTURE = 6

# scir_hip_precision
PREC_F32 = 0
PREC_F32_EXACT = 1
PREC_BF16X3 = 2
PREC_FP16 = 3
PREC_AUTO = 4
PREC_FFT32 = 5

PRECISIONS = {"f32": PREC_F32, "f32_exact": PREC_F32_EXACT, "bf16x3": PREC_BF16X3,
              "fp16": PREC_FP16, "auto": PREC_AUTO, "fft32": PREC_FFT32}

# Every function of include/scir_hip.h: name -> (restype, argtypes).
_c = ctypes
_P = _c.c_void_p
_I = _c.c_int
_I32 = _c.c_int32
_I64 = _c.c_int64
_U64 = _c.c_uint64
_F = _c.c_float
_Pi64 = _c.POINTER(_c.c_int64)
SIGNATURES = {
    "scir_hip_version": (_I, []),
    "scir_hip_init": (_I, []),
    "scir_hip_shutdown": (_I, []),
    "scir_hip_device_count": (_I, [_c.POINTER(_c.c_int)]),
    "scir_hip_last_error": (_c.c_char_p, []),
    "scir_hip_fir_f32_host": (_I, [_P, _I64, _P, _I32, _I64, _I64, _P, _I64, _I32]),
    "scir_hip_fir_f32_dev": (_I, [_P, _I64, _P, _I32, _I64, _I64, _P, _I64, _I32, _P]),
    "scir_hip_fir_f32_cpu": (_I, [_P, _I64, _P, _I32, _I64, _I64, _P, _I64, _I32]),
    "scir_hip_ragged_plan_create": (_I, [_Pi64, _Pi64, _I64, _I32, _c.POINTER(_P)]),
    "scir_hip_ragged_plan_execute": (_I, [_P, _P, _P, _P, _P]),
    "scir_hip_ragged_plan_execute_per_signal": (_I, [_P, _P, _P, _P, _P]),
    "scir_hip_ragged_plan_destroy": (_I, [_P]),
    "scir_hip_fir_ragged_f32_dev": (_I, [_P, _Pi64, _P, _Pi64, _I64, _P, _I32, _P]),
    "scir_hip_fir_vjp_dx_dev": (_I, [_P, _I64, _P, _I32, _I64, _I64, _P, _I64, _I32, _P]),
    "scir_hip_fir_vjp_db_dev": (_I, [_P, _I64, _P, _I64, _I32, _I64, _I64, _P, _P]),
    "scir_hip_fir_vjp_db_noalloc_dev": (_I, [_P, _I64, _P, _I64, _I32, _I64, _I64, _P, _P]),
    "scir_hip_graph_begin": (_I, [_P]),
    "scir_hip_graph_end": (_I, [_P, _c.POINTER(_P)]),
    "scir_hip_graph_launch": (_I, [_P, _P]),
    "scir_hip_graph_destroy": (_I, [_P]),
    "scir_hip_autotune": (_I, [_c.c_char_p, _I32]),
    "scir_hip_autotune_load": (_I, [_c.c_char_p]),
    "scir_hip_tune_set": (_I, [_c.c_char_p, _I32]),
    "scir_hip_synth_normal_f32": (_I, [_P, _I64, _I64, _I64, _U64, _F, _F, _P]),
    "scir_hip_sosfilt_f64_dev": (_I, [_P, _I64, _I64, _I64, _P, _I32, _P, _I64, _I32, _P]),
    "scir_hip_filtfilt_f64_dev": (_I, [_P, _I64, _I64, _I64, _P, _I32, _P, _I64, _I32, _P]),
    "scir_hip_sosfilt_f64_host": (_I, [_P, _I64, _I64, _I64, _P, _I32, _P, _I64, _I32]),
    "scir_hip_filtfilt_f64_host": (_I, [_P, _I64, _I64, _I64, _P, _I32, _P, _I64, _I32]),
    "scir_hip_sosfilt_f64_cpu": (_I, [_P, _I64, _I64, _I64, _P, _I32, _P, _I64, _I32]),
    "scir_hip_filtfilt_f64_cpu": (_I, [_P, _I64, _I64, _I64, _P, _I32, _P, _I64, _I32]),
    "scir_hip_fir_vjp_f32_host": (_I, [_P, _I64, _P, _I64, _P, _I32, _I64, _I64, _P, _I64, _P,
                                      _I32]),
    "scir_hip_buffer_create": (_I, [_c.c_size_t, _c.POINTER(_P)]),
    "scir_hip_buffer_upload": (_I, [_P, _P, _c.c_size_t]),
    "scir_hip_buffer_download": (_I, [_P, _P, _c.c_size_t]),
    "scir_hip_buffer_device_ptr": (_P, [_P]),
    "scir_hip_buffer_destroy": (_I, [_P]),
    "scir_hip_add_scalar_f32_dev": (_I, [_P, _F, _P, _I64, _P]),
    "scir_hip_mul_scalar_f32_dev": (_I, [_P, _F, _P, _I64, _P]),
    "scir_hip_add_f32_dev": (_I, [_P, _P, _P, _I64, _P]),
    "scir_hip_resample_len": (_I, [_I64, _I32, _I32, _I32, _Pi64]),
    "scir_hip_resample_poly_f64_dev": (_I, [_P, _I64, _I64, _I64, _I32, _I32, _P, _I32, _P,
                                            _I64, _P]),
    "scir_hip_resample_poly_f64_cpu": (_I, [_P, _I64, _I64, _I64, _I32, _I32, _P, _I32, _P,
                                            _I64, _I32]),
    "scir_hip_resample_poly_f64_host": (_I, [_P, _I64, _I64, _I64, _I32, _I32, _P, _I32, _P,
                                             _I64]),
}


class LibraryMissing(ImportError):
    """libscir_hip.so is not built (run ``python -m scir_amd.build``)."""


class HipError(RuntimeError):
    """A non-zero status from the C ABI, with the library's message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[scir_hip status {status}] {message}")
        self.status = status
        self.message = message


_lib = None


def lib():
    """The loaded library (loads on first use; raises LibraryMissing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `python -m scir_amd.build` "
                "(no CPU or torch fallback exists for the HIP path)")
        # One HIP runtime per process: torch ships its own libamdhip64 with
        # the same soname as the system one the library links.  Whichever is
        # loaded first serves both, and torch only works on its own build, so
        # when torch is importable it is loaded before the library (device
        # pointers and streams then come from the runtime the kernels use).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    """Raise HipError for a non-zero status."""
    if status != OK:
        msg = lib().scir_hip_last_error().decode(errors="replace")
        raise HipError(status, msg)


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().scir_hip_device_count(ctypes.byref(n)))
    return n.value


def precision_code(p) -> int:
    if isinstance(p, str):
        try:
            return PRECISIONS[p]
        except KeyError:
            raise ValueError(f"unknown precision {p!r}; one of {sorted(PRECISIONS)}") from None
    return int(p)
