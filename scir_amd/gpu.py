"""Host-side mirror of ``scir_gpu``'s FIR API over the HIP C ABI.

Mirrors, with the same names and argument meaning:

* ``Device`` / ``DType`` / ``GpuError`` -- crates/scir-gpu/src/lib.rs:15-35, 56-74;
  ``Device.Hip`` is the new variant beside ``Cpu``.
* ``fir1d_batched_f32(x, taps)`` -- the CPU FIR, crates/scir-gpu/src/lib.rs:1134-1152,
  here executed by the library's CPU arm (``scir_hip_fir_f32_cpu``), which
  keeps the reference's arithmetic order and is bit-identical to it.
* ``fir1d_batched_f32_hip(x, taps)`` -- the backend function that replaces
  ``cuda::fir1d_batched_f32_cuda`` (crates/scir-gpu/src/lib.rs:1036-1113):
  host arrays in, host array out, raises ``GpuError`` on failure.
* ``fir1d_batched_f32_auto(x, taps, device)`` -- the dispatcher,
  crates/scir-gpu/src/lib.rs:515-531.

``taps`` is in the reference's order everywhere in this module:
``y[i] = sum_{t=0}^{min(i,K-1)} taps[K-1-t] * x[i-t]``.  The C ABI takes the
lfilter-order ``b = taps[::-1]``; the reversal happens here.

Failure behaviour.  The reference's dispatcher silently recomputes on the CPU
when the CUDA call fails (crates/scir-gpu/src/lib.rs:520-523).  This mirror
raises ``GpuError`` instead, so a broken GPU path can never hide behind a
CPU result; set ``SCIR_HIP_CPU_FALLBACK=1`` to get the reference's fallback,
which then also writes one line to stderr per fallback.
"""
from __future__ import annotations

import ctypes
import enum
import os
import sys

import numpy as np

from . import _lib


class DType(enum.Enum):
    """Supported data types (crates/scir-gpu/src/lib.rs:15-22)."""
    F32 = "f32"
    F64 = "f64"


class Device(enum.Enum):
    """Execution device (crates/scir-gpu/src/lib.rs:24-35) with the HIP variant."""
    Cpu = "cpu"
    Hip = "hip"


class GpuError(RuntimeError):
    """GPU errors (crates/scir-gpu/src/lib.rs:56-74)."""


class BackendUnavailable(GpuError):
    """``GpuError::BackendUnavailable(String)``: the message names the failing call."""


class ShapeMismatch(GpuError):
    """``GpuError::ShapeMismatch``."""


def _as_x(x) -> np.ndarray:
    x = np.asarray(x)
    if x.ndim != 2:
        raise ShapeMismatch("x must be a 2-D array (batch, n)")
    # Standard layout fixes the reference's memory-order hazard
    # (crates/scir-gpu/src/lib.rs:1042 takes the raw vector of any layout).
    return np.ascontiguousarray(x, dtype=np.float32)


def _as_b(taps) -> np.ndarray:
    taps = np.asarray(taps, dtype=np.float32)
    if taps.ndim != 1:
        raise ShapeMismatch("taps must be a 1-D array")
    return np.ascontiguousarray(taps[::-1])


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def fir1d_batched_f32(x, taps) -> np.ndarray:
    """Causal FIR over each row of ``x`` (CPU, f32), same shape out.

    >>> import numpy as np
    >>> fir1d_batched_f32(np.array([[1., 2., 3., 4.], [0.5, 0., -0.5, -1.]]),
    ...                   np.array([0.25, 0.5, 0.25])).tolist()
    [[0.25, 1.0, 2.0, 3.0], [0.125, 0.25, 0.0, -0.5]]
    """
    xa = _as_x(x)
    b = _as_b(taps)
    B, N = xa.shape
    y = np.zeros((B, N), np.float32)
    status = _lib.lib().scir_hip_fir_f32_cpu(_ptr(xa), N, _ptr(b), b.size, B, N, _ptr(y), N, 0)
    if status != _lib.OK:
        raise GpuError(_lib.lib().scir_hip_last_error().decode())
    return y


def fir1d_batched_f32_hip(x, taps, precision="f32") -> np.ndarray:
    """HIP batched FIR, host arrays in and out (replaces
    ``cuda::fir1d_batched_f32_cuda``).  Raises ``GpuError`` on any failure.

    ``precision``: ``"f32"`` (default; FMA direct form within the K*eps bound),
    ``"f32_exact"`` (bit-identical to the CPU loop), ``"bf16x3"``, ``"fp16"``
    or ``"auto"``."""
    xa = _as_x(x)
    b = _as_b(taps)
    B, N = xa.shape
    y = np.zeros((B, N), np.float32)
    try:
        L = _lib.lib()
    except _lib.LibraryMissing as e:
        raise BackendUnavailable(str(e)) from None
    status = L.scir_hip_fir_f32_host(_ptr(xa), N, _ptr(b), b.size, B, N, _ptr(y), N,
                                     _lib.precision_code(precision))
    if status == _lib.ERR_INVALID_ARGUMENT:
        raise ShapeMismatch(L.scir_hip_last_error().decode())
    if status != _lib.OK:
        raise BackendUnavailable(L.scir_hip_last_error().decode())
    return y


def fir1d_batched_f32_auto(x, taps, device: Device) -> np.ndarray:
    """FIR over each row of ``x`` with device dispatch
    (crates/scir-gpu/src/lib.rs:515-531).

    >>> import numpy as np
    >>> fir1d_batched_f32_auto(np.array([[1., 2., 3., 4.]]), np.array([0.25, 0.5, 0.25]),
    ...                        Device.Cpu).shape
    (1, 4)
    """
    if device is Device.Cpu:
        return fir1d_batched_f32(x, taps)
    if device is Device.Hip:
        try:
            return fir1d_batched_f32_hip(x, taps)
        except GpuError as e:
            if os.environ.get("SCIR_HIP_CPU_FALLBACK") == "1":
                print(f"scir_amd: HIP FIR failed ({e}); recomputing on the CPU",
                      file=sys.stderr)
                return fir1d_batched_f32(x, taps)
            raise
    raise ValueError(f"unknown device {device!r}")
