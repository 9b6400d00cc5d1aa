"""Mirror of the ``scir_signal`` functions this library accelerates.

* ``fir1d_batched_f32(x, taps, device)`` -- the ``scir_signal::gpu``
  forwarder (crates/scir-signal/src/lib.rs:364-375): batched FIR with device
  selection.
* ``resample_poly(x, up, down, device)`` -- ``scir_signal::resample_poly``
  (crates/scir-signal/src/lib.rs:304-362), the reference's polyphase
  resampler, with the ``Device`` choice added: ``Device.Cpu`` runs the
  library's CPU arm in the reference's operation order (bit-identical to it),
  ``Device.Hip`` the batched float64 GPU kernel.  ``resample_poly_batched``
  is the batched, any-``up``/``down`` form behind both.
* ``sosfilt(sos, x, device)`` and ``filtfilt(sos, x, device)`` --
  ``scir_signal::sosfilt`` / ``filtfilt`` (crates/scir-signal/src/lib.rs:
  260-291) with the ``Device`` choice added (the same two arms);
  ``sosfilt_batched`` / ``filtfilt_batched`` filter every row of a 2-D array.
  ``sos`` is an [S, 6] array of {b0, b1, b2, a0, a1, a2} rows (scipy's
  layout and the #[repr(C)] layout of the reference's SosCoeffs); a0 is not
  read, as in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .gpu import BackendUnavailable, Device, ShapeMismatch, fir1d_batched_f32_auto

# The reference's fixed 31-tap filter for up = 2, down = 3
# (crates/scir-signal/src/lib.rs:315-346): twice the 31-tap Hamming low-pass
# at 1/3 of Nyquist (scir_amd.design.resample_filter(2, 3) reproduces it to
# within one ulp).  Kept as the same decimal values so that the CPU arm's
# output is bit-identical to the reference's.
RESAMPLE_H_2_3 = np.array([
    2.074505559623708e-18, 3.538600036880219e-3, 5.068355234374961e-3,
    -4.352628574659906e-18, -1.161274884414711e-2, -1.705404264360788e-2,
    1.031683205857806e-17, 3.382730607794791e-2, 4.62174845519145e-2,
    -1.768899299634197e-17, -8.47200828101812e-2, -1.166001508003416e-1,
    2.365319648026013e-17, 2.641261112200551e-1, 5.446004399706618e-1,
    6.652174560128868e-1, 5.446004399706618e-1, 2.641261112200551e-1,
    2.365319648026013e-17, -1.166001508003416e-1, -8.472008281018122e-2,
    -1.768899299634197e-17, 4.621748455191452e-2, 3.382730607794793e-2,
    1.031683205857806e-17, -1.70540426436079e-2, -1.161274884414712e-2,
    -4.352628574659906e-18, 5.068355234374964e-3, 3.538600036880219e-3,
    2.074505559623708e-18,
], dtype=np.float64)


def fir1d_batched_f32(x, taps, device: Device):
    """Batched FIR for f32 with device selection (forwards to the dispatcher)."""
    return fir1d_batched_f32_auto(x, taps, device)


def resample_len(n: int, up: int, down: int, K: int) -> int:
    """Number of outputs for n input samples (the reference's loop bound)."""
    m = ctypes.c_int64(0)
    _lib.check(_lib.lib().scir_hip_resample_len(int(n), int(up), int(down), int(K),
                                                 ctypes.byref(m)))
    return m.value


def resample_poly_batched(x, up: int, down: int, h, device: Device = Device.Cpu) -> np.ndarray:
    """Polyphase resampling of every row of ``x`` (float64): upsample by
    ``up``, filter with ``h``, keep every ``down``-th sample from
    ``(len(h)-1)//2`` -- the steps of crates/scir-signal/src/lib.rs:313-362
    for any up, down >= 1.  Returns ``[rows, resample_len(...)]``."""
    xa = np.asarray(x, dtype=np.float64)
    if xa.ndim != 2:
        raise ShapeMismatch("x must be a 2-D array (rows, n)")
    xa = np.ascontiguousarray(xa)
    ha = np.ascontiguousarray(np.asarray(h, dtype=np.float64).ravel())
    B, N = xa.shape
    M = resample_len(N, up, down, ha.size)
    y = np.zeros((B, M), np.float64)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    try:
        L = _lib.lib()
    except _lib.LibraryMissing as e:
        raise BackendUnavailable(str(e)) from None
    if device is Device.Cpu:
        status = L.scir_hip_resample_poly_f64_cpu(vp(xa), N, B, N, up, down, vp(ha), ha.size,
                                                  vp(y), max(M, 1), 0)
    elif device is Device.Hip:
        status = L.scir_hip_resample_poly_f64_host(vp(xa), N, B, N, up, down, vp(ha), ha.size,
                                                   vp(y), max(M, 1))
    else:
        raise ValueError(f"unknown device {device!r}")
    if status == _lib.ERR_INVALID_ARGUMENT:
        raise ShapeMismatch(L.scir_hip_last_error().decode())
    if status != _lib.OK:
        raise BackendUnavailable(L.scir_hip_last_error().decode())
    return y


def resample_poly(x, up: int, down: int, device: Device = Device.Cpu) -> np.ndarray:
    """``scir_signal::resample_poly(input, up, down)``: one float64 signal,
    ``up == 2 and down == 3`` only (the reference asserts it), the reference's
    31-tap filter.

    >>> import numpy as np
    >>> len(resample_poly(np.array([1.0, 2.0, 3.0, 4.0]), 2, 3)) > 0
    True
    """
    if not (up == 2 and down == 3):
        raise ValueError("resample_poly supports up == 2 and down == 3 (as the reference)")
    xa = np.asarray(x, dtype=np.float64).ravel()
    return resample_poly_batched(xa[None, :], up, down, RESAMPLE_H_2_3, device)[0]


def _sos_call(name, sos, x, device, reference_order):
    xa = np.asarray(x, dtype=np.float64)
    if xa.ndim != 2:
        raise ShapeMismatch("x must be a 2-D array (rows, n)")
    xa = np.ascontiguousarray(xa)
    sa = np.ascontiguousarray(np.asarray(sos, dtype=np.float64))
    if sa.ndim != 2 or sa.shape[1] != 6:
        raise ShapeMismatch("sos must be an [S, 6] array")
    rows, n = xa.shape
    y = np.empty_like(xa)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    try:
        L = _lib.lib()
    except _lib.LibraryMissing as e:
        raise BackendUnavailable(str(e)) from None
    if device is Device.Cpu:
        status = getattr(L, f"scir_hip_{name}_f64_cpu")(vp(xa), max(n, 1), rows, n, vp(sa),
                                                       sa.shape[0], vp(y), max(n, 1), 0)
    elif device is Device.Hip:
        flags = 1 if reference_order else 0  # SCIR_HIP_SOS_REFERENCE_ORDER
        status = getattr(L, f"scir_hip_{name}_f64_host")(vp(xa), max(n, 1), rows, n, vp(sa),
                                                        sa.shape[0], vp(y), max(n, 1), flags)
    else:
        raise ValueError(f"unknown device {device!r}")
    if status == _lib.ERR_INVALID_ARGUMENT:
        raise ShapeMismatch(L.scir_hip_last_error().decode())
    if status != _lib.OK:
        raise BackendUnavailable(L.scir_hip_last_error().decode())
    return y


def sosfilt_batched(sos, x, device: Device = Device.Cpu, reference_order: bool = False):
    """``sosfilt`` over every row of ``x`` ([rows, n] float64).  On
    ``Device.Hip`` rows are split in time and joined through the state
    transition (within a few ulps of the reference loop) unless
    ``reference_order`` asks for the reference loop's exact result."""
    return _sos_call("sosfilt", sos, x, device, reference_order)


def filtfilt_batched(sos, x, device: Device = Device.Cpu, reference_order: bool = False):
    """``filtfilt`` over every row of ``x`` ([rows, n] float64)."""
    return _sos_call("filtfilt", sos, x, device, reference_order)


def sosfilt(sos, x, device: Device = Device.Cpu) -> np.ndarray:
    """``scir_signal::sosfilt(&sos, &x)``: one float64 signal through the
    second-order-section cascade from zero state.

    >>> import numpy as np
    >>> sosfilt(np.array([[0.5, 0.0, 0.0, 1.0, -0.5, 0.0]]), np.array([1.0, 0.0, 0.0])).tolist()
    [0.5, 0.25, 0.125]
    """
    xa = np.asarray(x, dtype=np.float64).ravel()
    return sosfilt_batched(sos, xa[None, :], device)[0]


def filtfilt(sos, x, device: Device = Device.Cpu) -> np.ndarray:
    """``scir_signal::filtfilt(&sos, &x)``: the cascade forward, then a fresh
    cascade backward over that output (zero phase, no padding)."""
    xa = np.asarray(x, dtype=np.float64).ravel()
    return filtfilt_batched(sos, xa[None, :], device)[0]
