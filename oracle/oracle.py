"""CPU oracle for the batched-FIR path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py`` import this
module: as the checker (bench.py's error figures against FP64) or as the timed
single-thread CPU baseline; the product (``scir_amd``) never imports it and
nothing measured as GPU throughput runs through it.  It wraps ``oracle/libfir_oracle.so`` (built from
``oracle/fir_oracle.c`` by ``oracle/Makefile``) with numpy-facing helpers.

Functions and the reference lines they restate:

* :func:`fir_ref_f32`  -- crates/scir-gpu/src/lib.rs:1134-1152, bit-faithful.
* :func:`fir_ref_f64`  -- crates/scir-gpu/src/lib.rs:1166-1184.
* :func:`lfilter_f64`  -- ``y = lfilter(b, [1], x)`` in FP64 on float32 inputs
  (the convention the reference's function equals with ``b = taps[::-1]``).
* :func:`abs_bound`    -- ``sum_k |b_k| |x_{n-k}|`` for the error bound.
* :func:`vjp_f64`      -- FP64 reference gradient of ``sum(g * y)``.
* :func:`lfilter_ragged_f64` -- the CSR batch form of :func:`lfilter_f64`.
* :func:`resample_poly_f64` -- crates/scir-signal/src/lib.rs:293-362
  (``convolve`` + ``resample_poly``), for any up/down and taps.
* :func:`sosfilt_f64` / :func:`filtfilt_f64` -- crates/scir-signal/src/lib.rs:
  260-291 over crates/scir-iir-filters/src/filter.rs:47-62 (Direct Form II
  Transposed), bit-faithful.
* :func:`sosfilt_longdouble` -- the same cascade in x87 extended precision,
  the near-exact result the float64 loop's own error is measured against.

Every function is pinned by tests/test_oracle.py, tests/test_resample.py or
tests/test_sosfilt.py (see the C file header).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfir_oracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64


def build(force: bool = False) -> str:
    """Compile the oracle library with oracle/Makefile (gcc, no FP contraction)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"],
                       check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.oracle_fir_ref_f32.argtypes = [_f32p, _i64, _f32p, _i64, _i64, _i64, _f32p, _i64]
        lib.oracle_fir_ref_f64.argtypes = [_f64p, _i64, _f64p, _i64, _i64, _i64, _f64p, _i64]
        lib.oracle_lfilter_f64.argtypes = [_f32p, _i64, _f32p, _i64, _i64, _i64, _f64p, _i64]
        lib.oracle_fir_absbound.argtypes = [_f32p, _i64, _f32p, _i64, _i64, _i64, _f64p, _i64]
        lib.oracle_fir_vjp_f64.argtypes = [_f32p, _i64, _f32p, _i64, _f32p, _i64, _i64, _i64,
                                           _f64p, _i64, _f64p]
        lib.oracle_lfilter_ragged_f64.argtypes = [_f32p, _i64p, _f32p, _i64p, _i64, _f64p]
        for fn in (lib.oracle_fir_ref_f32, lib.oracle_fir_ref_f64, lib.oracle_lfilter_f64,
                   lib.oracle_fir_absbound, lib.oracle_fir_vjp_f64,
                   lib.oracle_lfilter_ragged_f64):
            fn.restype = None
        lib.oracle_resample_len.argtypes = [_i64, _i64, _i64, _i64]
        lib.oracle_resample_len.restype = _i64
        lib.oracle_resample_poly_f64.argtypes = [_f64p, _i64, _i64, _i64, _f64p, _i64, _f64p,
                                                 _f64p]
        lib.oracle_resample_poly_f64.restype = _i64
        lib.oracle_sosfilt_f64.argtypes = [_f64p, _i64, _f64p, _i64, _f64p, _f64p]
        lib.oracle_sosfilt_f64.restype = None
        lib.oracle_filtfilt_f64.argtypes = [_f64p, _i64, _f64p, _i64, _f64p, _f64p, _f64p]
        lib.oracle_filtfilt_f64.restype = None
        _lib = lib
    return _lib


def _as2d(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.ndim == 1:
        a = a[None, :]
    if a.ndim != 2:
        raise ValueError("expected a 1-D or 2-D array")
    return a


def fir_ref_f32(x, taps) -> np.ndarray:
    """Reference-order FIR in float32, bit-identical to the reference loop."""
    x = _as2d(x, np.float32)
    taps = np.ascontiguousarray(taps, dtype=np.float32).ravel()
    B, N = x.shape
    y = np.zeros((B, N), np.float32)
    _load().oracle_fir_ref_f32(x, N, taps, taps.size, B, N, y, N)
    return y


def fir_ref_f64(x, taps) -> np.ndarray:
    """Reference-order FIR in float64 (crates/scir-gpu/src/lib.rs:1166-1184)."""
    x = _as2d(x, np.float64)
    taps = np.ascontiguousarray(taps, dtype=np.float64).ravel()
    B, N = x.shape
    y = np.zeros((B, N), np.float64)
    _load().oracle_fir_ref_f64(x, N, taps, taps.size, B, N, y, N)
    return y


def lfilter_f64(x, b) -> np.ndarray:
    """FP64 ``lfilter(b, [1], x)`` along the last axis, float32 inputs."""
    x = _as2d(x, np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32).ravel()
    B, N = x.shape
    y = np.zeros((B, N), np.float64)
    _load().oracle_lfilter_f64(x, N, b, b.size, B, N, y, N)
    return y


def abs_bound(x, b) -> np.ndarray:
    """``sum_k |b_k| |x_{n-k}|`` per output (float64)."""
    x = _as2d(x, np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32).ravel()
    B, N = x.shape
    s = np.zeros((B, N), np.float64)
    _load().oracle_fir_absbound(x, N, b, b.size, B, N, s, N)
    return s


def fir_error_bound(x, b) -> np.ndarray:
    """Per-output bound ``(K+1) * 2^-23 * sum_k |b_k||x_{n-k}|`` on any FP32
    evaluation order of the K-term sum (first-order gamma_K, strict form)."""
    K = int(np.asarray(b).size)
    return (K + 1) * 2.0 ** -23 * abs_bound(x, b)


def vjp_f64(g, x, b):
    """FP64 reference gradient of ``L = sum(g * lfilter(b,[1],x))``.

    Returns ``(dx float64 [B,N], db float64 [K])``."""
    g = _as2d(g, np.float32)
    x = _as2d(x, np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32).ravel()
    if g.shape != x.shape:
        raise ValueError("g and x must have the same shape")
    B, N = x.shape
    dx = np.zeros((B, N), np.float64)
    db = np.zeros(b.size, np.float64)
    _load().oracle_fir_vjp_f64(g, N, x, N, b, b.size, B, N, dx, N, db)
    return dx, db


def lfilter_ragged_f64(x, offs, b, boffs) -> np.ndarray:
    """FP64 ragged lfilter: signal s is ``x[offs[s]:offs[s+1]]`` filtered by
    ``b[boffs[s]:boffs[s+1]]``."""
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    b = np.ascontiguousarray(b, dtype=np.float32).ravel()
    offs = np.ascontiguousarray(offs, dtype=np.int64)
    boffs = np.ascontiguousarray(boffs, dtype=np.int64)
    y = np.zeros(x.size, np.float64)
    _load().oracle_lfilter_ragged_f64(x, offs, b, boffs, offs.size - 1, y)
    return y


def resample_len(n: int, up: int, down: int, K: int) -> int:
    """Number of outputs of :func:`resample_poly_f64` for n input samples."""
    return int(_load().oracle_resample_len(n, up, down, K))


def resample_poly_f64(x, up: int, down: int, h) -> np.ndarray:
    """Polyphase resampling of one float64 signal, restating the reference's
    upsample / full convolution / decimate steps in their operation order
    (bitwise equal to that loop).  ``h`` is the anti-aliasing filter (the
    reference uses 31 taps, up=2, down=3)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    h = np.ascontiguousarray(h, dtype=np.float64).ravel()
    n, K = x.size, h.size
    work = np.zeros(max(1, n * up + K - 1), np.float64)
    y = np.zeros(max(1, resample_len(n, up, down, K)), np.float64)
    m = _load().oracle_resample_poly_f64(x, n, up, down, h, K, work, y)
    return y[:m]


def _sos_args(x, sos):
    x = np.ascontiguousarray(x, dtype=np.float64)
    sos = np.ascontiguousarray(sos, dtype=np.float64).reshape(-1, 6)
    return x, sos


def sosfilt_f64(sos, x) -> np.ndarray:
    """The reference's sosfilt (Direct Form II Transposed cascade, zero
    state, reference operation order) over the last axis of ``x`` (1-D or
    2-D); ``sos`` is [S, 6] {b0, b1, b2, a0, a1, a2}, a0 not read."""
    x, sos = _sos_args(x, sos)
    rows = x.reshape(-1, x.shape[-1]) if x.ndim else x.reshape(1, 1)
    y = np.empty_like(rows)
    state = np.zeros(max(1, 2 * sos.shape[0]), np.float64)
    for r in range(rows.shape[0]):
        src = np.ascontiguousarray(rows[r])
        out = np.empty_like(src)
        _load().oracle_sosfilt_f64(src, src.size, sos, sos.shape[0], state, out)
        y[r] = out
    return y.reshape(x.shape)


def filtfilt_f64(sos, x) -> np.ndarray:
    """The reference's filtfilt: the cascade forward, then a fresh cascade
    over that output backward (no padding), over the last axis of ``x``."""
    x, sos = _sos_args(x, sos)
    rows = x.reshape(-1, x.shape[-1])
    y = np.empty_like(rows)
    state = np.zeros(max(1, 2 * sos.shape[0]), np.float64)
    tmp = np.empty(max(1, rows.shape[1]), np.float64)
    for r in range(rows.shape[0]):
        src = np.ascontiguousarray(rows[r])
        out = np.empty_like(src)
        _load().oracle_filtfilt_f64(src, src.size, sos, sos.shape[0], state, tmp, out)
        y[r] = out
    return y.reshape(x.shape)


def sosfilt_longdouble(sos, x, zero_phase: bool = False) -> np.ndarray:
    """The same cascade in x87 extended precision (numpy longdouble, 64-bit
    mantissa) over the rows of a 2-D ``x``: the near-exact result against
    which the float64 reference loop's own rounding error is measured (for
    ill-conditioned filters that error is far above one ulp)."""
    sos = np.asarray(sos, dtype=np.longdouble).reshape(-1, 6)
    x = np.asarray(x, dtype=np.longdouble)
    rows, n = x.shape

    def run(src):
        s1 = np.zeros((rows, sos.shape[0]), np.longdouble)
        s2 = np.zeros_like(s1)
        out = np.empty_like(src)
        for t in range(n):
            v = src[:, t]
            for i in range(sos.shape[0]):
                b0, b1, b2, _, a1, a2 = sos[i]
                p = v
                v = b0 * p + s1[:, i]
                s1[:, i] = b1 * p - a1 * v + s2[:, i]
                s2[:, i] = b2 * p - a2 * v
            out[:, t] = v
        return out

    y = run(x)
    if zero_phase:
        y = run(y[:, ::-1].copy())[:, ::-1]
    return y


def rel_l2(y, y64) -> float:
    """``||y - y64||_2 / ||y64||_2`` computed in float64."""
    y = np.asarray(y, np.float64)
    y64 = np.asarray(y64, np.float64)
    den = np.linalg.norm(y64.ravel())
    return float(np.linalg.norm((y - y64).ravel()) / (den if den > 0 else 1.0))
