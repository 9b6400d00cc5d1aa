"""Device-tier wrappers for device-resident float32 torch tensors.

PyTorch only provides device memory and the current HIP stream here; all
filtering runs in libscir_hip.so.  Every function takes ``b`` in lfilter
order (``y[n] = sum_k b[k] x[n-k]``) like the C ABI and enqueues on torch's
current stream of the tensors' device.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _stream_ptr(t: torch.Tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _vp(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


def _check_rows(t: torch.Tensor, name: str):
    if t.dtype != torch.float32 or not t.is_cuda:
        raise TypeError(f"{name} must be a float32 tensor on a HIP device")
    if t.dim() != 2:
        raise ValueError(f"{name} must be 2-D (rows, samples)")
    if t.stride(1) != 1:
        raise ValueError(f"{name} rows must be contiguous (stride(1) == 1)")


def _check_taps(b: torch.Tensor):
    if b.dtype != torch.float32 or not b.is_cuda or b.dim() != 1 or not b.is_contiguous():
        raise TypeError("b must be a contiguous 1-D float32 tensor on a HIP device")


def fir(x: torch.Tensor, b: torch.Tensor, precision="f32", out: torch.Tensor | None = None):
    """``y = lfilter(b, [1], x)`` per row of ``x`` on the GPU."""
    _check_rows(x, "x")
    _check_taps(b)
    B, N = x.shape
    if out is None:
        out = torch.empty((B, N), dtype=torch.float32, device=x.device)
    _check_rows(out, "out")
    if tuple(out.shape) != (B, N):
        raise ValueError("out must have the shape of x")
    _lib.check(_lib.lib().scir_hip_fir_f32_dev(
        _vp(x), x.stride(0), _vp(b), b.numel(), B, N, _vp(out), out.stride(0),
        _lib.precision_code(precision), _stream_ptr(x)))
    return out


def vjp_dx(g: torch.Tensor, b: torch.Tensor, precision="f32", out=None):
    """``dx[r,m] = sum_k b[k] g[r,m+k]`` (gradient of sum(g*y) w.r.t. x)."""
    _check_rows(g, "g")
    _check_taps(b)
    B, N = g.shape
    if out is None:
        out = torch.empty((B, N), dtype=torch.float32, device=g.device)
    _check_rows(out, "out")
    _lib.check(_lib.lib().scir_hip_fir_vjp_dx_dev(
        _vp(g), g.stride(0), _vp(b), b.numel(), B, N, _vp(out), out.stride(0),
        _lib.precision_code(precision), _stream_ptr(g)))
    return out


def vjp_db(g: torch.Tensor, x: torch.Tensor, K: int, out=None):
    """``db[k] = sum_{r,n} g[r,n] x[r,n-k]`` (gradient of sum(g*y) w.r.t. b)."""
    _check_rows(g, "g")
    _check_rows(x, "x")
    if g.shape != x.shape:
        raise ValueError("g and x must have the same shape")
    B, N = x.shape
    if out is None:
        out = torch.empty(K, dtype=torch.float32, device=x.device)
    _lib.check(_lib.lib().scir_hip_fir_vjp_db_dev(
        _vp(g), g.stride(0), _vp(x), x.stride(0), int(K), B, N, _vp(out), _stream_ptr(x)))
    return out


def synth_normal(B: int, N: int, seed: int, outlier_frac: float = 0.0,
                 outlier_scale: float = 1.0, device="cuda", out=None) -> torch.Tensor:
    """Seeded N(0,1) float32 [B, N] from the library's counter-based generator."""
    if out is None:
        out = torch.empty((B, N), dtype=torch.float32, device=device)
    _check_rows(out, "out")
    _lib.check(_lib.lib().scir_hip_synth_normal_f32(
        _vp(out), out.stride(0), B, N, ctypes.c_uint64(seed & (2**64 - 1)),
        float(outlier_frac), float(outlier_scale), _stream_ptr(out)))
    return out


def sosfilt(x: torch.Tensor, sos: torch.Tensor, reference_order: bool = False,
            zero_phase: bool = False, out: torch.Tensor | None = None):
    """IIR cascade (``sosfilt``, or ``filtfilt`` with ``zero_phase``) over
    each row of a float64 tensor; ``sos`` is a contiguous [S, 6] float64
    tensor {b0, b1, b2, a0, a1, a2} on the same device (a0 not read).
    ``out`` may be ``x`` (in place)."""
    if x.dtype != torch.float64 or not x.is_cuda or x.dim() != 2 or x.stride(1) != 1:
        raise TypeError("x must be a 2-D float64 HIP tensor with contiguous rows")
    if (sos.dtype != torch.float64 or not sos.is_cuda or sos.dim() != 2 or sos.shape[1] != 6
            or not sos.is_contiguous()):
        raise TypeError("sos must be a contiguous [S, 6] float64 HIP tensor")
    rows, n = x.shape
    if out is None:
        out = torch.empty((rows, n), dtype=torch.float64, device=x.device)
    if out.dtype != torch.float64 or tuple(out.shape) != (rows, n) or out.stride(1) != 1:
        raise ValueError("out must be a float64 tensor of the shape of x")
    fn = _lib.lib().scir_hip_filtfilt_f64_dev if zero_phase else _lib.lib().scir_hip_sosfilt_f64_dev
    _lib.check(fn(_vp(x), max(x.stride(0), n), rows, n, _vp(sos), sos.shape[0], _vp(out),
                  max(out.stride(0), n), 1 if reference_order else 0, _stream_ptr(x)))
    return out
