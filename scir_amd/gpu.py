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


def fir1d_batched_f32_hip(x, taps, precision="auto") -> np.ndarray:
    """HIP batched FIR, host arrays in and out (replaces
    ``cuda::fir1d_batched_f32_cuda``).  Raises ``GpuError`` on any failure.

    ``precision``: ``"auto"`` (default: the FP32 direct form for short
    filters, BF16x3 on MFMA for long ones; both within (K+1)*2^-23*sum|b||x|
    of the exact sum and non-finite exactly where the CPU loop is),
    ``"f32"`` (FMA direct form), ``"f32_exact"`` (bit-identical to the CPU
    loop), ``"bf16x3"``, ``"fp16"`` or ``"fft32"``."""
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


def fir1d_batched_f32_vjp_hip(x, taps, g, precision="f32"):
    """Vector-Jacobian product of ``fir1d_batched_f32`` on the HIP device
    (host arrays in and out): for L = sum(g * fir1d_batched_f32(x, taps))
    returns ``(dL/dx, dL/dtaps)``, both in the caller's convention (dtaps in
    the reference's taps order).  No counterpart in the reference."""
    xa = _as_x(x)
    ga = _as_x(g)
    if ga.shape != xa.shape:
        raise ShapeMismatch("g must have the shape of x")
    b = _as_b(taps)
    B, N = xa.shape
    dx = np.zeros((B, N), np.float32)
    db = np.zeros(b.size, np.float32)
    try:
        L = _lib.lib()
    except _lib.LibraryMissing as e:
        raise BackendUnavailable(str(e)) from None
    status = L.scir_hip_fir_vjp_f32_host(_ptr(xa), max(N, 1), _ptr(ga), max(N, 1), _ptr(b), b.size,
                                         B, N, _ptr(dx), max(N, 1), _ptr(db),
                                         _lib.precision_code(precision))
    if status == _lib.ERR_INVALID_ARGUMENT:
        raise ShapeMismatch(L.scir_hip_last_error().decode())
    if status != _lib.OK:
        raise BackendUnavailable(L.scir_hip_last_error().decode())
    return dx, db[::-1].copy()


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


class DeviceArray:
    """``scir_gpu::DeviceArray`` (crates/scir-gpu/src/lib.rs:76-260) with a
    device-resident HIP arm.

    The reference keeps host storage whatever the device and ``to_device``
    only tags it.  Here ``to_device(Device.Hip)`` copies the data into a
    library-owned device buffer and later operations on the array --
    ``add_scalar``, ``mul_scalar``, ``add`` and :func:`fir1d_batched_f32_resident`
    -- run on the device without host transfers; ``to_cpu_vec`` and
    ``to_device(Device.Cpu)`` copy back.  Arrays on the HIP device hold
    float32 (the dtype of every HIP operation here).

    >>> import numpy as np
    >>> a = DeviceArray.from_cpu_slice([2, 2], DType.F32, [1.0, 2.0, 3.0, 4.0])
    >>> a.shape(), a.device()
    ([2, 2], <Device.Cpu: 'cpu'>)
    >>> a.add_scalar(1.0).to_cpu_vec().tolist()
    [2.0, 3.0, 4.0, 5.0]
    """

    def __init__(self, shape, dtype: DType, host: np.ndarray):
        self._shape = [int(s) for s in shape]
        self._dtype = dtype
        self._device = Device.Cpu
        self._host = host          # 1-D numpy storage when on the CPU
        self._buf = None           # scir_hip_buffer when on the HIP device

    @staticmethod
    def _np_dtype(dtype: DType):
        return np.float32 if dtype is DType.F32 else np.float64

    @classmethod
    def from_cpu_slice(cls, shape, dtype: DType, data) -> "DeviceArray":
        """Copy ``data`` (length = product of ``shape``) into a CPU array."""
        arr = np.array(data, dtype=cls._np_dtype(dtype)).ravel()
        if int(np.prod(shape, dtype=np.int64)) != arr.size:
            raise ValueError("shape does not match the data length")
        return cls(shape, dtype, arr)

    def shape(self):
        return list(self._shape)

    def dtype(self) -> DType:
        return self._dtype

    def device(self) -> Device:
        return self._device

    def _size(self) -> int:
        return int(np.prod(self._shape, dtype=np.int64))

    def _nbytes(self) -> int:
        return self._size() * np.dtype(self._np_dtype(self._dtype)).itemsize

    def _dev_ptr(self):
        return ctypes.c_void_p(_lib.lib().scir_hip_buffer_device_ptr(self._buf))

    def to_cpu_vec(self) -> np.ndarray:
        """The elements as a new 1-D host array."""
        if self._device is Device.Cpu:
            return self._host.copy()
        out = np.empty(self._size(), dtype=self._np_dtype(self._dtype))
        _check_gpu(_lib.lib().scir_hip_buffer_download(self._buf, _ptr(out), self._nbytes()))
        return out

    def to_device(self, device: Device) -> None:
        """Move the array (``Device.Hip``: allocate and copy; ``Device.Cpu``:
        copy back and free).  Raises ``GpuError`` on failure."""
        if device is self._device:
            return
        if device is Device.Cpu:
            self._host = self.to_cpu_vec()
            self._release()
            self._device = Device.Cpu
            return
        if device is not Device.Hip:
            raise ValueError(f"unknown device {device!r}")
        if self._dtype is not DType.F32:
            raise BackendUnavailable("arrays on the HIP device hold float32")
        try:
            L = _lib.lib()
        except _lib.LibraryMissing as e:
            raise BackendUnavailable(str(e)) from None
        buf = ctypes.c_void_p()
        _check_gpu(L.scir_hip_buffer_create(self._nbytes(), ctypes.byref(buf)))
        status = L.scir_hip_buffer_upload(buf, _ptr(self._host), self._nbytes())
        if status != _lib.OK:
            L.scir_hip_buffer_destroy(buf)
            _check_gpu(status)
        self._buf, self._host, self._device = buf, None, Device.Hip

    def _release(self):
        if self._buf is not None:
            _lib.lib().scir_hip_buffer_destroy(self._buf)
            self._buf = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def _like_on_device(self) -> "DeviceArray":
        out = DeviceArray(self._shape, self._dtype, None)
        buf = ctypes.c_void_p()
        _check_gpu(_lib.lib().scir_hip_buffer_create(self._nbytes(), ctypes.byref(buf)))
        out._buf, out._device = buf, Device.Hip
        return out

    def _elementwise(self, name, alpha=None, other=None) -> "DeviceArray":
        if self._device is Device.Cpu:
            a = self._host
            t = a.dtype.type
            res = {"add_scalar": lambda: a + t(alpha), "mul_scalar": lambda: a * t(alpha),
                   "add": lambda: a + other._host}[name]()
            return DeviceArray(self._shape, self._dtype, res.astype(a.dtype, copy=False))
        out = self._like_on_device()
        L = _lib.lib()
        n = self._size()
        if name == "add_scalar":
            st = L.scir_hip_add_scalar_f32_dev(self._dev_ptr(), float(alpha), out._dev_ptr(), n, None)
        elif name == "mul_scalar":
            st = L.scir_hip_mul_scalar_f32_dev(self._dev_ptr(), float(alpha), out._dev_ptr(), n, None)
        else:
            st = L.scir_hip_add_f32_dev(self._dev_ptr(), other._dev_ptr(), out._dev_ptr(), n, None)
        _check_gpu(st)
        return out

    def add_scalar(self, alpha) -> "DeviceArray":
        """Elementwise ``x + alpha`` on the array's device (lib.rs:206-221)."""
        return self._elementwise("add_scalar", alpha=alpha)

    def mul_scalar(self, alpha) -> "DeviceArray":
        """Elementwise ``x * alpha`` on the array's device (lib.rs:223-243)."""
        return self._elementwise("mul_scalar", alpha=alpha)

    def add(self, other: "DeviceArray") -> "DeviceArray":
        """Elementwise sum; ``ShapeMismatch`` for other shapes or devices
        (lib.rs:245-260)."""
        if other._shape != self._shape or other._device is not self._device:
            raise ShapeMismatch("add needs arrays of one shape on one device")
        return self._elementwise("add", other=other)

    # The reference's ``*_auto`` names (lib.rs:262-380) dispatch on the
    # array's device, as the methods above do.
    add_scalar_auto = add_scalar
    mul_scalar_auto = mul_scalar
    add_auto = add


def _check_gpu(status: int) -> None:
    if status == _lib.OK:
        return
    msg = _lib.lib().scir_hip_last_error().decode()
    if status == _lib.ERR_INVALID_ARGUMENT:
        raise ShapeMismatch(msg)
    raise BackendUnavailable(msg)


def fir1d_batched_f32_resident(x: DeviceArray, taps, precision="f32") -> DeviceArray:
    """Batched FIR of a ``[batch, n]`` array on the HIP device; the result
    stays on the device (no host transfer of x or y).  ``taps`` in the
    reference's order."""
    if x.device() is not Device.Hip or len(x.shape()) != 2:
        raise ShapeMismatch("x must be a 2-D DeviceArray on Device.Hip")
    B, N = x.shape()
    b = _as_b(taps)
    L = _lib.lib()
    out = x._like_on_device()
    bbuf = ctypes.c_void_p()
    _check_gpu(L.scir_hip_buffer_create(max(1, b.nbytes), ctypes.byref(bbuf)))
    try:
        _check_gpu(L.scir_hip_buffer_upload(bbuf, _ptr(b), b.nbytes))
        _check_gpu(L.scir_hip_fir_f32_dev(x._dev_ptr(), N, ctypes.c_void_p(
            L.scir_hip_buffer_device_ptr(bbuf)), b.size, B, N, out._dev_ptr(), N,
            _lib.precision_code(precision), None))
    finally:
        L.scir_hip_buffer_destroy(bbuf)  # waits for the device first
    return out
