"""scir_amd -- MI355X (gfx950) implementation of SciR's batched FIR path.

Layout:
  csrc/        HIP kernels and the C ABI (built into libscir_hip.so)
  _lib.py      ctypes binding of the C ABI (include/scir_hip.h)
  gpu.py       mirror of scir_gpu's FIR API (Device, fir1d_batched_f32_auto, ...)
  signal.py    mirror of the scir_signal::gpu forwarder
  device.py    device-tier calls on torch tensors (memory/streams only)
  ragged.py    CSR ragged batches (plans, persistent launch, graph replay)
  autograd.py  test-only torch.autograd.Function over the adjoint kernels
  design.py    windowed-sinc tap design (firwin conventions)
  configs.py   the benchmark workloads' synthetic inputs
  build.py     in-tree build of libscir_hip.so
"""
from .gpu import (BackendUnavailable, Device, DType, GpuError, ShapeMismatch,  # noqa: F401
                  fir1d_batched_f32, fir1d_batched_f32_auto, fir1d_batched_f32_hip)
