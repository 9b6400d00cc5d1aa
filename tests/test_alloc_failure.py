"""Allocation failures inside the library (injected with the environment
variable SCIR_HIP_FAIL_ALLOC) come back as a status and a message from every
entry point, never as a C++ exception crossing the C ABI (which would end the
caller's process).  The CPU arms need no device; the device entry points are
covered by the GPU test at the end."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from scir_amd import _lib


@pytest.fixture
def inject():
    os.environ["SCIR_HIP_FAIL_ALLOC"] = "1"
    try:
        yield
    finally:
        os.environ.pop("SCIR_HIP_FAIL_ALLOC", None)


def vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def cpu_calls():
    x = np.ones((4, 64), np.float32)
    y = np.empty_like(x)
    b = np.ones(5, np.float32)
    xd = np.ones((4, 64), np.float64)
    yd = np.empty((4, 64), np.float64)
    sos = np.array([[1.0, 0.0, 0.0, 1.0, -0.5, 0.0]])
    h = np.ones(7, np.float64)
    L = _lib.lib()
    return [
        ("fir_f32_cpu", lambda: L.scir_hip_fir_f32_cpu(vp(x), 64, vp(b), 5, 4, 64, vp(y), 64, 4)),
        ("sosfilt_f64_cpu", lambda: L.scir_hip_sosfilt_f64_cpu(vp(xd), 64, 4, 64, vp(sos), 1,
                                                               vp(yd), 64, 4)),
        ("filtfilt_f64_cpu", lambda: L.scir_hip_filtfilt_f64_cpu(vp(xd), 64, 4, 64, vp(sos), 1,
                                                                 vp(yd), 64, 4)),
        ("resample_poly_f64_cpu", lambda: L.scir_hip_resample_poly_f64_cpu(
            vp(xd), 64, 4, 64, 2, 3, vp(h), 7, vp(yd), 64, 4)),
    ]


@pytest.mark.parametrize("name,call", cpu_calls(), ids=lambda v: v if isinstance(v, str) else "")
def test_cpu_arms_report_host_allocation_failure(inject, name, call):
    L = _lib.lib()
    assert call() == _lib.ERR_OUT_OF_MEMORY
    assert b"host memory allocation failed" in L.scir_hip_last_error()
    assert L.scir_hip_last_error().startswith(b"scir_hip_")


@pytest.mark.parametrize("name,call", cpu_calls(), ids=lambda v: v if isinstance(v, str) else "")
def test_cpu_arms_recover_after_injection(name, call):
    os.environ.pop("SCIR_HIP_FAIL_ALLOC", None)
    assert call() == _lib.OK


def test_zero_disables_injection():
    os.environ["SCIR_HIP_FAIL_ALLOC"] = "0"
    try:
        name, call = cpu_calls()[0]
        assert call() == _lib.OK
    finally:
        os.environ.pop("SCIR_HIP_FAIL_ALLOC", None)


@pytest.mark.gpu
def test_device_entries_report_allocation_failure(hip, inject):
    """Entry points that always allocate (plan, buffer, graph handle) fail
    with ERR_OUT_OF_MEMORY and a message while injection is on, and work
    again afterwards; the library stays usable."""
    import torch
    L = hip.lib()
    offs = np.array([0, 100, 300], np.int64)
    boffs = np.array([0, 5, 12], np.int64)
    i64 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    plan = ctypes.c_void_p()
    assert L.scir_hip_ragged_plan_create(i64(offs), i64(boffs), 2, 0,
                                         ctypes.byref(plan)) == _lib.ERR_OUT_OF_MEMORY
    assert not plan.value
    assert L.scir_hip_last_error()
    buf = ctypes.c_void_p()
    assert L.scir_hip_buffer_create(1024, ctypes.byref(buf)) == _lib.ERR_OUT_OF_MEMORY
    assert b"SCIR_HIP_FAIL_ALLOC" in L.scir_hip_last_error() or \
        b"allocation failed" in L.scir_hip_last_error()
    s = torch.cuda.Stream()
    sp = ctypes.c_void_p(s.cuda_stream)
    assert L.scir_hip_graph_begin(sp) == _lib.OK
    g = ctypes.c_void_p()
    assert L.scir_hip_graph_end(sp, ctypes.byref(g)) == _lib.ERR_OUT_OF_MEMORY
    assert not g.value
    os.environ.pop("SCIR_HIP_FAIL_ALLOC", None)
    assert L.scir_hip_ragged_plan_create(i64(offs), i64(boffs), 2, 0, ctypes.byref(plan)) == _lib.OK
    assert L.scir_hip_ragged_plan_destroy(plan) == _lib.OK
    assert L.scir_hip_buffer_create(1024, ctypes.byref(buf)) == _lib.OK
    assert L.scir_hip_buffer_destroy(buf) == _lib.OK
    x = torch.randn(2, 1000, device="cuda")
    b = torch.ones(9, device="cuda")
    from scir_amd import device
    y = device.fir(x, b, "f32")
    torch.cuda.synchronize()
    assert torch.isfinite(y).all()
