"""The C-ABI library loads and exports exactly what include/scir_hip.h
declares; argument validation and the no-device path answer without a GPU.
No compute call is made here."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from scir_amd import _lib

HEADER = os.path.join(ROOT, "include", "scir_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"SCIR_HIP_API\s+(?:const\s+)?\w+\s*\*?\s*(scir_hip_\w+)\s*\(", text)))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    return sorted({l.split()[-1] for l in out.splitlines() if " T scir_hip_" in l})


def test_library_exports_every_declared_symbol():
    decl = declared_functions()
    assert len(decl) >= 20
    assert exported_symbols() == decl


def test_binding_covers_header():
    assert sorted(_lib.SIGNATURES) == declared_functions()
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name)


def test_version_and_device_count():
    L = _lib.lib()
    assert L.scir_hip_version() == 10100
    assert _lib.device_count() >= 0


def test_invalid_arguments_rejected_before_device_use():
    L = _lib.lib()
    x = np.zeros((2, 8), np.float32)
    p = x.ctypes.data_as(ctypes.c_void_p)
    st = L.scir_hip_fir_f32_host(p, 4, p, 3, 2, 8, p, 8, 0)  # ldx < N
    assert st == _lib.ERR_INVALID_ARGUMENT
    assert b"row stride" in L.scir_hip_last_error()
    st = L.scir_hip_fir_f32_host(p, 8, p, -1, 2, 8, p, 8, 0)
    assert st == _lib.ERR_INVALID_ARGUMENT
    st = L.scir_hip_fir_f32_host(None, 8, p, 3, 2, 8, p, 8, 0)
    assert st == _lib.ERR_INVALID_ARGUMENT
    # empty shapes are a no-op success, like the reference's empty loops
    assert L.scir_hip_fir_f32_host(p, 8, p, 3, 0, 8, p, 8, 0) == _lib.OK


def test_no_device_reports_backend_unavailable():
    if _lib.device_count() > 0:
        pytest.skip("a HIP device is visible")
    L = _lib.lib()
    x = np.ones((2, 8), np.float32)
    p = x.ctypes.data_as(ctypes.c_void_p)
    assert L.scir_hip_init() == _lib.ERR_NO_DEVICE
    assert L.scir_hip_fir_f32_host(p, 8, p, 3, 2, 8, p, 8, 0) == _lib.ERR_NO_DEVICE
    assert b"no HIP device" in L.scir_hip_last_error()
    assert L.scir_hip_shutdown() == _lib.OK


def test_error_message_is_thread_local():
    import threading
    L = _lib.lib()
    x = np.zeros((1, 4), np.float32)
    p = x.ctypes.data_as(ctypes.c_void_p)
    assert L.scir_hip_fir_f32_cpu(p, 2, p, 1, 1, 4, p, 4, 1) == _lib.ERR_INVALID_ARGUMENT
    seen = []
    t = threading.Thread(target=lambda: seen.append(L.scir_hip_last_error()))
    t.start(); t.join()
    assert seen == [b""]
    assert L.scir_hip_last_error() != b""


def test_every_tuning_key_is_documented():
    """scir_hip_tune_set accepts exactly the keys the header lists."""
    src = open(os.path.join(ROOT, "scir_amd", "csrc", "autotune.hip")).read()
    accepted = set(re.findall(r'strcmp\(key, "([a-z_.0-9]+)"\)', src))
    documented = set(re.findall(r'^ \*   "([a-z_.0-9]+)"', open(HEADER).read(), re.M))
    assert len(accepted) >= 15
    assert accepted == documented
