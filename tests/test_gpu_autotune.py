"""The autotuner runs on the device, writes a table that loads back, and the
tuned AUTO path stays within the FP32 bound."""
from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu


def test_autotune_roundtrip_and_auto_parity(hip, oracle, tmp_path):
    import torch
    from scir_amd import device
    L = hip.lib()
    path = tmp_path / "tune.json"
    hip.check(L.scir_hip_autotune(str(path).encode(), 60000))
    table = json.loads(path.read_text())
    assert table["arch"].startswith("gfx950")
    assert len(table["entries"]) >= 1 and table["auto_mfma_min_k"] >= 1
    assert table["auto_fft_min_k"] >= 1
    print("autotune:", json.dumps(table))
    hip.check(L.scir_hip_autotune_load(str(path).encode()))
    r = rng(21)
    for K in (5, 40, 200, 700):
        x = r.standard_normal((3, 20000)).astype(np.float32)
        b = r.standard_normal(K).astype(np.float32)
        y = device.fir(torch.tensor(x, device="cuda"), torch.tensor(b, device="cuda"), "auto")
        torch.cuda.synchronize()
        y64 = oracle.lfilter_f64(x, b)
        assert np.all(np.abs(y.cpu().numpy() - y64) <= oracle.fir_error_bound(x, b))


def test_auto_fft_opt_in(hip, oracle):
    """AUTO takes the FFT path only when auto.fft is set, from auto.fft_min_k
    up; its output then meets the FFT path's normwise bound, and without the
    key it keeps the elementwise FP32 bound."""
    import torch
    from scir_amd import device
    L = hip.lib()
    r = rng(22)
    K = 600
    x = r.standard_normal((2, 30000)).astype(np.float32)
    b = (r.standard_normal(K) / 25).astype(np.float32)
    xt, bt = torch.tensor(x, device="cuda"), torch.tensor(b, device="cuda")
    y64 = oracle.lfilter_f64(x, b)
    y_auto = device.fir(xt, bt, "auto").cpu().numpy()
    assert np.all(np.abs(y_auto - y64) <= oracle.fir_error_bound(x, b))
    try:
        hip.check(L.scir_hip_tune_set(b"auto.fft", 1))
        hip.check(L.scir_hip_tune_set(b"auto.fft_min_k", 512))
        y_fft = device.fir(xt, bt, "auto").cpu().numpy()
        assert np.array_equal(y_fft, device.fir(xt, bt, "fft32").cpu().numpy())
        hip.check(L.scir_hip_tune_set(b"auto.fft_min_k", 1000))  # above K: not taken
        assert np.array_equal(device.fir(xt, bt, "auto").cpu().numpy(), y_auto)
    finally:
        L.scir_hip_tune_set(b"auto.fft", -1)
        L.scir_hip_tune_set(b"auto.fft_min_k", -1)
