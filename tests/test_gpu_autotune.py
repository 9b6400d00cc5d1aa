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
