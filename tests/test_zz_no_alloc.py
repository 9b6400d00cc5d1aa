"""The allocation-free dL/db entry point (scir_hip_fir_vjp_db_noalloc_dev)
through the C ABI.

* Two calls in a row, the second smaller: each equals the allocating entry
  point scir_hip_fir_vjp_db_dev bit for bit (same kernels and grouping).
* A row length whose usual grouping would not fit the fixed array runs
  regrouped, within the FP64 bound of the dL/db tests.
* A batch with rows * K beyond the array is refused with a message.
* On a stream the library has never seen, the call allocates nothing (free
  device memory unchanged) and can be captured into a HIP graph, whose
  replays equal eager results.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

LIMIT = 1 << 20  # partial sums the entry point's array holds


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def db_call(fn, g, x, K, db, stream):
    B, N = x.shape
    return fn(_vp(g), g.stride(0), _vp(x), x.stride(0), K, B, N, _vp(db),
              ctypes.c_void_p(stream.cuda_stream))


def inputs(seed, B, N):
    import torch
    r = rng(seed)
    x = torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda")
    g = torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda")
    return x, g


def test_noalloc_db_equals_allocating_entry_twice(hip):
    import torch
    L = hip.lib()
    s = torch.cuda.current_stream()
    for B, N, K in ((64, 50000, 255), (8, 10000, 100)):  # the second call is smaller
        print(f"noalloc dL/db B={B} N={N} K={K}", flush=True)
        x, g = inputs(B + N + K, B, N)
        ref = torch.full((K,), float("nan"), device="cuda")
        got = torch.full((K,), float("nan"), device="cuda")
        hip.check(db_call(L.scir_hip_fir_vjp_db_dev, g, x, K, ref, s))
        hip.check(db_call(L.scir_hip_fir_vjp_db_noalloc_dev, g, x, K, got, s))
        torch.cuda.synchronize()
        assert torch.equal(got, ref), f"B={B} N={N} K={K}: differs from scir_hip_fir_vjp_db_dev"


def test_noalloc_db_regrouped_within_bound(hip, oracle):
    """4 rows of 300000 samples, K = 2000: the usual 588 workgroups x 2000
    partial sums exceed the array, so the entry point runs 296 longer
    workgroups; the result stays within the bound of the dL/db tests."""
    import torch
    L = hip.lib()
    B, N, K = 4, 300000, 2000
    print(f"noalloc dL/db regrouped B={B} N={N} K={K}", flush=True)
    x, g = inputs(77, B, N)
    got = torch.empty(K, device="cuda")
    hip.check(db_call(L.scir_hip_fir_vjp_db_noalloc_dev, g, x, K, got,
                      torch.cuda.current_stream()))
    torch.cuda.synchronize()
    xs, gs = x.cpu().numpy(), g.cpu().numpy()
    b = np.ones(K, np.float32)
    _, db64 = oracle.vjp_f64(gs, xs, b)
    _, dba = oracle.vjp_f64(np.abs(gs), np.abs(xs), b)
    bound = (N / 2048 + 200) * 2.0 ** -24 * dba
    assert np.all(np.abs(got.cpu().numpy() - db64) <= bound)


def test_noalloc_db_refuses_beyond_its_array(hip):
    import torch
    L = hip.lib()
    B, N, K = 2000, 100, 1000  # B * K = 2e6 partial sums > 2^20
    x, g = inputs(5, B, N)
    db = torch.zeros(K, device="cuda")
    st = db_call(L.scir_hip_fir_vjp_db_noalloc_dev, g, x, K, db, torch.cuda.current_stream())
    assert st == hip.ERR_INVALID_ARGUMENT
    msg = L.scir_hip_last_error()
    assert b"allocation-free dL/db holds 1048576 partial sums" in msg, msg


def test_noalloc_db_on_a_new_stream_allocates_nothing_and_captures(hip):
    """The allocating entry point needs per-stream scratch, so on a new stream
    it cannot be captured (SCIR_HIP_ERR_CAPTURE); the allocation-free one
    can, and its replays equal eager calls."""
    import torch
    from scir_amd.ragged import Graph
    L = hip.lib()
    B, N, K = 16, 40000, 255
    x, g = inputs(9, B, N)
    ref = torch.empty(K, device="cuda")
    got = torch.empty(K, device="cuda")
    hip.check(db_call(L.scir_hip_fir_vjp_db_dev, g, x, K, ref, torch.cuda.current_stream()))
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):  # the stream's first launch (queue set-up) before measuring
        warm = x[:1, :1] * 1.0
    s.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    hip.check(db_call(L.scir_hip_fir_vjp_db_noalloc_dev, g, x, K, got, s))
    s.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, f"free device memory changed by {free0 - free1} bytes"
    assert torch.equal(got, ref)
    s2 = torch.cuda.Stream()
    with pytest.raises(hip.HipError) as e:
        with Graph(s2):
            hip.check(db_call(L.scir_hip_fir_vjp_db_dev, g, x, K, got, s2))
    assert e.value.status == hip.ERR_CAPTURE
    s3 = torch.cuda.Stream()
    with Graph(s3) as gr:
        hip.check(db_call(L.scir_hip_fir_vjp_db_noalloc_dev, g, x, K, got, s3))
    for rep in range(3):
        x.mul_(-0.75)
        hip.check(db_call(L.scir_hip_fir_vjp_db_dev, g, x, K, ref, torch.cuda.current_stream()))
        got.fill_(float("nan"))
        torch.cuda.synchronize()
        gr.launch()
        s3.synchronize()
        assert torch.equal(got, ref), f"replay {rep} differs"
    gr.close()
