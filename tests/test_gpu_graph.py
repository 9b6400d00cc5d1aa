"""HIP-graph capture of the dense device tier (every precision, FFT32
included), the adjoints, sosfilt / filtfilt and the resampler.

The library's kernels take every buffer as an argument and clear memory
with kernels (never hipMemsetAsync), so a captured sequence replays the
eager results exactly.  Each replay here follows eager work and a
synchronisation, the inputs change between replays (the graph reads them
in place), and every replay is compared bit for bit with an eager call on
the same inputs.  A buffer that would have to grow while a stream captures
is refused with SCIR_HIP_ERR_CAPTURE rather than allocated.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,K", [("f32", 33), ("f32", 400), ("f32_exact", 33),
                                         ("bf16x3", 512), ("fp16", 512), ("auto", 129),
                                         ("fft32", 700)])
def test_dense_graph_replays_match_eager(hip, precision, K):
    import torch
    from scir_amd import device
    from scir_amd.ragged import Graph
    r = rng(K + len(precision))
    B, N = 6, 50000
    x = torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda")
    g = torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda")
    b = torch.tensor(r.standard_normal(K).astype(np.float32), device="cuda")
    s = torch.cuda.Stream()
    y, dx, db = torch.empty_like(x), torch.empty_like(x), torch.empty(K, device="cuda")
    # Eager calls on the capture stream first: the per-stream dL/db partials
    # are allocated here, outside the capture.
    with torch.cuda.stream(s):
        device.fir(x, b, precision, out=y)
        device.vjp_dx(g, b, precision, out=dx)
        device.vjp_db(g, x, K, out=db)
    s.synchronize()
    torch.cuda.synchronize()
    with Graph(s) as gr:
        device.fir(x, b, precision, out=y)
        device.vjp_dx(g, b, precision, out=dx)
        device.vjp_db(g, x, K, out=db)
    for rep in range(3):
        x.copy_(torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda"))
        g.mul_(-0.5)
        ye = device.fir(x, b, precision)
        dxe = device.vjp_dx(g, b, precision)
        dbe = device.vjp_db(g, x, K)
        for t in (y, dx, db):
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            gr.launch()
        s.synchronize()
        assert torch.equal(y, ye), f"{precision}: forward replay {rep} differs"
        assert torch.equal(dx, dxe), f"{precision}: dL/dx replay {rep} differs"
        assert torch.equal(db, dbe), f"{precision}: dL/db replay {rep} differs"
    gr.close()


def test_growth_during_capture_is_refused(hip):
    """dL/db on a fresh stream needs its partials buffer: inside a capture
    that is an error (SCIR_HIP_ERR_CAPTURE), not a hidden allocation."""
    import torch
    from scir_amd import _lib, device
    from scir_amd.ragged import Graph
    x = torch.ones((2, 70000), device="cuda")
    device.vjp_db(x, x, 64)  # eager, on the current stream
    s = torch.cuda.Stream()  # fresh stream: no partials buffer yet
    torch.cuda.synchronize()
    with pytest.raises(_lib.HipError) as ei:
        with Graph(s):
            device.vjp_db(x, x, 64)
    assert "captur" in str(ei.value).lower()
    torch.cuda.synchronize()


def test_sosfilt_and_resample_graph_replays(hip):
    """The float64 paths under capture: sosfilt (setup, chunk-state, scan and
    filtering launches, few rows so that rows split into chunks), filtfilt in
    place, and the resampler replay their eager results bit for bit."""
    import ctypes
    import torch
    from scipy import signal as ss
    from scir_amd import _lib, device, signal
    from scir_amd.ragged import Graph
    r = rng(91)
    rows, n = 3, 300000
    x = torch.tensor(r.standard_normal((rows, n)), device="cuda")
    sos = torch.tensor(ss.butter(6, 0.15, output="sos"), device="cuda")
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    h = torch.tensor(signal.RESAMPLE_H_2_3, device="cuda")
    m = signal.resample_len(n, 2, 3, h.numel())
    w = torch.empty((rows, m), dtype=torch.float64, device="cuda")
    L = _lib.lib()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    s = torch.cuda.Stream()

    def work():
        device.sosfilt(x, sos, out=y)
        z.copy_(x)
        device.sosfilt(z, sos, zero_phase=True, out=z)
        _lib.check(L.scir_hip_resample_poly_f64_dev(
            vp(x), n, rows, n, 2, 3, vp(h), h.numel(), vp(w), m,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    with torch.cuda.stream(s):
        work()  # eager first: the chunk-state scratch is allocated here
    s.synchronize()
    torch.cuda.synchronize()
    with Graph(s) as gr:
        with torch.cuda.stream(s):
            work()
    for rep in range(3):
        x.copy_(torch.tensor(r.standard_normal((rows, n)), device="cuda"))
        work()
        ye, ze, we = y.clone(), z.clone(), w.clone()
        for t in (y, z, w):
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            gr.launch()
        s.synchronize()
        assert torch.equal(y, ye), f"sosfilt replay {rep} differs"
        assert torch.equal(z, ze), f"filtfilt replay {rep} differs"
        assert torch.equal(w, we), f"resample replay {rep} differs"
    gr.close()


def test_graph_survives_scratch_growth_on_its_stream(hip):
    """A graph holds the addresses of the stream's scratch (the FFT filter
    spectrum, the dL/db partials).  An eager call with a longer filter on the
    same stream grows that scratch; the library keeps the outgrown blocks, so
    later replays still read and write valid memory and equal eager results
    bit for bit."""
    import torch
    from scir_amd import device
    from scir_amd.ragged import Graph
    r = rng(4242)
    B, N, K = 4, 60000, 512
    x = torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda")
    g = torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda")
    b = torch.tensor(r.standard_normal(K).astype(np.float32), device="cuda")
    s = torch.cuda.Stream()
    y, db = torch.empty_like(x), torch.empty(K, device="cuda")
    with torch.cuda.stream(s):  # scratch sized for K = 512 outside the capture
        device.fir(x, b, "fft32", out=y)
        device.vjp_db(g, x, K, out=db)
    s.synchronize()
    with Graph(s) as gr:
        device.fir(x, b, "fft32", out=y)
        device.vjp_db(g, x, K, out=db)
    # grow both scratch blocks on the capture stream (K = 2048: twice the
    # spectrum, four times the partials)
    K2 = 2048
    b2 = torch.tensor(r.standard_normal(K2).astype(np.float32), device="cuda")
    with torch.cuda.stream(s):
        y2 = device.fir(x, b2, "fft32")
        db2 = device.vjp_db(g, x, K2)
    s.synchronize()
    for rep in range(3):
        x.mul_(-1.25)
        ye = device.fir(x, b, "fft32")
        dbe = device.vjp_db(g, x, K)
        y.fill_(float("nan"))
        db.fill_(float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            gr.launch()
            # the larger eager calls between replays keep using the new blocks
            y2b = device.fir(x, b2, "fft32")
            db2b = device.vjp_db(g, x, K2)
        s.synchronize()
        assert torch.equal(y, ye), f"forward replay {rep} differs after scratch growth"
        assert torch.equal(db, dbe), f"dL/db replay {rep} differs after scratch growth"
        assert torch.equal(y2b, device.fir(x, b2, "fft32"))
        assert torch.equal(db2b, device.vjp_db(g, x, K2))
    gr.close()
    del y2, db2
