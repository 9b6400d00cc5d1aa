"""GPU parity of the FP32 direct-form kernel through the C ABI.

* ``f32_exact`` must equal the oracle's restatement of the reference loop
  (crates/scir-gpu/src/lib.rs:1134-1152) bit for bit;
* ``f32`` must lie within (K+1)*2^-23*sum|b||x| of the FP64 oracle;
* full-size rows of the [4096, 65536] workload are checked on a row subset
  against the oracle and, over every row, through size-independent
  properties (exact-vs-FMA bound, run-to-run determinism).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -23


def host_fir(hip, x, b, precision):
    x = np.ascontiguousarray(x, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    B, N = x.shape
    y = np.full((B, N), np.nan, np.float32)
    hip.check(hip.lib().scir_hip_fir_f32_host(
        x.ctypes.data_as(ctypes.c_void_p), N, b.ctypes.data_as(ctypes.c_void_p), b.size,
        B, N, y.ctypes.data_as(ctypes.c_void_p), N, hip.precision_code(precision)))
    return y


@pytest.fixture(params=[(0, 16, 1, 3), (0, 8, 1, 3), (1, 16, 8, 3), (1, 8, 3, 3), (1, 8, 1, 3),
                        (2, 8, 8, 3), (2, 8, 1, 3), (2, 8, 2, 2), (2, 8, 5, 4), (2, 16, 3, 3),
                        (3, 8, 4, 2, 1), (3, 8, 4, 2, 3), (3, 8, 4, 2, 8), (3, 8, 1, 2, 64)],
                ids=lambda p: f"loader{p[0]}-r{p[1]}-seg{p[2]}-bufs{p[3]}" +
                              (f"-strip{p[4]}" if len(p) > 4 else ""))
def variant(request, hip):
    """Each direct-kernel variant: one tile per workgroup or pipelined row
    segments of `seg` tiles (register prefetch, or LDS DMA with `bufs` tile
    buffers), 8 or 16 outputs per lane; loader 3 runs F32 on FP32 MFMA in
    strips of `strip` 256-output tiles per wave (K <= 257; F32_EXACT and
    longer filters take loader 2's rules)."""
    loader, r, seg, bufs = request.param[:4]
    strip = request.param[4] if len(request.param) > 4 else -1
    L = hip.lib()
    hip.check(L.scir_hip_tune_set(b"direct.loader", loader))
    hip.check(L.scir_hip_tune_set(b"direct.r", r))
    hip.check(L.scir_hip_tune_set(b"direct.seg_tiles", seg))
    hip.check(L.scir_hip_tune_set(b"direct.dma_bufs", bufs))
    hip.check(L.scir_hip_tune_set(b"direct.mfma_strip", strip))
    yield request.param
    for key in (b"direct.loader", b"direct.r", b"direct.seg_tiles", b"direct.dma_bufs",
                b"direct.mfma_strip"):
        hip.check(L.scir_hip_tune_set(key, -1))


def test_reference_kat(hip):
    x = np.array([[1.0, 2.0, 3.0, 4.0], [0.5, 0.0, -0.5, -1.0]], np.float32)
    taps = np.array([0.25, 0.5, 0.25], np.float32)
    for prec in ("f32_exact", "f32"):
        y = host_fir(hip, x, taps[::-1], prec)
        np.testing.assert_array_equal(y, [[0.25, 1.0, 2.0, 3.0], [0.125, 0.25, 0.0, -0.5]])


K_CASES = [1, 2, 3, 4, 5, 7, 8, 15, 16, 17, 31, 32, 33, 34, 47, 63, 64, 65, 127, 129, 255,
           256, 511, 1023, 2047, 2048, 2049, 2051, 4097]
N_CASES = [1, 3, 100, 4095, 4096, 4097, 9001]


@pytest.mark.parametrize("K", K_CASES)
def test_exact_bitwise_vs_reference_loop(hip, oracle, variant, K):
    r = rng(100 + K)
    for N in N_CASES:
        B = 3 if N < 5000 else 2
        x = r.standard_normal((B, N)).astype(np.float32)
        taps = (1.0 / (np.arange(K) + 1.0)).astype(np.float32)  # fir_bench taps (asymmetric)
        y = host_fir(hip, x, taps[::-1], "f32_exact")
        ref = oracle.fir_ref_f32(x, taps)
        mism = np.flatnonzero((y != ref).ravel())
        assert mism.size == 0, f"K={K} N={N}: {mism.size} mismatches, first at {mism[:5]}"


@pytest.mark.parametrize("K", [1, 5, 33, 129, 255, 400, 513, 1000, 2048, 3001])
def test_f32_within_bound_of_fp64(hip, oracle, variant, K):
    r = rng(200 + K)
    for N in (17, 4096, 10007):
        x = r.standard_normal((3, N)).astype(np.float32)
        x[1, ::53] *= 1000.0
        b = r.standard_normal(K).astype(np.float32)
        y = host_fir(hip, x, b, "f32")
        y64 = oracle.lfilter_f64(x, b)
        bound = oracle.fir_error_bound(x, b)
        excess = np.abs(y - y64) - bound
        assert np.all(excess <= 0), f"K={K} N={N}: max excess {excess.max()}"


def test_cpu_config_workload(hip, oracle):
    """The reference's CPU case through the drop-in host tier: x [64, 16384]
    ~ N(0,1) f32 from a seed, the 33-tap Hamming low-pass (cutoff 0.2 of
    Nyquist, unit DC gain).  f32_exact equals the reference loop bit for bit;
    f32 and AUTO (whichever kernel a loaded tuning table makes it take) lie
    within (K+1)*2^-23*sum|b||x| of FP64."""
    from scir_amd import design
    x = rng(0x5C120000).standard_normal((64, 16384)).astype(np.float32)
    b = design.lowpass(33, 0.2).astype(np.float32)
    np.testing.assert_array_equal(host_fir(hip, x, b, "f32_exact"),
                                  oracle.fir_ref_f32(x, b[::-1]))
    y64, bound = oracle.lfilter_f64(x, b), oracle.fir_error_bound(x, b)
    for prec in ("f32", "auto"):
        assert np.all(np.abs(host_fir(hip, x, b, prec) - y64) <= bound), prec


def test_k0_zero_and_empty_shapes(hip):
    x = rng(1).standard_normal((3, 50)).astype(np.float32)
    y = host_fir(hip, x, np.zeros(0, np.float32), "f32")
    assert not y.any()
    L = hip.lib()
    p = x.ctypes.data_as(ctypes.c_void_p)
    assert L.scir_hip_fir_f32_host(p, 50, p, 3, 0, 50, p, 50, 0) == hip.OK
    assert L.scir_hip_fir_f32_host(p, 50, p, 3, 3, 0, p, 50, 0) == hip.OK


def test_device_tier_strided_and_unaligned(hip, oracle, variant):
    import torch
    from scir_amd import device
    r = rng(3)
    K, B, N = 37, 5, 5000
    base = torch.tensor(r.standard_normal((B, N + 7)).astype(np.float32), device="cuda")
    x = base[:, 1:N + 1]           # ldx = N+7 (not a multiple of 4), base offset 1 float
    b_np = r.standard_normal(K).astype(np.float32)
    b = torch.tensor(b_np, device="cuda")
    out_base = torch.full((B, N + 3), float("nan"), device="cuda")
    out = out_base[:, 2:N + 2]
    device.fir(x, b, "f32_exact", out=out)
    torch.cuda.synchronize()
    ref = oracle.fir_ref_f32(x.cpu().numpy(), b_np[::-1])
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    ob = out_base.cpu().numpy()
    assert np.isnan(ob[:, :2]).all() and np.isnan(ob[:, N + 2:]).all()  # nothing outside y


def test_run_to_run_deterministic(hip):
    import torch
    from scir_amd import device
    x = device.synth_normal(64, 70000, seed=5)
    b = torch.tensor(rng(4).standard_normal(129).astype(np.float32), device="cuda")
    y1 = device.fir(x, b, "f32")
    y2 = device.fir(x, b, "f32")
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("K", [33, 129])
def test_full_size_bandwidth_workload(hip, oracle, K):
    """x [4096, 65536] ~ N(0,1), Hamming low-pass 0.2 (K = 33, 129)."""
    import torch
    from scir_amd import design, device
    B, N = 4096, 65536
    print(f"full-size direct K={K} B={B} N={N}", flush=True)
    x = device.synth_normal(B, N, seed=0x5C120001)
    b_np = design.lowpass(K)
    b = torch.tensor(b_np, device="cuda")
    y = device.fir(x, b, "f32")
    ye = device.fir(x, b, "f32_exact")
    # |b|*|x| on the GPU bounds both results over every row
    s = device.fir(x.abs(), b.abs(), "f32")
    torch.cuda.synchronize()
    lim = 2.0 * (K + 1) * EPS * s * 1.001 + 1e-30
    assert bool(((y - ye).abs() <= lim).all())
    rows = [0, 1, 777, 2048, 4095]
    xs = x[rows].cpu().numpy()
    np.testing.assert_array_equal(ye[rows].cpu().numpy(), oracle.fir_ref_f32(xs, b_np[::-1]))
    y64 = oracle.lfilter_f64(xs, b_np)
    assert np.all(np.abs(y[rows].cpu().numpy() - y64) <= oracle.fir_error_bound(xs, b_np))


@pytest.mark.parametrize("K", [3, 21, 300, 1023])
def test_signal_forwarder_hip_device(hip, oracle, K):
    """scir_signal::gpu::fir1d_batched_f32 (crates/scir-signal/src/lib.rs:371-374,
    its test :670-696) on Device::Hip: the forwarder reaches the HIP arm (AUTO:
    the direct form for short filters, BF16x3 on MFMA for long ones) and stays
    within the FP32 bound of the reference loop; the reference's 2x4 case
    gives its known answer exactly."""
    from scir_amd import Device, signal
    x = np.array([[1.0, 2.0, 3.0, 4.0], [0.5, 0.0, -0.5, -1.0]], np.float32)
    kat = signal.fir1d_batched_f32(x, np.array([0.25, 0.5, 0.25], np.float32), Device.Hip)
    np.testing.assert_array_equal(kat, [[0.25, 1.0, 2.0, 3.0], [0.125, 0.25, 0.0, -0.5]])
    r = rng(900 + K)
    x = r.standard_normal((5, 6000)).astype(np.float32)
    taps = (r.standard_normal(K) / np.sqrt(K)).astype(np.float32)
    y = signal.fir1d_batched_f32(x, taps, Device.Hip)
    assert y.shape == x.shape and y.dtype == np.float32
    y64 = oracle.fir_ref_f64(x.astype(np.float64), taps.astype(np.float64))
    assert np.all(np.abs(y - y64) <= oracle.fir_error_bound(x, taps[::-1]))
    y_cpu = signal.fir1d_batched_f32(x, taps, Device.Cpu)
    np.testing.assert_array_equal(y_cpu, oracle.fir_ref_f32(x, taps))


def test_auto_dispatch_hip_device(hip, oracle):
    import scir_amd
    r = rng(8)
    x = r.standard_normal((4, 3000)).astype(np.float32)
    taps = r.standard_normal(21).astype(np.float32)
    y = scir_amd.fir1d_batched_f32_auto(x, taps, scir_amd.Device.Hip)
    y64 = oracle.fir_ref_f64(x.astype(np.float64), taps.astype(np.float64))
    assert np.all(np.abs(y - y64) <= oracle.fir_error_bound(x, taps[::-1]))


@pytest.mark.parametrize("precision", ["f32", "f32_exact", "bf16x3", "fp16", "fft32"])
@pytest.mark.parametrize("ox,oy", [(1, 2), (3, 1), (2, 3), (1, 0)])
def test_unaligned_rows_equal_aligned(hip, precision, ox, oy):
    """Rows at any 4-byte offset and odd pitch give the same bits as the same
    rows stored aligned and packed (for the FP32 forms: the register-prefetch
    kernel, which takes unaligned rows, against the LDS-DMA kernel, which
    takes aligned ones), and nothing outside y is written."""
    import torch
    from scir_amd import device
    B, N, K = 6, 10000, 129
    print(f"unaligned rows {precision} offsets x={ox} y={oy}", flush=True)
    r = rng(ox * 10 + oy)
    xa = torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda")
    b = torch.tensor((r.standard_normal(K) / np.sqrt(K)).astype(np.float32), device="cuda")
    ya = device.fir(xa, b, precision)
    xbase = torch.full((B, N + 5), float("nan"), device="cuda")
    xbase[:, ox:ox + N] = xa
    ybase = torch.full((B, N + 7), float("nan"), device="cuda")
    yu = ybase[:, oy:oy + N]
    device.fir(xbase[:, ox:ox + N], b, precision, out=yu)
    torch.cuda.synchronize()
    assert torch.equal(ya, yu), precision
    yb = ybase.cpu().numpy()
    assert np.isnan(yb[:, :oy]).all() and np.isnan(yb[:, oy + N:]).all()


@pytest.mark.parametrize("strip", [1, 3, 8])
@pytest.mark.parametrize("K", [16, 17, 240, 241, 256, 257, 258, 497, 498, 513])
def test_f32_mfma_boundaries_both_directions(hip, oracle, strip, K):
    """F32 on FP32 MFMA (K <= 513) where the Toeplitz block gains a column
    group (NG = ceil((K-1)/16) + 1 steps up after K = 16, 240, 257 and 497;
    K = 257 is the last tap count whose history is one tile back, from 498 the
    step J = 32 reads the tile two back, 513 is the largest), in strips of 1,
    3 and 8 tiles, on rows that end inside a tile, on one and past a strip:
    forward, and back to front as dL/dx uses it (dx[m] = sum_k b[k] g[m+k]),
    both within (K+1) 2^-23 sum|b||x| of FP64."""
    import torch
    from scir_amd import device
    L = hip.lib()
    hip.check(L.scir_hip_tune_set(b"direct.mfma_strip", strip))
    try:
        for N in (3, 255, 256, 257, 2049, 6001):
            print(f"f32 mfma K={K} strip={strip} N={N}", flush=True)
            r = rng(K * 31 + N + strip)
            x = r.standard_normal((3, N)).astype(np.float32)
            x[1, ::41] *= 1000.0
            b = r.standard_normal(K).astype(np.float32)
            xt, bt = torch.tensor(x, device="cuda"), torch.tensor(b, device="cuda")
            y = device.fir(xt, bt, "f32").cpu().numpy()
            assert np.all(np.abs(y - oracle.lfilter_f64(x, b)) <= oracle.fir_error_bound(x, b)), N
            dx = device.vjp_dx(xt, bt, "f32").cpu().numpy()
            xr = np.ascontiguousarray(x[:, ::-1])
            dx64 = oracle.lfilter_f64(xr, b)[:, ::-1]
            bound = oracle.fir_error_bound(xr, b)[:, ::-1]
            assert np.all(np.abs(dx - dx64) <= bound), f"dL/dx N={N}"
    finally:
        hip.check(L.scir_hip_tune_set(b"direct.mfma_strip", -1))
