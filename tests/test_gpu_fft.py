"""FP32 FFT overlap-save FIR (precision "fft32") against the float64 oracle.

The FFT's rounding is normwise: over an NF-sample segment the output error is
about log2(NF) * 2^-24 * |x_seg|_2 * |b|_1.  The test bound per output is
c * log2(NF) * 2^-24 * |b|_1 * |x|_2 over the row (a superset of every
segment, so the bound is looser than the per-segment one) with c = 4, and the
relative L2 error of each case with rows of at least 4K samples must stay
under 1e-6 (the accuracy bar the long-filter path is held to).  Filters
longer than 8192 taps run as blocks of 8192 taps (partitioned overlap-save),
whose outputs are summed; the same bound holds."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from scir_amd import design, device
from scir_amd.gpu import fir1d_batched_f32_hip

pytestmark = pytest.mark.gpu
EPS32 = 2.0 ** -24


def _nf(K):
    return 4096 if 4 * K <= 4096 else (8192 if 4 * K <= 8192 else 16384)


def _check(y, x, b):
    y64 = orc.lfilter_f64(x, b)
    nf = _nf(len(b))
    bound = 4 * np.log2(nf) * EPS32 * np.sum(np.abs(b)) * np.linalg.norm(x, axis=1, keepdims=True)
    err = np.abs(y.astype(np.float64) - y64)
    assert np.all(err <= bound + 1e-30), float(np.max(err / (bound + 1e-30)))
    rel = orc.rel_l2(y, y64)
    if x.shape[1] >= 4 * len(b):
        # (on rows much shorter than the filter most outputs see only a few
        # taps and are small next to the energy the normwise error scales with)
        assert rel <= 1e-6, rel
    return rel


@pytest.mark.parametrize("K", [1, 2, 17, 512, 1024, 1025, 2048, 2049, 4096, 8192])
@pytest.mark.parametrize("shape", [(3, 1), (2, 700), (5, 20001), (1, 100003)])
def test_fft_matches_oracle(K, shape):
    print(f"fft K={K} shape={shape}", flush=True)
    rng = np.random.default_rng(K + shape[1])
    x = rng.standard_normal(shape).astype(np.float32)
    b = (rng.standard_normal(K) / np.sqrt(K)).astype(np.float32)
    y = device.fir(torch.tensor(x, device="cuda:0"), torch.tensor(b, device="cuda:0"), "fft32")
    torch.cuda.synchronize()
    _check(y.cpu().numpy(), x, b)


def test_fft_long_taps_with_outliers():
    # the long-filter workload's arithmetic: Kaiser band-pass, 1% outliers x1e3
    for K in (512, 2048):
        print(f"fft outliers K={K}", flush=True)
        x = device.synth_normal(4, 50000, seed=K, outlier_frac=0.01, outlier_scale=1000.0,
                                device="cuda:0")
        b_np = design.bandpass_kaiser(K)
        y = device.fir(x, torch.tensor(b_np, device="cuda:0"), "fft32").cpu().numpy()
        rel = _check(y, x.cpu().numpy(), b_np)
        print(f"  rel-L2 {rel:.3g}")


def test_fft_strided_rows_and_host_tier():
    rng = np.random.default_rng(3)
    big = torch.tensor(rng.standard_normal((6, 9001)).astype(np.float32), device="cuda:0")
    x = big[:, :9000]
    b_np = (rng.standard_normal(777) / 30).astype(np.float32)
    y = device.fir(x, torch.tensor(b_np, device="cuda:0"), "fft32")
    torch.cuda.synchronize()
    _check(y.cpu().numpy(), x.cpu().numpy(), b_np)
    xh = rng.standard_normal((3, 5000)).astype(np.float32)
    taps = rng.standard_normal(300).astype(np.float32) / 17
    yh = fir1d_batched_f32_hip(xh, taps, "fft32")   # reference-order taps
    _check(yh, xh, taps[::-1].copy())


def test_fft_adjoint_dx():
    # dL/dx = the anti-causal correlation, run back to front through the FFT
    rng = np.random.default_rng(4)
    B, N, K = 3, 30000, 1500
    g = rng.standard_normal((B, N)).astype(np.float32)
    x = rng.standard_normal((B, N)).astype(np.float32)
    b = (rng.standard_normal(K) / 40).astype(np.float32)
    dx = device.vjp_dx(torch.tensor(g, device="cuda:0"), torch.tensor(b, device="cuda:0"),
                       "fft32").cpu().numpy()
    dx64, _ = orc.vjp_f64(g, x, b)
    nf = _nf(K)
    bound = 4 * np.log2(nf) * EPS32 * np.sum(np.abs(b)) * np.linalg.norm(g, axis=1, keepdims=True)
    assert np.all(np.abs(dx - dx64) <= bound)
    assert orc.rel_l2(dx, dx64) <= 1e-6


@pytest.mark.parametrize("K", [8193, 12000, 16384, 20001])
@pytest.mark.parametrize("shape", [(2, 5000), (3, 40000), (1, 100003)])
def test_fft_partitioned_long_filters(K, shape):
    """K > 8192: blocks of 8192 taps, each on the input delayed by its first
    tap, summed into y; rows shorter than a block offset skip that block."""
    print(f"fft partitioned K={K} shape={shape}", flush=True)
    rng = np.random.default_rng(K + 7 * shape[1])
    x = rng.standard_normal(shape).astype(np.float32)
    b = (rng.standard_normal(K) / np.sqrt(K)).astype(np.float32)
    y = device.fir(torch.tensor(x, device="cuda:0"), torch.tensor(b, device="cuda:0"), "fft32")
    torch.cuda.synchronize()
    _check(y.cpu().numpy(), x, b)


def test_fft_partitioned_adjoint_dx():
    # the back-to-front run of a partitioned filter (dL/dx with K > 8192)
    rng = np.random.default_rng(5)
    B, N, K = 2, 60000, 10000
    g = rng.standard_normal((B, N)).astype(np.float32)
    x = rng.standard_normal((B, N)).astype(np.float32)
    b = (rng.standard_normal(K) / 100).astype(np.float32)
    dx = device.vjp_dx(torch.tensor(g, device="cuda:0"), torch.tensor(b, device="cuda:0"),
                       "fft32").cpu().numpy()
    dx64, _ = orc.vjp_f64(g, x, b)
    bound = 4 * np.log2(16384) * EPS32 * np.sum(np.abs(b)) * np.linalg.norm(g, axis=1,
                                                                           keepdims=True)
    assert np.all(np.abs(dx - dx64) <= bound)
    assert orc.rel_l2(dx, dx64) <= 1e-6


def test_fft_refusals():
    x = torch.zeros((1, 100), device="cuda:0")
    with pytest.raises(Exception):
        device.fir(x, torch.ones(5, device="cuda:0"), "fft32", out=x)
    # K = 0 gives zeros like every path
    y = device.fir(torch.ones((2, 50), device="cuda:0"), torch.zeros(0, device="cuda:0"), "fft32")
    assert float(y.abs().max()) == 0.0
