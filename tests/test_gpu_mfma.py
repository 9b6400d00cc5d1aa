"""GPU parity of the Toeplitz-block MFMA kernels (BF16x3 and FP16).

BF16x3 keeps x = hi+mid+lo exactly and drops only the mid*lo, lo*mid and
lo*lo products (each below 2^-24 of |b_k x_{n-k}|), so its error is held to
the same (K+1)*2^-23*sum|b||x| bound as the FP32 direct form.  FP16 rounds x
and b to 11 significant bits: its bound adds 2^-10*sum|b||x|.
Full-size [2048, 262144] rows with 1% of samples scaled by 1e3 must meet
rel-L2 <= 1e-6 against the FP64 oracle (BF16x3).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -23
K_CASES = [1, 2, 15, 16, 17, 31, 33, 64, 100, 129, 255, 512, 1000, 2048, 4097]
N_CASES = [1, 31, 100, 4095, 4097, 16385, 20000]


def host_fir(hip, x, b, precision):
    x = np.ascontiguousarray(x, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    B, N = x.shape
    y = np.full((B, N), np.nan, np.float32)
    hip.check(hip.lib().scir_hip_fir_f32_host(
        x.ctypes.data_as(ctypes.c_void_p), N, b.ctypes.data_as(ctypes.c_void_p), b.size,
        B, N, y.ctypes.data_as(ctypes.c_void_p), N, hip.precision_code(precision)))
    return y


@pytest.fixture(params=[1, 2, 4], ids=lambda t: f"tiles{t}")
def tiles(request, hip):
    L = hip.lib()
    hip.check(L.scir_hip_tune_set(b"mfma.tiles", request.param))
    yield request.param
    hip.check(L.scir_hip_tune_set(b"mfma.tiles", -1))


@pytest.mark.parametrize("K", K_CASES)
def test_bf16x3_within_fp32_bound(hip, oracle, tiles, K):
    r = rng(300 + K)
    worst = 0.0
    for N in N_CASES:
        x = r.standard_normal((2, N)).astype(np.float32)
        x[1, ::41] *= 1000.0
        b = r.standard_normal(K).astype(np.float32)
        y = host_fir(hip, x, b, "bf16x3")
        y64 = oracle.lfilter_f64(x, b)
        ratio = np.abs(y - y64) / (oracle.fir_error_bound(x, b) + 1e-300)
        worst = max(worst, float(ratio.max()))
        assert np.all(ratio <= 1.0), f"K={K} N={N} tiles={tiles}: err/bound {ratio.max():.3g}"
    print(f"K={K} tiles={tiles}: max err/bound {worst:.3g}")


@pytest.mark.parametrize("K", [1, 17, 129, 512, 2048])
def test_fp16_within_rounding_bound(hip, oracle, K):
    r = rng(400 + K)
    for N in (100, 4097, 20000):
        x = r.standard_normal((2, N)).astype(np.float32)
        b = r.standard_normal(K).astype(np.float32)
        y = host_fir(hip, x, b, "fp16")
        y64 = oracle.lfilter_f64(x, b)
        bound = (2.0 ** -10 + (K + 2) * EPS) * oracle.abs_bound(x, b)
        assert np.all(np.abs(y - y64) <= bound), f"K={K} N={N}"


def test_strided_unaligned_rows(hip, oracle):
    import torch
    from scir_amd import device
    r = rng(5)
    K, B, N = 300, 3, 9001
    base = torch.tensor(r.standard_normal((B, N + 9)).astype(np.float32), device="cuda")
    x = base[:, 3:N + 3]
    b_np = r.standard_normal(K).astype(np.float32)
    b = torch.tensor(b_np, device="cuda")
    ob = torch.full((B, N + 5), float("nan"), device="cuda")
    out = ob[:, 1:N + 1]
    device.fir(x, b, "bf16x3", out=out)
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    y64 = oracle.lfilter_f64(xs, b_np)
    assert np.all(np.abs(out.cpu().numpy() - y64) <= oracle.fir_error_bound(xs, b_np))
    o = ob.cpu().numpy()
    assert np.isnan(o[:, :1]).all() and np.isnan(o[:, N + 1:]).all()


def test_deterministic(hip):
    import torch
    from scir_amd import device
    x = device.synth_normal(16, 100000, seed=9)
    b = torch.tensor(rng(9).standard_normal(777).astype(np.float32), device="cuda")
    y1 = device.fir(x, b, "bf16x3")
    y2 = device.fir(x, b, "bf16x3")
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("K", [512, 2048])
def test_full_size_long_taps_rel_l2(hip, oracle, K):
    """x [2048, 262144] ~ N(0,1) with 1% of samples x1e3, Kaiser(8) band-pass
    0.1-0.3: BF16x3 rel-L2 <= 1e-6 vs FP64 (row subset), FP16 reported.
    Over every row, BF16x3 stays within 2*(K+1)*2^-23*sum|b||x| of the
    reference loop (f32_exact), with sum|b||x| computed on the GPU."""
    import torch
    from scir_amd import design, device
    B, N = 2048, 262144
    print(f"full-size mfma K={K} B={B} N={N}", flush=True)
    x = device.synth_normal(B, N, seed=0x5C120002, outlier_frac=0.01, outlier_scale=1000.0)
    b_np = design.bandpass_kaiser(K)
    b = torch.tensor(b_np, device="cuda")
    y3 = device.fir(x, b, "bf16x3")
    y16 = device.fir(x, b, "fp16")
    ye = device.fir(x, b, "f32_exact")
    sabs = device.fir(x.abs(), b.abs(), "f32")
    torch.cuda.synchronize()
    # every row: |bf16x3 - reference loop| <= 2 (K+1) 2^-23 sum|b||x| (each
    # side within (K+1) 2^-23 sum|b||x| of the exact sum; 0.1 % slack for
    # the rounding of sabs itself)
    lim = 2.0 * (K + 1) * EPS * sabs * 1.001 + 1e-30
    ratio = ((y3 - ye).abs() / lim).max().item()
    print(f"K={K}: every-row max error / bound {ratio:.3g} (<= 1 passes)")
    assert ratio <= 1.0
    del ye, sabs
    rows = [0, 1023, 2047]
    xs = x[rows].cpu().numpy()
    y64 = oracle.lfilter_f64(xs, b_np)
    rel3 = oracle.rel_l2(y3[rows].cpu().numpy(), y64)
    rel16 = oracle.rel_l2(y16[rows].cpu().numpy(), y64)
    print(f"K={K}: rel-L2 bf16x3 {rel3:.3g}, fp16 {rel16:.3g}")
    assert rel3 <= 1e-6
    assert np.all(np.abs(y3[rows].cpu().numpy() - y64) <= oracle.fir_error_bound(xs, b_np))


@pytest.fixture(params=[2, 5, 64], ids=lambda s: f"segs{s}")
def segs(request, hip):
    L = hip.lib()
    hip.check(L.scir_hip_tune_set(b"mfma.segs", request.param))
    yield request.param
    hip.check(L.scir_hip_tune_set(b"mfma.segs", -1))


@pytest.mark.parametrize("K", [17, 512, 2048, 4097])
def test_segments_per_workgroup(hip, oracle, tiles, segs, K):
    """One workgroup walking several row segments (next segment's samples
    prefetched during the current one's MFMAs): segments that cross row
    boundaries, a partial last segment, and K > 2049 (history longer than
    the prefetched part), forward and anti-causal (dL/dx)."""
    import torch
    from scir_amd import device
    r = rng(700 + K + segs)
    B, N = 7, 20000
    x = r.standard_normal((B, N)).astype(np.float32)
    x[3, ::37] *= 1000.0
    b = r.standard_normal(K).astype(np.float32)
    y64 = oracle.lfilter_f64(x, b)
    for prec in ("bf16x3", "fp16"):
        y = host_fir(hip, x, b, prec)
        if prec == "bf16x3":
            bound = oracle.fir_error_bound(x, b)
        else:
            bound = (2.0 ** -10 + (K + 2) * EPS) * oracle.abs_bound(x, b)
        assert np.all(np.abs(y - y64) <= bound), f"{prec} K={K} segs={segs} tiles={tiles}"
    w = r.standard_normal((B, N)).astype(np.float32)
    dx = device.vjp_dx(torch.tensor(w, device="cuda"), torch.tensor(b, device="cuda"), "bf16x3")
    torch.cuda.synchronize()
    dx64, _ = oracle.vjp_f64(w, x, b)
    dxa, _ = oracle.vjp_f64(np.abs(w), np.abs(x), np.abs(b))
    assert np.all(np.abs(dx.cpu().numpy() - dx64) <= (K + 1) * EPS * dxa)



def test_tap_limit_and_auto_beyond_it(hip, oracle):
    """BF16x3 takes K up to 6689 (the taps' tables and one tile per wave in
    LDS); beyond it the MFMA precision is refused and AUTO takes the FP32
    direct form, within the same bound."""
    r = rng(31)
    x = r.standard_normal((2, 20000)).astype(np.float32)
    for K, mfma_ok in ((6689, True), (6690, False), (9000, False)):
        b = (r.standard_normal(K) / np.sqrt(K)).astype(np.float32)
        y64, bound = oracle.lfilter_f64(x, b), oracle.fir_error_bound(x, b)
        if mfma_ok:
            assert np.all(np.abs(host_fir(hip, x, b, "bf16x3") - y64) <= bound)
        else:
            with pytest.raises(hip.HipError) as e:
                host_fir(hip, x, b, "bf16x3")
            assert e.value.status == hip.ERR_UNSUPPORTED
        assert np.all(np.abs(host_fir(hip, x, b, "auto") - y64) <= bound), K


@pytest.mark.parametrize("case", ["x_subnormal", "x_near_min", "b_tiny", "tiny_segment"])
def test_bf16x3_tiny_magnitudes_within_bound(hip, case):
    """Samples or taps below the range in which the bf16 split terms and the
    kept products stay normal (subnormal samples, samples near FLT_MIN, tiny
    taps, one tiny stretch inside normal data): the segments that stage them
    are recomputed from the reference loop's terms, so BF16x3 stays within
    2 (K+1) 2^-23 sum|b||x| (plus one subnormal unit per term, the reference
    loop's own underflow) of the reference loop (f32_exact), dense and ragged."""
    import torch
    from scir_amd import device
    from scir_amd.ragged import RaggedPlan
    print(f"bf16x3 tiny magnitudes {case}", flush=True)
    r = rng(len(case))
    B, N, K = 3, 20000, 200
    x = r.standard_normal((B, N)).astype(np.float32)
    b = (r.standard_normal(K) / np.sqrt(K)).astype(np.float32)
    if case == "x_subnormal":
        x *= np.float32(1e-39)
    elif case == "x_near_min":
        x *= np.float32(1e-36)
    elif case == "b_tiny":
        b *= np.float32(1e-38)
    else:
        x[1, 9000:9100] *= np.float32(1e-41)
    xt, bt = torch.tensor(x, device="cuda"), torch.tensor(b, device="cuda")
    # Compared on the host in float64 (which keeps every FP32 subnormal).
    ye = device.fir(xt, bt, "f32_exact").cpu().numpy().astype(np.float64)
    s = device.fir(xt.abs(), bt.abs(), "f32_exact").cpu().numpy().astype(np.float64)
    lim = 2.0 * (K + 1) * (EPS * s + 2.0 ** -149)
    y = device.fir(xt, bt, "bf16x3").cpu().numpy().astype(np.float64)
    ratio = np.max(np.abs(y - ye) / lim)
    assert ratio <= 1.0, f"dense {case}: max error / bound {ratio}"
    offs = np.array([0, N, 2 * N, 3 * N], dtype=np.int64)
    boffs = np.array([0, K, 2 * K, 3 * K], dtype=np.int64)
    plan = RaggedPlan(offs, boffs, "bf16x3")
    yr = plan.execute(xt.reshape(-1), bt.repeat(3)).cpu().numpy().astype(np.float64)
    plan.close()
    ratio = np.max(np.abs(yr - ye.reshape(-1)) / lim.reshape(-1))
    assert ratio <= 1.0, f"ragged {case}: max error / bound {ratio}"
