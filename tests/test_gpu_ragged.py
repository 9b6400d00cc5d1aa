"""GPU parity of the ragged (CSR) persistent kernel and its graph replay.

Reference: per-signal FP64 oracle (oracle.lfilter_ragged_f64) within the
(K_s+1)*2^-23*sum|b||x| bound; the f32_exact mode bit for bit against the
reference loop per signal; the persistent launch against one dense launch
per signal of the same arithmetic, the VALU direct form (loader 2; dense
F32 calls with K <= 257 otherwise run on FP32 MFMA): identical bits.  The MFMA precisions (bf16x3,
fp16) run a persistent Toeplitz contraction whose per-output arithmetic is
that of the dense MFMA kernel, so they too equal the per-signal dense launches
bit for bit; bf16x3 is held to relative L2 error <= 1e-6 against FP64."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu


def make_case(seed, S, max_len, max_k, allow_empty=True):
    r = rng(seed)
    lens = r.integers(0 if allow_empty else 1, max_len + 1, S)
    if S >= 4:
        lens[:4] = [1, 2047, 2048, 2049][: min(4, S)]
    ks = r.integers(0 if allow_empty else 1, max_k + 1, S)
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    boffs = np.concatenate(([0], np.cumsum(ks))).astype(np.int64)
    x = r.standard_normal(int(offs[-1])).astype(np.float32)
    b = r.standard_normal(max(int(boffs[-1]), 1)).astype(np.float32)
    return offs, boffs, x, b


class valu_direct_form:
    """Dense F32 calls on the VALU direct form (direct.loader 2), the
    arithmetic of the ragged persistent F32 kernel, while in effect."""

    def __enter__(self):
        from scir_amd import _lib
        _lib.check(_lib.lib().scir_hip_tune_set(b"direct.loader", 2))

    def __exit__(self, *exc):
        from scir_amd import _lib
        _lib.check(_lib.lib().scir_hip_tune_set(b"direct.loader", -1))


def run(offs, boffs, x, b, precision, per_signal=False):
    import torch
    from scir_amd.ragged import RaggedPlan
    plan = RaggedPlan(offs, boffs, precision)
    xt = torch.tensor(x, device="cuda")
    bt = torch.tensor(b, device="cuda")
    yt = torch.full_like(xt, float("nan"))
    (plan.execute_per_signal if per_signal else plan.execute)(xt, bt, yt)
    torch.cuda.synchronize()
    plan.close()
    return yt.cpu().numpy()


def ragged_bound(oracle, x, offs, b, boffs):
    out = np.zeros(x.size)
    for s in range(offs.size - 1):
        seg = slice(offs[s], offs[s + 1])
        k = b[boffs[s]:boffs[s + 1]]
        if offs[s + 1] > offs[s] and k.size:
            out[seg] = oracle.fir_error_bound(x[seg], k)[0]
    return out


@pytest.mark.parametrize("seed,S,max_len,max_k", [(1, 1, 10, 3), (2, 17, 5000, 40),
                                                  (3, 60, 20000, 1100), (4, 200, 3000, 300)])
def test_ragged_within_bound(hip, oracle, seed, S, max_len, max_k):
    offs, boffs, x, b = make_case(seed, S, max_len, max_k)
    y = run(offs, boffs, x, b, "f32")
    y64 = oracle.lfilter_ragged_f64(x, offs, b, boffs)
    assert np.all(np.abs(y - y64) <= ragged_bound(oracle, x, offs, b, boffs))


@pytest.mark.parametrize("seed", [5, 6])
def test_ragged_exact_bitwise(hip, oracle, seed):
    offs, boffs, x, b = make_case(seed, 40, 9000, 700)
    y = run(offs, boffs, x, b, "f32_exact")
    for s in range(offs.size - 1):
        seg = slice(offs[s], offs[s + 1])
        if offs[s + 1] == offs[s]:
            continue
        taps = b[boffs[s]:boffs[s + 1]][::-1]
        ref = oracle.fir_ref_f32(x[seg], taps)[0] if taps.size else np.zeros(offs[s + 1] - offs[s])
        np.testing.assert_array_equal(y[seg], ref)


def test_persistent_equals_per_signal(hip):
    offs, boffs, x, b = make_case(7, 80, 12000, 900)
    for prec in ("f32", "f32_exact"):
        with valu_direct_form():
            y_sig = run(offs, boffs, x, b, prec, per_signal=True)
        np.testing.assert_array_equal(run(offs, boffs, x, b, prec), y_sig)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16"])
@pytest.mark.parametrize("seed,S,max_len,max_k", [(11, 1, 10, 3), (12, 30, 20000, 40),
                                                  (13, 60, 40000, 1100), (14, 200, 3000, 300)])
def test_mfma_ragged_equals_dense_per_signal(hip, seed, S, max_len, max_k, precision):
    """Empty signals, K = 0, lengths 1 / 2047 / 2048 / 2049 and segments cut
    mid-item: the persistent MFMA launch equals one dense MFMA launch per
    signal, bit for bit."""
    offs, boffs, x, b = make_case(seed, S, max_len, max_k)
    y = run(offs, boffs, x, b, precision)
    assert not np.isnan(y).any()
    np.testing.assert_array_equal(y, run(offs, boffs, x, b, precision, per_signal=True))


@pytest.mark.parametrize("precision,tol", [("bf16x3", 1e-6), ("fp16", 2e-3)])
def test_mfma_ragged_rel_l2(hip, oracle, precision, tol):
    offs, boffs, x, b = make_case(15, 40, 30000, 1023, allow_empty=False)
    y = run(offs, boffs, x, b, precision)
    y64 = oracle.lfilter_ragged_f64(x, offs, b, boffs)
    assert oracle.rel_l2(y, y64) <= tol


def test_auto_precision_resolves_per_plan(hip, oracle):
    """AUTO on a ragged plan: BF16x3 MFMA when the batch's multiply-add
    weighted K reaches the cut-over (output equal to an explicit bf16x3
    plan), the FP32 direct form for short filters (equal to an f32 plan);
    both within the elementwise FP32 bound."""
    for seed, max_k, expect in ((17, 900, "bf16x3"), (18, 20, "f32")):
        offs, boffs, x, b = make_case(seed, 30, 20000, max_k, allow_empty=False)
        y = run(offs, boffs, x, b, "auto")
        np.testing.assert_array_equal(y, run(offs, boffs, x, b, expect))
        y64 = oracle.lfilter_ragged_f64(x, offs, b, boffs)
        assert np.all(np.abs(y - y64) <= ragged_bound(oracle, x, offs, b, boffs)), expect


def test_mfma_ragged_unaligned_signals(hip, oracle):
    """Signals starting at odd sample offsets take the scalar load/store path."""
    offs = np.array([0, 3, 9004, 9005, 30001], dtype=np.int64)
    boffs = np.array([0, 5, 600, 601, 1700], dtype=np.int64)
    r = rng(16)
    x = r.standard_normal(int(offs[-1])).astype(np.float32)
    b = r.standard_normal(int(boffs[-1])).astype(np.float32)
    y = run(offs, boffs, x, b, "bf16x3")
    np.testing.assert_array_equal(y, run(offs, boffs, x, b, "bf16x3", per_signal=True))
    assert oracle.rel_l2(y, oracle.lfilter_ragged_f64(x, offs, b, boffs)) <= 1e-6


def test_fft_precision_rejected_for_ragged(hip):
    from scir_amd import _lib
    from scir_amd.ragged import RaggedPlan
    with pytest.raises(_lib.HipError) as e:
        RaggedPlan(np.array([0, 5]), np.array([0, 2]), "fft32")
    assert e.value.status == _lib.ERR_UNSUPPORTED


def test_mfma_ragged_tap_limit(hip):
    """BF16x3 ragged plans take taps up to K = 5601 (the LDS of one CU)."""
    from scir_amd import _lib
    from scir_amd.ragged import RaggedPlan
    RaggedPlan(np.array([0, 10]), np.array([0, 5601]), "bf16x3").close()
    with pytest.raises(_lib.HipError) as e:
        RaggedPlan(np.array([0, 10]), np.array([0, 5602]), "bf16x3")
    assert e.value.status == _lib.ERR_UNSUPPORTED


def test_invalid_offsets_rejected(hip):
    from scir_amd import _lib
    from scir_amd.ragged import RaggedPlan
    with pytest.raises(_lib.HipError):
        RaggedPlan(np.array([0, 5, 3]), np.array([0, 1, 2]))
    with pytest.raises(_lib.HipError):
        RaggedPlan(np.array([1, 5]), np.array([0, 1]))


@pytest.mark.parametrize("precision", ["f32", "bf16x3"])
def test_graph_replay_matches_eager(hip, precision):
    """Capture reset + persistent launch; replay several times with eager
    work and synchronisation between replays; every replay must equal the
    eager result bit for bit."""
    import torch
    from scir_amd.ragged import Graph, RaggedPlan
    offs, boffs, x, b = make_case(8, 50, 10000, 500, allow_empty=False)
    plan = RaggedPlan(offs, boffs, precision)
    xt = torch.tensor(x, device="cuda")
    bt = torch.tensor(b, device="cuda")
    eager = plan.execute(xt, bt)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    yg = torch.zeros_like(xt)
    torch.cuda.synchronize()
    with Graph(s) as g:
        plan.execute(xt, bt, yg)
    for rep in range(4):
        yg.fill_(float("nan"))
        torch.cuda.synchronize()
        other = torch.randn(1000, device="cuda").sum()  # eager work between replays
        with torch.cuda.stream(s):
            g.launch()
        s.synchronize()
        assert torch.equal(yg, eager), f"replay {rep} differs"
        float(other)
    g.close()
    plan.close()


def test_full_size_ragged_workload(hip, oracle):
    """10000 signals, lengths log-uniform in [256, 2^20], K in [15, 1023],
    windowed-sinc low-passes: a subset of signals vs the FP64 oracle, the
    persistent launch vs per-signal launches over every signal, and every
    sample of every signal (FMA direct form and BF16x3) within
    2 (K_s + 1) 2^-23 sum|b||x| of the reference loop (f32_exact), with
    sum|b||x| computed on the GPU."""
    import torch
    from scir_amd import configs, device
    from scir_amd.ragged import RaggedPlan
    offs, boffs, taps = configs.ragged_layout()
    total = int(offs[-1])
    print(f"full-size ragged S={offs.size - 1} samples={total}", flush=True)
    x = device.synth_normal(1, total, seed=configs.SEED_BASE + 3).view(-1)
    bt = torch.tensor(taps, device="cuda")
    plan = RaggedPlan(offs, boffs, "f32")
    y = plan.execute(x, bt)
    with valu_direct_form():
        y2 = plan.execute_per_signal(x, bt)
    sabs = plan.execute(x.abs(), bt.abs())
    torch.cuda.synchronize()
    assert torch.equal(y, y2)
    del y2
    pe = RaggedPlan(offs, boffs, "f32_exact")
    ye = pe.execute(x, bt)
    pe.close()
    kp1 = torch.repeat_interleave(
        torch.tensor(np.diff(boffs) + 1, dtype=torch.float32, device="cuda"),
        torch.tensor(np.diff(offs), device="cuda"))
    # 0.1 % slack for the rounding of sabs itself
    lim = 2.0 * kp1 * 2.0 ** -23 * sabs * 1.001 + 1e-30
    del kp1, sabs
    ratio_f32 = ((y - ye).abs() / lim).max().item()
    print(f"ragged f32: every-sample max error / bound {ratio_f32:.3g} (<= 1 passes)")
    assert ratio_f32 <= 1.0
    lens = np.diff(offs)
    pick = sorted(set([int(np.argmin(lens)), int(np.argmax(lens))] +
                      list(rng(9).integers(0, offs.size - 1, 10))))
    for s in pick:
        seg = slice(int(offs[s]), int(offs[s + 1]))
        xs = x[seg].cpu().numpy()
        k = taps[boffs[s]:boffs[s + 1]]
        y64 = oracle.lfilter_f64(xs, k)[0]
        assert np.all(np.abs(y[seg].cpu().numpy() - y64) <= oracle.fir_error_bound(xs, k)[0])
    plan.close()
    # The same workload on MFMA (BF16x3): persistent vs per-signal launches,
    # and relative L2 error <= 1e-6 against FP64 on the same signals.
    plan = RaggedPlan(offs, boffs, "bf16x3")
    y = plan.execute(x, bt)
    y2 = plan.execute_per_signal(x, bt)
    torch.cuda.synchronize()
    assert torch.equal(y, y2)
    del y2
    ratio_b3 = ((y - ye).abs() / lim).max().item()
    print(f"ragged bf16x3: every-sample max error / bound {ratio_b3:.3g} (<= 1 passes)")
    assert ratio_b3 <= 1.0
    del ye, lim
    for s in pick:
        seg = slice(int(offs[s]), int(offs[s + 1]))
        xs = x[seg].cpu().numpy()
        y64 = oracle.lfilter_f64(xs, taps[boffs[s]:boffs[s + 1]])[0]
        assert oracle.rel_l2(y[seg].cpu().numpy(), y64) <= 1e-6, s
    plan.close()
