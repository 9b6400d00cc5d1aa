"""Pins the CPU oracle (oracle/fir_oracle.c) before anything is checked
against it: the reference's known-answer test, an independent float32
emulation (bitwise), closed forms, scipy golden vectors, brute force, and --
for the adjoint -- finite differences and torch float64 autograd."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden, rng


def numpy_ref_f32(x, taps):
    """Independent float32 emulation of crates/scir-gpu/src/lib.rs:1141-1148,
    vectorised across outputs: for t ascending, acc = f32(acc + f32(tap*x)).
    numpy rounds every float32 operation and never fuses, so this matches a
    correct restatement bit for bit and catches an FMA-contracted one."""
    x = np.asarray(x, np.float32)
    taps = np.asarray(taps, np.float32)
    B, N = x.shape
    K = taps.size
    acc = np.zeros((B, N), np.float32)
    for t in range(K):
        shifted = np.zeros((B, N), np.float32)
        if t < N:
            shifted[:, t:] = x[:, :N - t]
        prod = (taps[K - 1 - t] * shifted).astype(np.float32)
        # terms past the row start are absent in the reference loop: skip them
        acc[:, t:] = (acc[:, t:] + prod[:, t:]).astype(np.float32)
    return acc


def test_reference_kat(oracle):
    # crates/scir-gpu/src/lib.rs:1249-1261
    x = np.array([[1.0, 2.0, 3.0, 4.0], [0.5, 0.0, -0.5, -1.0]], np.float32)
    taps = np.array([0.25, 0.5, 0.25], np.float32)
    y = oracle.fir_ref_f32(x, taps)
    np.testing.assert_array_equal(y, [[0.25, 1.0, 2.0, 3.0], [0.125, 0.25, 0.0, -0.5]])
    np.testing.assert_array_equal(oracle.fir_ref_f64(x, taps), y.astype(np.float64))


@pytest.mark.parametrize("K", [1, 2, 5, 31, 64])
def test_f32_bitwise_vs_numpy_emulation(oracle, K):
    r = rng(K)
    x = r.standard_normal((3, 97)).astype(np.float32)
    taps = (1.0 / (np.arange(K) + 1.0)).astype(np.float32)  # fir_bench taps, asymmetric
    np.testing.assert_array_equal(oracle.fir_ref_f32(x, taps), numpy_ref_f32(x, taps))


def test_planted_fma_is_caught(oracle):
    """An FMA-contracted loop (a plausible restatement mistake) differs from
    the bitwise emulation, so the test above can fail."""
    r = rng(7)
    x = r.standard_normal((2, 64)).astype(np.float32)
    taps = r.standard_normal(9).astype(np.float32)
    ref = numpy_ref_f32(x, taps)
    fused = np.zeros_like(ref)
    for i in range(64):
        acc = np.float64(0.0)
        for t in range(min(i, 8) + 1):
            acc = np.float32(acc + np.float64(taps[8 - t]) * np.float64(x[:, i - t][0]))
        fused[0, i] = acc
    assert np.any(fused[0] != ref[0])


def poison(x, taps):
    """+Inf, -Inf and NaN in x and in the taps (the first taps in the
    reference's order, i.e. the last b[k], are met only by outputs i >= K-1)."""
    x = x.copy()
    taps = taps.copy()
    x[0, 5] = np.inf
    x[1, 40] = -np.inf
    x[2, 70] = np.nan
    x[0, 80] = np.inf
    x[0, 81] = -np.inf
    taps[0] = np.inf
    taps[2] = np.nan
    taps[-1] = -np.inf
    return x, taps


@pytest.mark.parametrize("K", [1, 3, 8, 31])
def test_f32_nonfinite_bitwise_vs_numpy_emulation(oracle, K):
    """Inf and NaN reach exactly the outputs whose reference terms hold them:
    the oracle equals the emulation NaN for NaN."""
    r = rng(50 + K)
    x, taps = poison(r.standard_normal((3, 97)).astype(np.float32),
                     r.standard_normal(max(K, 3)).astype(np.float32))
    taps = taps[:K]
    with np.errstate(invalid="ignore", over="ignore"):
        ref = numpy_ref_f32(x, taps)
    y = oracle.fir_ref_f32(x, taps)
    assert np.array_equal(y, ref, equal_nan=True)
    assert np.isnan(y).any() and np.isinf(y).any()


def test_planted_ramp_in_term_is_caught(oracle):
    """A restatement that also forms the terms before the row start
    (taps * 0 for t > i, finite for finite taps) differs from the oracle as
    soon as a tap is infinite, so the non-finite test above can fail."""
    r = rng(9)
    x = r.standard_normal((1, 40)).astype(np.float32)
    taps = r.standard_normal(6).astype(np.float32)
    taps[0] = np.inf  # applied to x[i-5]: before the row start for i < 5
    K = taps.size
    padded = np.concatenate([np.zeros((1, K - 1), np.float32), x], axis=1)
    with np.errstate(invalid="ignore"):
        wrong = np.zeros_like(x)
        for t in range(K):
            prod = (taps[K - 1 - t] * padded[:, K - 1 - t:K - 1 - t + 40]).astype(np.float32)
            wrong = (wrong + prod).astype(np.float32)
    y = oracle.fir_ref_f32(x, taps)
    assert np.isnan(wrong[0, :5]).all() and np.isfinite(y[0, :5]).all()


def test_impulse_step_dc_closed_forms(oracle):
    K = 7
    taps = rng(1).standard_normal(K).astype(np.float32)
    x = np.zeros((1, 20), np.float32)
    x[0, 0] = 1.0
    y = oracle.fir_ref_f32(x, taps)[0]
    np.testing.assert_array_equal(y[:K], taps[::-1])   # y[i] = taps[K-1-i]
    np.testing.assert_array_equal(y[K:], 0.0)
    step = np.ones((1, 20), np.float64)
    ys = oracle.fir_ref_f64(step, taps.astype(np.float64))[0]
    np.testing.assert_allclose(ys[:K], np.cumsum(taps[::-1].astype(np.float64)), rtol=1e-14)
    np.testing.assert_allclose(ys[K:], taps.astype(np.float64).sum(), rtol=1e-14)  # DC gain


@pytest.mark.parametrize("K", [1, 2, 3, 17, 33, 129])
def test_lfilter_golden(oracle, K):
    x = load_golden("lfilter_x.npy")
    b = load_golden(f"lfilter_b{K}.npy")
    y_sp = load_golden(f"lfilter_y{K}.npy")
    y64 = oracle.lfilter_f64(x, b)
    np.testing.assert_allclose(y64, y_sp, rtol=1e-12, atol=1e-12 * np.abs(y_sp).max())
    # reference-order f64 function == lfilter with reversed taps
    yr = oracle.fir_ref_f64(x.astype(np.float64), b[::-1].astype(np.float64))
    np.testing.assert_allclose(yr, y_sp, rtol=1e-12, atol=1e-12 * np.abs(y_sp).max())
    # float32 reference loop within the stated bound of the exact sum
    y32 = oracle.fir_ref_f32(x, b[::-1])
    bound = oracle.fir_error_bound(x, b)
    assert np.all(np.abs(y32 - y_sp) <= bound + 1e-300)


def test_error_bound_is_tight_enough_to_fail(oracle):
    """Dropping one tap (a plausible indexing mistake) violates the bound."""
    x = load_golden("lfilter_x.npy")
    b = load_golden("lfilter_b33.npy")
    y_sp = load_golden("lfilter_y33.npy")
    wrong = oracle.fir_ref_f32(x, b[:-1][::-1])
    assert np.any(np.abs(wrong - y_sp) > oracle.fir_error_bound(x, b))


def test_brute_force_tiny(oracle):
    r = rng(3)
    for B, N, K in [(1, 1, 1), (2, 5, 9), (4, 64, 17), (3, 17, 17)]:
        x = r.standard_normal((B, N)).astype(np.float32)
        b = r.standard_normal(K).astype(np.float32)
        brute = np.zeros((B, N))
        for row in range(B):
            for n in range(N):
                brute[row, n] = sum(float(b[k]) * float(x[row, n - k]) for k in range(K) if n - k >= 0)
        np.testing.assert_allclose(oracle.lfilter_f64(x, b), brute, rtol=1e-13, atol=1e-13)


def test_absbound_closed_form(oracle):
    x = -np.ones((1, 10), np.float32)
    b = np.array([2.0, -3.0], np.float32)
    s = oracle.abs_bound(x, b)[0]
    np.testing.assert_array_equal(s, [2.0] + [5.0] * 9)


def test_vjp_finite_differences(oracle):
    """L = sum(w*y) is bilinear in (x, b): central differences are exact up
    to rounding, so they pin both gradients."""
    r = rng(11)
    B, N, K = 2, 40, 6
    x = r.standard_normal((B, N)).astype(np.float32)
    b = r.standard_normal(K).astype(np.float32)
    w = r.standard_normal((B, N)).astype(np.float32)
    dx, db = oracle.vjp_f64(w, x, b)

    def loss(xx, bb):
        return float(np.sum(w.astype(np.float64) * oracle.lfilter_f64(xx, bb)))

    h = 1e-3
    for (row, n) in [(0, 0), (0, 17), (1, 39), (1, 35)]:
        xp, xm = x.astype(np.float64), x.astype(np.float64)
        xp = xp.copy(); xm = xm.copy()
        xp[row, n] += h; xm[row, n] -= h
        # evaluate in float64 via the ref-order f64 oracle on reversed taps
        lp = float(np.sum(w * oracle.fir_ref_f64(xp, b[::-1].astype(np.float64))))
        lm = float(np.sum(w * oracle.fir_ref_f64(xm, b[::-1].astype(np.float64))))
        assert abs((lp - lm) / (2 * h) - dx[row, n]) <= 1e-9 * max(1.0, abs(dx[row, n]))
    for k in range(K):
        bp = b.astype(np.float64).copy(); bm = bp.copy()
        bp[k] += h; bm[k] -= h
        lp = float(np.sum(w * oracle.fir_ref_f64(x.astype(np.float64), bp[::-1].copy())))
        lm = float(np.sum(w * oracle.fir_ref_f64(x.astype(np.float64), bm[::-1].copy())))
        assert abs((lp - lm) / (2 * h) - db[k]) <= 1e-9 * max(1.0, abs(db[k]))
    assert loss(x, b) != 0.0


def test_vjp_vs_torch_autograd_f64(oracle):
    torch = pytest.importorskip("torch")
    r = rng(12)
    B, N, K = 3, 50, 9
    x = r.standard_normal((B, N)).astype(np.float32)
    b = r.standard_normal(K).astype(np.float32)
    w = r.standard_normal((B, N)).astype(np.float32)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    # y[n] = sum_k b[k] x[n-k]: conv1d is a correlation, so flip b and left-pad
    y = torch.nn.functional.conv1d(torch.nn.functional.pad(xt[:, None, :], (K - 1, 0)),
                                   bt.flip(0)[None, None, :])[:, 0, :]
    (y * torch.tensor(w, dtype=torch.float64)).sum().backward()
    dx, db = oracle.vjp_f64(w, x, b)
    np.testing.assert_allclose(dx, xt.grad.numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_ragged_equals_per_signal(oracle):
    r = rng(5)
    lens = [1, 7, 64, 130]
    ks = [3, 1, 17, 5]
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    boffs = np.concatenate(([0], np.cumsum(ks))).astype(np.int64)
    x = r.standard_normal(offs[-1]).astype(np.float32)
    b = r.standard_normal(boffs[-1]).astype(np.float32)
    y = oracle.lfilter_ragged_f64(x, offs, b, boffs)
    for s in range(len(lens)):
        seg = slice(offs[s], offs[s + 1])
        np.testing.assert_array_equal(y[seg], oracle.lfilter_f64(x[seg], b[boffs[s]:boffs[s + 1]])[0])
