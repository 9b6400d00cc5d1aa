"""Non-finite inputs and taps (Inf, -Inf, NaN) on every GPU path.

The reference loop (crates/scir-gpu/src/lib.rs:1141-1148) forms only the
terms taps[K-1-t] * x[i-t] with t <= min(i, K-1), so an Inf or NaN reaches
exactly the outputs whose terms include it.  The kernels stage zeros before
the row start (direct form), multiply zero taps of Toeplitz blocks (MFMA) and
transform whole segments (FFT); every output that comes out non-finite is
recomputed from the reference loop's terms, so:

* ``f32_exact`` equals the oracle's restatement of the reference loop bit for
  bit, NaN for NaN (every direct-kernel variant, dense and ragged, forward
  and back to front);
* ``f32``, ``bf16x3``, ``fp16``, ``fft32`` and ``auto`` are non-finite at
  exactly the reference loop's positions, with the same infinities, and
  their finite outputs keep each path's usual error bound.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng
from test_gpu_direct import host_fir

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[(0, 8, 1, 3), (0, 16, 1, 3), (1, 8, 3, 3), (2, 8, 4, 2), (2, 16, 3, 3),
                        (3, 8, 4, 2, 1), (3, 8, 4, 2, 6)],
                ids=lambda p: f"loader{p[0]}-r{p[1]}-seg{p[2]}-bufs{p[3]}" +
                              (f"-strip{p[4]}" if len(p) > 4 else ""))
def variant(request, hip):
    """Direct-kernel variants: one tile per workgroup, register prefetch and
    LDS DMA, 8 or 16 outputs per lane; F32 on FP32 MFMA (loader 3) in strips
    of 1 and 6 tiles."""
    loader, r, seg, bufs = request.param[:4]
    strip = request.param[4] if len(request.param) > 4 else -1
    L = hip.lib()
    for key, v in ((b"direct.loader", loader), (b"direct.r", r), (b"direct.seg_tiles", seg),
                   (b"direct.dma_bufs", bufs), (b"direct.mfma_strip", strip)):
        hip.check(L.scir_hip_tune_set(key, v))
    yield request.param
    for key in (b"direct.loader", b"direct.r", b"direct.seg_tiles", b"direct.dma_bufs",
                b"direct.mfma_strip"):
        hip.check(L.scir_hip_tune_set(key, -1))


EPS = 2.0 ** -23


def poisoned(seed, B, N, K, where):
    """x [B, N] ~ N(0,1) and lfilter-order taps b[K], with +Inf, -Inf and NaN
    planted in x ('x'), in b ('b': the last taps too, which only the first
    K-1 outputs of a row would meet through the staged zeros) or both."""
    r = rng(seed)
    x = r.standard_normal((B, N)).astype(np.float32)
    b = (r.standard_normal(K) / np.sqrt(K)).astype(np.float32)
    if "x" in where:
        for row in range(B):
            pos = r.choice(N, size=min(N, 3), replace=False)
            x[row, pos[0]] = np.inf
            if pos.size > 1:
                x[row, pos[1]] = -np.inf
            if pos.size > 2:
                x[row, pos[2]] = np.nan
        x[0, 0] = np.nan
    if "b" in where and K > 1:
        b[K - 1] = np.inf
        if K > 3:
            b[K // 2] = -np.inf
        if K > 5:
            b[K - 2] = np.nan
    return x, b


def finite_bound(oracle, x, b, K, scale):
    """scale * (K+1) * 2^-23 * sum|b||x| with non-finite values counted as 0
    (a finite reference output has only finite terms)."""
    xa = np.where(np.isfinite(x), np.abs(x), 0.0)
    ba = np.where(np.isfinite(b), np.abs(b), 0.0)
    return scale * (K + 1) * EPS * oracle.lfilter_f64(xa, ba)


def check_positions(y, ref, what):
    fy, fr = np.isfinite(y), np.isfinite(ref)
    bad = np.flatnonzero((fy != fr).ravel())
    assert bad.size == 0, f"{what}: non-finite positions differ at {bad[:8]}"
    inf = np.isinf(ref)
    assert np.array_equal(y[inf], ref[inf]), f"{what}: infinities differ"
    assert np.all(np.isnan(y[np.isnan(ref)])), f"{what}: NaN positions differ"


WHERE = ["x", "b", "xb"]
K_DIRECT = [1, 2, 5, 33, 129, 300, 513, 1000, 2049]


@pytest.mark.parametrize("where", WHERE)
@pytest.mark.parametrize("K", K_DIRECT)
def test_exact_bitwise_nonfinite(hip, oracle, variant, K, where):
    for N in (3, 4096, 9001):
        x, b = poisoned(300 + K + N, 3, N, K, where)
        y = host_fir(hip, x, b, "f32_exact")
        ref = oracle.fir_ref_f32(x, b[::-1].copy())
        same = (y == ref) | (np.isnan(y) & np.isnan(ref))
        assert same.all(), f"K={K} N={N} {where}: {np.count_nonzero(~same)} outputs differ"



@pytest.mark.parametrize("where", WHERE)
@pytest.mark.parametrize("K", K_DIRECT)
def test_f32_nonfinite_positions(hip, oracle, variant, K, where):
    for N in (17, 4096, 9001):
        x, b = poisoned(400 + K + N, 3, N, K, where)
        y = host_fir(hip, x, b, "f32")
        ref = oracle.fir_ref_f32(x, b[::-1].copy())
        check_positions(y, ref, f"f32 K={K} N={N} {where}")
        fin = np.isfinite(ref)
        bound = finite_bound(oracle, x, b, K, 2.0)
        assert np.all(np.abs(y[fin] - ref[fin]) <= bound[fin])


@pytest.mark.parametrize("prec", ["bf16x3", "fp16", "fft32", "auto"])
@pytest.mark.parametrize("where", WHERE)
@pytest.mark.parametrize("K", [2, 33, 100, 512, 1023, 2048])
def test_other_paths_nonfinite_positions(hip, oracle, prec, K, where):
    for N in (100, 20000):
        x, b = poisoned(500 + K + N, 3, N, K, where)
        y = host_fir(hip, x, b, prec)
        ref = oracle.fir_ref_f32(x, b[::-1].copy())
        check_positions(y, ref, f"{prec} K={K} N={N} {where}")
        fin = np.isfinite(ref)
        if prec in ("bf16x3", "auto"):
            bound = finite_bound(oracle, x, b, K, 4.0)
            assert np.all(np.abs(y[fin] - ref[fin]) <= bound[fin]), f"{prec} K={K} N={N}"


def test_partitioned_fft_nonfinite_positions(hip, oracle):
    """Filters longer than 8192 taps: the block that writes an output last
    checks it (and recomputes it with the whole filter)."""
    K, N = 9000, 20000
    for where in WHERE:
        x, b = poisoned(600, 2, N, K, where)
        y = host_fir(hip, x, b, "fft32")
        ref = oracle.fir_ref_f32(x, b[::-1].copy())
        check_positions(y, ref, f"fft32 K={K} {where}")


@pytest.mark.parametrize("prec", ["f32_exact", "f32", "bf16x3", "fft32"])
def test_adjoint_dx_nonfinite(hip, oracle, prec):
    """dL/dx runs the kernels back to front: dx[m] = sum_k b[k] g[m+k] is the
    reference loop on the reversed row."""
    import torch
    from scir_amd import device
    K, N = 129, 9001
    for where in WHERE:
        g, b = poisoned(700 + K, 3, N, K, where)
        dx = device.vjp_dx(torch.tensor(g, device="cuda"), torch.tensor(b, device="cuda"),
                           prec).cpu().numpy()
        ref = oracle.fir_ref_f32(np.ascontiguousarray(g[:, ::-1]), b[::-1].copy())[:, ::-1]
        if prec == "f32_exact":
            assert np.array_equal(dx, ref, equal_nan=True), f"dx {prec} {where}"
        else:
            check_positions(dx, ref, f"dx {prec} {where}")


@pytest.mark.parametrize("prec", ["f32_exact", "f32", "bf16x3", "fp16"])
def test_ragged_nonfinite(hip, oracle, prec):
    import torch
    from scir_amd.ragged import RaggedPlan
    r = rng(800)
    lens = [0, 5, 31, 4000, 9001, 20000, 3]
    ks = [7, 1, 40, 33, 600, 129, 2]
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    boffs = np.concatenate([[0], np.cumsum(ks)]).astype(np.int64)
    x = r.standard_normal(offs[-1]).astype(np.float32)
    taps = (r.standard_normal(boffs[-1]) * 0.1).astype(np.float32)
    for s in range(len(lens)):
        if lens[s] > 2:
            x[offs[s] + r.integers(0, lens[s])] = [np.inf, -np.inf, np.nan][s % 3]
        if ks[s] > 1:
            taps[boffs[s + 1] - 1] = [np.nan, np.inf, -np.inf][s % 3]
    plan = RaggedPlan(offs, boffs, prec)
    try:
        y = torch.full((offs[-1],), 7.0, device="cuda")
        plan.execute(torch.tensor(x, device="cuda"), torch.tensor(taps, device="cuda"), y)
        y = y.cpu().numpy()
    finally:
        plan.close()
    for s in range(len(lens)):
        seg = slice(offs[s], offs[s + 1])
        if lens[s] == 0:
            continue
        bs = taps[boffs[s]:boffs[s + 1]]
        ref = oracle.fir_ref_f32(x[seg][None, :], bs[::-1].copy())[0]
        if prec == "f32_exact":
            assert np.array_equal(y[seg], ref, equal_nan=True), f"signal {s}"
        else:
            check_positions(y[seg], ref, f"ragged {prec} signal {s}")
