"""GPU parity of the adjoint kernels (dL/dx, dL/db) through the C ABI.

L = sum(w * y), y = lfilter(b,[1],x) per row.  References: the FP64 oracle
gradient (oracle.vjp_f64), central finite differences of the FP64 loss
(exact up to rounding for a bilinear loss), and torch float64 autograd
(through the test-only autograd wrapper).  Bounds:
  dx: |dx - dx64| <= (K+1) * 2^-23 * sum_k |b_k||w_{m+k}|
  db: |db - db64| <= (L + 200) * 2^-24 * sum_{r,n} |w_{r,n}||x_{r,n-k}|,
      L = tiles per workgroup (<= samples/2048): the longest FP32 chain of
      the two-stage reduction is 128 + L + 16 additions, the rest is Kahan.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -23


def _t(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), device="cuda")


def gpu_grads(x, b, w, precision="f32"):
    import torch
    from scir_amd import device
    xt, bt, wt = _t(x), _t(b), _t(w)
    dx = device.vjp_dx(wt, bt, precision)
    db = device.vjp_db(wt, xt, b.size)
    torch.cuda.synchronize()
    return dx.cpu().numpy(), db.cpu().numpy()


def bounds(oracle, x, b, w):
    dxa, dba = oracle.vjp_f64(np.abs(w), np.abs(x), np.abs(b))
    N = x.shape[1]
    K = b.size
    return (K + 1) * EPS * dxa, (N / 2048 + 200) * 2.0 ** -24 * dba


@pytest.mark.parametrize("B,N,K", [(1, 1, 1), (2, 5, 9), (3, 100, 1), (2, 2047, 33), (2, 2049, 255),
                                   (1, 9000, 300), (3, 5000, 1024), (1, 7000, 3001), (4, 65536, 255)])
@pytest.mark.parametrize("precision", ["f32", "bf16x3"])
def test_grads_vs_fp64(hip, oracle, B, N, K, precision):
    r = rng(B * 7 + N + K)
    x = r.standard_normal((B, N)).astype(np.float32)
    b = r.standard_normal(K).astype(np.float32)
    w = r.standard_normal((B, N)).astype(np.float32)
    dx, db = gpu_grads(x, b, w, precision)
    dx64, db64 = oracle.vjp_f64(w, x, b)
    bdx, bdb = bounds(oracle, x, b, w)
    assert np.all(np.abs(dx - dx64) <= bdx), f"dx max excess {np.max(np.abs(dx - dx64) - bdx)}"
    assert np.all(np.abs(db - db64) <= bdb), f"db max excess {np.max(np.abs(db - db64) - bdb)}"


@pytest.mark.parametrize("mfma", [0, 1])
@pytest.mark.parametrize("B,N,K", [(1, 1, 1), (2, 2049, 255), (1, 9000, 300), (2, 5000, 600),
                                   (3, 70000, 255)])
def test_db_forms_vs_fp64(hip, oracle, B, N, K, mfma):
    """Both stage-1 forms of dL/db (FP32 MFMA, FP32 VALU) within the bound,
    including K beyond one 256-lag pass and rows of several workgroups."""
    L = hip.lib()
    hip.check(L.scir_hip_tune_set(b"vjp.db_mfma", mfma))
    try:
        r = rng(B * 11 + N + K + mfma)
        x = r.standard_normal((B, N)).astype(np.float32)
        b = r.standard_normal(K).astype(np.float32)
        w = r.standard_normal((B, N)).astype(np.float32)
        _, db = gpu_grads(x, b, w)
        _, db64 = oracle.vjp_f64(w, x, b)
        _, bdb = bounds(oracle, x, b, w)
        assert np.all(np.abs(db - db64) <= bdb), f"db max excess {np.max(np.abs(db - db64) - bdb)}"
    finally:
        hip.check(L.scir_hip_tune_set(b"vjp.db_mfma", -1))


def test_finite_differences(hip, oracle):
    """64 seeded coordinates (x and b), h = 1e-3, FP64 loss via the oracle."""
    r = rng(64)
    B, N, K = 3, 3000, 255
    x = r.standard_normal((B, N)).astype(np.float32)
    b = r.standard_normal(K).astype(np.float32)
    w = r.standard_normal((B, N)).astype(np.float32)
    dx, db = gpu_grads(x, b, w)
    w64 = w.astype(np.float64)

    def loss_rows(xx, bb, rows=slice(None)):
        return float(np.sum(w64[rows] * oracle.fir_ref_f64(xx, bb[::-1].copy())))

    h = 1e-3
    x64 = x.astype(np.float64)
    b64 = b.astype(np.float64)
    for _ in range(48):
        row, m = int(r.integers(B)), int(r.integers(N))
        xp = x64.copy(); xm = x64.copy()
        xp[row, m] += h; xm[row, m] -= h
        # only row `row` of the loss depends on x[row, m]
        one = slice(row, row + 1)
        fd = (loss_rows(xp[one], b64, one) - loss_rows(xm[one], b64, one)) / (2 * h)
        assert abs(fd - dx[row, m]) <= 1e-5 * (1.0 + abs(fd)), (row, m, fd, dx[row, m])
    for _ in range(16):
        k = int(r.integers(K))
        bp = b64.copy(); bm = b64.copy()
        bp[k] += h; bm[k] -= h
        fd = (loss_rows(x64, bp) - loss_rows(x64, bm)) / (2 * h)
        assert abs(fd - db[k]) <= 1e-5 * (1.0 + abs(fd)), (k, fd, db[k])


def test_autograd_wrapper_vs_torch_f64(hip):
    import torch
    from scir_amd import autograd
    r = rng(3)
    B, N, K = 2, 4000, 129
    x = r.standard_normal((B, N)).astype(np.float32)
    b = r.standard_normal(K).astype(np.float32)
    w = r.standard_normal((B, N)).astype(np.float32)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    y = autograd.fir(xt, bt, "f32")
    (y * torch.tensor(w, device="cuda")).sum().backward()
    xc = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    bc = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    yc = torch.nn.functional.conv1d(torch.nn.functional.pad(xc[:, None, :], (K - 1, 0)),
                                    bc.flip(0)[None, None, :])[:, 0, :]
    np.testing.assert_allclose(y.detach().cpu().numpy(), yc.detach().numpy(), rtol=1e-4, atol=1e-4)
    (yc * torch.tensor(w, dtype=torch.float64)).sum().backward()
    np.testing.assert_allclose(xt.grad.cpu().numpy(), xc.grad.numpy(), rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(bt.grad.cpu().numpy(), bc.grad.numpy(), rtol=1e-4, atol=1e-3)


def test_empty_and_strided(hip, oracle):
    import torch
    from scir_amd import device
    r = rng(4)
    K = 17
    b = torch.tensor(r.standard_normal(K).astype(np.float32), device="cuda")
    x0 = torch.zeros((0, 10), device="cuda")
    db = device.vjp_db(x0, x0, K)
    torch.cuda.synchronize()
    assert not db.cpu().numpy().any()
    base_g = torch.tensor(r.standard_normal((3, 1003)).astype(np.float32), device="cuda")
    base_x = torch.tensor(r.standard_normal((3, 1005)).astype(np.float32), device="cuda")
    g, x = base_g[:, 1:1000], base_x[:, 3:1002]
    dx = device.vjp_dx(g, b)
    dbv = device.vjp_db(g, x, K)
    torch.cuda.synchronize()
    dx64, db64 = oracle.vjp_f64(g.cpu().numpy(), x.cpu().numpy(), b.cpu().numpy())
    bdx, bdb = bounds(oracle, x.cpu().numpy(), b.cpu().numpy(), g.cpu().numpy())
    assert np.all(np.abs(dx.cpu().numpy() - dx64) <= bdx)
    assert np.all(np.abs(dbv.cpu().numpy() - db64) <= bdb)


def test_full_size_differentiation_workload(hip, oracle):
    """x [1024, 65536], 255-tap low-pass, w ~ N(0,1): dx on a row subset vs
    FP64; db over all rows checked by additivity over row halves and on a
    row subset vs FP64."""
    import torch
    from scir_amd import design, device
    B, N, K = 1024, 65536, 255
    print(f"full-size vjp B={B} N={N} K={K}", flush=True)
    x = device.synth_normal(B, N, seed=0x5C120004)
    w = device.synth_normal(B, N, seed=0x5C120005)
    b_np = design.lowpass(K)
    b = torch.tensor(b_np, device="cuda")
    dx = device.vjp_dx(w, b, "bf16x3")
    db_all = device.vjp_db(w, x, K)
    db_lo = device.vjp_db(w[:512], x[:512], K)
    db_hi = device.vjp_db(w[512:], x[512:], K)
    rows = [0, 511, 1023]
    db_sub = device.vjp_db(w[rows], x[rows], K)
    torch.cuda.synchronize()
    xs, ws = x[rows].cpu().numpy(), w[rows].cpu().numpy()
    dx64, db64 = oracle.vjp_f64(ws, xs, b_np)
    bdx, bdb = bounds(oracle, xs, b_np, ws)
    assert np.all(np.abs(dx[rows].cpu().numpy() - dx64) <= bdx)
    assert np.all(np.abs(db_sub.cpu().numpy() - db64) <= bdb)
    # additivity over rows; scale from |w||x| of the full problem
    sabs = device.vjp_db(w.abs(), x.abs(), K)
    torch.cuda.synchronize()
    lim = (N / 2048 + 200) * 2.0 ** -24 * 2.0 * sabs.cpu().numpy()
    assert np.all(np.abs(db_all.cpu().numpy() - (db_lo + db_hi).cpu().numpy()) <= lim)


@pytest.mark.parametrize("precision", ["f32", "bf16x3", "fft32"])
def test_host_tier_vjp(hip, oracle, precision):
    """The host-tier entry (the Rust vjp API's): both gradients in the
    caller's convention, equal to the device-tier kernels' results."""
    from scir_amd.gpu import fir1d_batched_f32_vjp_hip
    r = rng(77)
    B, N, K = 3, 9001, 257
    x = r.standard_normal((B, N)).astype(np.float32)
    taps = r.standard_normal(K).astype(np.float32)
    w = r.standard_normal((B, N)).astype(np.float32)
    dx, dtaps = fir1d_batched_f32_vjp_hip(x, taps, w, precision)
    b = taps[::-1].copy()                     # lfilter order
    dx_dev, db_dev = gpu_grads(x, b, w, precision)
    assert np.array_equal(dx, dx_dev)
    assert np.array_equal(dtaps, db_dev[::-1])
    dx64, db64 = oracle.vjp_f64(w, x, b)
    bdx, bdb = bounds(oracle, x, b, w)
    if precision != "fft32":
        assert np.all(np.abs(dx - dx64) <= bdx)
    assert oracle.rel_l2(dx, dx64) <= 1e-6
    assert np.all(np.abs(dtaps - db64[::-1]) <= bdb[::-1])
    # empty rows: dtaps = 0, dx empty
    dx0, dt0 = fir1d_batched_f32_vjp_hip(np.zeros((2, 0), np.float32), taps,
                                         np.zeros((2, 0), np.float32))
    assert dx0.shape == (2, 0) and not np.any(dt0)


@pytest.mark.parametrize("og,ox", [(1, 2), (3, 0), (2, 3)])
def test_unaligned_rows_equal_aligned(hip, og, ox):
    """g and x rows at 4-byte offsets with odd pitches give the same bits as
    aligned packed copies, for dL/dx (FP32, BF16x3) and dL/db (FP32 MFMA form)."""
    import torch
    from scir_amd import device
    B, N, K = 5, 7001, 65
    print(f"vjp unaligned rows offsets g={og} x={ox}", flush=True)
    r = rng(og * 10 + ox)
    ga = torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda")
    xa = torch.tensor(r.standard_normal((B, N)).astype(np.float32), device="cuda")
    b = torch.tensor((r.standard_normal(K) / np.sqrt(K)).astype(np.float32), device="cuda")
    gbase = torch.full((B, N + 5), float("nan"), device="cuda")
    gbase[:, og:og + N] = ga
    xbase = torch.full((B, N + 3), float("nan"), device="cuda")
    xbase[:, ox:ox + N] = xa
    gu, xu = gbase[:, og:og + N], xbase[:, ox:ox + N]
    for precision in ("f32", "bf16x3"):
        ybase = torch.full((B, N + 7), float("nan"), device="cuda")
        device.vjp_dx(gu, b, precision, out=ybase[:, ox:ox + N])
        torch.cuda.synchronize()
        assert torch.equal(device.vjp_dx(ga, b, precision), ybase[:, ox:ox + N]), precision
        yb = ybase.cpu().numpy()
        assert np.isnan(yb[:, :ox]).all() and np.isnan(yb[:, ox + N:]).all()
    assert torch.equal(device.vjp_db(ga, xa, K), device.vjp_db(gu, xu, K))
