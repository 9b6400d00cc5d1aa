"""GPU parity of the batched float64 polyphase resampler against the oracle
(the reference's resample_poly steps, crates/scir-signal/src/lib.rs:293-362).

Any FMA summation order of the at most ceil(K/up) terms of an output stays
within (ceil(K/up) + 1) * 2^-52 * sum |h_k| |x_s| of the reference's
sequentially rounded sum, which is the bound checked per output."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng
from scir_amd import design, signal
from scir_amd.gpu import Device

pytestmark = pytest.mark.gpu

EPS64 = 2.0 ** -52


def bound(oracle, x, up, down, h):
    tmax = -(-h.size // up)
    return (tmax + 1) * EPS64 * np.stack([oracle.resample_poly_f64(np.abs(r), up, down, np.abs(h))
                                          for r in np.atleast_2d(x)])


@pytest.mark.parametrize("up,down,K", [(2, 3, 31), (3, 2, 31), (1, 4, 17), (5, 1, 40), (4, 4, 9),
                                       (2, 3, 1), (7, 3, 64), (2, 3, 1023)])
@pytest.mark.parametrize("N", [1, 2, 3, 100, 5003])
def test_gpu_within_bound_of_oracle(hip, oracle, up, down, K, N):
    r = rng(up * 1000 + down * 100 + K + N)
    x = r.standard_normal((3, N))
    x[1, ::17] *= 1e3
    h = r.standard_normal(K)
    y = signal.resample_poly_batched(x, up, down, h, Device.Hip)
    ref = np.stack([oracle.resample_poly_f64(x[i], up, down, h) for i in range(3)])
    assert y.shape == ref.shape
    assert np.all(np.abs(y - ref) <= bound(oracle, x, up, down, h) + 1e-300)


def test_reference_contract_on_hip(hip, oracle):
    x = rng(3).standard_normal(1000)
    y_hip = signal.resample_poly(x, 2, 3, Device.Hip)
    y_cpu = signal.resample_poly(x, 2, 3, Device.Cpu)  # bit-identical to the reference loop
    assert np.all(np.abs(y_hip - y_cpu) <= bound(oracle, x, 2, 3, signal.RESAMPLE_H_2_3)[0])


def test_full_size_rows(hip, oracle):
    """x [4096, 65536] float64, up 2, down 3, the reference's filter: row
    subset against the oracle, every row against the CPU arm on a sample."""
    import torch
    from scir_amd import _lib
    import ctypes
    B, N = 4096, 65536
    print(f"full-size resample B={B} N={N}", flush=True)
    xd = torch.randn((B, N), dtype=torch.float64, device="cuda",
                     generator=torch.Generator(device="cuda").manual_seed(11))
    h = signal.RESAMPLE_H_2_3
    hd = torch.tensor(h, device="cuda")
    M = signal.resample_len(N, 2, 3, h.size)
    yd = torch.empty((B, M), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.lib().scir_hip_resample_poly_f64_dev(
        ctypes.c_void_p(xd.data_ptr()), N, B, N, 2, 3, ctypes.c_void_p(hd.data_ptr()), h.size,
        ctypes.c_void_p(yd.data_ptr()), M, ctypes.c_void_p(st)))
    torch.cuda.synchronize()
    rows = [0, 1, 2047, 4095]
    xs = xd[rows].cpu().numpy()
    ys = yd[rows].cpu().numpy()
    for i, r in enumerate(rows):
        ref = oracle.resample_poly_f64(xs[i], 2, 3, h)
        assert np.all(np.abs(ys[i] - ref) <= bound(oracle, xs[i], 2, 3, h)[0])


@pytest.mark.parametrize("ox,oy", [(1, 0), (1, 1), (0, 1)])
def test_unaligned_rows_equal_aligned(hip, ox, oy):
    """Rows of x and y at odd 8-byte offsets with odd pitches give the same
    bits as aligned packed rows, and nothing outside y is written."""
    import torch
    import ctypes
    from scir_amd import _lib
    B, N, up, down = 7, 6001, 2, 3
    h = signal.RESAMPLE_H_2_3
    M = signal.resample_len(N, up, down, h.size)
    print(f"resample unaligned rows offsets x={ox} y={oy}", flush=True)
    xa = torch.tensor(rng(ox * 2 + oy).standard_normal((B, N)), device="cuda")
    hd = torch.tensor(h, device="cuda")
    st = torch.cuda.current_stream().cuda_stream

    def call(x, ldx, y, ldy):
        _lib.check(_lib.lib().scir_hip_resample_poly_f64_dev(
            ctypes.c_void_p(x.data_ptr()), ldx, B, N, up, down, ctypes.c_void_p(hd.data_ptr()),
            h.size, ctypes.c_void_p(y.data_ptr()), ldy, ctypes.c_void_p(st)))

    ya = torch.empty((B, M), dtype=torch.float64, device="cuda")
    call(xa, N, ya, M)
    xbase = torch.full((B, N + 3), float("nan"), dtype=torch.float64, device="cuda")
    xbase[:, ox:ox + N] = xa
    ybase = torch.full((B, M + 5), float("nan"), dtype=torch.float64, device="cuda")
    call(xbase[:, ox:], N + 3, ybase[:, oy:], M + 5)
    torch.cuda.synchronize()
    assert torch.equal(ya, ybase[:, oy:oy + M])
    yb = ybase.cpu().numpy()
    assert np.isnan(yb[:, :oy]).all() and np.isnan(yb[:, oy + M:]).all()


@pytest.mark.parametrize("B,N", [(3, 1), (3, 2), (3, 3), (5, 100), (4, 5003), (2, 70001)])
def test_fixed_ratio_kernel_equals_general(hip, B, N):
    """up 2, down 3 with 31 taps (the reference's one case) runs the
    register-window kernel; the general kernel (tuning key resample.fixed = 0)
    sums the same terms in the same order: the same bits, row tails included."""
    from scir_amd import _lib
    print(f"resample fixed vs general B={B} N={N}", flush=True)
    r = rng(N + 7)
    x = r.standard_normal((B, N))
    x[0, ::13] *= 1e3
    h = r.standard_normal(31)
    y_fixed = signal.resample_poly_batched(x, 2, 3, h, Device.Hip)
    _lib.check(_lib.lib().scir_hip_tune_set(b"resample.fixed", 0))
    try:
        y_general = signal.resample_poly_batched(x, 2, 3, h, Device.Hip)
    finally:
        _lib.check(_lib.lib().scir_hip_tune_set(b"resample.fixed", -1))
    assert y_fixed.shape == y_general.shape
    assert np.array_equal(y_fixed, y_general)
