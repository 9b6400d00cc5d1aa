"""Host-side mirror of scir_gpu's FIR API on the CPU: the reference's own
tests for this path, restated (crates/scir-gpu/src/lib.rs:1249-1298,
crates/scir-signal/src/lib.rs:670-696), plus the CPU arm against the oracle
bit for bit."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import rng
import scir_amd
from scir_amd import Device, GpuError, signal


def test_fir_batched_matches_naive_f32():
    # crates/scir-gpu/src/lib.rs:1249-1261
    x = np.array([[1.0, 2.0, 3.0, 4.0], [0.5, 0.0, -0.5, -1.0]], np.float32)
    taps = np.array([0.25, 0.5, 0.25], np.float32)
    y = scir_amd.fir1d_batched_f32(x, taps)
    np.testing.assert_allclose(y[0].astype(np.float64), [0.25, 1.0, 2.0, 3.0], atol=1e-7, rtol=1e-7)
    np.testing.assert_allclose(y[1].astype(np.float64), [0.125, 0.25, 0.0, -0.5], atol=1e-7, rtol=1e-7)


def test_auto_cpu_matches_naive_bitwise(oracle):
    # crates/scir-signal/src/lib.rs:670-696: Device::Cpu vs the naive loop, tol 0
    x = np.array([[1.0, 2.0, 3.0, 4.0], [0.5, 0.0, -0.5, -1.0]], np.float32)
    taps = np.array([0.25, 0.5, 0.25], np.float32)
    y = signal.fir1d_batched_f32(x, taps, Device.Cpu)
    np.testing.assert_array_equal(y, oracle.fir_ref_f32(x, taps))


@pytest.mark.parametrize("B,N,K", [(3, 32, 5), (1, 1, 1), (2, 5, 9), (4, 1000, 31),
                                   (2, 333, 129), (1, 64, 64), (5, 70, 0)])
def test_cpu_arm_bitwise_vs_oracle(oracle, B, N, K):
    r = rng(B * 1000 + N + K)
    x = r.uniform(-1, 1, (B, N)).astype(np.float32)
    taps = (1.0 / (np.arange(K) + 1.0)).astype(np.float32)  # fir_bench taps
    y = scir_amd.fir1d_batched_f32(x, taps)
    assert y.shape == (B, N) and y.dtype == np.float32
    if K == 0:
        assert not y.any()
    else:
        np.testing.assert_array_equal(y, oracle.fir_ref_f32(x, taps))


@pytest.mark.parametrize("K", [1, 4, 33])
def test_cpu_arm_nonfinite_bitwise_vs_oracle(oracle, K):
    """Inf and NaN in x and taps: the CPU arm forms only the reference
    loop's terms, so it equals the oracle NaN for NaN."""
    r = rng(60 + K)
    x = r.standard_normal((3, 200)).astype(np.float32)
    taps = r.standard_normal(K).astype(np.float32)
    x[0, 3], x[1, 100], x[2, 150] = np.inf, -np.inf, np.nan
    taps[0] = np.inf
    if K > 2:
        taps[1] = np.nan
    y = scir_amd.fir1d_batched_f32(x, taps)
    assert np.array_equal(y, oracle.fir_ref_f32(x, taps), equal_nan=True)


def test_cpu_config_workload_cpu_arm(oracle):
    """The reference's CPU case (x [64, 16384] ~ N(0,1) f32, 33-tap Hamming
    low-pass, cutoff 0.2 of Nyquist) on the library's CPU arm: bit for bit
    the reference loop, and within (K+1)*2^-23*sum|b||x| of FP64."""
    from scir_amd import design
    x = rng(0x5C120000).standard_normal((64, 16384)).astype(np.float32)
    b = design.lowpass(33, 0.2).astype(np.float32)
    y = scir_amd.fir1d_batched_f32(x, b[::-1].copy())
    np.testing.assert_array_equal(y, oracle.fir_ref_f32(x, b[::-1]))
    assert np.all(np.abs(y - oracle.lfilter_f64(x, b)) <= oracle.fir_error_bound(x, b))


def test_cpu_arm_accepts_non_contiguous_input(oracle):
    r = rng(9)
    base = r.standard_normal((6, 50)).astype(np.float32)
    x = base[::2, 3:40]  # strided rows and an offset column window
    taps = r.standard_normal(7).astype(np.float32)
    np.testing.assert_array_equal(scir_amd.fir1d_batched_f32(x, taps),
                                  oracle.fir_ref_f32(np.ascontiguousarray(x), taps))
    xt = np.asfortranarray(r.standard_normal((4, 30)).astype(np.float32))
    np.testing.assert_array_equal(scir_amd.fir1d_batched_f32(xt, taps),
                                  oracle.fir_ref_f32(np.ascontiguousarray(xt), taps))


def test_shapes_like_doctests():
    y = scir_amd.fir1d_batched_f32_auto(np.array([[1.0, 2.0, 3.0, 4.0]]),
                                        np.array([0.25, 0.5, 0.25]), Device.Cpu)
    assert y.shape == (1, 4)


def test_shape_errors():
    with pytest.raises(scir_amd.ShapeMismatch):
        scir_amd.fir1d_batched_f32(np.ones(4), np.ones(3))
    with pytest.raises(scir_amd.ShapeMismatch):
        scir_amd.fir1d_batched_f32(np.ones((2, 4)), np.ones((3, 1)))


def test_hip_without_device_raises_not_falls_back(monkeypatch):
    from scir_amd import _lib
    if _lib.device_count() > 0:
        pytest.skip("a HIP device is visible")
    monkeypatch.delenv("SCIR_HIP_CPU_FALLBACK", raising=False)
    with pytest.raises(GpuError):
        scir_amd.fir1d_batched_f32_auto(np.ones((2, 4)), np.ones(3), Device.Hip)


def test_opt_in_fallback_matches_reference_behaviour(monkeypatch, capsys, oracle):
    from scir_amd import _lib
    if _lib.device_count() > 0:
        pytest.skip("a HIP device is visible")
    monkeypatch.setenv("SCIR_HIP_CPU_FALLBACK", "1")
    x = rng(2).standard_normal((2, 16)).astype(np.float32)
    taps = np.array([0.25, 0.5, 0.25], np.float32)
    y = scir_amd.fir1d_batched_f32_auto(x, taps, Device.Hip)
    np.testing.assert_array_equal(y, oracle.fir_ref_f32(x, taps))
    assert "recomputing on the CPU" in capsys.readouterr().err


def test_doctests():
    import doctest
    import scir_amd.gpu
    res = doctest.testmod(scir_amd.gpu)
    assert res.failed == 0 and res.attempted >= 2
