"""DeviceArray on the HIP device: data resident in a library buffer,
elementwise results bit-identical to the CPU arm, FIR on resident data equal
to the host-tier FIR."""
import numpy as np
import pytest

from scir_amd import design
from scir_amd.gpu import (BackendUnavailable, DType, Device, DeviceArray, ShapeMismatch,
                          fir1d_batched_f32_hip, fir1d_batched_f32_resident)

pytestmark = pytest.mark.gpu


def _arr(shape, seed, dtype=DType.F32):
    rng = np.random.default_rng(seed)
    data = rng.standard_normal(int(np.prod(shape))).astype(np.float32)
    return data, DeviceArray.from_cpu_slice(shape, dtype, data)


@pytest.mark.parametrize("n", [1, 3, 257, 100003])
def test_roundtrip_and_elementwise_bitwise(n):
    print(f"device array n={n}", flush=True)
    data, cpu = _arr([n], seed=n)
    _, hip = _arr([n], seed=n)
    hip.to_device(Device.Hip)
    assert hip.device() is Device.Hip and hip.shape() == [n]
    assert np.array_equal(hip.to_cpu_vec(), data)
    for name, arg in (("add_scalar", 0.1), ("mul_scalar", -2.5)):
        got = getattr(hip, name)(arg)
        assert got.device() is Device.Hip
        assert np.array_equal(got.to_cpu_vec(), getattr(cpu, name)(arg).to_cpu_vec()), name
    _, other_cpu = _arr([n], seed=n + 1)
    _, other = _arr([n], seed=n + 1)
    other.to_device(Device.Hip)
    assert np.array_equal(hip.add(other).to_cpu_vec(), cpu.add(other_cpu).to_cpu_vec())


def test_reference_examples_on_hip():
    a = DeviceArray.from_cpu_slice([3], DType.F32, [1.0, 2.0, 3.0])
    b = DeviceArray.from_cpu_slice([3], DType.F32, [0.5, 1.5, 2.5])
    a.to_device(Device.Hip)
    b.to_device(Device.Hip)
    assert a.add_scalar_auto(1.0).to_cpu_vec().tolist() == [2.0, 3.0, 4.0]
    assert a.mul_scalar_auto(2.0).to_cpu_vec().tolist() == [2.0, 4.0, 6.0]
    assert a.add_auto(b).to_cpu_vec().tolist() == [1.5, 3.5, 5.5]


def test_chain_stays_on_device_and_back():
    data, arr = _arr([4, 1000], seed=7)
    arr.to_device(Device.Hip)
    out = arr.mul_scalar(3.0).add_scalar(-1.0)
    want = (data * np.float32(3.0)) + np.float32(-1.0)
    out.to_device(Device.Cpu)
    assert out.device() is Device.Cpu
    assert np.array_equal(out.to_cpu_vec(), want)


def test_add_across_devices_refused():
    _, a = _arr([8], seed=1)
    _, b = _arr([8], seed=2)
    a.to_device(Device.Hip)
    with pytest.raises(ShapeMismatch):
        a.add(b)


def test_f64_stays_on_cpu():
    _, a = _arr([8], seed=3, dtype=DType.F64)
    with pytest.raises(BackendUnavailable):
        a.to_device(Device.Hip)


@pytest.mark.parametrize("prec", ["f32", "f32_exact", "bf16x3", "auto"])
def test_resident_fir_equals_host_tier(prec):
    print(f"resident fir prec={prec}", flush=True)
    B, N, K = 6, 5000, 65
    data, arr = _arr([B, N], seed=11)
    b = design.lowpass(K)
    arr.to_device(Device.Hip)
    y = fir1d_batched_f32_resident(arr, b, prec)
    assert y.device() is Device.Hip and y.shape() == [B, N]
    want = fir1d_batched_f32_hip(data.reshape(B, N), b, prec)
    assert np.array_equal(y.to_cpu_vec().reshape(B, N), want)
    # and the result feeds the next operation without leaving the device
    z = y.mul_scalar(2.0).to_cpu_vec().reshape(B, N)
    assert np.array_equal(z, want * np.float32(2.0))
