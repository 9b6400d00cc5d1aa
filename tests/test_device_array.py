"""DeviceArray on the CPU device: the reference's own cases
(crates/scir-gpu/src/lib.rs:1193-1247 and the doc examples at lib.rs:86-260)."""
import numpy as np
import pytest

from scir_amd.gpu import DType, Device, DeviceArray, ShapeMismatch


def test_roundtrip():
    data = [1.0, 2.0, 3.0, 4.0]
    arr = DeviceArray.from_cpu_slice([2, 2], DType.F32, data)
    assert arr.shape() == [2, 2]
    assert arr.dtype() is DType.F32
    assert arr.device() is Device.Cpu
    assert arr.to_cpu_vec().tolist() == data
    assert arr.to_cpu_vec().dtype == np.float32


def test_shape_must_match_length():
    with pytest.raises(ValueError):
        DeviceArray.from_cpu_slice([2, 3], DType.F32, [1.0, 2.0])


def test_elementwise_f64_baseline():
    arr = DeviceArray.from_cpu_slice([4], DType.F64, [1.0, 2.0, 3.0, 4.0])
    assert arr.add_scalar(1.0).to_cpu_vec().tolist() == [2.0, 3.0, 4.0, 5.0]
    assert arr.mul_scalar(2.0).to_cpu_vec().tolist() == [2.0, 4.0, 6.0, 8.0]
    assert arr.add_scalar(1.0).to_cpu_vec().dtype == np.float64


def test_add_arrays_and_auto_names():
    a = DeviceArray.from_cpu_slice([3], DType.F32, [1.0, 2.0, 3.0])
    b = DeviceArray.from_cpu_slice([3], DType.F32, [0.5, 1.5, 2.5])
    assert a.add(b).to_cpu_vec().tolist() == [1.5, 3.5, 5.5]
    assert a.add_auto(b).to_cpu_vec().tolist() == [1.5, 3.5, 5.5]
    assert a.add_scalar_auto(1.0).to_cpu_vec().tolist() == [2.0, 3.0, 4.0]
    assert a.mul_scalar_auto(2.0).to_cpu_vec().tolist() == [2.0, 4.0, 6.0]


def test_add_shape_mismatch():
    a = DeviceArray.from_cpu_slice([3], DType.F32, [1.0, 2.0, 3.0])
    b = DeviceArray.from_cpu_slice([1, 3], DType.F32, [1.0, 2.0, 3.0])
    with pytest.raises(ShapeMismatch):
        a.add(b)


def test_f32_rounding_is_per_element():
    # float32 storage: one rounding per operation, as the reference's f32 loop
    x = np.array([1.0, 3.0, 1e-8, 16777216.0], np.float32)
    arr = DeviceArray.from_cpu_slice([4], DType.F32, x)
    want = x + np.float32(0.1)
    assert np.array_equal(arr.add_scalar(0.1).to_cpu_vec(), want)


def test_to_device_cpu_is_a_no_op():
    arr = DeviceArray.from_cpu_slice([4], DType.F32, [1.0, 2.0, 3.0, 4.0])
    arr.to_device(Device.Cpu)
    assert arr.device() is Device.Cpu
    assert arr.to_cpu_vec().tolist() == [1.0, 2.0, 3.0, 4.0]
