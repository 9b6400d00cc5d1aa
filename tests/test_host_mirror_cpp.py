"""Builds and runs the C++ host-mirror test program (tests/cpp) against
libscir_hip.so.  Without a device the HIP arm must raise GpuError."""
from __future__ import annotations

import os
import subprocess

import pytest

from conftest import ROOT
from scir_amd import _lib


def build_program(tmp_path):
    exe = str(tmp_path / "host_mirror_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                    os.path.join(ROOT, "tests", "cpp", "host_mirror_test.cpp"),
                    "-L" + os.path.dirname(_lib.LIB_PATH), "-lscir_hip",
                    "-Wl,-rpath," + os.path.dirname(_lib.LIB_PATH), "-o", exe], check=True)
    return exe


def test_cpp_mirror_without_device(tmp_path):
    if _lib.device_count() > 0:
        pytest.skip("a HIP device is visible")
    exe = build_program(tmp_path)
    env = dict(os.environ)
    env.pop("SCIR_HIP_CPU_FALLBACK", None)
    r = subprocess.run([exe, "--expect-no-device"], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host mirror: ok" in r.stdout


@pytest.mark.gpu
def test_cpp_mirror_with_device(tmp_path, hip):
    exe = build_program(tmp_path)
    r = subprocess.run([exe, "--expect-device"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
