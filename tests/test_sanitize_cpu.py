"""The library's host-only code (CPU arms, tuning-table reader and writer,
ragged work-item builder, the C ABI's exception guard and allocation-failure
injection) built with g++ under AddressSanitizer and UndefinedBehaviorSanitizer
and run on random and edge-case inputs (tests/cpp/sanitize_test.cpp).  No
device code is involved: the HIP translation units are not part of this
build."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "scir_amd", "csrc")
HOST_SOURCES = ["cpu_fir.cpp", "errors.cpp", "tune_table.cpp", "ragged_items.cpp"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_host_code_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_test")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "sanitize_test.cpp"),
           *[os.path.join(CSRC, s) for s in HOST_SOURCES], "-lpthread", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ)
    env.pop("SCIR_HIP_FAIL_ALLOC", None)
    # a preloaded library of the environment may precede the ASan runtime
    env["ASAN_OPTIONS"] = "detect_leaks=1:verify_asan_link_order=0:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "sanitize: 0 failures" in r.stdout
