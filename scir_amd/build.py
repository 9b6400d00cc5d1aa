"""Build libscir_hip.so in-tree (scir_amd/libscir_hip.so) for gfx950.

Every ``*.hip`` under scir_amd/csrc is compiled by hipcc for gfx950 only;
the host-only ``*.cpp`` files (the CPU arms, error reporting, the tuning
table parser, the ragged plan builder) by g++ with FP contraction off, since
the CPU arms must keep the reference's separately rounded multiply and add.  Objects go
to scir_amd/_build/ and the shared library is linked with hipcc.

Run as ``python -m scir_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libscir_hip.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=fast-honor-pragmas", "-munsafe-fp-atomics",
    # Packed FP32 (v_pk_*_f32) is kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-ffp-contract=off",
             "-fno-fast-math", "-Wall", "-I" + os.path.join(ROOT, "include")]


def _sources():
    names = sorted(os.listdir(CSRC))
    return ([os.path.join(CSRC, n) for n in names if n.endswith(".hip")],
            [os.path.join(CSRC, n) for n in names if n.endswith(".cpp")])


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that note is noise.
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "not a recognized feature for this target" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hips, cpps = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    cmds = []
    objs = []
    for src in hips:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        objs.append(obj)
    for src in cpps:  # host-only code (no HIP headers): g++
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmds.append([CXX, *CXX_FLAGS, "-c", src, "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp, *objs,
          "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
