"""Scratch memory of the built kernels, read from the gfx950 code objects that
libscir_hip.so carries (offload bundles, AMDGPU metadata notes).  A kernel
that starts using scratch (private segment) has spilled registers or lost a
register array to the stack: the BF16x3 back-to-front kernel did once, after
a change in how staging formed its sample pairs, and ran at two thirds of its
speed with no other sign.  No GPU is needed."""
from __future__ import annotations

import os
import struct

import msgpack
import pytest

from scir_amd import _lib

# Kernels allowed a private segment, and why.
ALLOWED_SCRATCH = {
    # sosfilt's chunk-state scan (run only when few rows are cut into chunks;
    # one short pass over the chunks' end states) keeps the 2S-element state
    # vectors of 8 sections (and of 9 to 16, SMAX = 16) partly in scratch.
    # The per-sample kernels use none.
    "sos_scan_kernel<8>": "_ZN4scir12_GLOBAL__N_115sos_scan_kernelILi8EEEvPdPKdlli",
    "sos_scan_kernel<16>": "_ZN4scir12_GLOBAL__N_115sos_scan_kernelILi16EEEvPdPKdlli",
}


def _code_objects(blob: bytes):
    """(target id, ELF bytes) of every device code object in the library."""
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    i = blob.find(magic)
    while i >= 0:
        (n,) = struct.unpack_from("<Q", blob, i + len(magic))
        p = i + len(magic) + 8
        for _ in range(n):
            off, size, idl = struct.unpack_from("<QQQ", blob, p)
            p += 24
            ident = blob[p:p + idl].decode()
            p += idl
            if "amdgcn" in ident and size:
                yield ident, blob[i + off:i + off + size]
        i = blob.find(magic, i + len(magic))


def _amdgpu_metadata(elf: bytes):
    """The NT_AMDGPU_METADATA note (msgpack) of one ELF64 code object."""
    assert elf[:4] == b"\x7fELF" and elf[4] == 2 and elf[5] == 1
    (shoff,) = struct.unpack_from("<Q", elf, 0x28)
    shentsize, shnum = struct.unpack_from("<HH", elf, 0x3A)
    for k in range(shnum):
        sh = shoff + k * shentsize
        (stype,) = struct.unpack_from("<I", elf, sh + 4)
        if stype != 7:  # SHT_NOTE
            continue
        off, size = struct.unpack_from("<QQ", elf, sh + 0x18)
        p = off
        while p < off + size:
            namesz, descsz, ntype = struct.unpack_from("<III", elf, p)
            p += 12
            name = elf[p:p + namesz].rstrip(b"\0")
            p += (namesz + 3) & ~3
            desc = elf[p:p + descsz]
            p += (descsz + 3) & ~3
            if name == b"AMDGPU" and ntype == 32:
                return msgpack.unpackb(desc, raw=False, strict_map_key=False)
    raise AssertionError("code object without AMDGPU metadata")


@pytest.fixture(scope="module")
def kernels():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip(f"{_lib.LIB_PATH} is not built")
    blob = open(_lib.LIB_PATH, "rb").read()
    out = {}
    for ident, elf in _code_objects(blob):
        assert ident.endswith("gfx950"), ident
        for k in _amdgpu_metadata(elf)["amdhsa.kernels"]:
            out[k[".name"]] = k
    return out


def test_every_translation_unit_has_kernels(kernels):
    names = " ".join(kernels)
    for part in ("fir_mfma_kernel", "fir_mfma_ragged_kernel", "fir_direct_mfma_kernel",
                 "fir_direct_dma_kernel", "fir_fft_kernel", "fir_vjp_db_mfma_kernel",
                 "resample_poly_kernel", "sos_wave_kernel"):
        assert part in names, part


def test_no_kernel_uses_scratch(kernels):
    allowed = set(ALLOWED_SCRATCH.values())
    offenders = {n: (k[".private_segment_fixed_size"], k.get(".vgpr_spill_count", 0))
                 for n, k in kernels.items()
                 if (k[".private_segment_fixed_size"] or k.get(".vgpr_spill_count", 0))
                 and n not in allowed}
    assert not offenders, offenders


def test_allowed_scratch_list_is_current(kernels):
    # Each listed kernel exists and still needs its exception.
    for label, sym in ALLOWED_SCRATCH.items():
        assert sym in kernels, label
        assert kernels[sym][".private_segment_fixed_size"] > 0, label
