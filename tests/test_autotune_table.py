"""The tuning-table reader of libscir_hip.so (no device needed): a table in
the autotuner's format loads, malformed ones are rejected, overrides are
validated."""
from __future__ import annotations

from scir_amd import _lib

TABLE = """{"arch": "gfx950", "auto_mfma_min_k": 128,
 "entries": [
  {"k_lo": 1, "k_hi": 16, "n_lo": 0, "n_hi": 8192, "direct_r": 8, "direct_loader": 1, "direct_seg_tiles": 2, "mfma_tiles": 2, "mfma_tiles_fp16": 4},
  {"k_lo": 17, "k_hi": 64, "n_lo": 8193, "n_hi": 4611686018427387904, "direct_r": 16, "direct_loader": 0, "direct_seg_tiles": 1, "mfma_tiles": 1, "mfma_tiles_fp16": 2, "mfma_segs": 4}]}
"""


def test_load_valid_table(tmp_path):
    p = tmp_path / "tune.json"
    p.write_text(TABLE)
    L = _lib.lib()
    assert L.scir_hip_autotune_load(str(p).encode()) == _lib.OK


def test_reject_malformed(tmp_path):
    L = _lib.lib()
    p = tmp_path / "bad.json"
    p.write_text('{"arch": "gfx950", "entries": [{"k_lo": 1}]}')
    assert L.scir_hip_autotune_load(str(p).encode()) == _lib.ERR_INVALID_ARGUMENT
    assert b"malformed" in L.scir_hip_last_error()
    assert L.scir_hip_autotune_load(str(tmp_path / "missing.json").encode()) == _lib.ERR_INVALID_ARGUMENT
    p.write_text('{"entries": []}')
    assert L.scir_hip_autotune_load(str(p).encode()) == _lib.ERR_INVALID_ARGUMENT


def test_overrides_validated():
    L = _lib.lib()
    assert L.scir_hip_tune_set(b"direct.r", 12) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"mfma.tiles", 3) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"mfma.segs", 0) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"direct.dma_bufs", 5) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"direct.loader", 3) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"direct.dma_bufs", 2) == _lib.OK
    assert L.scir_hip_tune_set(b"direct.dma_bufs", -1) == _lib.OK
    assert L.scir_hip_tune_set(b"mfma.segs", 8) == _lib.OK
    assert L.scir_hip_tune_set(b"mfma.segs", -1) == _lib.OK
    assert L.scir_hip_tune_set(b"auto.fft", 2) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"auto.fft", 1) == _lib.OK
    assert L.scir_hip_tune_set(b"auto.fft", -1) == _lib.OK
    assert L.scir_hip_tune_set(b"auto.fft_min_k", 300) == _lib.OK
    assert L.scir_hip_tune_set(b"auto.fft_min_k", -1) == _lib.OK
    assert L.scir_hip_tune_set(b"vjp.db_mfma", 2) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"vjp.db_mfma", 0) == _lib.OK
    assert L.scir_hip_tune_set(b"vjp.db_mfma", -1) == _lib.OK
    assert L.scir_hip_tune_set(b"no.such.key", 1) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"direct.r", 16) == _lib.OK
    assert L.scir_hip_tune_set(b"direct.r", -1) == _lib.OK
