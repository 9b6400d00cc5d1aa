"""The tuning-table reader of libscir_hip.so (no device needed): a table in
the autotuner's format loads, malformed ones are rejected, overrides are
validated."""
from __future__ import annotations

from scir_amd import _lib

TABLE = """{"arch": "gfx950", "auto_mfma_min_k": 128,
 "entries": [
  {"k_lo": 1, "k_hi": 16, "n_lo": 0, "n_hi": 8192, "direct_r": 8, "direct_loader": 1, "direct_seg_tiles": 2, "mfma_tiles": 2, "mfma_tiles_fp16": 4},
  {"k_lo": 17, "k_hi": 64, "n_lo": 8193, "n_hi": 4611686018427387904, "direct_r": 16, "direct_loader": 0, "direct_seg_tiles": 1, "mfma_tiles": 1, "mfma_tiles_fp16": 2, "mfma_segs": 4},
  {"k_lo": 65, "k_hi": 256, "n_lo": 0, "n_hi": 4611686018427387904, "direct_r": 8, "direct_loader": 3, "direct_seg_tiles": 4, "direct_mfma_strip": 32, "mfma_tiles": 2, "mfma_tiles_fp16": 4}]}
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
    assert L.scir_hip_tune_set(b"direct.loader", 4) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"direct.loader", 3) == _lib.OK
    assert L.scir_hip_tune_set(b"direct.loader", -1) == _lib.OK
    assert L.scir_hip_tune_set(b"direct.mfma_strip", 0) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"direct.mfma_strip", 4097) == _lib.ERR_INVALID_ARGUMENT
    assert b"direct.mfma_strip" in L.scir_hip_last_error()
    assert L.scir_hip_tune_set(b"direct.mfma_strip", 16) == _lib.OK
    assert L.scir_hip_tune_set(b"direct.mfma_strip", -1) == _lib.OK
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
    assert L.scir_hip_tune_set(b"resample.fixed", 2) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"resample.fixed", 0) == _lib.OK
    assert L.scir_hip_tune_set(b"resample.fixed", -1) == _lib.OK
    assert L.scir_hip_tune_set(b"no.such.key", 1) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"direct.r", 16) == _lib.OK
    assert L.scir_hip_tune_set(b"direct.r", -1) == _lib.OK
    assert L.scir_hip_tune_set(b"host.chunk_kib", 0) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"host.copy_threads", 0) == _lib.ERR_INVALID_ARGUMENT
    assert b"host.copy_threads" in L.scir_hip_last_error()
    assert L.scir_hip_tune_set(b"host.copy_threads", 65) == _lib.ERR_INVALID_ARGUMENT
    assert L.scir_hip_tune_set(b"host.copy_threads", 64) == _lib.OK
    assert L.scir_hip_tune_set(b"host.copy_threads", -1) == _lib.OK


RAGGED = """{"arch": "gfx950", "auto_mfma_min_k": 128, "auto_fft_min_k": 384,
 "ragged": {"direct_wg_per_cu": 4, "direct_min_item_kmacs": 2048, "mfma_segs_per_item": 16,
            "mfma_wg_per_cu": 1},
 "entries": []}
"""


def test_load_table_with_ragged_shapes(tmp_path):
    p = tmp_path / "tune.json"
    p.write_text(RAGGED)
    L = _lib.lib()
    assert L.scir_hip_autotune_load(str(p).encode()) == _lib.OK
    p.write_text('{"arch": "gfx950", "ragged": {"direct_wg_per_cu": 4, "entries": []')
    assert L.scir_hip_autotune_load(str(p).encode()) == _lib.ERR_INVALID_ARGUMENT
    assert b'malformed table ("ragged")' in L.scir_hip_last_error()


def test_malformed_message_names_the_missing_key(tmp_path):
    L = _lib.lib()
    p = tmp_path / "t.json"
    p.write_text('{"arch": "gfx950", "entries": [{"k_lo": 1}]}')
    assert L.scir_hip_autotune_load(str(p).encode()) == _lib.ERR_INVALID_ARGUMENT
    assert b'entry without "k_hi"' in L.scir_hip_last_error()
    p.write_text('{"arch": "gfx950", "entries": [{"k_lo": 9, "k_hi": 2, "n_lo": 0, "n_hi": 1, '
                 '"direct_r": 8, "direct_loader": 1, "direct_seg_tiles": 2, "mfma_tiles": 2, '
                 '"mfma_tiles_fp16": 4}]}')
    assert L.scir_hip_autotune_load(str(p).encode()) == _lib.ERR_INVALID_ARGUMENT
    assert b"empty bucket" in L.scir_hip_last_error()


def test_truncated_tables_never_crash(tmp_path):
    """Every prefix of a valid table either loads or is rejected with a
    message (the reader never reads past the text)."""
    L = _lib.lib()
    p = tmp_path / "t.json"
    for cut in range(0, len(TABLE), 7):
        p.write_text(TABLE[:cut])
        st = L.scir_hip_autotune_load(str(p).encode())
        assert st in (_lib.OK, _lib.ERR_INVALID_ARGUMENT)
        if st != _lib.OK:
            assert L.scir_hip_last_error()
    p.write_text(TABLE)
    assert L.scir_hip_autotune_load(str(p).encode()) == _lib.OK


def test_ragged_overrides_validated():
    L = _lib.lib()
    for key, bad, good in ((b"ragged.direct_wg_per_cu", 0, 8), (b"ragged.direct_wg_per_cu", 33, 4),
                           (b"ragged.direct_min_item_kmacs", 0, 256),
                           (b"ragged.mfma_segs_per_item", 0, 4),
                           (b"ragged.mfma_segs_per_item", 2000, 32),
                           (b"ragged.mfma_wg_per_cu", 9, 1)):
        assert L.scir_hip_tune_set(key, bad) == _lib.ERR_INVALID_ARGUMENT
        assert L.scir_hip_tune_set(key, good) == _lib.OK
        assert L.scir_hip_tune_set(key, -1) == _lib.OK
