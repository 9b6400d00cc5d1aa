// Host-only internals of libscir_hip.so shared by the HIP translation units
// and the plain C++ ones (the CPU arms, the tuning-table parser, the ragged
// plan builder), which build without HIP headers: error reporting, the guard
// that keeps C++ exceptions inside the library, allocation-failure
// injection, the tuning choice and the ragged work items.  Nothing here is
// part of the C ABI.
#pragma once

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/scir_hip.h"

namespace scir {

// Records `code` and a printf-style message as this thread's last error and
// returns `code`, so call sites read `return fail(...)`.  Never allocates.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Allocation-failure injection for tests of the error paths: while the
// environment variable SCIR_HIP_FAIL_ALLOC is set to a non-zero integer,
// every allocation the library makes for itself (device memory, pinned host
// memory and the host structures behind plans, graphs and buffers) fails.
bool alloc_failure_injected();
// Throws std::bad_alloc while injection is on; called where the library is
// about to allocate host memory.
void host_alloc_point();

// Runs the body of a C ABI entry point: no C++ exception leaves the library.
// std::bad_alloc becomes SCIR_HIP_ERR_OUT_OF_MEMORY, any other exception
// SCIR_HIP_ERR_BACKEND, each with a message naming the entry point.
template <class F>
int guarded(const char *entry, F &&body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return fail(SCIR_HIP_ERR_OUT_OF_MEMORY, "%s: host memory allocation failed", entry);
    } catch (const std::exception &e) {
        return fail(SCIR_HIP_ERR_BACKEND, "%s: %s", entry, e.what());
    } catch (...) {
        return fail(SCIR_HIP_ERR_BACKEND, "%s: unknown C++ exception", entry);
    }
}

#define SCIR_TRY(expr)              \
    do {                            \
        int s_ = (expr);            \
        if (s_ != SCIR_HIP_OK)      \
            return s_;              \
    } while (0)

// ---- tuning ---------------------------------------------------------------
// Autotuned choices (defaults until a table is loaded or a search ran).
struct TuneChoice {
    int direct_r = 8;         // outputs per lane in the direct kernel
    int direct_loader = 3;    // 0: one tile per workgroup, 1: register prefetch, 2: LDS DMA,
                              // 3: F32 mode on FP32 MFMA (K <= 513; otherwise as 2)
    int direct_mfma_strip = 0; // 256-output tiles per wave of loader 3 (0: by K)
    int direct_seg_tiles = 4; // tiles per workgroup of the pipelined variants
    int direct_dma_bufs = 2;  // LDS buffers of the DMA loader (tiles in flight + 1), 2..4
    int mfma_tiles = 2;       // 32x32 output tiles per wave, BF16x3 MFMA kernel
    int mfma_tiles_fp16 = 4;  // the same for the FP16 MFMA kernel
    int mfma_segs = 16;       // row segments per workgroup of the MFMA kernel
    bool mfma_segs_forced = false;  // set through scir_hip_tune_set: no grid-size cap
    int auto_mfma_min_k = 130; // AUTO uses BF16x3 MFMA from this K up (F32 on MFMA below)
    // AUTO may take the FFT path (normwise error) from auto_fft_min_k up,
    // when enabled with the tuning key auto.fft = 1 (off by default).
    bool auto_fft = false;
    bool vjp_db_mfma = true;  // dL/db on FP32 MFMA (false: the FP32 VALU kernel)
    bool resample_fixed = true;  // up 2, down 3, 31 taps: the register-window kernel
    int auto_fft_min_k = 512; // measured: FFT32 beats BF16x3 MFMA from here (K <= 8192)
};

// Launch shapes of ragged (CSR) batches: one set for the whole batch,
// chosen by the autotuner (tuning keys ragged.*).
struct RaggedTune {
    int direct_wg_per_cu = 8;         // persistent workgroups per CU, FP32 direct form
    int direct_min_item_kmacs = 512;  // multiply-adds per direct work item, at least (x1024)
    int mfma_segs_per_item = 8;       // segments per MFMA work item, at most
    int mfma_wg_per_cu = 0;           // persistent workgroups per CU, MFMA (0: occupancy)
    int auto_mfma_min_k = 56;         // AUTO plans take BF16x3 from this weighted mean K up
};

// One row of the tuning table: a (K, N) bucket and its launch shapes.
struct TuneEntry {
    int k_lo, k_hi;
    int64_t n_lo, n_hi;
    TuneChoice c;
};

// A parsed tuning table (the JSON written by scir_hip_autotune).
struct TuneTable {
    std::string arch;
    std::vector<TuneEntry> entries;
    int64_t auto_mfma_min_k = -1;  // -1: absent
    int64_t auto_fft_min_k = -1;   // -1: absent
    bool has_ragged = false;
    RaggedTune ragged;
};

// Parses `text`; on a malformed table returns SCIR_HIP_ERR_INVALID_ARGUMENT
// with a message naming `source`.
int parse_tune_table(const std::string &text, const char *source, TuneTable *out);
std::string tune_table_json(const TuneTable &t);

// ---- ragged (CSR) work items ------------------------------------------------
constexpr int kRaggedDirectTile = 2048;  // outputs per tile of the direct form (256 lanes x 8)
constexpr int kMfmaRaggedSegLen = 8192;  // outputs per segment (4 waves x 2 tiles x 32 x 32)

// Direct form: consecutive tiles of one signal.
struct RaggedItem {
    int64_t x0;     // first sample of the signal in x / y
    int64_t t0;     // first output of the item within the signal
    int64_t len;    // signal length
    int64_t b0;     // first tap of the signal in b
    int32_t K;      // taps of the signal
    int32_t ntiles; // consecutive tiles in the item
};

// MFMA: segments seg0 .. seg0+nseg-1 (kMfmaRaggedSegLen outputs each) of the
// signal x[x0 .. x0+len) with taps b[b0 .. b0+K) (lfilter order).
struct MfmaRaggedItem {
    int64_t x0, len, b0;
    int32_t K, seg0, nseg, pad;
};

// Checks CSR offsets (offs[S+1], b_offs[S+1], host arrays).
int validate_csr(const int64_t *offs, const int64_t *b_offs, int64_t S);

// Cuts the signals of a validated CSR batch into work items, longest first
// (LPT order by multiply-adds, or MFMA half-steps x segments): direct-form
// items when `mfma` is false, MFMA items otherwise.  `kmax` receives the
// longest filter of a non-empty signal.
int build_ragged_items(const int64_t *offs, const int64_t *b_offs, int64_t S, bool mfma,
                       const RaggedTune &tune, std::vector<RaggedItem> *items,
                       std::vector<MfmaRaggedItem> *mitems, int *kmax);

// Mean tap count of a batch weighted by multiply-adds, its mean length and
// its longest filter (of a non-empty signal): the inputs of AUTO's choice.
void ragged_batch_stats(const int64_t *offs, const int64_t *b_offs, int64_t S, int *kmean,
                        int64_t *nmean, int *kmax);

}  // namespace scir
