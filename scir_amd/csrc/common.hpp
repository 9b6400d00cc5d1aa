// Shared internals of libscir_hip.so: error reporting, per-device state and
// library-owned scratch buffers.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>

#include "../../include/scir_hip.h"

namespace scir {

// Records `code` and a printf-style message as this thread's last error and
// returns `code`, so call sites read `return fail(...)`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define SCIR_HIP_TRY(call)                                                            \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::scir::fail(e_ == hipErrorOutOfMemory ? SCIR_HIP_ERR_OUT_OF_MEMORY \
                                                          : SCIR_HIP_ERR_BACKEND,     \
                                "%s failed: %s", #call, hipGetErrorString(e_));       \
    } while (0)

#define SCIR_TRY(expr)              \
    do {                            \
        int s_ = (expr);            \
        if (s_ != SCIR_HIP_OK)      \
            return s_;              \
    } while (0)

// A grow-only device allocation owned by the library.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// Scratch that one stream's calls share.  Calls enqueued on the same stream
// are ordered by the stream, so they may reuse the buffers; calls on other
// streams get their own set.
struct StreamScratch {
    DeviceBuffer partials;    // per-workgroup dL/db partial sums
    DeviceBuffer sos_states;  // per-chunk IIR states and the chunk transition
    DeviceBuffer fft_spectrum;  // the filter's spectrum for the FFT path
};

struct DeviceState {
    int device = -1;
    bool ready = false;
    int cu_count = 0;
    int xcd_count = 8;
    size_t lds_per_block = 0;
    std::string arch;
    std::mutex mu;  // guards everything below
    std::map<void *, StreamScratch> scratch;
    // Host tier: device copies of the caller's host arrays.
    DeviceBuffer host_x, host_y, host_b;
    DeviceBuffer host_z;  // third matrix of the host-tier adjoint (dL/dx)
    hipStream_t host_stream = nullptr;
};

// Initialises (once) and returns the state of the current device.
int current_device_state(DeviceState **out);

// Makes `buf` hold at least `bytes`.  Growing frees the old block; a growth
// requested while `stream` is capturing fails with SCIR_HIP_ERR_CAPTURE.
int ensure_buffer(DeviceBuffer &buf, size_t bytes, hipStream_t stream);

// Scratch of `stream` on device state `st` (caller holds st->mu).
StreamScratch &scratch_for(DeviceState *st, hipStream_t stream);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Validates the dense-call arguments shared by the FIR entry points.
int check_dense_args(const void *x, int64_t ldx, const void *b, int32_t K, int64_t B,
                     int64_t N, const void *y, int64_t ldy);

// Launch-status check after a kernel launch.
int check_launch(const char *what);
// Lets kernel `fn` take `bytes` of dynamic LDS on the current device (set
// once per device and kernel; thread-safe).
int allow_dynamic_lds(const void *fn, int bytes);

// ---- dense entry points shared across translation units -----------------
// `rev` runs the filter back to front: y[N-1-i] = sum_k b[k] x[N-1-i+k], the
// anti-causal correlation that gives dL/dx.
int fir_direct_dev(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                   float *y, int64_t ldy, bool exact, bool rev, hipStream_t stream,
                   DeviceState *st);
int fir_fft_dev(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                float *y, int64_t ldy, bool rev, hipStream_t stream, DeviceState *st);
int fir_mfma_dev(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                 float *y, int64_t ldy, int precision, bool rev, hipStream_t stream,
                 DeviceState *st);
// Whether the MFMA kernel's LDS holds K taps (one tile per wave).
bool fir_mfma_takes(int precision, int K);
int fir_dense_dispatch(const float *x, int64_t ldx, const float *b, int32_t K, int64_t B,
                       int64_t N, float *y, int64_t ldy, int32_t precision, bool rev,
                       hipStream_t stream, DeviceState *st);
int fill_zero_2d(float *y, int64_t ldy, int64_t B, int64_t N, hipStream_t stream);

// ---- ragged (CSR) batches on MFMA ----------------------------------------
// One work item: segments seg0 .. seg0+nseg-1 (8192 outputs each) of the
// signal x[x0 .. x0+len) with taps b[b0 .. b0+K) (lfilter order).
struct MfmaRaggedItem {
    int64_t x0, len, b0;
    int32_t K, seg0, nseg, pad;
};
constexpr int kMfmaRaggedSegLen = 8192;  // outputs per segment (4 waves x 2 tiles x 32 x 32)
// Dynamic LDS bytes of the ragged MFMA kernel for taps up to kmax; fails with
// SCIR_HIP_ERR_UNSUPPORTED when that exceeds the CU's LDS.
int fir_mfma_ragged_lds(int precision, int kmax, size_t *lds);
// Resident workgroups per CU of the ragged MFMA kernel at `lds` bytes.
int fir_mfma_ragged_occupancy(int precision, size_t lds, int *per_cu);
// Persistent launch over `items` (pulled through `counter`, which must read 0).
int fir_mfma_ragged_launch(const float *x, const float *b, float *y, const MfmaRaggedItem *items,
                           int n_items, unsigned int *counter, int precision, int kmax,
                           size_t lds, int grid, hipStream_t stream);

// Autotuned choices (defaults until a table is loaded or a search ran).
struct TuneChoice {
    int direct_r = 8;         // outputs per lane in the direct kernel
    int direct_loader = 2;    // 0: one tile per workgroup, 1: register prefetch, 2: LDS DMA
    int direct_seg_tiles = 4; // tiles per workgroup of the pipelined variants
    int direct_dma_bufs = 2;  // LDS buffers of the DMA loader (tiles in flight + 1), 2..4
    int mfma_tiles = 2;       // 32x32 output tiles per wave, BF16x3 MFMA kernel
    int mfma_tiles_fp16 = 4;  // the same for the FP16 MFMA kernel
    int mfma_segs = 16;       // row segments per workgroup of the MFMA kernel
    bool mfma_segs_forced = false;  // set through scir_hip_tune_set: no grid-size cap
    int auto_mfma_min_k = 56; // AUTO uses BF16x3 MFMA from this K up
    // AUTO may take the FFT path (normwise error) from auto_fft_min_k up,
    // when enabled with the tuning key auto.fft = 1 (off by default).
    bool auto_fft = false;
    bool vjp_db_mfma = true;  // dL/db on FP32 MFMA (false: the FP32 VALU kernel)
    int auto_fft_min_k = 512; // measured: FFT32 beats BF16x3 MFMA from here (K <= 8192)
};
TuneChoice tune_lookup(int K, int64_t N);

}  // namespace scir
