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
#include <vector>

#include "host_common.hpp"

namespace scir {

#define SCIR_HIP_TRY(call)                                                            \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::scir::fail(e_ == hipErrorOutOfMemory ? SCIR_HIP_ERR_OUT_OF_MEMORY \
                                                          : SCIR_HIP_ERR_BACKEND,     \
                                "%s failed: %s", #call, hipGetErrorString(e_));       \
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
    // Scratch blocks outgrown while a captured graph may still hold their
    // addresses: kept until scir_hip_shutdown.
    std::vector<void *> retired;
    DeviceBuffer fft_twiddles[3];  // FFT path tables for NF = 2^12, 2^13, 2^14
    // Host tier: device copies of the caller's host arrays, and pinned host
    // staging buffers for the chunked copies.
    DeviceBuffer host_x, host_y, host_b;
    DeviceBuffer host_z;  // third matrix of the host-tier adjoint (dL/dx)
    // Device address of the allocation-free dL/db scratch array (looked up
    // once when the state is made; null, with the reason, if that failed).
    float *vjp_noalloc_scratch = nullptr;
    std::string vjp_noalloc_error;
    void *pinned[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t pinned_bytes = 0;
    hipStream_t host_stream = nullptr;
    hipStream_t host_copy_stream[2] = {nullptr, nullptr};
};

// Initialises (once) and returns the state of the current device.
int current_device_state(DeviceState **out);

// hipMalloc of library-owned device memory (fails while allocation-failure
// injection is on).
int device_alloc(void **ptr, size_t bytes, const char *what);

// Makes `buf` hold at least `bytes` of a host-tier buffer (never captured):
// growing frees the old block.
int ensure_buffer(DeviceBuffer &buf, size_t bytes, hipStream_t stream);

// Makes the scratch block `buf` of device state `st` hold at least `bytes`.
// A HIP graph captured earlier may hold the old block's address, so growing
// retires it (freed at scir_hip_shutdown) instead of freeing it; sizes grow
// at least twofold, so retired blocks add up to less than the live one.  A
// growth requested while `stream` is capturing fails with
// SCIR_HIP_ERR_CAPTURE.  The caller holds st->mu.
int ensure_scratch(DeviceState *st, DeviceBuffer &buf, size_t bytes, hipStream_t stream);

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
// Partial sums (floats) the allocation-free dL/db entry point holds: 4 MiB.
constexpr size_t kVjpNoAllocPartials = (size_t)1 << 20;
// Device address of that entry point's scratch array on the current device.
int vjp_noalloc_scratch_address(void **ptr);

// Whether the MFMA kernel's LDS holds K taps (one tile per wave).
bool fir_mfma_takes(int precision, int K);
int fir_dense_dispatch(const float *x, int64_t ldx, const float *b, int32_t K, int64_t B,
                       int64_t N, float *y, int64_t ldy, int32_t precision, bool rev,
                       hipStream_t stream, DeviceState *st);
int fill_zero_2d(float *y, int64_t ldy, int64_t B, int64_t N, hipStream_t stream);

// ---- ragged (CSR) batches on MFMA ----------------------------------------
// Dynamic LDS bytes of the ragged MFMA kernel for taps up to kmax; fails with
// SCIR_HIP_ERR_UNSUPPORTED when that exceeds the CU's LDS.
int fir_mfma_ragged_lds(int precision, int kmax, size_t *lds);
// Resident workgroups per CU of the ragged MFMA kernel at `lds` bytes.
int fir_mfma_ragged_occupancy(int precision, size_t lds, int *per_cu);
// Persistent launch over `items` (pulled through `counter`, which must read 0).
int fir_mfma_ragged_launch(const float *x, const float *b, float *y, const MfmaRaggedItem *items,
                           int n_items, unsigned int *counter, int precision, int kmax,
                           size_t lds, int grid, hipStream_t stream);

TuneChoice tune_lookup(int K, int64_t N);
RaggedTune ragged_tune_lookup();
// Bytes of samples per chunk of the pipelined host tier (tuning key
// host.chunk_kib; 16 MiB by default).
size_t host_chunk_bytes();
// Threads per direction for the host tier's staging copies through pinned
// memory (tuning key host.copy_threads; 4 by default).
int host_copy_threads();

}  // namespace scir
