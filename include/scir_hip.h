/*
 * scir_hip.h -- C ABI of the MI355X (gfx950) batched-FIR library
 *               (scir_amd/libscir_hip.so).
 *
 * This is the drop-in boundary for SciR's batched FIR path:
 *   scir_gpu::fir1d_batched_f32_auto(x, taps, Device)   crates/scir-gpu/src/lib.rs:515-531
 *     -> cuda::fir1d_batched_f32_cuda(x, taps)          crates/scir-gpu/src/lib.rs:1036-1113
 *     -> CPU fir1d_batched_f32(x, taps)                 crates/scir-gpu/src/lib.rs:1134-1152
 * A Rust `-sys` crate binds these symbols one to one; the `Device::Hip` arm of
 * the dispatcher calls scir_hip_fir_f32_host().
 *
 * Conventions shared by every entry point
 *   - All functions return an int status (scir_hip_status); 0 is success.  On
 *     failure scir_hip_last_error() returns a message naming the failing call
 *     (the thread's last error), like GpuError::BackendUnavailable(String)
 *     (crates/scir-gpu/src/lib.rs:583-589).  No C++ exception crosses the ABI:
 *     a failed host allocation inside the library returns
 *     SCIR_HIP_ERR_OUT_OF_MEMORY, any other internal exception
 *     SCIR_HIP_ERR_BACKEND.  For tests of these paths, every allocation the
 *     library makes for itself fails while the environment variable
 *     SCIR_HIP_FAIL_ALLOC is set to a non-zero integer.
 *   - Non-finite values: every precision gives a non-finite output exactly
 *     where the reference loop (crates/scir-gpu/src/lib.rs:1141-1148, which
 *     forms taps[K-1-t] * x[i-t] for t <= min(i, K-1) only) gives one, with
 *     the same infinities; F32_EXACT equals it bit for bit, NaN for NaN.  The
 *     kernels stage zeros before the row start (direct form), multiply zero
 *     taps of Toeplitz blocks (MFMA) or transform whole segments (FFT), which
 *     spreads an Inf or NaN to neighbouring outputs; each output that comes
 *     out non-finite is recomputed from the reference loop's terms.
 *   - Taps are passed as `b` in lfilter order: y[n] = sum_k b[k] * x[n-k] with
 *     x[<0] = 0 (zero initial state), output length == input length.  The
 *     reference's `taps` argument is b reversed (b[k] == taps[K-1-k]); the
 *     host-side dispatcher reverses it before the call.
 *   - 2-D arrays are row-major with an explicit row stride in elements
 *     (ldx >= N); rows are independent signals.
 *   - `*_host` entry points take and return host memory and are synchronous.
 *     `*_dev` entry points take device pointers and a hipStream_t passed as
 *     void* (NULL = the null stream); they enqueue work and return without
 *     synchronising.  Callers own all memory they pass in.
 *   - Thread safety: every entry point may be called from any thread.  Work
 *     that shares library-owned scratch (the dL/db partial sums, the
 *     host-tier staging buffers) is serialised per device; the dL/db partials
 *     are kept per stream.
 */
#ifndef SCIR_HIP_H
#define SCIR_HIP_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define SCIR_HIP_API __attribute__((visibility("default")))
#else
#define SCIR_HIP_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
typedef enum scir_hip_status {
    SCIR_HIP_OK = 0,
    SCIR_HIP_ERR_INVALID_ARGUMENT = 1, /* shapes/strides/pointers (GpuError::ShapeMismatch) */
    SCIR_HIP_ERR_NO_DEVICE = 2,        /* no gfx950 HIP device (GpuError::BackendUnavailable) */
    SCIR_HIP_ERR_BACKEND = 3,          /* a HIP runtime call failed; message names it */
    SCIR_HIP_ERR_OUT_OF_MEMORY = 4,    /* device or pinned-host allocation failed */
    SCIR_HIP_ERR_UNSUPPORTED = 5,      /* valid request this build does not implement */
    SCIR_HIP_ERR_CAPTURE = 6           /* the call would allocate during stream capture */
} scir_hip_status;

/* ---- arithmetic selection --------------------------------------------- */
typedef enum scir_hip_precision {
    /* FP32 arithmetic, any summation order; error within
     * (K+1) * 2^-23 * sum_k |b_k||x_{n-k}| of the exact sum.  K <= 513 runs
     * as an FP32 Toeplitz contraction on the matrix cores
     * (v_mfma_f32_16x16x4_f32, twice the non-packed VALU FMA rate), longer
     * filters as the direct form with fused multiply-adds on the VALU.  Both
     * recompute any non-finite output from the reference loop's terms, so
     * outputs are non-finite exactly where the reference loop's are. */
    SCIR_HIP_PREC_F32 = 0,
    /* FP32 direct form with a separately rounded multiply and add, taps
     * accumulated in the reference's order (k = 0, 1, ..., K-1 starting from
     * +0.0f): bit-identical to crates/scir-gpu/src/lib.rs:1134-1152. */
    SCIR_HIP_PREC_F32_EXACT = 1,
    /* Toeplitz-block contraction on MFMA: x and b split into three bf16 terms
     * (hi+mid+lo), six cross products kept, FP32 accumulation; within
     * (K+1) * 2^-23 * sum_k |b_k||x_{n-k}| of the exact sum like F32.
     * K <= 6689 (FP16: K <= 23457); longer filters give
     * SCIR_HIP_ERR_UNSUPPORTED.  Values beyond the bf16 (fp16) range and
     * non-finite values take the recomputation described above; so does
     * every output of a segment of 8192 whose samples or taps include a
     * nonzero magnitude below 2^-110 (or whose smallest sample and tap
     * magnitudes multiply to below 2^-100), where the split terms would
     * leave the normal range: there the bound holds with FP32's subnormal
     * floor, (K+1) * 2^-149 added. */
    SCIR_HIP_PREC_BF16X3 = 2,
    /* Toeplitz-block contraction on MFMA with one fp16 product (x and b
     * rounded to fp16), FP32 accumulation.  Reported, not accuracy-bound. */
    SCIR_HIP_PREC_FP16 = 3,
    /* Fastest of F32 / BF16X3 for the shape (autotuned table when loaded,
     * else built-in cut-over; F32 for K > 6689, beyond the MFMA kernel's
     * LDS); never FP16, and FFT32 only when the "auto.fft" key is set.  Either
     * choice keeps the F32 bound and the reference loop's non-finite
     * outputs, so AUTO is the drop-in default. */
    SCIR_HIP_PREC_AUTO = 4,
    /* FP32 FFT overlap-save: cost O(log K) per output instead of O(K) for
     * K <= 8192; longer filters run as blocks of 8192 taps (partitioned
     * overlap-save, one pass over the rows per block, outputs summed).  The
     * error is normwise (about log2(NF) * 2^-24 * |x|_2 |b|_1 over each
     * NF = 4096, 8192 or 16384 sample segment), not elementwise.  x and y
     * must not be the same array. */
    SCIR_HIP_PREC_FFT32 = 5
} scir_hip_precision;

/* ---- runtime ------------------------------------------------------------ */
/* Version of this ABI (major*10000 + minor*100 + patch): 10100. */
SCIR_HIP_API int scir_hip_version(void);
/* Initialise the runtime on the current HIP device (idempotent).  Replaces the
 * per-call CudaCtx::create_default + cuModuleLoadData of
 * crates/scir-gpu/src/lib.rs:601-622, 814-824: state is created once and kept. */
SCIR_HIP_API int scir_hip_init(void);
/* Release library-owned buffers on every device (idempotent).  HIP graphs
 * that captured library calls must be destroyed first. */
SCIR_HIP_API int scir_hip_shutdown(void);
/* Number of visible HIP devices (0 when none), like cuda_available
 * (crates/scir-gpu/src/lib.rs:591-599); never fails for a missing device. */
SCIR_HIP_API int scir_hip_device_count(int *count);
/* Message of this thread's last failing call ("" if none).  Valid until the
 * thread's next call into the library. */
SCIR_HIP_API const char *scir_hip_last_error(void);

/* ---- dense batched FIR ---------------------------------------------------- */
/* Host tier: drop-in for cuda::fir1d_batched_f32_cuda
 * (crates/scir-gpu/src/lib.rs:1036-1113).  x[B, ldx] and y[B, ldy] in host
 * memory, b[K] host; copies in, filters, copies out, synchronises.  B or N == 0
 * is a no-op; K == 0 writes zeros (the CPU loop's empty sum).  Batches of
 * three chunks of rows or more (tuning key host.chunk_kib) are staged through
 * pinned host memory by host.copy_threads threads per direction, with the
 * copy in of one chunk, the kernels of another and the copy out of a third
 * overlapping on three streams; the result does not depend on the chunking.
 * The library keeps four pinned blocks of one chunk each until shutdown. */
SCIR_HIP_API int scir_hip_fir_f32_host(const float *x, int64_t ldx, const float *b, int32_t K,
                          int64_t B, int64_t N, float *y, int64_t ldy,
                          int32_t precision);

/* Device tier of the same operation (PTX kernel fir1d_batched_f32,
 * crates/scir-gpu/src/lib.rs:727-811, and its launch at :1083-1099), for
 * device-resident callers (torch tensors, benchmarks, graph capture). */
SCIR_HIP_API int scir_hip_fir_f32_dev(const float *x, int64_t ldx, const float *b, int32_t K,
                         int64_t B, int64_t N, float *y, int64_t ldy,
                         int32_t precision, void *stream);

/* CPU arm of the dispatcher: the reference's CPU fir1d_batched_f32
 * (crates/scir-gpu/src/lib.rs:1134-1152) in the same arithmetic order (so the
 * result is bit-identical), with rows spread over `nthreads` host threads
 * (<= 0: all hardware threads).  Needs no device. */
SCIR_HIP_API int scir_hip_fir_f32_cpu(const float *x, int64_t ldx, const float *b, int32_t K,
                         int64_t B, int64_t N, float *y, int64_t ldy,
                         int32_t nthreads);

/* ---- device-resident arrays -------------------------------------------------
 * The HIP arm of scir_gpu::DeviceArray (crates/scir-gpu/src/lib.rs:76-189),
 * whose to_device only tags host storage: a buffer owns device memory on the
 * current device until destroyed, so chained FIR and elementwise calls on it
 * need no host transfers.  Upload/download are synchronous (null stream). */
typedef struct scir_hip_buffer_s *scir_hip_buffer;
SCIR_HIP_API int scir_hip_buffer_create(size_t bytes, scir_hip_buffer *buf);
SCIR_HIP_API int scir_hip_buffer_upload(scir_hip_buffer buf, const void *host, size_t bytes);
SCIR_HIP_API int scir_hip_buffer_download(scir_hip_buffer buf, void *host, size_t bytes);
/* Device address of the buffer, for the *_dev entry points (NULL for NULL). */
SCIR_HIP_API void *scir_hip_buffer_device_ptr(scir_hip_buffer buf);
/* Waits for the device, then frees. */
SCIR_HIP_API int scir_hip_buffer_destroy(scir_hip_buffer buf);
/* Elementwise float32 on device pointers, n elements, enqueued on `stream`:
 * DeviceArray::add_scalar / mul_scalar / add (crates/scir-gpu/src/lib.rs:206-260);
 * one IEEE operation per element, so results equal the reference's bit for bit. */
SCIR_HIP_API int scir_hip_add_scalar_f32_dev(const float *a, float alpha, float *out, int64_t n,
                                             void *stream);
SCIR_HIP_API int scir_hip_mul_scalar_f32_dev(const float *a, float alpha, float *out, int64_t n,
                                             void *stream);
SCIR_HIP_API int scir_hip_add_f32_dev(const float *a, const float *b, float *out, int64_t n,
                                      void *stream);

/* ---- polyphase resampling (float64) -----------------------------------------
 * Replaces the reference's CPU-only resample_poly
 * (crates/scir-signal/src/lib.rs:313-362: x upsampled by `up` with zeros,
 * fully convolved with a K-tap filter h, every `down`-th sample kept from
 * off = (K-1)/2 while below len(conv) - off), generalised to batches of rows
 * and to any up, down >= 1 and h (the reference fixes up = 2, down = 3 and a
 * 31-tap table).  Row r of y holds M outputs,
 *   y[r, m] = sum_{k = n mod up, n = off + down*m; k < K; k += up} h[k] * x[r, (n-k)/up]
 * with x[r, s] = 0 outside [0, N).  M is given by scir_hip_resample_len. */
SCIR_HIP_API int scir_hip_resample_len(int64_t N, int32_t up, int32_t down, int32_t K,
                                       int64_t *M);
/* Device tier: x [B, ldx >= N], y [B, ldy >= M] float64 device pointers,
 * h [K] float64 device pointer (K <= 8192), enqueued on `stream`. */
SCIR_HIP_API int scir_hip_resample_poly_f64_dev(const double *x, int64_t ldx, int64_t B,
                                                int64_t N, int32_t up, int32_t down,
                                                const double *h, int32_t K, double *y,
                                                int64_t ldy, void *stream);
/* Host tier: host arrays in and out (library-owned device buffers, copies
 * and launch on the library's stream, synchronous), like scir_hip_fir_f32_host. */
SCIR_HIP_API int scir_hip_resample_poly_f64_host(const double *x, int64_t ldx, int64_t B,
                                                 int64_t N, int32_t up, int32_t down,
                                                 const double *h, int32_t K, double *y,
                                                 int64_t ldy);
/* CPU arm: the reference's operation order (each output accumulates h[k] *
 * x terms in ascending x index, multiply and add rounded separately), so
 * for the reference's filter table the result is bit-identical to
 * resample_poly; rows spread over `nthreads` host threads (<= 0: all). */
SCIR_HIP_API int scir_hip_resample_poly_f64_cpu(const double *x, int64_t ldx, int64_t B,
                                                int64_t N, int32_t up, int32_t down,
                                                const double *h, int32_t K, double *y,
                                                int64_t ldy, int32_t nthreads);

/* ---- IIR filtering by second-order sections (float64) ----------------------
 * Replaces the reference's CPU-only sosfilt and filtfilt
 * (crates/scir-signal/src/lib.rs:260-291), batched over rows.  Each row is
 * filtered by the Direct Form II Transposed cascade of
 * crates/scir-iir-filters/src/filter.rs:47-62 from zero state.  sos holds S
 * sections of six doubles {b0, b1, b2, a0, a1, a2} (the #[repr(C)] layout
 * of SosCoeffs, and scipy's sos rows); a0 is not read (sections are taken as
 * normalised, as in the reference).  filtfilt runs the cascade forward, then
 * backward over that output from zero state (the reference's zero-phase
 * filter, no padding).  0 <= S <= SCIR_HIP_SOS_MAX_SECTIONS; S = 0 copies.
 * flags: 0 splits rows into chunks joined through the state transition
 * (parallel in time; within a few ulps of the reference loop), or
 * SCIR_HIP_SOS_REFERENCE_ORDER: one chunk per row, every output equal to the
 * reference loop bit for bit (parallel over rows only). */
#define SCIR_HIP_SOS_MAX_SECTIONS 16
#define SCIR_HIP_SOS_REFERENCE_ORDER 1
/* Device tier: x [rows, ldx >= n], y [rows, ldy >= n] float64 device
 * pointers (y == x allowed), sos [S * 6] device pointer, on `stream`. */
SCIR_HIP_API int scir_hip_sosfilt_f64_dev(const double *x, int64_t ldx, int64_t rows, int64_t n,
                                          const double *sos, int32_t S, double *y, int64_t ldy,
                                          int32_t flags, void *stream);
SCIR_HIP_API int scir_hip_filtfilt_f64_dev(const double *x, int64_t ldx, int64_t rows, int64_t n,
                                           const double *sos, int32_t S, double *y, int64_t ldy,
                                           int32_t flags, void *stream);
/* Host tier: host arrays in and out, synchronous (like scir_hip_fir_f32_host). */
SCIR_HIP_API int scir_hip_sosfilt_f64_host(const double *x, int64_t ldx, int64_t rows, int64_t n,
                                           const double *sos, int32_t S, double *y, int64_t ldy,
                                           int32_t flags);
SCIR_HIP_API int scir_hip_filtfilt_f64_host(const double *x, int64_t ldx, int64_t rows,
                                            int64_t n, const double *sos, int32_t S, double *y,
                                            int64_t ldy, int32_t flags);
/* CPU arm: the reference loop per row (bit-identical to sosfilt / filtfilt),
 * rows spread over `nthreads` host threads (<= 0: all). */
SCIR_HIP_API int scir_hip_sosfilt_f64_cpu(const double *x, int64_t ldx, int64_t rows, int64_t n,
                                          const double *sos, int32_t S, double *y, int64_t ldy,
                                          int32_t nthreads);
SCIR_HIP_API int scir_hip_filtfilt_f64_cpu(const double *x, int64_t ldx, int64_t rows, int64_t n,
                                           const double *sos, int32_t S, double *y, int64_t ldy,
                                           int32_t nthreads);

/* ---- ragged (CSR) batches -------------------------------------------------
 * No counterpart in the reference (its FIR takes one dense Array2).  Signal s
 * is x[offs[s] .. offs[s+1]) with taps b[b_offs[s] .. b_offs[s+1]); y has the
 * same layout as x.  A plan holds the host-built, cost-sorted work list and
 * the device work queue; executing it is one persistent launch.  Executing a
 * plan does not allocate, so it may be captured in a HIP graph.
 * precision: F32 (FP32 direct form), F32_EXACT (the reference loop's order,
 * bitwise), AUTO (resolved when the plan is made: BF16X3 when the batch's
 * multiply-add weighted tap count reaches the tuned MFMA cut-over and the
 * longest filter fits, F32 otherwise), or BF16X3 / FP16 (the MFMA Toeplitz
 * contraction with the dense MFMA kernel's arithmetic, segments of 8192
 * outputs; taps up to the LDS limit: K <= 5601 for BF16X3).  FFT32 gives
 * SCIR_HIP_ERR_UNSUPPORTED.  Launch shapes (persistent workgroups per CU,
 * work per item) come from the tuning keys ragged.* when the plan is made.
 * Each plan owns one device work-queue counter: at most one execution of a
 * plan may be in flight at a time (executions on one stream are ordered;
 * on several streams, or a graph replay beside an eager execute, they race
 * and lose or repeat items), and a plan must outlive every graph that
 * captured its execution. */
typedef struct scir_hip_ragged_plan_s *scir_hip_ragged_plan;
/* offs[S+1], b_offs[S+1]: HOST arrays (int64, non-decreasing, offs[0] ==
 * b_offs[0] == 0). */
SCIR_HIP_API int scir_hip_ragged_plan_create(const int64_t *offs, const int64_t *b_offs, int64_t S,
                                int32_t precision, scir_hip_ragged_plan *plan);
SCIR_HIP_API int scir_hip_ragged_plan_execute(scir_hip_ragged_plan plan, const float *x,
                                 const float *b, float *y, void *stream);
SCIR_HIP_API int scir_hip_ragged_plan_destroy(scir_hip_ragged_plan plan);
/* One-shot convenience: create + execute + destroy (synchronises the stream
 * before freeing).  offs/b_offs are HOST arrays; x, b, y are device pointers. */
SCIR_HIP_API int scir_hip_fir_ragged_f32_dev(const float *x, const int64_t *offs, const float *b,
                                const int64_t *b_offs, int64_t S, float *y,
                                int32_t precision, void *stream);
/* Launch one plan per signal instead (S launches of the dense kernel) -- the
 * baseline a persistent launch is measured against. */
SCIR_HIP_API int scir_hip_ragged_plan_execute_per_signal(scir_hip_ragged_plan plan, const float *x,
                                            const float *b, float *y, void *stream);

/* ---- adjoints (vector-Jacobian products) ----------------------------------
 * No counterpart in the reference.  For L = sum_{r,n} g[r,n] * y[r,n] with
 * y = lfilter(b,[1],x) per row:
 *   dx[r,m] = sum_k b[k] * g[r,m+k]           (terms with m+k < N)
 *   db[k]   = sum_{r,n} g[r,n] * x[r,n-k]     (terms with n-k >= 0)
 * In the reference's taps order the tap gradient is db reversed. */
SCIR_HIP_API int scir_hip_fir_vjp_dx_dev(const float *g, int64_t ldg, const float *b, int32_t K,
                            int64_t B, int64_t N, float *dx, int64_t lddx,
                            int32_t precision, void *stream);
/* db[K] is overwritten (device pointer).  Two-stage deterministic reduction:
 * per-tile FP32 partial sums, then a fixed-order compensated sum. */
SCIR_HIP_API int scir_hip_fir_vjp_db_dev(const float *g, int64_t ldg, const float *x, int64_t ldx,
                            int32_t K, int64_t B, int64_t N, float *db, void *stream);
/* The same dL/db for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Nothing is allocated: the partial sums
 * live in a fixed 4 MiB device array (1048576 floats) that the library holds
 * on every GPU it runs on until the process ends; call scir_hip_init() on the
 * device once first (outside capture).  Limit: B * K <= 1048576, else
 * SCIR_HIP_ERR_INVALID_ARGUMENT with a message.  The result equals
 * scir_hip_fir_vjp_db_dev's bit for bit whenever that call's B * segments * K
 * partial sums fit; larger rows use fewer, longer workgroups (the same
 * deterministic two-stage reduction in another grouping).  Every call on a
 * GPU shares that GPU's array, which keeps the last call's contents: calls
 * that overlap on one GPU (on different streams, or a graph replay beside an
 * eager call) race and must be ordered by the caller. */
SCIR_HIP_API int scir_hip_fir_vjp_db_noalloc_dev(const float *g, int64_t ldg, const float *x,
                                                 int64_t ldx, int32_t K, int64_t B, int64_t N,
                                                 float *db, void *stream);

/* Host tier of both adjoints (the Rust vjp API's entry): host arrays in and
 * out, synchronous, like scir_hip_fir_f32_host.  dx [B, lddx >= N] and db[K]
 * are overwritten; dx uses `precision` (as scir_hip_fir_vjp_dx_dev), db the
 * deterministic FP32 reduction. */
SCIR_HIP_API int scir_hip_fir_vjp_f32_host(const float *x, int64_t ldx, const float *g,
                                           int64_t ldg, const float *b, int32_t K, int64_t B,
                                           int64_t N, float *dx, int64_t lddx, float *db,
                                           int32_t precision);

/* ---- HIP graphs ------------------------------------------------------------
 * Capture any sequence of *_dev / ragged-plan calls enqueued on `stream`
 * between begin and end, then replay it.  Calls that would need to grow
 * library scratch during capture fail with SCIR_HIP_ERR_CAPTURE: run the same
 * call once eagerly first.  A graph keeps the addresses of the scratch its
 * calls used (the dL/db partials, the FFT filter spectrum, the IIR chunk
 * states of the stream): when a later eager call on that stream needs more
 * scratch, the library allocates a larger block and keeps the old one
 * until scir_hip_shutdown, so the graph stays valid.  A replay and an eager
 * call on the same stream are ordered; calls on other streams use their own
 * scratch. */
typedef struct scir_hip_graph_s *scir_hip_graph;
SCIR_HIP_API int scir_hip_graph_begin(void *stream);
SCIR_HIP_API int scir_hip_graph_end(void *stream, scir_hip_graph *graph);
SCIR_HIP_API int scir_hip_graph_launch(scir_hip_graph graph, void *stream);
SCIR_HIP_API int scir_hip_graph_destroy(scir_hip_graph graph);

/* ---- autotuner -------------------------------------------------------------
 * Grid-search tile / wavefront shapes and the F32-vs-MFMA cut-over on the
 * current device for a set of (K, N) buckets, timing each candidate with HIP
 * events, and write the winners as JSON to cache_path (NULL: keep in memory
 * only).  budget_ms bounds the total search time (<= 0: default). */
SCIR_HIP_API int scir_hip_autotune(const char *cache_path, int32_t budget_ms);
/* Load a table written by scir_hip_autotune (entries for another arch are
 * ignored). */
SCIR_HIP_API int scir_hip_autotune_load(const char *cache_path);
/* Force one tuning parameter for every later call in this process, overriding
 * the table; value < 0 removes the override.  Keys:
 *   "direct.r"        outputs per lane of the direct kernel (8 or 16)
 *   "direct.loader"   0 = one tile per workgroup, 1 = pipelined row segments
 *                     (register prefetch), 2 = pipelined row segments copied
 *                     into LDS by DMA two tiles ahead (forward, 16-byte aligned
 *                     rows with N % 4 == 0; other calls use 1), 3 (default) =
 *                     F32 on FP32 MFMA for K <= 513 (F32_EXACT and longer
 *                     filters as 2)
 *   "direct.mfma_strip" 256-output tiles per wave of loader 3 (1 .. 4096;
 *                     default 4 for K <= 8, 8 to K = 40, 16 to K = 56, 32 to
 *                     K = 257, else 64)
 *   "direct.seg_tiles" tiles per workgroup of the pipelined direct kernel
 *   "direct.dma_bufs" LDS tile buffers of loader 2 (2, 3 or 4: 1 to 3 tiles in flight)
 *   "mfma.tiles"      32x32 output tiles per wave of the MFMA kernel (1, 2, 4)
 *   "mfma.segs"       row segments per workgroup of the MFMA kernel (>= 1); the
 *                     next segment's samples load while the current one computes
 *   "vjp.db_mfma"     1 (default): dL/db on FP32 MFMA; 0: the FP32 VALU kernel
 *   "resample.fixed"  1 (default): up 2, down 3 with 31 taps (the reference's
 *                     only resample_poly case) runs the register-window
 *                     kernel; 0: the general polyphase kernel (same sums)
 *   "auto.mfma_min_k" smallest K for which AUTO picks BF16x3 MFMA (default
 *                     130: F32 on FP32 MFMA is faster below; ragged plans,
 *                     whose F32 kernel is the VALU form: 56, by the batch's
 *                     multiply-add weighted mean K)
 *   "auto.fft"        1: AUTO may pick the FFT32 path (normwise error) from
 *                     auto.fft_min_k up (x != y); 0 (default): never
 *   "auto.fft_min_k"  smallest K for which AUTO (with auto.fft) picks FFT32;
 *                     the autotuner measures it (the K from which FFT32 beats
 *                     BF16x3 MFMA on long rows)
 *   "host.chunk_kib"  KiB of samples per chunk of the pipelined host tier
 *                     (default 16384): batches of three chunks or more move
 *                     in, through the kernels and out in overlapping chunks
 *   "host.copy_threads" threads per direction that copy the host tier's
 *                     chunks between the caller's arrays and pinned staging
 *                     memory (1 .. 64, default 4)
 *   "ragged.direct_wg_per_cu"      persistent workgroups per CU of ragged
 *                                  batches in the FP32 direct form (1 .. 32)
 *   "ragged.direct_min_item_kmacs" multiply-adds per direct-form work item,
 *                                  at least, in units of 1024
 *   "ragged.mfma_segs_per_item"    8192-output segments per MFMA work item,
 *                                  at most (1 .. 1024)
 *   "ragged.mfma_wg_per_cu"        persistent workgroups per CU of ragged
 *                                  MFMA batches (0: as many as fit, 1 .. 8)
 * The ragged keys apply to plans made after the call.  The autotuner's
 * search uses the same mechanism. */
SCIR_HIP_API int scir_hip_tune_set(const char *key, int32_t value);

/* ---- synthetic inputs -------------------------------------------------------
 * Fill x[B, ldx] (device) with N(0,1) float32 samples from a counter-based
 * generator (row r uses the stream seed ^ (r * 0x9E3779B97F4A7C15)); a
 * fraction `outlier_frac` of samples, chosen by an independent hash bit
 * stream, is multiplied by `outlier_scale`. */
SCIR_HIP_API int scir_hip_synth_normal_f32(float *x, int64_t ldx, int64_t B, int64_t N, uint64_t seed,
                              float outlier_frac, float outlier_scale, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SCIR_HIP_H */
