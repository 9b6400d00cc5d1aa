// Ragged (CSR) batches: signals of uneven length, each with its own taps,
// filtered by one persistent launch.
//
// Signal s is x[offs[s] .. offs[s+1]) with taps b[b_offs[s] .. b_offs[s+1])
// (lfilter order) and output y[offs[s] .. offs[s+1]).  A plan is built once
// on the host from the offsets (ragged_items.cpp): signals are cut into work
// items (runs of 2048-output tiles for the direct form, of 8192-output
// segments for MFMA), sorted longest first (LPT order); the item list and a
// work-queue counter live in device memory owned by the plan.
// Executing the plan launches a one-thread kernel that zeroes the counter
// (a kernel rather than a memset, so graph replays behave like eager runs)
// and then the persistent kernel: a tuned number of workgroups per CU, each
// taking the next item index with one atomic add until the queue is drained
// -- every workgroup leaves the loop on its first fetch past the end.  Items
// are filtered with the direct-form tile code of the dense kernels (same
// arithmetic, so the EXACT mode is bit-identical to the reference loop).
// With an MFMA precision (BF16x3 or FP16) the persistent kernel is the MFMA
// Toeplitz contraction of fir_mfma.hip, with the dense kernel's arithmetic.
// The launch shapes (workgroups per CU, work per item) come from the tuning
// keys ragged.* that scir_hip_autotune measures.
#include "direct_tiles.hpp"

#include <algorithm>
#include <memory>
#include <vector>

namespace scir {
namespace {

using namespace direct;

constexpr int kR = 8;                  // outputs per lane
constexpr int kT = kThreads * kR;      // outputs per tile
static_assert(kT == kRaggedDirectTile, "direct work items are cut into the kernel's tiles");

__global__ void ragged_reset_kernel(unsigned int *counter) { *counter = 0u; }

template <bool EXACT>
__global__ __launch_bounds__(kThreads) void fir_ragged_kernel(
    const float *__restrict__ x, const float *__restrict__ b, float *__restrict__ y,
    const RaggedItem *__restrict__ items, int n_items, unsigned int *__restrict__ counter)
{
    extern __shared__ __attribute__((aligned(16))) float xs[];
    __shared__ int s_item;
    const int lane_base = threadIdx.x * kR;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = s_item;
        __syncthreads();  // every lane has read s_item before it is written again
        if (it >= n_items) break;
        const RaggedItem w = items[it];
        const float *__restrict__ xrow = x + w.x0;
        float *__restrict__ yrow = y + w.x0;
        for (int tile = 0; tile < w.ntiles; ++tile) {
            const int64_t t0 = w.t0 + (int64_t)tile * kT;
            float acc[kR];
#pragma unroll
            for (int r = 0; r < kR; ++r) acc[r] = 0.0f;
            for (int ks = 0; ks < w.K; ks += kSuper) {
                const int kcount = min(kSuper, w.K - ks);
                const int H = (kcount + 6) & ~3;
                __syncthreads();  // previous stage finished with xs
                stage_row(xs, xrow, t0 - ks - H, kT + H + 4, w.len, false);
                __syncthreads();
                taps_from_lds<kR, EXACT>(acc, xs, H, lane_base, kcount, b + w.b0 + ks);
            }
            fix_ramp_in<kR, EXACT>(acc, xrow, w.len, false, b + w.b0, w.K, t0 + lane_base);
            store_outputs<kR>(acc, yrow, t0 + lane_base, w.len, false);
        }
    }
}

}  // namespace
}  // namespace scir

using namespace scir;

struct scir_hip_ragged_plan_s {
    int device = -1;
    int32_t precision = SCIR_HIP_PREC_F32;
    int64_t S = 0;
    std::vector<int64_t> offs, boffs;  // host copies (per-signal launches)
    int n_items = 0;
    RaggedItem *items = nullptr;       // device (direct form)
    MfmaRaggedItem *mitems = nullptr;  // device (MFMA)
    unsigned int *counter = nullptr;   // device
    int grid = 0;
    int kmax = 0;
    size_t lds = 0;

    ~scir_hip_ragged_plan_s()
    {
        if (items) (void)hipFree(items);
        if (mitems) (void)hipFree(mitems);
        if (counter) (void)hipFree(counter);
    }
};

namespace {

enum RaggedMode { kDirect, kDirectExact, kMfma };

// Kernel family of a precision: the FP32 direct form (f32 / auto, or the
// reference order with f32_exact) or the MFMA contraction (bf16x3 / fp16).
int ragged_mode(int32_t p, RaggedMode *mode)
{
    switch (p) {
    case SCIR_HIP_PREC_F32:
    case SCIR_HIP_PREC_AUTO: *mode = kDirect; return SCIR_HIP_OK;
    case SCIR_HIP_PREC_F32_EXACT: *mode = kDirectExact; return SCIR_HIP_OK;
    case SCIR_HIP_PREC_BF16X3:
    case SCIR_HIP_PREC_FP16: *mode = kMfma; return SCIR_HIP_OK;
    case SCIR_HIP_PREC_FFT32:
        return fail(SCIR_HIP_ERR_UNSUPPORTED,
                    "ragged batches run the direct form (f32 / f32_exact / auto) or the MFMA "
                    "contraction (bf16x3 / fp16)");
    default: return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "unknown precision %d", p);
    }
}

// Device copy of a host array (library-owned).
template <class T>
int upload(const std::vector<T> &h, T **d)
{
    if (h.empty()) return SCIR_HIP_OK;
    void *p = nullptr;
    SCIR_TRY(device_alloc(&p, sizeof(T) * h.size(), "ragged plan items"));
    *d = static_cast<T *>(p);
    SCIR_HIP_TRY(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return SCIR_HIP_OK;
}

}  // namespace

extern "C" {

int scir_hip_ragged_plan_create(const int64_t *offs, const int64_t *b_offs, int64_t S,
                                int32_t precision, scir_hip_ragged_plan *plan)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!plan) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "plan is null");
        *plan = nullptr;
        SCIR_TRY(validate_csr(offs, b_offs, S));
        if (precision == SCIR_HIP_PREC_AUTO) {
            // BF16x3 MFMA when the batch's taps (mean K weighted by
            // multiply-adds) reach the ragged cut-over and its LDS holds the
            // longest filter; the FP32 direct form otherwise.  Both keep the
            // elementwise FP32 bound.  The ragged F32 kernel is the VALU form,
            // so its cut-over is lower than the dense one (whose F32 runs on
            // FP32 MFMA up to K = 257).
            int kmean = 0, kmax_all = 0;
            int64_t nmean = 0;
            ragged_batch_stats(offs, b_offs, S, &kmean, &nmean, &kmax_all);
            size_t lds_need = 0;
            const bool fits = kmax_all > 0 &&
                              fir_mfma_ragged_lds(SCIR_HIP_PREC_BF16X3, kmax_all, &lds_need) ==
                                  SCIR_HIP_OK;
            clear_error();
            precision = (fits && kmean >= ragged_tune_lookup().auto_mfma_min_k)
                            ? SCIR_HIP_PREC_BF16X3
                            : SCIR_HIP_PREC_F32;
        }
        RaggedMode mode = kDirect;
        SCIR_TRY(ragged_mode(precision, &mode));
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        const RaggedTune rt = ragged_tune_lookup();

        std::vector<RaggedItem> items;
        std::vector<MfmaRaggedItem> mitems;
        int kmax = 0;
        SCIR_TRY(build_ragged_items(offs, b_offs, S, mode == kMfma, rt, &items, &mitems, &kmax));
        const size_t n_items = mode == kMfma ? mitems.size() : items.size();

        size_t lds = 0;
        int per_cu = rt.direct_wg_per_cu;
        if (mode == kMfma) {
            SCIR_TRY(fir_mfma_ragged_lds(precision, kmax, &lds));
            int fit = 1;
            SCIR_TRY(fir_mfma_ragged_occupancy(precision, lds, &fit));
            per_cu = rt.mfma_wg_per_cu > 0 ? std::min(rt.mfma_wg_per_cu, fit) : fit;
        } else {
            const int hmax = (std::min(std::max(kmax, 1), kSuper) + 6) & ~3;
            const int staged = kT + hmax + 4;
            lds = (size_t)(staged + (staged / 16) * 4 + 4) * sizeof(float);
        }

        host_alloc_point();
        std::unique_ptr<scir_hip_ragged_plan_s> p(new scir_hip_ragged_plan_s);
        p->device = st->device;
        p->precision = precision;
        p->S = S;
        p->offs.assign(offs, offs + S + 1);
        p->boffs.assign(b_offs, b_offs + S + 1);
        p->n_items = (int)n_items;
        p->kmax = kmax;
        p->lds = lds;
        p->grid = std::max(1, std::min(p->n_items, st->cu_count * std::max(1, per_cu)));
        void *counter = nullptr;
        SCIR_TRY(device_alloc(&counter, sizeof(unsigned int), "ragged plan queue counter"));
        p->counter = static_cast<unsigned int *>(counter);
        if (mode == kMfma)
            SCIR_TRY(upload(mitems, &p->mitems));
        else
            SCIR_TRY(upload(items, &p->items));
        *plan = p.release();
        return SCIR_HIP_OK;
    });
}

int scir_hip_ragged_plan_execute(scir_hip_ragged_plan plan, const float *x, const float *b,
                                 float *y, void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!plan) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "plan is null");
        if (plan->n_items == 0) return SCIR_HIP_OK;
        if (!x || !y || !b) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
        int dev = -1;
        SCIR_HIP_TRY(hipGetDevice(&dev));
        if (dev != plan->device)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "plan was made on device %d, current is %d",
                        plan->device, dev);
        RaggedMode mode = kDirect;
        SCIR_TRY(ragged_mode(plan->precision, &mode));
        hipStream_t s = as_stream(stream);
        hipLaunchKernelGGL(ragged_reset_kernel, dim3(1), dim3(1), 0, s, plan->counter);
        SCIR_TRY(check_launch("ragged_reset_kernel"));
        if (mode == kMfma) {
            DeviceState *st = nullptr;
            SCIR_TRY(current_device_state(&st));
            std::lock_guard<std::mutex> lk(st->mu);
            return fir_mfma_ragged_launch(x, b, y, plan->mitems, plan->n_items, plan->counter,
                                          plan->precision, plan->kmax, plan->lds, plan->grid, s);
        }
        if (mode == kDirectExact)
            hipLaunchKernelGGL(fir_ragged_kernel<true>, dim3(plan->grid), dim3(kThreads), plan->lds,
                               s, x, b, y, plan->items, plan->n_items, plan->counter);
        else
            hipLaunchKernelGGL(fir_ragged_kernel<false>, dim3(plan->grid), dim3(kThreads), plan->lds,
                               s, x, b, y, plan->items, plan->n_items, plan->counter);
        return check_launch("fir_ragged_kernel");
    });
}

int scir_hip_ragged_plan_execute_per_signal(scir_hip_ragged_plan plan, const float *x,
                                            const float *b, float *y, void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!plan) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "plan is null");
        if (!x || !y || !b) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
        RaggedMode mode = kDirect;
        SCIR_TRY(ragged_mode(plan->precision, &mode));
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        std::lock_guard<std::mutex> lk(st->mu);
        hipStream_t s = as_stream(stream);
        for (int64_t i = 0; i < plan->S; ++i) {
            const int64_t len = plan->offs[i + 1] - plan->offs[i];
            const int K = (int)(plan->boffs[i + 1] - plan->boffs[i]);
            if (len == 0) continue;
            if (mode == kMfma)
                SCIR_TRY(fir_mfma_dev(x + plan->offs[i], len, b + plan->boffs[i], K, 1, len,
                                      y + plan->offs[i], len, plan->precision, false, s, st));
            else
                SCIR_TRY(fir_direct_dev(x + plan->offs[i], len, b + plan->boffs[i], K, 1, len,
                                        y + plan->offs[i], len, mode == kDirectExact, false, s, st));
        }
        return SCIR_HIP_OK;
    });
}

int scir_hip_ragged_plan_destroy(scir_hip_ragged_plan plan)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!plan) return SCIR_HIP_OK;
        delete plan;  // frees the device item list and counter
        return SCIR_HIP_OK;
    });
}

int scir_hip_fir_ragged_f32_dev(const float *x, const int64_t *offs, const float *b,
                                const int64_t *b_offs, int64_t S, float *y, int32_t precision,
                                void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        scir_hip_ragged_plan plan = nullptr;
        SCIR_TRY(scir_hip_ragged_plan_create(offs, b_offs, S, precision, &plan));
        int st = scir_hip_ragged_plan_execute(plan, x, b, y, stream);
        if (st == SCIR_HIP_OK) {
            hipError_t e = hipStreamSynchronize(as_stream(stream));
            if (e != hipSuccess)
                st = fail(SCIR_HIP_ERR_BACKEND, "hipStreamSynchronize failed: %s",
                          hipGetErrorString(e));
        }
        std::string keep = st == SCIR_HIP_OK ? std::string() : std::string(scir_hip_last_error());
        scir_hip_ragged_plan_destroy(plan);
        if (st != SCIR_HIP_OK) return fail(st, "%s", keep.c_str());
        return SCIR_HIP_OK;
    });
}

}  // extern "C"
