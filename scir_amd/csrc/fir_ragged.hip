// Ragged (CSR) batches: signals of uneven length, each with its own taps,
// filtered by one persistent launch.
//
// Signal s is x[offs[s] .. offs[s+1]) with taps b[b_offs[s] .. b_offs[s+1])
// (lfilter order) and output y[offs[s] .. offs[s+1]).  A plan is built once
// on the host from the offsets:
//   * each signal is cut into work items of one or more consecutive tiles of
//     T = 2048 outputs (enough tiles per item that an item carries at least
//     kMinItemWork multiply-adds, so the queue overhead stays small);
//   * items are sorted by cost (outputs x taps) in decreasing order, so the
//     longest work starts first and short items fill the tail (LPT order);
//   * the item list and a work-queue counter live in device memory owned by
//     the plan.
// Executing the plan launches a one-thread kernel that zeroes the counter
// (a kernel rather than a memset, so graph replays behave like eager runs)
// and then the persistent kernel: about 8 workgroups per CU, each taking the
// next item index with one atomic add until the queue is drained -- every
// workgroup leaves the loop on its first fetch past the end.  Items are
// filtered with the direct-form tile code of the dense kernels (same
// arithmetic, so the EXACT mode is bit-identical to the reference loop).
// With an MFMA precision (BF16x3 or FP16) the items are runs of 8192-output
// segments and the persistent kernel is the MFMA Toeplitz contraction of
// fir_mfma.hip, with the same split arithmetic as the dense MFMA kernel.
#include "direct_tiles.hpp"

#include <algorithm>
#include <numeric>
#include <vector>

namespace scir {
namespace {

using namespace direct;

constexpr int kR = 8;                  // outputs per lane
constexpr int kT = kThreads * kR;      // outputs per tile
constexpr int64_t kMinItemWork = (int64_t)kT * 256;  // multiply-adds per item, at least

struct RaggedItem {
    int64_t x0;     // first sample of the signal in x / y
    int64_t t0;     // first output of the item within the signal
    int64_t len;    // signal length
    int64_t b0;     // first tap of the signal in b
    int32_t K;      // taps of the signal
    int32_t ntiles; // consecutive tiles in the item
};

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
        const bool vx = (reinterpret_cast<uintptr_t>(xrow) & 15) == 0;
        const bool vy = (reinterpret_cast<uintptr_t>(yrow) & 15) == 0;
        for (int tile = 0; tile < w.ntiles; ++tile) {
            const int64_t t0 = w.t0 + (int64_t)tile * kT;
            float acc[kR];
#pragma unroll
            for (int r = 0; r < kR; ++r) acc[r] = 0.0f;
            for (int ks = 0; ks < w.K; ks += kSuper) {
                const int kcount = min(kSuper, w.K - ks);
                const int H = (kcount + 6) & ~3;
                __syncthreads();  // previous stage finished with xs
                stage_row(xs, xrow, t0 - ks - H, kT + H + 4, w.len, vx, false);
                __syncthreads();
                taps_from_lds<kR, EXACT>(acc, xs, H, lane_base, kcount, b + w.b0 + ks);
            }
            store_outputs<kR>(acc, yrow, t0 + lane_base, w.len, vy, false);
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
};

namespace {

int validate_csr(const int64_t *offs, const int64_t *b_offs, int64_t S)
{
    if (S < 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative signal count");
    if (!offs || !b_offs) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null offsets");
    if (offs[0] != 0 || b_offs[0] != 0)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "offsets must start at 0");
    for (int64_t s = 0; s < S; ++s) {
        if (offs[s + 1] < offs[s] || b_offs[s + 1] < b_offs[s])
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "offsets decrease at signal %lld",
                        (long long)s);
        if (b_offs[s + 1] - b_offs[s] > 0x7fffffff)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "signal %lld has too many taps",
                        (long long)s);
    }
    return SCIR_HIP_OK;
}

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

constexpr int kMfmaSegsPerItem = 8;  // segments of 8192 outputs per MFMA work item, at most

}  // namespace

extern "C" {

int scir_hip_ragged_plan_create(const int64_t *offs, const int64_t *b_offs, int64_t S,
                                int32_t precision, scir_hip_ragged_plan *plan)
{
    clear_error();
    if (!plan) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "plan is null");
    *plan = nullptr;
    SCIR_TRY(validate_csr(offs, b_offs, S));
    if (precision == SCIR_HIP_PREC_AUTO) {
        // As for dense batches: BF16x3 MFMA when the batch's taps (mean K
        // weighted by multiply-adds) reach the tuned cut-over and its LDS holds
        // the longest filter; the FP32 direct form otherwise.  Both keep the
        // elementwise FP32 bound.
        double macs = 0.0, kmacs = 0.0, nmean = 0.0;
        int kmax_all = 0;
        for (int64_t s = 0; s < S; ++s) {
            const double len = (double)(offs[s + 1] - offs[s]);
            const double K = (double)(b_offs[s + 1] - b_offs[s]);
            macs += len * K;
            kmacs += len * K * K;
            nmean += len;
            if (len > 0) kmax_all = std::max(kmax_all, (int)(b_offs[s + 1] - b_offs[s]));
        }
        const int kmean = macs > 0.0 ? (int)(kmacs / macs) : 0;
        size_t lds_need = 0;
        const bool fits = kmax_all > 0 &&
                          fir_mfma_ragged_lds(SCIR_HIP_PREC_BF16X3, kmax_all, &lds_need) ==
                              SCIR_HIP_OK;
        clear_error();
        const TuneChoice tc = tune_lookup(std::max(kmean, 1), S > 0 ? (int64_t)(nmean / S) : 0);
        precision = (fits && kmean >= tc.auto_mfma_min_k) ? SCIR_HIP_PREC_BF16X3
                                                           : SCIR_HIP_PREC_F32;
    }
    RaggedMode mode = kDirect;
    SCIR_TRY(ragged_mode(precision, &mode));
    DeviceState *st = nullptr;
    SCIR_TRY(current_device_state(&st));

    // Work items and their costs (multiply-adds, or MFMA half-steps x segments).
    std::vector<RaggedItem> items;
    std::vector<MfmaRaggedItem> mitems;
    std::vector<int64_t> cost;
    int kmax = 0;
    for (int64_t s = 0; s < S; ++s) {
        const int64_t len = offs[s + 1] - offs[s];
        const int K = (int)(b_offs[s + 1] - b_offs[s]);
        if (len == 0) continue;
        kmax = std::max(kmax, K);
        if (mode == kMfma) {
            const int64_t segs = (len + kMfmaRaggedSegLen - 1) / kMfmaRaggedSegLen;
            if (segs > 0x7fffffff)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "signal %lld is too long",
                            (long long)s);
            const int64_t per_seg = (K + 14) / 16 + 2 + 2;  // half-steps of the loop
            for (int64_t g = 0; g < segs; g += kMfmaSegsPerItem) {
                MfmaRaggedItem it{};
                it.x0 = offs[s];
                it.len = len;
                it.b0 = b_offs[s];
                it.K = K;
                it.seg0 = (int32_t)g;
                it.nseg = (int32_t)std::min<int64_t>(kMfmaSegsPerItem, segs - g);
                mitems.push_back(it);
                cost.push_back((int64_t)it.nseg * per_seg);
            }
            continue;
        }
        const int64_t tiles = (len + kT - 1) / kT;
        const int64_t per_tile = (int64_t)kT * std::max(K, 1);
        const int64_t tpi = std::max<int64_t>(1, std::min<int64_t>(tiles, kMinItemWork / per_tile));
        for (int64_t t = 0; t < tiles; t += tpi) {
            RaggedItem it;
            it.x0 = offs[s];
            it.t0 = t * kT;
            it.len = len;
            it.b0 = b_offs[s];
            it.K = K;
            it.ntiles = (int)std::min<int64_t>(tpi, tiles - t);
            items.push_back(it);
            cost.push_back((int64_t)it.ntiles * per_tile);
        }
    }
    const size_t n_items = mode == kMfma ? mitems.size() : items.size();
    if (n_items > 0x7fffffff) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "too many items");
    // Longest first (LPT order).
    std::vector<size_t> order(n_items);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t c) { return cost[a] > cost[c]; });

    size_t lds = 0;
    int per_cu = 8;
    if (mode == kMfma) {
        SCIR_TRY(fir_mfma_ragged_lds(precision, kmax, &lds));
        SCIR_TRY(fir_mfma_ragged_occupancy(precision, lds, &per_cu));
    } else {
        const int hmax = (std::min(std::max(kmax, 1), kSuper) + 6) & ~3;
        const int staged = kT + hmax + 4;
        lds = (size_t)(staged + (staged / 16) * 4 + 4) * sizeof(float);
    }

    auto *p = new scir_hip_ragged_plan_s;
    p->device = st->device;
    p->precision = precision;
    p->S = S;
    p->offs.assign(offs, offs + S + 1);
    p->boffs.assign(b_offs, b_offs + S + 1);
    p->n_items = (int)n_items;
    p->kmax = kmax;
    p->lds = lds;
    p->grid = std::max(1, std::min(p->n_items, st->cu_count * per_cu));
    hipError_t e = hipMalloc(&p->counter, sizeof(unsigned int));
    if (e == hipSuccess && n_items > 0) {
        if (mode == kMfma) {
            std::vector<MfmaRaggedItem> sorted(n_items);
            for (size_t i = 0; i < n_items; ++i) sorted[i] = mitems[order[i]];
            e = hipMalloc(&p->mitems, sizeof(MfmaRaggedItem) * n_items);
            if (e == hipSuccess)
                e = hipMemcpy(p->mitems, sorted.data(), sizeof(MfmaRaggedItem) * n_items,
                              hipMemcpyHostToDevice);
        } else {
            std::vector<RaggedItem> sorted(n_items);
            for (size_t i = 0; i < n_items; ++i) sorted[i] = items[order[i]];
            e = hipMalloc(&p->items, sizeof(RaggedItem) * n_items);
            if (e == hipSuccess)
                e = hipMemcpy(p->items, sorted.data(), sizeof(RaggedItem) * n_items,
                              hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) {
        scir_hip_ragged_plan_destroy(p);
        return fail(e == hipErrorOutOfMemory ? SCIR_HIP_ERR_OUT_OF_MEMORY : SCIR_HIP_ERR_BACKEND,
                    "ragged plan upload failed: %s", hipGetErrorString(e));
    }
    *plan = p;
    return SCIR_HIP_OK;
}

int scir_hip_ragged_plan_execute(scir_hip_ragged_plan plan, const float *x, const float *b,
                                 float *y, void *stream)
{
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
}

int scir_hip_ragged_plan_execute_per_signal(scir_hip_ragged_plan plan, const float *x,
                                            const float *b, float *y, void *stream)
{
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
}

int scir_hip_ragged_plan_destroy(scir_hip_ragged_plan plan)
{
    clear_error();
    if (!plan) return SCIR_HIP_OK;
    if (plan->items) (void)hipFree(plan->items);
    if (plan->mitems) (void)hipFree(plan->mitems);
    if (plan->counter) (void)hipFree(plan->counter);
    delete plan;
    return SCIR_HIP_OK;
}

int scir_hip_fir_ragged_f32_dev(const float *x, const int64_t *offs, const float *b,
                                const int64_t *b_offs, int64_t S, float *y, int32_t precision,
                                void *stream)
{
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
}

}  // extern "C"
