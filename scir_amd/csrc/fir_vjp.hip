// Adjoint (vector-Jacobian product) kernels of the batched FIR.
//
// For L = sum_{r,n} g[r,n] * y[r,n], y = lfilter(b, [1], x) per row:
//   dx[r,m] = sum_k b[k] * g[r,m+k]            (terms with m+k < N)
//   db[k]   = sum_{r,n} g[r,n] * x[r,n-k]      (terms with n-k >= 0)
//
// dx is the forward filter run back to front (logical sample i == physical
// N-1-i), so it reuses the dense direct/MFMA kernels through their `rev` mode
// and inherits every precision mode.
//
// db is a correlation reduced over all rows and samples.  Stage 1 runs on
// FP32 MFMA by default (fir_vjp_db_mfma_kernel, below); the tuning key
// "vjp.db_mfma" = 0 selects the FP32 VALU form: a
// workgroup walks `tiles_per_wg` tiles of T = 2048 samples of one row,
// staging g and x (with K samples of history) in LDS.  Lane (kg, tg) owns 16
// lags k0 + kg*16 + (0..15) and 128 samples tg*128 + (0..127) of a tile; it
// slides a 32-sample x window against 16 g values (256 FMAs per 12 LDS
// reads), keeps a per-tile FP32 sum and adds it to a per-workgroup FP32
// total.  The 16 time groups are reduced in a fixed order through LDS into
// partial[wg][k].  Stage 2 sums the partials over workgroups with Kahan
// compensation in a fixed order.  Nothing is atomic: results are
// deterministic.  Lags beyond 256 are handled in passes of 256.
#include "common.hpp"

#include <algorithm>

namespace scir {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
// Four floats at any 4-byte aligned global address (unaligned access mode).
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

constexpr int kThreads = 256;
constexpr int kTile = 2048;   // samples per tile
constexpr int kLagPass = 256; // lags per pass (16 lag groups x 16 lags)

__device__ __forceinline__ int pad16(int s) { return s + ((s >> 4) << 2); }

__device__ __forceinline__ f32x4 load4g(const float *__restrict__ row, int64_t i, int64_t N)
{
    if (i >= 0 && i + 3 < N) return *reinterpret_cast<const f32x4u *>(row + i);
    f32x4 v;
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = (i + c >= 0 && i + c < N) ? row[i + c] : 0.0f;
    return v;
}

__global__ __launch_bounds__(kThreads) void fir_vjp_db_partial_kernel(
    const float *__restrict__ g, int64_t ldg, const float *__restrict__ x, int64_t ldx, int K,
    int64_t N, int segs_per_row, int tiles_per_wg, int hist, float *__restrict__ partial)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    float *gs = sm;                                 // kTile floats
    float *xs = gs + kTile;                         // pad16(kTile + hist) floats
    float *red = xs + pad16(kTile + hist + 32) + 4; // 16 x kLagPass floats

    const int64_t row = blockIdx.x / segs_per_row;
    const int64_t tile0 = (int64_t)(blockIdx.x % segs_per_row) * tiles_per_wg;
    const float *__restrict__ grow = g + row * ldg;
    const float *__restrict__ xrow = x + row * ldx;
    const int kg = threadIdx.x & 15, tg = threadIdx.x >> 4;
    const int64_t tiles_in_row = (N + kTile - 1) / kTile;
    const int ntiles = (int)min<int64_t>(tiles_per_wg, tiles_in_row - tile0);

    for (int k_pass = 0; k_pass < K; k_pass += kLagPass) {
        const int k0 = k_pass + kg * 16;
        float tot[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) tot[c] = 0.0f;
        for (int it = 0; it < ntiles; ++it) {
            const int64_t t0 = (tile0 + it) * kTile;
            // hist >= k_pass + kLagPass + 16 samples of history before t0
            const int64_t xbase = t0 - hist;
            __syncthreads();  // previous tile done with gs/xs
            for (int v = threadIdx.x; v < kTile / 4; v += kThreads)
                *reinterpret_cast<f32x4 *>(gs + 4 * v) = load4g(grow, t0 + 4 * v, N);
            for (int v = threadIdx.x; v < (kTile + hist) / 4; v += kThreads)
                *reinterpret_cast<f32x4 *>(xs + pad16(4 * v)) =
                    load4g(xrow, xbase + 4 * v, N);
            __syncthreads();
            float acc[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) acc[c] = 0.0f;
            if (k0 < K) {
#pragma unroll 2
                for (int w = 0; w < 8; ++w) {
                    const int n0 = tg * 128 + w * 16;
                    float gv[16], xv[32];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const f32x4 a = *reinterpret_cast<const f32x4 *>(gs + n0 + 4 * j);
                        gv[4 * j] = a.x; gv[4 * j + 1] = a.y; gv[4 * j + 2] = a.z; gv[4 * j + 3] = a.w;
                    }
                    // xv[j] = x[t0 + n0 - k0 - 16 + j] = xs[hist + n0 - k0 - 16 + j]
                    const int xb = hist + n0 - k0 - 16;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const f32x4 a = *reinterpret_cast<const f32x4 *>(xs + pad16(xb + 4 * j));
                        xv[4 * j] = a.x; xv[4 * j + 1] = a.y; xv[4 * j + 2] = a.z; xv[4 * j + 3] = a.w;
                    }
                    // acc[c] += sum_a g[n0+a] * x[n0+a-k0-c] = gv[a] * xv[a - c + 16]
#pragma unroll
                    for (int a = 0; a < 16; ++a)
#pragma unroll
                        for (int c = 0; c < 16; ++c)
                            acc[c] = __builtin_fmaf(gv[a], xv[a - c + 16], acc[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) tot[c] += acc[c];
        }
        // Fixed-order reduction over the 16 time groups.
#pragma unroll
        for (int c = 0; c < 16; ++c) red[tg * kLagPass + kg * 16 + c] = tot[c];
        __syncthreads();
        {
            const int lag = k_pass + threadIdx.x;
            float s = 0.0f;
#pragma unroll
            for (int q = 0; q < 16; ++q) s += red[q * kLagPass + threadIdx.x];
            if (lag < K) partial[(int64_t)blockIdx.x * K + lag] = s;
        }
        __syncthreads();
    }
}

// The same stage 1 on FP32 MFMA (v_mfma_f32_16x16x4_f32).  With lags
// k = kp + i + 16j (i, j < 16) one 16x16x4 product takes four samples m..m+3
// against all 256 lags of a pass:
//   C[i][j] += sum_l x[m + l - i - kp] * g[m + l + 16j],
// A[i][l] = x[m + l - i - kp] (a Toeplitz slice of x), B[l][j] =
// g[m + l + 16j]; summed over every m of the row, C[i][j] is
// db[kp + i + 16j] (FP32 operands, FP32 accumulation; no product is wasted
// for K <= 256).  A tile stages g[t0, t0 + 2304) and x[t0 - kp - 16,
// t0 - kp + 2064) in LDS with 4 pad floats per 16, which puts the 64
// B-operand reads of a product (16 lanes 16 samples apart) in distinct
// banks; the next tile's samples are loaded into registers while the current
// tile's products run.  Wave w takes the tile's samples 512w .. 512w + 511
// with four accumulators in turn (32 products each per tile, as long as the
// VALU kernel's per-tile chains); tile sums, the four waves' sums (fixed
// order) and partial[wg][k] follow, which the final kernel reduces as before.
constexpr int kMfmaLook = 256;  // g samples staged past the tile (16j + l < 256)

__global__ __launch_bounds__(kThreads) void fir_vjp_db_mfma_kernel(
    const float *__restrict__ g, int64_t ldg, const float *__restrict__ x, int64_t ldx, int K,
    int64_t N, int segs_per_row, int tiles_per_wg, float *__restrict__ partial)
{
    constexpr int GS = kTile + kMfmaLook;  // staged g samples
    constexpr int XS = kTile + 32;         // staged x samples
    __shared__ __attribute__((aligned(16))) float gs[GS + GS / 4];
    __shared__ __attribute__((aligned(16))) float xs[XS + XS / 4];
    __shared__ float red[kThreads / 64][kLagPass];

    const int64_t row = blockIdx.x / segs_per_row;
    const int64_t tile0 = (int64_t)(blockIdx.x % segs_per_row) * tiles_per_wg;
    const float *__restrict__ grow = g + row * ldg;
    const float *__restrict__ xrow = x + row * ldx;
    const int64_t tiles_in_row = (N + kTile - 1) / kTile;
    const int ntiles = (int)min<int64_t>(tiles_per_wg, tiles_in_row - tile0);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lq = lane >> 4;  // A: (row li, col lq); B: (row lq, col li)
    const int a_off = 16 + lq - li;  // x sample m + lq - li - kp, staged from t0 - kp - 16
    const int b_off = lq + 16 * li;  // g sample m + lq + 16 li, staged from t0

    constexpr int NG = (GS / 4 + kThreads - 1) / kThreads, NX = (XS / 4 + kThreads - 1) / kThreads;
    f32x4 pg[NG], px[NX];
    auto fetch = [&](int64_t t0, int kp) {
#pragma unroll
        for (int u = 0; u < NG; ++u) {
            const int v = threadIdx.x + u * kThreads;
            if (v < GS / 4) pg[u] = load4g(grow, t0 + 4 * v, N);
        }
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int v = threadIdx.x + u * kThreads;
            if (v < XS / 4) px[u] = load4g(xrow, t0 - kp - 16 + 4 * v, N);
        }
    };
    for (int kp = 0; kp < K; kp += kLagPass) {
        f32x4 tot = {0.0f, 0.0f, 0.0f, 0.0f};
        fetch(tile0 * kTile, kp);
        for (int it = 0; it < ntiles; ++it) {
            __syncthreads();  // previous tile done with gs/xs
#pragma unroll
            for (int u = 0; u < NG; ++u) {
                const int v = threadIdx.x + u * kThreads;
                if (v < GS / 4) *reinterpret_cast<f32x4 *>(gs + pad16(4 * v)) = pg[u];
            }
#pragma unroll
            for (int u = 0; u < NX; ++u) {
                const int v = threadIdx.x + u * kThreads;
                if (v < XS / 4) *reinterpret_cast<f32x4 *>(xs + pad16(4 * v)) = px[u];
            }
            __syncthreads();
            if (it + 1 < ntiles) fetch((tile0 + it + 1) * kTile, kp);
            // Four accumulators in turn (independent MFMA chains of 32 products
            // of 4 samples each per tile), added into the pass total per tile.
            // Sample group 4(4t + u) of the wave: each 16-sample step moves a
            // padded index by exactly 20 floats, so after the four bases the
            // fully unrolled loop's LDS reads carry constant offsets.
            const int m0 = 512 * wave;
            const float *xa0 = xs + pad16(m0 + a_off), *xa1 = xs + pad16(m0 + 4 + a_off);
            const float *xa2 = xs + pad16(m0 + 8 + a_off), *xa3 = xs + pad16(m0 + 12 + a_off);
            const float *gb0 = gs + pad16(m0 + b_off), *gb1 = gs + pad16(m0 + 4 + b_off);
            const float *gb2 = gs + pad16(m0 + 8 + b_off), *gb3 = gs + pad16(m0 + 12 + b_off);
            f32x4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = c0, c2 = c0, c3 = c0;
#pragma unroll
            for (int t = 0; t < kTile / 4 / 16; ++t) {  // 32 steps of 16 samples per wave
                c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa0[20 * t], gb0[20 * t], c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa1[20 * t], gb1[20 * t], c1, 0, 0, 0);
                c2 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa2[20 * t], gb2[20 * t], c2, 0, 0, 0);
                c3 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa3[20 * t], gb3[20 * t], c3, 0, 0, 0);
            }
            tot += (c0 + c1) + (c2 + c3);
        }
        // C[i][j] in lane (j = li, i = 4 lq + v): lag kp + 4 lq + v + 16 li.
#pragma unroll
        for (int v = 0; v < 4; ++v) red[wave][4 * lq + v + 16 * li] = tot[v];
        __syncthreads();
        {
            const int lag = kp + threadIdx.x;
            const float sum = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
                              red[3][threadIdx.x];
            if (lag < K) partial[(int64_t)blockIdx.x * K + lag] = sum;
        }
        __syncthreads();  // red is rewritten by the next pass
    }
}

constexpr int kFinalLags = 4;                       // lags per final workgroup
constexpr int kFinalSplit = kThreads / kFinalLags;  // workgroup residues per lag

__device__ __forceinline__ void kahan_add(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// db[k] = sum_w partial[w][k], Kahan-compensated, fixed order: lane (lag, q)
// sums workgroups w == q (mod 64) in increasing w (loads issued eight at a
// time), then the 64 sums are added in q order.
__global__ __launch_bounds__(kThreads) void fir_vjp_db_final_kernel(
    const float *__restrict__ partial, int64_t nwg, int K, float *__restrict__ db)
{
    __shared__ float sums[kFinalSplit][kFinalLags + 1];
    const int lag_l = threadIdx.x % kFinalLags, q = threadIdx.x / kFinalLags;
    const int lag = blockIdx.x * kFinalLags + lag_l;
    float s = 0.0f, c = 0.0f;
    if (lag < K) {
        int64_t w = q;
        for (; w + 7 * kFinalSplit < nwg; w += 8 * kFinalSplit) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partial[(w + u * kFinalSplit) * K + lag];
#pragma unroll
            for (int u = 0; u < 8; ++u) kahan_add(s, c, v[u]);
        }
        for (; w < nwg; w += kFinalSplit) kahan_add(s, c, partial[w * K + lag]);
    }
    sums[q][lag_l] = s;
    __syncthreads();
    if (q == 0 && lag < K) {
        float tot = 0.0f, cc = 0.0f;
        for (int j = 0; j < kFinalSplit; ++j) kahan_add(tot, cc, sums[j][lag_l]);
        db[lag] = tot;
    }
}

// Partial sums of the allocation-free dL/db entry point: a file-scope device
// array, one copy per GPU, held until the process ends.  Host code never
// passes `scratch` itself to a kernel (its name is a host-side stand-in);
// the device address comes from hipGetSymbolAddress, once per device.
__device__ __align__(256) float scratch[kVjpNoAllocPartials];

}  // namespace

int vjp_noalloc_scratch_address(void **ptr)
{
    *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(SCIR_HIP_ERR_BACKEND, "hipGetSymbolAddress(dL/db scratch) failed: %s",
                    hipGetErrorString(e));
    }
    return SCIR_HIP_OK;
}

int vjp_db_args(const float *g, int64_t ldg, const float *x, int64_t ldx, int32_t K, int64_t B,
                int64_t N, const float *db)
{
    if (K < 0 || B < 0 || N < 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative size");
    if (K == 0) return SCIR_HIP_OK;
    if (!db) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "db is null");
    if (B > 0 && N > 0) {
        if (!g || !x) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
        if (ldg < N || ldx < N)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row stride smaller than the row length");
    }
    return SCIR_HIP_OK;
}

// dL/db on stream s (caller holds st->mu; arguments checked).  With `fixed`,
// the partial sums go to that device buffer of `fixed_floats` floats instead
// of the stream's library scratch (nothing is allocated): the FP32 MFMA
// stage 1 always runs, and when the usual workgroup count would not fit,
// fewer workgroups per row walk longer runs of tiles.
int vjp_db_run(const float *g, int64_t ldg, const float *x, int64_t ldx, int32_t K, int64_t B,
               int64_t N, float *db, hipStream_t s, DeviceState *st, float *fixed = nullptr,
               size_t fixed_floats = 0)
{
    if (B == 0 || N == 0) {
        // Empty sum: db = 0 (a one-row, zero-work pass of the final kernel).
        hipLaunchKernelGGL(fir_vjp_db_final_kernel, dim3((K + kFinalLags - 1) / kFinalLags),
                           dim3(kThreads), 0, s, static_cast<const float *>(nullptr), (int64_t)0,
                           (int)K, db);
        return check_launch("fir_vjp_db_final_kernel");
    }
    const int64_t tiles_in_row = (N + kTile - 1) / kTile;
    // Enough workgroups to fill the device (about 8 per CU), each walking a
    // contiguous run of tiles of one row.
    const int64_t want = (int64_t)st->cu_count * 8;
    int64_t segs = std::max<int64_t>(1, std::min<int64_t>(tiles_in_row, (want + B - 1) / B));
    if (fixed) {
        // B rows of K partial sums at least (one workgroup per row)
        const int64_t rows_fit = (int64_t)(fixed_floats / (size_t)K);
        if (B > rows_fit)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT,
                        "the allocation-free dL/db holds %zu partial sums: rows * K must not "
                        "exceed it (rows=%lld, K=%d)",
                        fixed_floats, (long long)B, K);
        segs = std::max<int64_t>(1, std::min<int64_t>(segs, rows_fit / B));
    }
    const int64_t tiles_per_wg = (tiles_in_row + segs - 1) / segs;
    segs = (tiles_in_row + tiles_per_wg - 1) / tiles_per_wg;
    const int64_t nwg = B * segs;
    if (nwg > 0x7fffffff) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "batch too large");
    const TuneChoice tc = tune_lookup(K, N);
    if (tc.vjp_db_mfma || fixed) {
        float *partial = fixed;
        if (!fixed) {
            StreamScratch &sc = scratch_for(st, s);
            SCIR_TRY(ensure_scratch(st, sc.partials, (size_t)nwg * K * sizeof(float), s));
            partial = static_cast<float *>(sc.partials.ptr);
        }
        hipLaunchKernelGGL(fir_vjp_db_mfma_kernel, dim3((unsigned)nwg), dim3(kThreads), 0, s, g,
                           ldg, x, ldx, (int)K, N, (int)segs, (int)tiles_per_wg, partial);
        SCIR_TRY(check_launch("fir_vjp_db_mfma_kernel"));
        hipLaunchKernelGGL(fir_vjp_db_final_kernel, dim3((K + kFinalLags - 1) / kFinalLags),
                           dim3(kThreads), 0, s, partial, nwg, (int)K, db);
        return check_launch("fir_vjp_db_final_kernel");
    }
    const int passes = (K + kLagPass - 1) / kLagPass;
    const int hist = ((passes * kLagPass + 16) + 3) & ~3;  // history samples before a tile
    const size_t lds_floats =
        kTile + (size_t)(hist + kTile + 32) + ((hist + kTile + 32) / 16) * 4 + 4 + 16 * kLagPass;
    const size_t lds = lds_floats * sizeof(float);
    if (lds > 160 * 1024)
        return fail(SCIR_HIP_ERR_UNSUPPORTED, "K=%d is too long for the dL/db kernel", K);
    SCIR_TRY(allow_dynamic_lds(reinterpret_cast<const void *>(&fir_vjp_db_partial_kernel),
                               160 * 1024));
    StreamScratch &sc = scratch_for(st, s);
    SCIR_TRY(ensure_scratch(st, sc.partials, (size_t)nwg * K * sizeof(float), s));
    float *partial = static_cast<float *>(sc.partials.ptr);
    hipLaunchKernelGGL(fir_vjp_db_partial_kernel, dim3((unsigned)nwg), dim3(kThreads), lds, s, g,
                       ldg, x, ldx, (int)K, N, (int)segs, (int)tiles_per_wg, hist, partial);
    SCIR_TRY(check_launch("fir_vjp_db_partial_kernel"));
    hipLaunchKernelGGL(fir_vjp_db_final_kernel, dim3((K + kFinalLags - 1) / kFinalLags),
                       dim3(kThreads), 0, s, partial, nwg, (int)K, db);
    return check_launch("fir_vjp_db_final_kernel");
}

}  // namespace scir

using namespace scir;

extern "C" {

int scir_hip_fir_vjp_dx_dev(const float *g, int64_t ldg, const float *b, int32_t K, int64_t B,
                            int64_t N, float *dx, int64_t lddx, int32_t precision, void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        SCIR_TRY(check_dense_args(g, ldg, b, K, B, N, dx, lddx));
        if (B == 0 || N == 0) return SCIR_HIP_OK;
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        std::lock_guard<std::mutex> lk(st->mu);
        return fir_dense_dispatch(g, ldg, b, K, B, N, dx, lddx, precision, true, as_stream(stream), st);
    });
}

int scir_hip_fir_vjp_db_dev(const float *g, int64_t ldg, const float *x, int64_t ldx, int32_t K,
                            int64_t B, int64_t N, float *db, void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        SCIR_TRY(vjp_db_args(g, ldg, x, ldx, K, B, N, db));
        if (K == 0) return SCIR_HIP_OK;
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        std::lock_guard<std::mutex> lk(st->mu);
        return vjp_db_run(g, ldg, x, ldx, K, B, N, db, as_stream(stream), st);
    });
}

int scir_hip_fir_vjp_db_noalloc_dev(const float *g, int64_t ldg, const float *x, int64_t ldx,
                                    int32_t K, int64_t B, int64_t N, float *db, void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        SCIR_TRY(vjp_db_args(g, ldg, x, ldx, K, B, N, db));
        if (K == 0) return SCIR_HIP_OK;
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        if (!st->vjp_noalloc_scratch)
            return fail(SCIR_HIP_ERR_BACKEND, "the dL/db scratch array has no device address: %s",
                        st->vjp_noalloc_error.c_str());
        std::lock_guard<std::mutex> lk(st->mu);
        return vjp_db_run(g, ldg, x, ldx, K, B, N, db, as_stream(stream), st,
                          st->vjp_noalloc_scratch, kVjpNoAllocPartials);
    });
}

int scir_hip_fir_vjp_f32_host(const float *x, int64_t ldx, const float *g, int64_t ldg,
                              const float *b, int32_t K, int64_t B, int64_t N, float *dx,
                              int64_t lddx, float *db, int32_t precision)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        SCIR_TRY(check_dense_args(g, ldg, b, K, B, N, dx, lddx));
        SCIR_TRY(vjp_db_args(g, ldg, x, ldx, K, B, N, db));
        if (precision < SCIR_HIP_PREC_F32 || precision > SCIR_HIP_PREC_FFT32)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "unknown precision %d", precision);
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        std::lock_guard<std::mutex> lk(st->mu);
        if (!st->host_stream)
            SCIR_HIP_TRY(hipStreamCreateWithFlags(&st->host_stream, hipStreamNonBlocking));
        hipStream_t s = st->host_stream;
        // Device rows padded to a multiple of 4 floats (16-byte aligned rows).
        const int64_t ldd = (N + 3) & ~int64_t(3);
        const size_t mat_bytes = (size_t)std::max<int64_t>(B, 1) * ldd * sizeof(float);
        const size_t row_bytes = (size_t)N * sizeof(float);
        SCIR_TRY(ensure_buffer(st->host_x, mat_bytes, s));
        SCIR_TRY(ensure_buffer(st->host_y, mat_bytes, s));
        SCIR_TRY(ensure_buffer(st->host_z, mat_bytes, s));
        SCIR_TRY(ensure_buffer(st->host_b, (size_t)std::max(2 * K, 1) * sizeof(float), s));
        float *dxx = static_cast<float *>(st->host_x.ptr);
        float *dg = static_cast<float *>(st->host_y.ptr);
        float *ddx = static_cast<float *>(st->host_z.ptr);
        float *dbuf = static_cast<float *>(st->host_b.ptr);
        if (B > 0 && N > 0) {
            SCIR_HIP_TRY(hipMemcpy2DAsync(dxx, ldd * sizeof(float), x, ldx * sizeof(float), row_bytes,
                                          B, hipMemcpyHostToDevice, s));
            SCIR_HIP_TRY(hipMemcpy2DAsync(dg, ldd * sizeof(float), g, ldg * sizeof(float), row_bytes,
                                          B, hipMemcpyHostToDevice, s));
        }
        if (K > 0)
            SCIR_HIP_TRY(hipMemcpyAsync(dbuf, b, (size_t)K * sizeof(float), hipMemcpyHostToDevice, s));
        if (B > 0 && N > 0)
            SCIR_TRY(fir_dense_dispatch(dg, ldd, dbuf, K, B, N, ddx, ldd, precision, true, s, st));
        if (K > 0) SCIR_TRY(vjp_db_run(dg, ldd, dxx, ldd, K, B, N, dbuf + K, s, st));
        if (B > 0 && N > 0)
            SCIR_HIP_TRY(hipMemcpy2DAsync(dx, lddx * sizeof(float), ddx, ldd * sizeof(float), row_bytes,
                                          B, hipMemcpyDeviceToHost, s));
        if (K > 0)
            SCIR_HIP_TRY(hipMemcpyAsync(db, dbuf + K, (size_t)K * sizeof(float),
                                        hipMemcpyDeviceToHost, s));
        SCIR_HIP_TRY(hipStreamSynchronize(s));
        return SCIR_HIP_OK;
    });
}

}  // extern "C"
