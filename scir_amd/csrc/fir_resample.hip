// Batched polyphase resampler in float64 (the GPU form of the reference's
// resample_poly, crates/scir-signal/src/lib.rs:313-362).
//
// The reference upsamples x by `up` (inserting up-1 zeros after each
// sample), convolves with the K-tap filter h and keeps every `down`-th
// sample of the full convolution from off = (K-1)/2.  Only every up-th
// upsampled sample is non-zero, so output m needs only the taps of one
// phase:
//   n = off + down*m,  ph = n mod up,  q = n / up,
//   y[m] = sum_{t >= 0, ph + up*t < K} h[ph + up*t] * x[q - t]
// (x[s] = 0 outside [0, N)), about K/up multiply-adds instead of K.
//
// One workgroup computes T = 256*J consecutive outputs of one row; lane l
// owns outputs m0 + l + 256*j (consecutive lanes, consecutive outputs: the
// stores coalesce).  The x samples the tile needs and the taps are staged in
// LDS once.
#include "common.hpp"

#include <algorithm>

namespace scir {
namespace {

constexpr int kThreads = 256;
constexpr int kJ = 4;                     // outputs per lane
constexpr int kT = kThreads * kJ;         // outputs per workgroup
constexpr int kMaxTaps = 8192;

__global__ __launch_bounds__(kThreads) void resample_poly_kernel(
    const double *__restrict__ x, int64_t ldx, int64_t N, double *__restrict__ y, int64_t ldy,
    int64_t M, const double *__restrict__ h, int K, int up, int down, int tiles_per_row,
    int span)
{
    extern __shared__ __attribute__((aligned(16))) double lds_d[];
    double *hs = lds_d;           // K taps
    double *xs = lds_d + K;       // span samples: x[s0 .. s0 + span)
    const int64_t row = blockIdx.x / tiles_per_row;
    const int64_t m0 = (int64_t)(blockIdx.x % tiles_per_row) * kT;
    const int64_t off = (K - 1) / 2;
    const int tmax = (K + up - 1) / up;   // taps of the longest phase
    // First sample the tile reads: q(m0) - (tmax - 1).
    const int64_t s0 = (off + down * m0) / up - (tmax - 1);
    const double *__restrict__ xr = x + row * ldx;
    for (int k = threadIdx.x; k < K; k += kThreads) hs[k] = h[k];
    for (int i = threadIdx.x; i < span; i += kThreads) {
        const int64_t s = s0 + i;
        xs[i] = (s >= 0 && s < N) ? xr[s] : 0.0;
    }
    __syncthreads();
    double *__restrict__ yr = y + row * ldy;
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
        const int64_t m = m0 + threadIdx.x + kThreads * j;
        if (m >= M) break;
        const int64_t n = off + down * m;
        const int ph = (int)(n % up);
        const int qi = (int)(n / up - s0);  // index of x[q] in xs
        double acc = 0.0;
        for (int k = ph, t = 0; k < K; k += up, ++t) acc = __builtin_fma(hs[k], xs[qi - t], acc);
        yr[m] = acc;
    }
}

}  // namespace
}  // namespace scir

using namespace scir;

extern "C" {

int scir_hip_resample_len(int64_t N, int32_t up, int32_t down, int32_t K, int64_t *M)
{
    clear_error();
    if (!M) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "M is null");
    if (N < 0 || up < 1 || down < 1 || K < 1)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "need N >= 0, up >= 1, down >= 1, K >= 1");
    const int64_t len = N * up + K - 1, off = (K - 1) / 2, end = len - off;
    *M = end > off ? (end - off + down - 1) / down : 0;
    return SCIR_HIP_OK;
}

}  // extern "C"

namespace {

// Launch on `s`; the caller holds st->mu.
int resample_launch(const double *x, int64_t ldx, int64_t B, int64_t N, int32_t up, int32_t down,
                    const double *h, int32_t K, double *y, int64_t ldy, int64_t M, hipStream_t s)
{
    // Samples a tile reads: q(m0 + T - 1) - q(m0) + tmax <= down*(T-1)/up + 1 + tmax.
    const int tmax = (K + up - 1) / up;
    const int span = (int)(((int64_t)down * (kT - 1)) / up + 2 + tmax);
    const size_t lds = ((size_t)K + span) * sizeof(double);
    if (lds > 160 * 1024)
        return fail(SCIR_HIP_ERR_UNSUPPORTED, "K=%d, up=%d, down=%d need too much LDS", K, up,
                    down);
    static bool attr_set = false;  // guarded by the device mutex
    if (!attr_set) {
        SCIR_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&resample_poly_kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
    }
    const int64_t tiles = (M + kT - 1) / kT;
    if (tiles > 0x7fffffff) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row too long");
    const int64_t rows_per_launch = std::max<int64_t>(1, 0x7fffffff / tiles);
    for (int64_t r0 = 0; r0 < B; r0 += rows_per_launch) {
        const int64_t rows = std::min(rows_per_launch, B - r0);
        hipLaunchKernelGGL(resample_poly_kernel, dim3((unsigned)(rows * tiles)), dim3(kThreads),
                           lds, s, x + r0 * ldx, ldx, N, y + r0 * ldy, ldy, M, h, K, up, down,
                           (int)tiles, span);
        SCIR_TRY(check_launch("resample_poly_kernel"));
    }
    return SCIR_HIP_OK;
}

int resample_args(const double *x, int64_t ldx, int64_t B, int64_t N, int32_t up, int32_t down,
                  const double *h, int32_t K, double *y, int64_t ldy, int64_t *M)
{
    SCIR_TRY(scir_hip_resample_len(N, up, down, K, M));
    if (B < 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative row count");
    if (K > kMaxTaps) return fail(SCIR_HIP_ERR_UNSUPPORTED, "K=%d above %d taps", K, kMaxTaps);
    if (B == 0 || *M == 0) return SCIR_HIP_OK;
    if (!x || !y || !h) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (ldx < N || ldy < *M)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row stride smaller than the row length");
    return SCIR_HIP_OK;
}

}  // namespace

extern "C" {

int scir_hip_resample_poly_f64_dev(const double *x, int64_t ldx, int64_t B, int64_t N, int32_t up,
                                   int32_t down, const double *h, int32_t K, double *y,
                                   int64_t ldy, void *stream)
{
    int64_t M = 0;
    SCIR_TRY(resample_args(x, ldx, B, N, up, down, h, K, y, ldy, &M));
    if (B == 0 || M == 0) return SCIR_HIP_OK;
    DeviceState *st = nullptr;
    SCIR_TRY(current_device_state(&st));
    std::lock_guard<std::mutex> lk(st->mu);
    return resample_launch(x, ldx, B, N, up, down, h, K, y, ldy, M, as_stream(stream));
}

int scir_hip_resample_poly_f64_host(const double *x, int64_t ldx, int64_t B, int64_t N,
                                    int32_t up, int32_t down, const double *h, int32_t K,
                                    double *y, int64_t ldy)
{
    int64_t M = 0;
    SCIR_TRY(resample_args(x, ldx, B, N, up, down, h, K, y, ldy, &M));
    if (B == 0 || M == 0) return SCIR_HIP_OK;
    DeviceState *st = nullptr;
    SCIR_TRY(current_device_state(&st));
    std::lock_guard<std::mutex> lk(st->mu);
    if (!st->host_stream)
        SCIR_HIP_TRY(hipStreamCreateWithFlags(&st->host_stream, hipStreamNonBlocking));
    hipStream_t s = st->host_stream;
    SCIR_TRY(ensure_buffer(st->host_x, (size_t)B * N * sizeof(double), s));
    SCIR_TRY(ensure_buffer(st->host_y, (size_t)B * M * sizeof(double), s));
    SCIR_TRY(ensure_buffer(st->host_b, (size_t)K * sizeof(double), s));
    double *dx = static_cast<double *>(st->host_x.ptr);
    double *dy = static_cast<double *>(st->host_y.ptr);
    double *dh = static_cast<double *>(st->host_b.ptr);
    if (ldx == N)
        SCIR_HIP_TRY(hipMemcpyAsync(dx, x, (size_t)B * N * sizeof(double), hipMemcpyHostToDevice, s));
    else
        SCIR_HIP_TRY(hipMemcpy2DAsync(dx, N * sizeof(double), x, ldx * sizeof(double),
                                      N * sizeof(double), B, hipMemcpyHostToDevice, s));
    SCIR_HIP_TRY(hipMemcpyAsync(dh, h, (size_t)K * sizeof(double), hipMemcpyHostToDevice, s));
    SCIR_TRY(resample_launch(dx, N, B, N, up, down, dh, K, dy, M, M, s));
    if (ldy == M)
        SCIR_HIP_TRY(hipMemcpyAsync(y, dy, (size_t)B * M * sizeof(double), hipMemcpyDeviceToHost, s));
    else
        SCIR_HIP_TRY(hipMemcpy2DAsync(y, ldy * sizeof(double), dy, M * sizeof(double),
                                      M * sizeof(double), B, hipMemcpyDeviceToHost, s));
    SCIR_HIP_TRY(hipStreamSynchronize(s));
    return SCIR_HIP_OK;
}

}  // extern "C"
