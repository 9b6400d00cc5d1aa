// CPU arm of the dispatcher (Device::Cpu): the reference's CPU FIR
// (crates/scir-gpu/src/lib.rs:1134-1152) in the same arithmetic order, spread
// over host threads by row.
//
// Each output is acc = 0.0f; acc = acc + (taps[K-1-t] * x[i-t]) for t = 0..,
// i.e. acc = acc + b[t]*x[i-t] with b the lfilter-order taps, with the
// multiply and the add rounded separately (this file is compiled with
// -ffp-contract=off).  Blocks of 8 consecutive outputs are computed together
// so the compiler can vectorise across outputs; each output keeps its own
// sequential order, so the result is bit-identical to the reference loop.
#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

#include "host_common.hpp"

using scir::guarded;

namespace {

// Runs work(0) .. work(nt-1) on nt threads (work(0) on the caller's).  If a
// thread cannot be started, the ones already running are joined before the
// exception propagates (a joinable std::thread must never be destroyed).
template <class F>
void run_threads(int nt, F &work)
{
    scir::host_alloc_point();
    std::vector<std::thread> pool;
    pool.reserve(nt);
    try {
        for (int w = 1; w < nt; ++w) pool.emplace_back(work, w);
    } catch (...) {
        for (auto &t : pool) t.join();
        throw;
    }
    work(0);
    for (auto &t : pool) t.join();
}

}  // namespace

namespace {

void fir_row(const float *xr, const float *b, int64_t K, int64_t N, float *yr)
{
    // Ramp-in: outputs whose history is shorter than K taps.
    const int64_t ramp = std::min<int64_t>(N, K - 1);
    for (int64_t i = 0; i < ramp; ++i) {
        float acc = 0.0f;
        for (int64_t t = 0; t <= i; ++t) acc = acc + b[t] * xr[i - t];
        yr[i] = acc;
    }
    constexpr int V = 8;
    int64_t i = ramp;
    for (; i + V <= N; i += V) {
        float acc[V] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t t = 0; t < K; ++t) {
            const float bt = b[t];
            const float *xp = xr + i - t;
            for (int j = 0; j < V; ++j) acc[j] = acc[j] + bt * xp[j];
        }
        for (int j = 0; j < V; ++j) yr[i + j] = acc[j];
    }
    for (; i < N; ++i) {
        float acc = 0.0f;
        for (int64_t t = 0; t < K; ++t) acc = acc + b[t] * xr[i - t];
        yr[i] = acc;
    }
}

}  // namespace

extern "C" int scir_hip_fir_f32_cpu(const float *x, int64_t ldx, const float *b, int32_t K,
                                    int64_t B, int64_t N, float *y, int64_t ldy,
                                    int32_t nthreads)
{
    return guarded(__func__, [&]() -> int {
        scir::clear_error();
        if (B < 0 || N < 0 || K < 0)
            return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative size");
        if (B == 0 || N == 0) return SCIR_HIP_OK;
        if (!x || !y || (K > 0 && !b))
            return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
        if (ldx < N || ldy < N)
            return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row stride smaller than the row length");
        if (K == 0) {
            for (int64_t r = 0; r < B; ++r) std::fill(y + r * ldy, y + r * ldy + N, 0.0f);
            return SCIR_HIP_OK;
        }
        int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
        nt = (int)std::min<int64_t>(nt, B);
        auto work = [&](int w) {
            for (int64_t r = w; r < B; r += nt) fir_row(x + r * ldx, b, K, N, y + r * ldy);
        };
        if (nt <= 1) {
            work(0);
            return SCIR_HIP_OK;
        }
        run_threads(nt, work);
        return SCIR_HIP_OK;
    });
}

// CPU arm of the polyphase resampler, in the reference's operation order
// (crates/scir-signal/src/lib.rs:293-362): output m is entry n = off + down*m
// of the full convolution of the zero-stuffed signal with h, which the
// reference accumulates as acc = acc + x[s] * h[n - up*s] over s ascending
// (its inner loop also adds the products of the inserted zeros; adding +-0
// to a sum that is never -0.0 leaves every bit unchanged, so they are
// skipped here).
namespace {

void resample_row(const double *xr, int64_t N, int64_t up, int64_t down, const double *h,
                  int64_t K, double *yr, int64_t M)
{
    const int64_t off = (K - 1) / 2;
    for (int64_t m = 0; m < M; ++m) {
        const int64_t n = off + down * m;
        // s from ceil((n - K + 1) / up) (but >= 0) to min(n / up, N - 1)
        const int64_t lo_num = n - K + 1;
        int64_t s_lo = lo_num <= 0 ? 0 : (lo_num + up - 1) / up;
        const int64_t s_hi = std::min<int64_t>(n / up, N - 1);
        double acc = 0.0;
        for (int64_t s = s_lo; s <= s_hi; ++s) {
            const double p = xr[s] * h[n - up * s];
            acc = acc + p;
        }
        yr[m] = acc;
    }
}

}  // namespace

// Output length of resample_poly (crates/scir-signal/src/lib.rs:313-362):
// every `down`-th sample of the full convolution from off = (K-1)/2 while
// below len(conv) - off.
extern "C" int scir_hip_resample_len(int64_t N, int32_t up, int32_t down, int32_t K, int64_t *M)
{
    scir::clear_error();
    if (!M) return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "M is null");
    if (N < 0 || up < 1 || down < 1 || K < 1)
        return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "need N >= 0, up >= 1, down >= 1, K >= 1");
    if (N > ((int64_t)1 << 40))
        return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "N=%lld is too long", (long long)N);
    const int64_t len = N * up + K - 1, off = (K - 1) / 2, end = len - off;
    *M = end > off ? (end - off + down - 1) / down : 0;
    return SCIR_HIP_OK;
}

extern "C" int scir_hip_resample_poly_f64_cpu(const double *x, int64_t ldx, int64_t B, int64_t N,
                                              int32_t up, int32_t down, const double *h,
                                              int32_t K, double *y, int64_t ldy,
                                              int32_t nthreads)
{
    return guarded(__func__, [&]() -> int {
        int64_t M = 0;
        const int st = scir_hip_resample_len(N, up, down, K, &M);
        if (st != SCIR_HIP_OK) return st;
        if (B < 0) return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative row count");
        if (B == 0 || M == 0) return SCIR_HIP_OK;
        if (!x || !y || !h)
            return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
        if (ldx < N || ldy < M)
            return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row stride smaller than the row length");
        int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
        nt = (int)std::min<int64_t>(nt, B);
        auto work = [&](int w) {
            for (int64_t r = w; r < B; r += nt)
                resample_row(x + r * ldx, N, up, down, h, K, y + r * ldy, M);
        };
        if (nt <= 1) {
            work(0);
            return SCIR_HIP_OK;
        }
        run_threads(nt, work);
        return SCIR_HIP_OK;
    });
}

// CPU arm of sosfilt / filtfilt (crates/scir-signal/src/lib.rs:260-291): the
// Direct Form II Transposed cascade of crates/scir-iir-filters/src/filter.rs
// :47-62 per row, in its operation order with separate roundings (this file
// is compiled with -ffp-contract=off), so the output is bit-identical.
namespace {

constexpr int kMaxSos = SCIR_HIP_SOS_MAX_SECTIONS;

// Filters positions first, first+step, ... (n of them) of xr into yr from
// zero state; yr may be xr.
void sos_row(const double *xr, double *yr, int64_t n, int64_t first, int64_t step,
             const double *sos, int S)
{
    double s1[kMaxSos] = {0}, s2[kMaxSos] = {0};
    for (int64_t k = 0, p = first; k < n; ++k, p += step) {
        double v = xr[p];
        for (int i = 0; i < S; ++i) {
            const double *c = sos + 6 * i;
            const double prev = v;
            v = c[0] * prev + s1[i];
            s1[i] = c[1] * prev - c[4] * v + s2[i];
            s2[i] = c[2] * prev - c[5] * v;
        }
        yr[p] = v;
    }
}

int sos_cpu(const double *x, int64_t ldx, int64_t rows, int64_t N, const double *sos, int32_t S,
            double *y, int64_t ldy, int32_t nthreads, bool zero_phase)
{
    scir::clear_error();
    if (rows < 0 || N < 0) return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative size");
    if (S < 0 || S > kMaxSos)
        return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "sections S=%d outside [0, %d]", S,
                          kMaxSos);
    if (rows == 0 || N == 0) return SCIR_HIP_OK;
    if (!x || !y || (S > 0 && !sos))
        return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (ldx < N || ldy < N)
        return scir::fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row stride smaller than the row length");
    int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = (int)std::min<int64_t>(nt, rows);
    auto work = [&](int w) {
        for (int64_t r = w; r < rows; r += nt) {
            double *yr = y + r * ldy;
            sos_row(x + r * ldx, yr, N, 0, 1, sos, S);
            if (zero_phase) sos_row(yr, yr, N, N - 1, -1, sos, S);
        }
    };
    if (nt <= 1) {
        work(0);
        return SCIR_HIP_OK;
    }
    run_threads(nt, work);
    return SCIR_HIP_OK;
}

}  // namespace

extern "C" int scir_hip_sosfilt_f64_cpu(const double *x, int64_t ldx, int64_t rows, int64_t n,
                                        const double *sos, int32_t S, double *y, int64_t ldy,
                                        int32_t nthreads)
{
    return guarded(__func__, [&]() -> int {
        return sos_cpu(x, ldx, rows, n, sos, S, y, ldy, nthreads, false);
    });
}

extern "C" int scir_hip_filtfilt_f64_cpu(const double *x, int64_t ldx, int64_t rows, int64_t n,
                                         const double *sos, int32_t S, double *y, int64_t ldy,
                                         int32_t nthreads)
{
    return guarded(__func__, [&]() -> int {
        return sos_cpu(x, ldx, rows, n, sos, S, y, ldy, nthreads, true);
    });
}
