// Batched IIR filtering by cascaded second-order sections in float64: the
// GPU form of the reference's sosfilt / filtfilt
// (crates/scir-signal/src/lib.rs:260-291), whose per-sample step is the
// Direct Form II Transposed cascade of
// crates/scir-iir-filters/src/filter.rs:47-62:
//   for each section i:  p = v
//                        v = b0*p + s1[i]
//                        s1[i] = b1*p - a1*v + s2[i]
//                        s2[i] = b2*p - a2*v
// (a0 is not read: sections are taken as normalised, as in the reference).
//
// The recursion is sequential in time, so a row is cut into chunks of L
// samples and the filter's linearity joins them.  With the state
// z = (s1[0], s2[0], ..., s1[S-1], s2[S-1]) after a chunk,
//   z_end = T z_start + e,   T = A^L,  e = end state of the chunk from z = 0,
// where A is the cascade's zero-input state transition.  Three launches:
//   1. sos_chunk_kernel<STATE>: one lane per chunk runs it from zero state
//      and keeps e (x read once); one extra workgroup runs L zero-input steps
//      from each unit state, which gives the columns of T.
//   2. sos_scan_kernel: per row, an exclusive scan over the chunks,
//      z_c = T z_{c-1} + e_{c-1} (Kogge-Stone over a wave with T^(2^k)).
//   3. sos_chunk_kernel<RUN>: one lane per chunk runs it again from z_c and
//      writes y (x read again, y written once).
// Every sample step is the reference's arithmetic in its order (no fused
// multiply-adds), so chunk 0 of each row, and with one chunk per row
// (SCIR_HIP_SOS_REFERENCE_ORDER) every output, equals the reference's loop
// bit for bit; later chunks differ only by the rounding of z_c.
//
// A lane streams its chunk: sixteen samples (128 bytes) per step with
// 16-byte vector loads, the next group loaded while the current one is
// filtered.  Lanes of a wave hold consecutive chunks of one row.
#include "common.hpp"

#include <algorithm>

namespace scir {
namespace {

constexpr int kThreads = 256;
constexpr int kGroup = 16;        // samples per streamed group
constexpr int kMaxSections = 16;  // SCIR_HIP_SOS_MAX_SECTIONS

enum { kState = 0, kRun = 1 };

// Kernels are built for S = 1..8 sections exactly (SMAX == S, no run-time
// guards in the sample loop) and for 9..16 with SMAX = 16 and guards.
template <int SMAX>
constexpr bool kExactS = SMAX <= 8;

// The section coefficients the step uses, in registers.
template <int SMAX>
struct Coefs {
    double b0[SMAX], b1[SMAX], b2[SMAX], a1[SMAX], a2[SMAX];
};

template <int SMAX>
__device__ __forceinline__ void load_coefs(Coefs<SMAX> &c, const double *__restrict__ sos, int S)
{
#pragma unroll
    for (int i = 0; i < SMAX; ++i) {
        const bool on = kExactS<SMAX> || i < S;
        c.b0[i] = on ? sos[6 * i + 0] : 0.0;
        c.b1[i] = on ? sos[6 * i + 1] : 0.0;
        c.b2[i] = on ? sos[6 * i + 2] : 0.0;
        c.a1[i] = on ? sos[6 * i + 4] : 0.0;
        c.a2[i] = on ? sos[6 * i + 5] : 0.0;
    }
}

// One sample through the first S sections, in the reference's operation
// order and roundings (contraction off).
template <int SMAX>
__device__ __forceinline__ double df2t_step(double v, double (&s1)[SMAX], double (&s2)[SMAX],
                                            const Coefs<SMAX> &c, int S)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < SMAX; ++i) {
        if (kExactS<SMAX> || i < S) {
            const double p = v;
            v = c.b0[i] * p + s1[i];
            s1[i] = c.b1[i] * p - c.a1[i] * v + s2[i];
            s2[i] = c.b2[i] * p - c.a2[i] * v;
        }
    }
    return v;
}

__device__ __forceinline__ void load_group(double (&g)[kGroup], const double *p,
                                           bool vec)
{
    if (vec) {
#pragma unroll
        for (int k = 0; k < kGroup / 2; ++k) {
            const double2 v = *reinterpret_cast<const double2 *>(p + 2 * k);
            g[2 * k] = v.x;
            g[2 * k + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kGroup; ++k) g[k] = p[k];
    }
}

__device__ __forceinline__ void store_group(double *p, const double (&g)[kGroup],
                                            bool vec)
{
    if (vec) {
#pragma unroll
        for (int k = 0; k < kGroup / 2; ++k)
            *reinterpret_cast<double2 *>(p + 2 * k) = make_double2(g[2 * k], g[2 * k + 1]);
    } else {
#pragma unroll
        for (int k = 0; k < kGroup; ++k) p[k] = g[k];
    }
}

// MODE kState: chunks 0..nC-2 of every row from zero state, end states to
// zs[(row*nC + c) * 2S]; the grid's last workgroup writes T (2S x 2S,
// row-major) to tmat.  MODE kRun: every chunk from zs[...] (zero when zs is
// null), outputs to y.  REV runs each row back to front.
template <int SMAX, int MODE, bool REV>
__global__ __launch_bounds__(kThreads) void sos_chunk_kernel(
    const double *x, int64_t ldx, double *y, int64_t ldy, int64_t rows,
    int64_t N, const double *__restrict__ sos, int S, int64_t L, int64_t nC,
    double *__restrict__ zs, double *__restrict__ tmat, int vec)
{
    Coefs<SMAX> cf;
    load_coefs(cf, sos, S);
    double s1[SMAX], s2[SMAX];
#pragma unroll
    for (int i = 0; i < SMAX; ++i) s1[i] = s2[i] = 0.0;

    if (MODE == kState && blockIdx.x == gridDim.x - 1) {
        // Columns of T = A^L: L zero-input steps from each unit state.
        const int j = threadIdx.x;
        if (j >= 2 * S) return;
#pragma unroll
        for (int i = 0; i < SMAX; ++i) {
            s1[i] = (2 * i == j) ? 1.0 : 0.0;
            s2[i] = (2 * i + 1 == j) ? 1.0 : 0.0;
        }
        for (int64_t t = 0; t < L; ++t) (void)df2t_step<SMAX>(0.0, s1, s2, cf, S);
#pragma unroll
        for (int i = 0; i < SMAX; ++i) {
            if (i < S) {
                tmat[(2 * i) * (2 * S) + j] = s1[i];
                tmat[(2 * i + 1) * (2 * S) + j] = s2[i];
            }
        }
        return;
    }

    const int64_t item = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (item >= rows * nC) return;
    const int64_t row = item / nC;
    const int64_t c = item - row * nC;
    if (MODE == kState && c == nC - 1) return;  // its end state is never used
    // Chunk c covers positions [p0, p1): c*L onwards forward; back to front
    // the short chunk (when L does not divide N) is the first in time, so
    // every chunk starts at a multiple of L (16-byte aligned groups) and
    // every chunk whose end state is used is L long.
    const int64_t p0 = (REV ? nC - 1 - c : c) * L;
    const int64_t p1 = std::min<int64_t>(p0 + L, N);
    if (MODE == kRun && zs != nullptr) {
        const double *__restrict__ z = zs + item * (2 * S);
#pragma unroll
        for (int i = 0; i < SMAX; ++i) {
            if (i < S) {
                s1[i] = z[2 * i];
                s2[i] = z[2 * i + 1];
            }
        }
    }
    // x and y may be the same array (filtfilt's second pass runs in place):
    // a lane reads each group before it writes it and never reads it again.
    const double *xr = x + row * ldx;
    double *yr = (MODE == kRun) ? y + row * ldy : nullptr;
    const bool v = vec != 0;
    double g[kGroup];

    const int64_t nfull = (p1 - p0) / kGroup;
    if (!REV) {
        int64_t t = p0;
        const int64_t full_end = p0 + nfull * kGroup;
        if (t < full_end) {
            double nx[kGroup];
            load_group(nx, xr + t, v);
            for (; t < full_end; t += kGroup) {
#pragma unroll
                for (int k = 0; k < kGroup; ++k) g[k] = nx[k];
                if (t + kGroup < full_end) load_group(nx, xr + t + kGroup, v);
#pragma unroll
                for (int k = 0; k < kGroup; ++k) g[k] = df2t_step<SMAX>(g[k], s1, s2, cf, S);
                if (MODE == kRun) store_group(yr + t, g, v);
            }
        }
        for (; t < p1; ++t) {
            const double o = df2t_step<SMAX>(xr[t], s1, s2, cf, S);
            if (MODE == kRun) yr[t] = o;
        }
    } else {
        // Descending positions: the samples above the last whole group
        // first, then whole groups at descending addresses.
        const int64_t p_grp_top = p0 + nfull * kGroup;  // groups cover [p0, p_grp_top)
        for (int64_t p = p1 - 1; p >= p_grp_top; --p) {
            const double o = df2t_step<SMAX>(xr[p], s1, s2, cf, S);
            if (MODE == kRun) yr[p] = o;
        }
        if (nfull > 0) {
            int64_t base = p_grp_top - kGroup;
            double nx[kGroup];
            load_group(nx, xr + base, v);
            for (; base >= p0; base -= kGroup) {
#pragma unroll
                for (int k = 0; k < kGroup; ++k) g[k] = nx[k];
                if (base - kGroup >= p0) load_group(nx, xr + base - kGroup, v);
#pragma unroll
                for (int k = kGroup - 1; k >= 0; --k) g[k] = df2t_step<SMAX>(g[k], s1, s2, cf, S);
                if (MODE == kRun) store_group(yr + base, g, v);
            }
        }
    }

    if (MODE == kState) {
        double *__restrict__ z = zs + item * (2 * S);
#pragma unroll
        for (int i = 0; i < SMAX; ++i) {
            if (i < S) {
                z[2 * i] = s1[i];
                z[2 * i + 1] = s2[i];
            }
        }
    }
}

// Exclusive scan over the chunks of each row, in place on zs:
// z_0 = 0, z_c = T z_{c-1} + e_{c-1}.  One wave per row, 64 chunks per
// step: Kogge-Stone with T^(2^k) (k = 0..5), the previous step's last
// inclusive state folded into the first chunk.
template <int SMAX>
__global__ __launch_bounds__(kThreads) void sos_scan_kernel(double *__restrict__ zs,
                                                            const double *__restrict__ tmat,
                                                            int64_t rows, int64_t nC, int S)
{
    constexpr int D = 2 * SMAX;
    __shared__ double tp[6][D][D];
    const int d = 2 * S;
    // T^(2^k) by repeated squaring.
    for (int e = threadIdx.x; e < D * D; e += kThreads) {
        const int i = e / D, j = e % D;
        tp[0][i][j] = (i < d && j < d) ? tmat[i * d + j] : 0.0;
    }
    __syncthreads();
    for (int k = 1; k < 6; ++k) {
        for (int e = threadIdx.x; e < D * D; e += kThreads) {
            const int i = e / D, j = e % D;
            double acc = 0.0;
            for (int m = 0; m < d; ++m) acc = fma(tp[k - 1][i][m], tp[k - 1][m][j], acc);
            tp[k][i][j] = acc;
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (row >= rows) return;
    double carry[D];
#pragma unroll
    for (int i = 0; i < D; ++i) carry[i] = 0.0;
    double *__restrict__ zr = zs + row * nC * d;
    for (int64_t base = 0; base < nC; base += 64) {
        const int64_t c = base + lane;
        double v[D];
#pragma unroll
        for (int i = 0; i < D; ++i) v[i] = (i < d && c < nC) ? zr[c * d + i] : 0.0;
        if (lane == 0) {  // fold the carry: v_0 += T carry
            double w[D];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                double acc = 0.0;
#pragma unroll
                for (int m = 0; m < D; ++m) acc = fma(tp[0][i][m], carry[m], acc);
                w[i] = acc;
            }
#pragma unroll
            for (int i = 0; i < D; ++i) v[i] += w[i];
        }
        // Inclusive: v_c = sum_{j <= c} T^(c-j) e_j over the step.
#pragma unroll 1
        for (int k = 0; k < 6; ++k) {
            const int off = 1 << k;
            double w[D];
#pragma unroll
            for (int i = 0; i < D; ++i) w[i] = __shfl_up(v[i], off, 64);
            if (lane >= off) {
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    double acc = v[i];
#pragma unroll
                    for (int m = 0; m < D; ++m) acc = fma(tp[k][i][m], w[m], acc);
                    v[i] = acc;
                }
            }
        }
        // Exclusive result of chunk c is the inclusive one of chunk c-1.
        double ex[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const double up = __shfl_up(v[i], 1, 64);
            ex[i] = lane == 0 ? carry[i] : up;
        }
        const int last = (int)std::min<int64_t>(63, nC - 1 - base);
#pragma unroll
        for (int i = 0; i < D; ++i) carry[i] = __shfl(v[i], last, 64);
        if (c < nC) {
#pragma unroll
            for (int i = 0; i < D; ++i)
                if (i < d) zr[c * d + i] = ex[i];
        }
    }
}

bool aligned_rows(const double *p, int64_t ld)
{
    return (reinterpret_cast<uintptr_t>(p) & 15) == 0 && (ld & 1) == 0;
}

template <int SMAX, bool REV>
int sos_launch_t(const double *x, int64_t ldx, double *y, int64_t ldy, int64_t rows, int64_t N,
                 const double *sos, int S, int64_t L, int64_t nC, double *zs, double *tmat,
                 hipStream_t s)
{
    const int vec = aligned_rows(x, ldx) && aligned_rows(y, ldy);
    const int64_t items = rows * nC;
    const int64_t blocks = (items + kThreads - 1) / kThreads;
    if (nC > 1) {
        hipLaunchKernelGGL((sos_chunk_kernel<SMAX, kState, REV>), dim3((unsigned)(blocks + 1)),
                           dim3(kThreads), 0, s, x, ldx, y, ldy, rows, N, sos, S, L, nC, zs,
                           tmat, vec);
        SCIR_TRY(check_launch("sos_chunk_kernel(state)"));
        hipLaunchKernelGGL((sos_scan_kernel<SMAX>), dim3((unsigned)((rows + 3) / 4)),
                           dim3(kThreads), 0, s, zs, tmat, rows, nC, S);
        SCIR_TRY(check_launch("sos_scan_kernel"));
    }
    hipLaunchKernelGGL((sos_chunk_kernel<SMAX, kRun, REV>), dim3((unsigned)blocks),
                       dim3(kThreads), 0, s, x, ldx, y, ldy, rows, N, sos, S, L, nC,
                       nC > 1 ? zs : nullptr, tmat, vec);
    return check_launch("sos_chunk_kernel(run)");
}

template <bool REV>
int sos_launch_rev(const double *x, int64_t ldx, double *y, int64_t ldy, int64_t rows, int64_t N,
                   const double *sos, int S, int64_t L, int64_t nC, double *zs, double *tmat,
                   hipStream_t s)
{
#define SCIR_SOS_CASE(n) \
    case n: return sos_launch_t<n, REV>(x, ldx, y, ldy, rows, N, sos, S, L, nC, zs, tmat, s);
    switch (S) {
        SCIR_SOS_CASE(1) SCIR_SOS_CASE(2) SCIR_SOS_CASE(3) SCIR_SOS_CASE(4)
        SCIR_SOS_CASE(5) SCIR_SOS_CASE(6) SCIR_SOS_CASE(7) SCIR_SOS_CASE(8)
        default: return sos_launch_t<kMaxSections, REV>(x, ldx, y, ldy, rows, N, sos, S, L, nC,
                                                        zs, tmat, s);
    }
#undef SCIR_SOS_CASE
}

}  // namespace

// Chunk length: enough chunks for ~4 waves per SIMD, at least 256 samples
// (the chunk states stay small next to the data) and at most 4096 chunks a
// row (the scan walks them 64 at a time).  One chunk per row in reference
// order.
int64_t sos_chunk_length(int64_t rows, int64_t N, int cu_count, bool reference_order)
{
    if (reference_order || N <= 256) return std::max<int64_t>(N, 1);
    const int64_t target_lanes = (int64_t)cu_count * 4 * 4 * 64;
    auto up16 = [](int64_t v) { return (v + kGroup - 1) / kGroup * kGroup; };
    int64_t L = up16((rows * N + target_lanes - 1) / target_lanes);
    L = std::max<int64_t>(L, 256);
    L = std::max<int64_t>(L, up16((N + 4095) / 4096));
    return std::min<int64_t>(L, N);
}

// Validates arguments, sizes the chunk-state scratch and runs one direction
// (caller holds st->mu).
int sos_run(const double *x, int64_t ldx, double *y, int64_t ldy, int64_t rows, int64_t N,
            const double *sos, int S, bool reference_order, bool rev, hipStream_t s,
            DeviceState *st)
{
    if (S == 0) {  // no sections: the cascade passes its input through
        if (x != y)
            SCIR_HIP_TRY(hipMemcpy2DAsync(y, ldy * sizeof(double), x, ldx * sizeof(double),
                                          N * sizeof(double), rows, hipMemcpyDeviceToDevice, s));
        return SCIR_HIP_OK;
    }
    const int64_t L = sos_chunk_length(rows, N, st->cu_count, reference_order);
    const int64_t nC = (N + L - 1) / L;
    if (rows * nC / kThreads + 2 > 0x7fffffffLL)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "too many rows");
    double *zs = nullptr, *tmat = nullptr;
    if (nC > 1) {
        StreamScratch &sc = scratch_for(st, s);
        const size_t zbytes = (size_t)(rows * nC * 2 * S) * sizeof(double);
        const size_t tbytes = (size_t)(4 * S * S) * sizeof(double);
        SCIR_TRY(ensure_buffer(sc.sos_states, zbytes + tbytes + 256, s));
        zs = static_cast<double *>(sc.sos_states.ptr);
        tmat = zs + (zbytes / sizeof(double) + 31) / 32 * 32;
    }
    if (rev) return sos_launch_rev<true>(x, ldx, y, ldy, rows, N, sos, S, L, nC, zs, tmat, s);
    return sos_launch_rev<false>(x, ldx, y, ldy, rows, N, sos, S, L, nC, zs, tmat, s);
}

int sos_args(const double *x, int64_t ldx, int64_t rows, int64_t N, const double *sos, int32_t S,
             const double *y, int64_t ldy, int32_t flags)
{
    clear_error();
    if (rows < 0 || N < 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative size");
    if (S < 0 || S > kMaxSections)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "sections S=%d outside [0, %d]", S,
                    kMaxSections);
    if (flags & ~SCIR_HIP_SOS_REFERENCE_ORDER)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "unknown flags 0x%x", flags);
    if (rows == 0 || N == 0) return SCIR_HIP_OK;
    if (!x || !y || (S > 0 && !sos))
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (ldx < N || ldy < N)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row stride smaller than the row length");
    return SCIR_HIP_OK;
}

}  // namespace scir

using namespace scir;

namespace {

int sos_dev(const double *x, int64_t ldx, int64_t rows, int64_t N, const double *sos, int32_t S,
            double *y, int64_t ldy, int32_t flags, bool zero_phase, void *stream)
{
    SCIR_TRY(sos_args(x, ldx, rows, N, sos, S, y, ldy, flags));
    if (rows == 0 || N == 0) return SCIR_HIP_OK;
    DeviceState *st = nullptr;
    SCIR_TRY(current_device_state(&st));
    std::lock_guard<std::mutex> lk(st->mu);
    const bool ref = (flags & SCIR_HIP_SOS_REFERENCE_ORDER) != 0;
    hipStream_t s = as_stream(stream);
    SCIR_TRY(sos_run(x, ldx, y, ldy, rows, N, sos, S, ref, false, s, st));
    if (zero_phase) SCIR_TRY(sos_run(y, ldy, y, ldy, rows, N, sos, S, ref, true, s, st));
    return SCIR_HIP_OK;
}

int sos_host(const double *x, int64_t ldx, int64_t rows, int64_t N, const double *sos, int32_t S,
             double *y, int64_t ldy, int32_t flags, bool zero_phase)
{
    SCIR_TRY(sos_args(x, ldx, rows, N, sos, S, y, ldy, flags));
    if (rows == 0 || N == 0) return SCIR_HIP_OK;
    DeviceState *st = nullptr;
    SCIR_TRY(current_device_state(&st));
    std::lock_guard<std::mutex> lk(st->mu);
    if (!st->host_stream)
        SCIR_HIP_TRY(hipStreamCreateWithFlags(&st->host_stream, hipStreamNonBlocking));
    hipStream_t s = st->host_stream;
    const size_t row_bytes = (size_t)N * sizeof(double);
    SCIR_TRY(ensure_buffer(st->host_x, (size_t)rows * row_bytes, s));
    SCIR_TRY(ensure_buffer(st->host_b, (size_t)std::max(S, 1) * 6 * sizeof(double), s));
    double *dx = static_cast<double *>(st->host_x.ptr);
    double *dsos = static_cast<double *>(st->host_b.ptr);
    SCIR_HIP_TRY(hipMemcpy2DAsync(dx, row_bytes, x, ldx * sizeof(double), row_bytes, rows,
                                  hipMemcpyHostToDevice, s));
    if (S > 0)
        SCIR_HIP_TRY(hipMemcpyAsync(dsos, sos, (size_t)S * 6 * sizeof(double),
                                    hipMemcpyHostToDevice, s));
    // In place on the device copy: each chunk is read before it is written.
    const bool ref = (flags & SCIR_HIP_SOS_REFERENCE_ORDER) != 0;
    SCIR_TRY(sos_run(dx, N, dx, N, rows, N, dsos, S, ref, false, s, st));
    if (zero_phase) SCIR_TRY(sos_run(dx, N, dx, N, rows, N, dsos, S, ref, true, s, st));
    SCIR_HIP_TRY(hipMemcpy2DAsync(y, ldy * sizeof(double), dx, row_bytes, row_bytes, rows,
                                  hipMemcpyDeviceToHost, s));
    SCIR_HIP_TRY(hipStreamSynchronize(s));
    return SCIR_HIP_OK;
}

}  // namespace

extern "C" {

int scir_hip_sosfilt_f64_dev(const double *x, int64_t ldx, int64_t rows, int64_t n,
                             const double *sos, int32_t S, double *y, int64_t ldy, int32_t flags,
                             void *stream)
{
    return sos_dev(x, ldx, rows, n, sos, S, y, ldy, flags, false, stream);
}

int scir_hip_filtfilt_f64_dev(const double *x, int64_t ldx, int64_t rows, int64_t n,
                              const double *sos, int32_t S, double *y, int64_t ldy, int32_t flags,
                              void *stream)
{
    return sos_dev(x, ldx, rows, n, sos, S, y, ldy, flags, true, stream);
}

int scir_hip_sosfilt_f64_host(const double *x, int64_t ldx, int64_t rows, int64_t n,
                              const double *sos, int32_t S, double *y, int64_t ldy, int32_t flags)
{
    return sos_host(x, ldx, rows, n, sos, S, y, ldy, flags, false);
}

int scir_hip_filtfilt_f64_host(const double *x, int64_t ldx, int64_t rows, int64_t n,
                               const double *sos, int32_t S, double *y, int64_t ldy,
                               int32_t flags)
{
    return sos_host(x, ldx, rows, n, sos, S, y, ldy, flags, true);
}

}  // extern "C"
