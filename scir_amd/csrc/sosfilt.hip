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
// The recursion is sequential in time; linearity parallelises it.  With the
// cascade state z = (s1[0], s2[0], ..., s1[S-1], s2[S-1]) and its zero-input
// transition A, a stretch of n samples maps
//   z_end = A^n z_start + e,   y = y0 + C z_start,
// where e and y0 are the stretch's end state and outputs from zero state and
// C holds the zero-input outputs from unit states.
//
// sos_wave_kernel: one wave filters one chunk of a row, 1024 samples per
// step (16 consecutive samples a lane, so the wave reads and writes 8 KB of
// contiguous memory per step).  Each lane runs its 16 samples from zero
// state (the reference's arithmetic), the wave scans the lanes' end states
// (Kogge-Stone with (A^16)^(2^k)) to get every lane's start state, and each
// lane adds C z_start to its outputs; the last lane's state carries into the
// next step.  With enough rows each row is one chunk: one pass, x read and
// y written once.  With few rows, rows are cut into chunks of L samples:
//   1. sos_wave_kernel<STATE>: each chunk's end state from zero state,
//   2. sos_scan_kernel: per row, z_c = A^L z_{c-1} + e_{c-1},
//   3. sos_wave_kernel<RUN>: each chunk from z_c.
// sos_setup_kernel computes C, (A^16)^(2^k) and A^L first.  The first 16
// outputs of every chunk run from its exact start state (the reference's
// arithmetic, bit for bit in chunk 0); the others carry the rounding of the
// joins (measured ~1e-15 relative).  SCIR_HIP_SOS_REFERENCE_ORDER, and rows
// of at most 1024 samples, use sos_row_kernel instead: one lane a row, the
// reference loop exactly.
#include "common.hpp"

#include <algorithm>

namespace scir {
namespace {

constexpr int kThreads = 256;
constexpr int kGroup = 16;             // samples a lane filters per step
constexpr int kStep = 64 * kGroup;     // samples a wave filters per step
constexpr int kMaxSections = 16;       // SCIR_HIP_SOS_MAX_SECTIONS
constexpr int kPowers = 6;             // (A^16)^(2^k), k < 6: lane offsets 1..32

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

// Layout of the matrices sos_setup_kernel writes, D = 2*SMAX (d = 2S <= D;
// entries beyond d are zero):
//   C [16][D]            zero-input output k from unit state j
//   P [kPowers][D][D]    (A^16)^(2^k)
//   T [d][d]             A^L (the chunk transition; when there are chunks)
// The wave kernel reads C and P with uniform indices, which the compiler
// turns into scalar loads: the matrices sit in the scalar cache and reach
// the multiply-adds as scalar operands.
__host__ __device__ constexpr int64_t mats_p(int D) { return (int64_t)kGroup * D; }
__host__ __device__ constexpr int64_t mats_t(int D)
{
    return (int64_t)kGroup * D + (int64_t)kPowers * D * D;
}
__host__ __device__ constexpr int64_t mats_size(int D, int d) { return mats_t(D) + (int64_t)d * d; }

// One workgroup.  q = L/16 when a chunk transition is needed, else 0.
template <int SMAX>
__global__ __launch_bounds__(kThreads) void sos_setup_kernel(const double *__restrict__ sos, int S,
                                                             int64_t q, double *__restrict__ mats)
{
    constexpr int D = 2 * SMAX;
    __shared__ double base[D][D], res[D][D], tpr[D][D], tsq[D][D];
    const int d = 2 * S;
    const int tid = threadIdx.x;
    if (tid >= d && tid < D) {  // padding columns of C
        for (int k = 0; k < kGroup; ++k) mats[k * D + tid] = 0.0;
    }
    if (tid < d) {
        // 16 zero-input steps from unit state tid: C's column and A^16's
        Coefs<SMAX> cf;
        load_coefs(cf, sos, S);
        double s1[SMAX], s2[SMAX];
#pragma unroll
        for (int i = 0; i < SMAX; ++i) {
            s1[i] = (2 * i == tid) ? 1.0 : 0.0;
            s2[i] = (2 * i + 1 == tid) ? 1.0 : 0.0;
        }
        for (int k = 0; k < kGroup; ++k)
            mats[k * D + tid] = df2t_step<SMAX>(0.0, s1, s2, cf, S);
#pragma unroll
        for (int i = 0; i < SMAX; ++i) {
            if (i < S) {
                base[2 * i][tid] = s1[i];
                base[2 * i + 1][tid] = s2[i];
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < d * d; e += kThreads) res[e / d][e % d] = (e / d == e % d) ? 1.0 : 0.0;
    for (int e = tid; e < D * D; e += kThreads) {
        const int i = e / D, j = e % D;
        mats[mats_p(D) + e] = (i < d && j < d) ? base[i][j] : 0.0;
    }
    __syncthreads();
    // base runs through (A^16)^(2^k); res collects A^L = (A^16)^q bit by bit.
    for (int k = 0;; ++k) {
        const bool bit = q > 0 && ((q >> k) & 1);
        const bool more = (k + 1 < kPowers) || ((q >> (k + 1)) > 0);
        for (int e = tid; e < d * d; e += kThreads) {
            const int i = e / d, j = e % d;
            double pr = 0.0, sq = 0.0;
            for (int m = 0; m < d; ++m) {
                pr = fma(res[i][m], base[m][j], pr);
                sq = fma(base[i][m], base[m][j], sq);
            }
            tpr[i][j] = pr;
            tsq[i][j] = sq;
        }
        __syncthreads();
        for (int e = tid; e < d * d; e += kThreads) {
            const int i = e / d, j = e % d;
            if (bit) res[i][j] = tpr[i][j];
            if (more) base[i][j] = tsq[i][j];
        }
        if (more && k + 1 < kPowers) {  // padded to D x D with zeros
            for (int e = tid; e < D * D; e += kThreads) {
                const int i = e / D, j = e % D;
                mats[mats_p(D) + (int64_t)(k + 1) * D * D + e] = (i < d && j < d) ? tsq[i][j] : 0.0;
            }
        }
        __syncthreads();
        if (!more) break;
    }
    if (q > 0)
        for (int e = tid; e < d * d; e += kThreads) mats[mats_t(D) + e] = res[e / d][e % d];
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive scan over the wave of the lanes' end states: lane l ends with
// sum_{j <= l} (A^16)^(l-j) e_j.
template <int D>
__device__ __forceinline__ void wave_scan(double (&e)[D], const double *__restrict__ pw, int lane)
{
#pragma unroll 1
    for (int k = 0; k < kPowers; ++k) {
        const int off = 1 << k;
        double w[D];
#pragma unroll
        for (int i = 0; i < D; ++i) w[i] = __shfl_up(e[i], off, 64);
        if (lane >= off) {
#pragma unroll
            for (int i = 0; i < D; ++i) {
                double acc = e[i];
#pragma unroll
                for (int m = 0; m < D; ++m) acc = fma(pw[(k * D + i) * D + m], w[m], acc);
                e[i] = acc;
            }
        }
    }
}

// Loads a lane's group (cnt samples from xr[lo]); missing samples are 0.
__device__ __forceinline__ void load_lane(double (&v)[kGroup], const double *xr, int64_t lo,
                                          int64_t cnt, bool vec)
{
    if (vec && cnt == kGroup) {
#pragma unroll
        for (int k = 0; k < kGroup / 2; ++k) {
            const double2 t = *reinterpret_cast<const double2 *>(xr + lo + 2 * k);
            v[2 * k] = t.x;
            v[2 * k + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kGroup; ++k) v[k] = (k < cnt) ? xr[lo + k] : 0.0;
    }
}

// Group gidx (in time order) of the chunk [p0, p1): positions [lo, lo+cnt).
// Back to front, group 0 is the one holding p1-1 and only it can be short.
template <bool REV>
__device__ __forceinline__ void group_span(int64_t gidx, int64_t ngroups, int64_t p0, int64_t p1,
                                           int64_t *lo, int64_t *cnt)
{
    if (gidx >= ngroups) {
        *lo = p0;
        *cnt = 0;
    } else if (!REV) {
        *lo = p0 + gidx * kGroup;
        *cnt = std::min<int64_t>(kGroup, p1 - *lo);
    } else {
        const int64_t m = (p1 - 1) / kGroup - gidx;
        *lo = std::max<int64_t>(p0, m * kGroup);
        *cnt = std::min<int64_t>(p1, m * kGroup + kGroup) - *lo;
    }
}

// One wave per chunk (see the top of the file).  MODE kState: chunks
// 0..nC-2 of every row from zero state, end states to zs[(row*nC + c)*2S].
// MODE kRun: every chunk from zs[...] (zero when zs is null), outputs to y.
// REV runs each row back to front; chunk c is c-th in time, and back to
// front the short chunk (when L does not divide N) comes first, so every
// chunk starts at a multiple of L (16-byte aligned groups) and only the
// first group in time can be short; every chunk after the first in time is
// L long, so A^L is its transition.
template <int SMAX, int MODE, bool REV>
__global__ __launch_bounds__(kThreads) void sos_wave_kernel(
    const double *x, int64_t ldx, double *y, int64_t ldy, int64_t rows, int64_t N,
    const double *__restrict__ sos, int S, int64_t L, int64_t nC,
    const double *__restrict__ mats, double *__restrict__ zs, int vec)
{
    constexpr int D = 2 * SMAX;
    __shared__ __attribute__((aligned(16))) double tiles[kThreads / 64][kStep];
    const int d = 2 * S;

    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (wave >= rows * nC) return;  // the whole wave
    const int64_t row = wave / nC;
    const int64_t c = wave - row * nC;
    if (MODE == kState && c == nC - 1) return;  // its end state is never used
    const int64_t p0 = (REV ? nC - 1 - c : c) * L;
    const int64_t p1 = std::min<int64_t>(p0 + L, N);
    const int64_t ngroups = REV ? (p1 - 1) / kGroup - p0 / kGroup + 1
                                : (p1 - p0 + kGroup - 1) / kGroup;
    const int64_t nsteps = (ngroups + 63) / 64;
    Coefs<SMAX> cf;
    load_coefs(cf, sos, S);
    double zin[D];  // the wave's state at the start of the step (same in every lane)
#pragma unroll
    for (int i = 0; i < D; ++i)
        zin[i] = (MODE == kRun && zs != nullptr && i < d) ? zs[(row * nC + c) * d + i] : 0.0;
    // x and y may be the same array (filtfilt's second pass runs in place):
    // a step's samples are all read before any is written.
    const double *xr = x + row * ldx;
    double *yr = (MODE == kRun) ? y + row * ldy : nullptr;
    const bool vb = vec != 0;
    double *tile = tiles[threadIdx.x >> 6];
    // A step whose 64 groups are all whole covers 8 KB of contiguous memory:
    // 64 lines of 128 bytes, line lam = lane forward (lane 63 - lam back to
    // front).  It moves through the wave's LDS tile: memory instruction k
    // carries 16-byte pieces of lines 8k..8k+7 (a permutation of 1 KB, so
    // it is coalesced), and piece j of line lam sits in tile slot
    //   64*(lam/8) + 8*(lam%8) + (j + lam%8) % 8     (slots of 16 bytes),
    // which makes both the instruction-ordered and the line-ordered LDS
    // accesses at most 2-way bank conflicted.
    const int lam = REV ? 63 - lane : lane;
    const int own_base = 64 * (lam >> 3) + 8 * (lam & 7);  // + (j + lam%8) % 8
    // this lane's piece in memory instruction k: line 8k + lane/8, piece pj
    const int mline = lane >> 3;
    const int pj = ((lane & 7) - mline) & 7;

    for (int64_t st = 0; st < nsteps; ++st) {
        // The matrices are re-read (scalar loads) every step: an opaque zero
        // offset keeps the compiler from holding all of them in registers.
        int zoff = 0;
        asm volatile("" : "+s"(zoff));
        const double *__restrict__ cmv = mats + zoff;
        const double *__restrict__ pwv = mats + mats_p(D) + zoff;
        int64_t lo, cnt;
        group_span<REV>(st * 64 + lane, ngroups, p0, p1, &lo, &cnt);
        const bool full = cnt == kGroup;
        const bool tiled = vb && __all(full);
        // first sample of the step's 8 KB block when tiled
        const int64_t blk = REV ? ((p1 - 1) / kGroup - st * 64 - 63) * kGroup : p0 + st * kStep;
        double v[kGroup];
        if (tiled) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double2 t = *reinterpret_cast<const double2 *>(
                    xr + blk + (8 * k + mline) * kGroup + 2 * pj);
                *reinterpret_cast<double2 *>(tile + 2 * (64 * k + lane)) = t;
            }
            wave_sync();
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const double2 t = *reinterpret_cast<const double2 *>(
                    tile + 2 * (own_base + ((j + lam) & 7)));
                v[2 * j] = t.x;
                v[2 * j + 1] = t.y;
            }
            wave_sync();
        } else {
            load_lane(v, xr, lo, cnt, vb);
        }
        // The lane's samples from zero state, in time order.
        double s1[SMAX], s2[SMAX];
#pragma unroll
        for (int i = 0; i < SMAX; ++i) s1[i] = s2[i] = 0.0;
        if (__all(full)) {
            if (!REV) {
#pragma unroll
                for (int k = 0; k < kGroup; ++k) v[k] = df2t_step<SMAX>(v[k], s1, s2, cf, S);
            } else {
#pragma unroll
                for (int k = kGroup - 1; k >= 0; --k) v[k] = df2t_step<SMAX>(v[k], s1, s2, cf, S);
            }
        } else {
            if (!REV) {
#pragma unroll
                for (int k = 0; k < kGroup; ++k)
                    if (k < cnt) v[k] = df2t_step<SMAX>(v[k], s1, s2, cf, S);
            } else {
#pragma unroll
                for (int k = kGroup - 1; k >= 0; --k)
                    if (k < cnt) v[k] = df2t_step<SMAX>(v[k], s1, s2, cf, S);
            }
        }
        double e[D];
#pragma unroll
        for (int i = 0; i < SMAX; ++i) {
            e[2 * i] = s1[i];
            e[2 * i + 1] = s2[i];
        }
        if (lane == 0) {  // the carried state enters through the first lane
#pragma unroll
            for (int i = 0; i < D; ++i) {
                double acc = e[i];
#pragma unroll
                for (int m = 0; m < D; ++m) acc = fma(pwv[i * D + m], zin[m], acc);
                e[i] = acc;
            }
        }
        wave_scan<D>(e, pwv, lane);
        double z[D];  // the lane's start state
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const double up = __shfl_up(e[i], 1, 64);
            z[i] = lane == 0 ? zin[i] : up;
        }
        // carry: the state after the step's last group (back to front the
        // first chunk in time is the short one, and its end state is used)
        const int last = (int)std::min<int64_t>(63, ngroups - 1 - st * 64);
#pragma unroll
        for (int i = 0; i < D; ++i) zin[i] = __shfl(e[i], last, 64);
        if (MODE == kRun) {
            // y = y0 + C z; the chunk's very first lane starts from zero
            // (and keeps the reference's bits, signed zeros included).
            if (!(st == 0 && lane == 0 && c == 0)) {
#pragma unroll
                for (int k = 0; k < kGroup; ++k) {
                    const int tk = REV ? (int)cnt - 1 - k : k;  // time index of v[k]
                    if (k < cnt) {
                        double acc = v[k];
#pragma unroll
                        for (int m = 0; m < D; ++m) acc = fma(cmv[tk * D + m], z[m], acc);
                        v[k] = acc;
                    }
                }
            }
            if (tiled) {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    *reinterpret_cast<double2 *>(tile + 2 * (own_base + ((j + lam) & 7))) =
                        make_double2(v[2 * j], v[2 * j + 1]);
                wave_sync();
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const double2 t = *reinterpret_cast<const double2 *>(tile + 2 * (64 * k + lane));
                    *reinterpret_cast<double2 *>(yr + blk + (8 * k + mline) * kGroup + 2 * pj) = t;
                }
                wave_sync();
            } else if (vb && full) {
#pragma unroll
                for (int k = 0; k < kGroup / 2; ++k)
                    *reinterpret_cast<double2 *>(yr + lo + 2 * k) = make_double2(v[2 * k], v[2 * k + 1]);
            } else {
#pragma unroll
                for (int k = 0; k < kGroup; ++k)
                    if (k < cnt) yr[lo + k] = v[k];
            }
        }
    }
    if (MODE == kState && lane == 0) {
#pragma unroll
        for (int i = 0; i < D; ++i)
            if (i < d) zs[(row * nC + c) * d + i] = zin[i];
    }
}

// One lane per row, the reference loop exactly (reference order, and rows
// too short to split).  A wave serves `rpw` rows (its other lanes idle), so
// that few rows still spread over many CUs, and each lane keeps kRowAhead
// groups of 16 samples in flight (the recursion itself is sequential).
constexpr int kRowAhead = 4;

template <int SMAX, bool REV>
__global__ __launch_bounds__(64) void sos_row_kernel(const double *x, int64_t ldx, double *y,
                                                     int64_t ldy, int64_t rows, int64_t N,
                                                     const double *__restrict__ sos, int S,
                                                     int vec, int rpw)
{
    const int lane = threadIdx.x;
    if (lane >= rpw) return;
    const int64_t row = (int64_t)blockIdx.x * rpw + lane;
    if (row >= rows) return;
    Coefs<SMAX> cf;
    load_coefs(cf, sos, S);
    double s1[SMAX], s2[SMAX];
#pragma unroll
    for (int i = 0; i < SMAX; ++i) s1[i] = s2[i] = 0.0;
    const double *xr = x + row * ldx;
    double *yr = y + row * ldy;
    // Whole groups [16m, 16m+16) for m < nfull; back to front the samples
    // above the last whole group come first in time.
    const int64_t nfull = vec ? N / kGroup : 0;
    if (REV)
        for (int64_t t = N - 1; t >= nfull * kGroup; --t) yr[t] = df2t_step<SMAX>(xr[t], s1, s2, cf, S);
    if (nfull > 0) {
        // ring[j] holds group number (in time) g + j
        double2 ring[kRowAhead][kGroup / 2];
        auto gbase = [&](int64_t g) { return (REV ? nfull - 1 - g : g) * kGroup; };
#pragma unroll
        for (int j = 0; j < kRowAhead; ++j)
            if (j < nfull) {
#pragma unroll
                for (int k = 0; k < kGroup / 2; ++k)
                    ring[j][k] = *reinterpret_cast<const double2 *>(xr + gbase(j) + 2 * k);
            }
        for (int64_t g = 0; g < nfull; g += kRowAhead) {
#pragma unroll
            for (int j = 0; j < kRowAhead; ++j) {
                if (g + j < nfull) {
                    double v[kGroup];
#pragma unroll
                    for (int k = 0; k < kGroup / 2; ++k) {
                        v[2 * k] = ring[j][k].x;
                        v[2 * k + 1] = ring[j][k].y;
                    }
                    // refill this slot with the group kRowAhead later (read
                    // before any write of it: in place is safe)
                    if (g + j + kRowAhead < nfull) {
#pragma unroll
                        for (int k = 0; k < kGroup / 2; ++k)
                            ring[j][k] = *reinterpret_cast<const double2 *>(
                                xr + gbase(g + j + kRowAhead) + 2 * k);
                    }
                    if (!REV) {
#pragma unroll
                        for (int k = 0; k < kGroup; ++k) v[k] = df2t_step<SMAX>(v[k], s1, s2, cf, S);
                    } else {
#pragma unroll
                        for (int k = kGroup - 1; k >= 0; --k)
                            v[k] = df2t_step<SMAX>(v[k], s1, s2, cf, S);
                    }
                    const int64_t base = gbase(g + j);
#pragma unroll
                    for (int k = 0; k < kGroup / 2; ++k)
                        *reinterpret_cast<double2 *>(yr + base + 2 * k) =
                            make_double2(v[2 * k], v[2 * k + 1]);
                }
            }
        }
    }
    if (!REV)
        for (int64_t t = nfull * kGroup; t < N; ++t) yr[t] = df2t_step<SMAX>(xr[t], s1, s2, cf, S);
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
                 const double *sos, int S, bool row_mode, int64_t L, int64_t nC, double *mats,
                 double *zs, int cu_count, hipStream_t s)
{
    const int vec = aligned_rows(x, ldx) && aligned_rows(y, ldy);
    if (row_mode) {
        // rows per wave: spread the rows over the CUs (one wave each) first
        const int rpw = (int)std::min<int64_t>(64, std::max<int64_t>(1, (rows + cu_count - 1) / cu_count));
        hipLaunchKernelGGL((sos_row_kernel<SMAX, REV>), dim3((unsigned)((rows + rpw - 1) / rpw)),
                           dim3(64), 0, s, x, ldx, y, ldy, rows, N, sos, S, vec, rpw);
        return check_launch("sos_row_kernel");
    }
    hipLaunchKernelGGL((sos_setup_kernel<SMAX>), dim3(1), dim3(kThreads), 0, s, sos, S,
                       nC > 1 ? L / kGroup : (int64_t)0, mats);
    SCIR_TRY(check_launch("sos_setup_kernel"));
    const int64_t waves = rows * nC;
    const unsigned blocks = (unsigned)((waves + kThreads / 64 - 1) / (kThreads / 64));
    if (nC > 1) {
        hipLaunchKernelGGL((sos_wave_kernel<SMAX, kState, REV>), dim3(blocks), dim3(kThreads), 0, s,
                           x, ldx, y, ldy, rows, N, sos, S, L, nC, mats, zs, vec);
        SCIR_TRY(check_launch("sos_wave_kernel(state)"));
        hipLaunchKernelGGL((sos_scan_kernel<SMAX>), dim3((unsigned)((rows + 3) / 4)),
                           dim3(kThreads), 0, s, zs, mats + mats_t(2 * SMAX), rows, nC, S);
        SCIR_TRY(check_launch("sos_scan_kernel"));
    }
    hipLaunchKernelGGL((sos_wave_kernel<SMAX, kRun, REV>), dim3(blocks), dim3(kThreads), 0, s, x,
                       ldx, y, ldy, rows, N, sos, S, L, nC, mats, nC > 1 ? zs : nullptr, vec);
    return check_launch("sos_wave_kernel(run)");
}

template <bool REV>
int sos_launch_rev(const double *x, int64_t ldx, double *y, int64_t ldy, int64_t rows, int64_t N,
                   const double *sos, int S, bool row_mode, int64_t L, int64_t nC, double *mats,
                   double *zs, int cu_count, hipStream_t s)
{
#define SCIR_SOS_CASE(n)                                                                     \
    case n:                                                                                  \
        return sos_launch_t<n, REV>(x, ldx, y, ldy, rows, N, sos, S, row_mode, L, nC, mats, \
                                    zs, cu_count, s);
    switch (S) {
        SCIR_SOS_CASE(1) SCIR_SOS_CASE(2) SCIR_SOS_CASE(3) SCIR_SOS_CASE(4)
        SCIR_SOS_CASE(5) SCIR_SOS_CASE(6) SCIR_SOS_CASE(7) SCIR_SOS_CASE(8)
        default:
            return sos_launch_t<kMaxSections, REV>(x, ldx, y, ldy, rows, N, sos, S, row_mode, L,
                                                   nC, mats, zs, cu_count, s);
    }
#undef SCIR_SOS_CASE
}

}  // namespace

// Work split.  Reference order, and rows of at most one wave step, run one
// lane per row.  Otherwise one wave per chunk: whole rows when there are
// enough of them for ~4 waves per SIMD, else chunks of L samples (a
// multiple of one wave step, at most 4096 chunks a row: the chunk scan
// walks them 64 at a time).
void sos_plan(int64_t rows, int64_t N, int cu_count, bool reference_order, bool *row_mode,
              int64_t *L, int64_t *nC)
{
    *row_mode = reference_order || N <= kStep;
    *L = N;
    *nC = 1;
    if (*row_mode) return;
    const int64_t target_waves = (int64_t)cu_count * 16;
    if (rows >= target_waves) return;
    auto up = [](int64_t v) { return (v + kStep - 1) / kStep * kStep; };
    int64_t l = up((rows * N + target_waves - 1) / target_waves);
    l = std::max<int64_t>(l, up((N + 4095) / 4096));
    if (l >= N) return;
    *L = l;
    *nC = (N + l - 1) / l;
}

// Runs one direction (caller holds st->mu).
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
    bool row_mode = false;
    int64_t L = 0, nC = 0;
    sos_plan(rows, N, st->cu_count, reference_order, &row_mode, &L, &nC);
    if ((row_mode ? rows : rows * nC / (kThreads / 64)) + 2 > 0x7fffffffLL)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "too many rows");
    double *mats = nullptr, *zs = nullptr;
    if (!row_mode) {
        StreamScratch &sc = scratch_for(st, s);
        const int d = 2 * S;
        const int Dk = 2 * (S <= 8 ? S : kMaxSections);  // the kernels' padded dimension
        const int64_t mdoubles = (mats_size(Dk, d) + 31) / 32 * 32;
        const int64_t zdoubles = nC > 1 ? rows * nC * d : 0;
        SCIR_TRY(ensure_scratch(st, sc.sos_states, (size_t)(mdoubles + zdoubles) * sizeof(double), s));
        mats = static_cast<double *>(sc.sos_states.ptr);
        zs = mats + mdoubles;
    }
    if (rev)
        return sos_launch_rev<true>(x, ldx, y, ldy, rows, N, sos, S, row_mode, L, nC, mats, zs,
                                    st->cu_count, s);
    return sos_launch_rev<false>(x, ldx, y, ldy, rows, N, sos, S, row_mode, L, nC, mats, zs,
                                 st->cu_count, s);
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
    return guarded(__func__, [&]() -> int {
        return sos_dev(x, ldx, rows, n, sos, S, y, ldy, flags, false, stream);
    });
}

int scir_hip_filtfilt_f64_dev(const double *x, int64_t ldx, int64_t rows, int64_t n,
                              const double *sos, int32_t S, double *y, int64_t ldy, int32_t flags,
                              void *stream)
{
    return guarded(__func__, [&]() -> int {
        return sos_dev(x, ldx, rows, n, sos, S, y, ldy, flags, true, stream);
    });
}

int scir_hip_sosfilt_f64_host(const double *x, int64_t ldx, int64_t rows, int64_t n,
                              const double *sos, int32_t S, double *y, int64_t ldy, int32_t flags)
{
    return guarded(__func__, [&]() -> int {
        return sos_host(x, ldx, rows, n, sos, S, y, ldy, flags, false);
    });
}

int scir_hip_filtfilt_f64_host(const double *x, int64_t ldx, int64_t rows, int64_t n,
                               const double *sos, int32_t S, double *y, int64_t ldy,
                               int32_t flags)
{
    return guarded(__func__, [&]() -> int {
        return sos_host(x, ldx, rows, n, sos, S, y, ldy, flags, true);
    });
}

}  // extern "C"
