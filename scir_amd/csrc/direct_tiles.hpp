// Device-side building blocks of the FP32 direct-form FIR, shared by the
// dense kernels (fir_direct.hip) and the ragged persistent kernel
// (fir_ragged.hip).
//
// A tile of T = 256*R consecutive outputs is computed by one 256-lane
// workgroup: the tile's inputs plus the (K-1)-sample history sit in LDS, lane
// l owns outputs l*R .. l*R+R-1 and slides a register window over the LDS
// copy, applying taps 32 at a time from scalar registers; remainders use
// chunks of 16, 8, 4, 2, 1 taps so no tap is padded.
//
// LDS layout: linear, sample s of the staged window at float s.  A window
// read (16 lanes 32 bytes apart, 16-byte loads) meets 2-way bank conflicts,
// but every read address is the chunk's base plus a constant offset, which
// saves the VALU instructions a padded or swizzled address costs: 257
// instead of 296 VALU instructions per 32 taps and 8 outputs, measured
// 384 vs ~395 us at K = 33 and 97.7 % of the non-packed FP32 peak at
// K >= 1000.  Tiles staged by LDS DMA (64 consecutive 16-byte chunks per
// wave instruction) cannot be padded at all.
//
// EXACT = true keeps the reference's arithmetic: separately rounded multiply
// and add, taps accumulated in the order k = 0..K-1 from +0.0f
// (crates/scir-gpu/src/lib.rs:1141-1148 adds taps[K-1-t]*x[i-t] for t
// ascending, and b[t] == taps[K-1-t]).  For outputs i < K-1 the tile loop
// also forms b[k]*0 for k > i (the zeros staged before the row start), which
// the reference loop never forms.  With a finite b[k] such a term is +-0 and
// leaves every bit unchanged (the running sum is never -0.0); with b[k] = Inf
// or NaN it makes the output NaN, so every non-finite output i < K-1 is
// recomputed from the reference loop's terms alone (fix_ramp_in).
#pragma once

#include "common.hpp"

namespace scir {

// Output i (logical index) of a row, from the reference loop's terms only:
// sum_{k=0}^{min(i, K-1)} b[k] * x[i-k] in ascending k from +0.0f, with a
// separately rounded multiply and add (EXACT, the order of
// crates/scir-gpu/src/lib.rs:1141-1148) or a fused multiply-add.  x is read
// from global memory; with `rev` logical sample t is physical N-1-t.  This is
// the slow path every kernel takes for an output that came out non-finite:
// zero padding (the direct tiles' ramp-in, a Toeplitz block's zero taps, an
// FFT segment) spreads an Inf or NaN to outputs whose reference terms are
// all finite, and this recomputation gives the reference loop's value class.
template <bool EXACT>
__device__ float direct_output(const float *__restrict__ xrow, int64_t N, bool rev,
                               const float *__restrict__ b, int K, int64_t i)
{
    float acc = 0.0f;
    const int64_t kend = i < (int64_t)K - 1 ? i : (int64_t)K - 1;
#pragma clang loop unroll(disable)
    for (int64_t k = 0; k <= kend; ++k) {
        const int64_t t = i - k;
        const float xv = xrow[rev ? N - 1 - t : t];
        if (EXACT) {
#pragma clang fp contract(off)
            const float p = b[k] * xv;
            acc = acc + p;
        } else {
            acc = __builtin_fmaf(b[k], xv, acc);
        }
    }
    return acc;
}

// direct_output for an output of the ramp-in (i < K-1, i < N); any other
// output keeps the value v the tile loop computed.  Not inlined: the slow
// path stays out of the register allocation of the kernels' tile loops.
template <bool EXACT>
__device__ __noinline__ float ramp_in_output(float v, const float *__restrict__ xrow, int64_t N,
                                             bool rev, const float *__restrict__ b, int K,
                                             int64_t i)
{
    if (i >= (int64_t)K - 1 || i >= N) return v;
    return direct_output<EXACT>(xrow, N, rev, b, K, i);
}

namespace direct {

constexpr int kThreads = 256;
constexpr int kSuper = 2048;  // taps per LDS super-chunk (multiple of 32)
constexpr int kChunk = 32;    // taps per fully unrolled step

typedef float f32x4 __attribute__((ext_vector_type(4)));
// Four floats at any 4-byte aligned global address (the hardware runs in
// unaligned access mode): rows are read and written with 16-byte accesses
// whatever their offset or pitch.
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <bool EXACT>
__device__ __forceinline__ float mac(float acc, float t, float xv);

template <>
__device__ __forceinline__ float mac<true>(float acc, float t, float xv)
{
#pragma clang fp contract(off)
    float p = t * xv;
    return acc + p;
}

template <>
__device__ __forceinline__ float mac<false>(float acc, float t, float xv)
{
    return __builtin_fmaf(t, xv, acc);
}

// acc[r] += sum_{kk < C} bk[kk] * x[i_r - k0 - kk], where the LDS window
// starts OFF floats below the 4-aligned index `wa`:
//   x[i_r - k0 - kk] == xs[wa + OFF + r + C - kk].
// With FIRST the chunk starts the sums: its first tap is accumulated onto 0
// (the same arithmetic as zeroed accumulators, with no zeroing instructions).
template <int R, int C, int OFF, bool EXACT, bool FIRST = false>
__device__ __forceinline__ void taps_chunk(float (&acc)[R], const float *xs, int wa,
                                           const float *__restrict__ bk)
{
    constexpr int W4 = (R + C + OFF + 3) / 4;
    float w[W4 * 4];
#pragma unroll
    for (int j = 0; j < W4; ++j) {
        f32x4 v = *reinterpret_cast<const f32x4 *>(xs + wa + 4 * j);
        w[4 * j + 0] = v.x;
        w[4 * j + 1] = v.y;
        w[4 * j + 2] = v.z;
        w[4 * j + 3] = v.w;
    }
#pragma unroll
    for (int kk = 0; kk < C; ++kk) {
        const float t = bk[kk];
#pragma unroll
        for (int r = 0; r < R; ++r)
            acc[r] = mac<EXACT>(FIRST && kk == 0 ? 0.0f : acc[r], t, w[OFF + r + C - kk]);
    }
}

// Logical samples g..g+3 of a row as a float4, zero outside [0, N).  With
// `rev` the row is read back to front (logical i == physical N-1-i), which
// turns the causal filter into the anti-causal correlation of the adjoint.
// One 16-byte load when the four samples are inside the row.
__device__ __forceinline__ f32x4 load4(const float *__restrict__ xrow, int64_t g, int64_t N,
                                       bool rev)
{
    f32x4 v;
    if (g >= 0 && g + 3 < N) {
        if (!rev) {
            v = __builtin_nontemporal_load(reinterpret_cast<const f32x4u *>(xrow + g));
        } else {
            f32x4 t = __builtin_nontemporal_load(reinterpret_cast<const f32x4u *>(xrow + (N - 4 - g)));
            v = f32x4{t.w, t.z, t.y, t.x};
        }
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int64_t i = g + c;
            v[c] = (i >= 0 && i < N) ? xrow[rev ? N - 1 - i : i] : 0.0f;
        }
    }
    return v;
}

// Stages logical samples g0 .. g0 + count (count and g0 multiples of 4) into
// xs (padded layout).
__device__ __forceinline__ void stage_row(float *xs, const float *__restrict__ xrow, int64_t g0,
                                          int count, int64_t N, bool rev)
{
    for (int v = threadIdx.x; v < count / 4; v += kThreads)
        *reinterpret_cast<f32x4 *>(xs + 4 * v) = load4(xrow, g0 + 4 * (int64_t)v, N, rev);
}

// acc[r] += sum_{k' < kcount} bs[k'] * x[i_r - k'] with the window in LDS:
// x[i_r - k'] == xs[H + lane_base + r - k'].  Chunks run in ascending k'.
// With INIT the sums start from zero instead (acc[r] = sum ...).
template <int R, bool EXACT, bool INIT = false>
__device__ __forceinline__ void taps_from_lds(float (&acc)[R], const float *xs, int H,
                                              int lane_base, int kcount,
                                              const float *__restrict__ bs)
{
    // INIT: acc holds no partial sums on entry; the first 32-tap chunk (or,
    // for fewer taps, explicit zeros) starts them.
    int kc = 0;
    if (INIT) {
        if (kcount >= kChunk) {
            taps_chunk<R, kChunk, 0, EXACT, true>(acc, xs, H + lane_base - kChunk, bs);
            kc = kChunk;
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = 0.0f;
        }
    }
    for (; kc + kChunk <= kcount; kc += kChunk)
        taps_chunk<R, kChunk, 0, EXACT>(acc, xs, H + lane_base - kc - kChunk, bs + kc);
    const int rem = kcount - kc;
    if (rem & 16) {
        taps_chunk<R, 16, 0, EXACT>(acc, xs, H + lane_base - kc - 16, bs + kc);
        kc += 16;
    }
    if (rem & 8) {
        taps_chunk<R, 8, 0, EXACT>(acc, xs, H + lane_base - kc - 8, bs + kc);
        kc += 8;
    }
    if (rem & 4) {
        taps_chunk<R, 4, 0, EXACT>(acc, xs, H + lane_base - kc - 4, bs + kc);
        kc += 4;
    }
    // kc % 4 == 0 here.  The last one or two taps read windows that start 1,
    // 2 or 3 below a 4-aligned index; each case has its own branch, so no
    // window registers are merged (copied) between the branches.
    if (rem & 2) {  // window starts 2 below the aligned index
        taps_chunk<R, 2, 2, EXACT>(acc, xs, H + lane_base - kc - 4, bs + kc);
        if (rem & 1)  // kc + 2: window starts 1 below the aligned index
            taps_chunk<R, 1, 1, EXACT>(acc, xs, H + lane_base - kc - 4, bs + kc + 2);
    } else if (rem & 1) {  // window starts 3 below the aligned index
        taps_chunk<R, 1, 3, EXACT>(acc, xs, H + lane_base - kc - 4, bs + kc);
    }
}

// A lane's outputs i0 .. i0+R-1 (logical): each non-finite one below K-1 --
// the only outputs whose tile sums include staged zeros from before the row
// start -- is recomputed from the reference loop's terms (direct_output).
template <int R, bool EXACT>
__device__ __forceinline__ void fix_ramp_in(float (&acc)[R], const float *__restrict__ xrow,
                                            int64_t N, bool rev, const float *__restrict__ b,
                                            int K, int64_t i0)
{
    if (i0 >= (int64_t)K - 1) return;
    unsigned bad = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) bad |= __builtin_isfinite(acc[r]) ? 0u : 1u << r;
    if (bad) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (bad & (1u << r)) acc[r] = ramp_in_output<EXACT>(acc[r], xrow, N, rev, b, K, i0 + r);
    }
}

template <int R>
__device__ __forceinline__ void store_outputs(const float (&acc)[R], float *__restrict__ yrow,
                                              int64_t i0, int64_t N, bool rev)
{
    if (i0 + R <= N) {
#pragma unroll
        for (int r = 0; r < R; r += 4) {
            if (!rev) {
                f32x4 v = {acc[r], acc[r + 1], acc[r + 2], acc[r + 3]};
                *reinterpret_cast<f32x4u *>(yrow + i0 + r) = v;
            } else {
                f32x4 v = {acc[r + 3], acc[r + 2], acc[r + 1], acc[r]};
                *reinterpret_cast<f32x4u *>(yrow + (N - 4 - i0 - r)) = v;
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t i = i0 + r;
            if (i < N) yrow[rev ? N - 1 - i : i] = acc[r];
        }
    }
}

}  // namespace direct
}  // namespace scir
