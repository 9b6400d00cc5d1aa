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
// ascending, and b[t] == taps[K-1-t]).  Extra terms before the row start add
// +-0, which leaves every bit unchanged (the running sum is never -0.0).
#pragma once

#include "common.hpp"

namespace scir {
namespace direct {

constexpr int kThreads = 256;
constexpr int kSuper = 2048;  // taps per LDS super-chunk (multiple of 32)
constexpr int kChunk = 32;    // taps per fully unrolled step

typedef float f32x4 __attribute__((ext_vector_type(4)));
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
// `vec` promises 16-byte aligned rows (and N % 4 == 0 when rev is set).
__device__ __forceinline__ f32x4 load4(const float *__restrict__ xrow, int64_t g, int64_t N,
                                       bool vec, bool rev)
{
    f32x4 v;
    if (vec && g >= 0 && g + 3 < N) {
        if (!rev) {
            v = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(xrow + g));
        } else {
            f32x4 t = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(xrow + (N - 4 - g)));
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
                                          int count, int64_t N, bool vec_ok, bool rev)
{
    for (int v = threadIdx.x; v < count / 4; v += kThreads)
        *reinterpret_cast<f32x4 *>(xs + 4 * v) =
            load4(xrow, g0 + 4 * (int64_t)v, N, vec_ok, rev);
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

template <int R>
__device__ __forceinline__ void store_outputs(const float (&acc)[R], float *__restrict__ yrow,
                                              int64_t i0, int64_t N, bool vec_y, bool rev)
{
    if (vec_y && i0 + R <= N) {
#pragma unroll
        for (int r = 0; r < R; r += 4) {
            if (!rev) {
                f32x4 v = {acc[r], acc[r + 1], acc[r + 2], acc[r + 3]};
                *reinterpret_cast<f32x4 *>(yrow + i0 + r) = v;
            } else {
                f32x4 v = {acc[r + 3], acc[r + 2], acc[r + 1], acc[r]};
                *reinterpret_cast<f32x4 *>(yrow + (N - 4 - i0 - r)) = v;
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
