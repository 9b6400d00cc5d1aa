// Direct-form FP32 batched FIR on gfx950, and the dense entry points.
//
// Replaces the reference's one-thread-per-output PTX kernel
// (crates/scir-gpu/src/lib.rs:727-811: two global loads per tap, integer
// div/rem per thread, no reuse) with a tiled design:
//
//   * one workgroup (256 lanes, 4 waves) = one tile of T = 256*R consecutive
//     outputs of one row;
//   * the tile's inputs plus the (K-1)-sample history are staged once into
//     LDS with 16-byte loads (zero-filled before the row start and past its
//     end: zero initial state, output length == input length);
//   * each lane owns R consecutive outputs and slides a register window over
//     the LDS copy, so each staged sample is read from LDS ~(R+32)/R times
//     and each tap feeds R multiply-adds from one scalar register;
//   * taps are read with uniform (scalar) loads and applied 32 at a time;
//     remainders use chunks of 16, 8, 4, 2, 1 taps, so no tap is padded;
//   * a non-finite output among the first K-1 of a row is recomputed from
//     the reference loop's terms alone (fix_ramp_in), so Inf and NaN in x or
//     b give non-finite outputs exactly where the reference loop does;
//   * long filters are processed in super-chunks of 2048 taps so the LDS
//     footprint stays bounded for any K.
//
// LDS layout: linear (see direct_tiles.hpp).
//
// Two arithmetic modes (template EXACT):
//   EXACT = false: fused multiply-add in any order -- error within
//                  (K+1)*2^-23*sum|b||x| of the exact sum;
//   EXACT = true : separately rounded multiply and add, taps accumulated in
//                  the order k = 0..K-1 from +0.0f, which is the reference
//                  CPU loop's order (crates/scir-gpu/src/lib.rs:1141-1148:
//                  taps[K-1-t]*x[i-t] for t ascending == b[t]*x[i-t]).  Extra
//                  terms past the row start add +-0 to the sum for finite
//                  taps, which leaves every bit unchanged (the sum never
//                  becomes -0.0); outputs they make non-finite are
//                  recomputed without them.  The output is bit-identical to
//                  the reference for any input, Inf and NaN included.
#include "direct_tiles.hpp"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace scir {
namespace {

using namespace direct;

// One workgroup = one tile; any K (taps in LDS super-chunks of kSuper).
template <int R, bool EXACT>
__global__ __launch_bounds__(kThreads) void fir_direct_kernel(
    const float *__restrict__ x, int64_t ldx, const float *__restrict__ b, int K, int64_t N,
    float *__restrict__ y, int64_t ldy, int64_t row0, int tiles_per_row, int rev)
{
    extern __shared__ __attribute__((aligned(16))) float xs[];
    constexpr int T = kThreads * R;
    const int64_t row = row0 + blockIdx.x / tiles_per_row;
    const int64_t t0 = (int64_t)(blockIdx.x % tiles_per_row) * T;
    const float *__restrict__ xrow = x + row * ldx;
    const int lane_base = threadIdx.x * R;

    float acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0f;

    for (int ks = 0; ks < K; ks += kSuper) {
        const int kcount = min(kSuper, K - ks);
        const int H = (kcount + 6) & ~3;  // >= kcount + 3, multiple of 4
        if (ks > 0) __syncthreads();      // previous super-chunk done with xs
        stage_row(xs, xrow, t0 - ks - H, T + H + 4, N, rev != 0);
        __syncthreads();
        taps_from_lds<R, EXACT>(acc, xs, H, lane_base, kcount, b + ks);
    }
    fix_ramp_in<R, EXACT>(acc, xrow, N, rev != 0, b, K, t0 + lane_base);
    store_outputs<R>(acc, y + row * ldy, t0 + lane_base, N, rev != 0);
}

// Pipelined variant for K <= kPipeMaxK: one workgroup walks `tiles_per_seg`
// consecutive tiles of a row.  The next tile's samples are fetched into NS
// float4 registers per lane while the current tile is computed from one of
// two LDS buffers, so HBM reads stay in flight across the compute phase.
// One barrier per tile: buffer (it & 1) is written at the top of iteration it
// and was last read in iteration it-2, which every lane finished before
// passing iteration it-1's barrier.
template <int R, bool EXACT, int NS>
__global__ __launch_bounds__(kThreads) void fir_direct_pipe_kernel(
    const float *__restrict__ x, int64_t ldx, const float *__restrict__ b, int K, int64_t N,
    float *__restrict__ y, int64_t ldy, int64_t row0, int segs_per_row, int tiles_per_seg,
    int64_t tiles_per_row, int buf_floats, int rev)
{
    extern __shared__ __attribute__((aligned(16))) float lds[];
    constexpr int T = kThreads * R;
    const int H = (K + 6) & ~3;
    const int count4 = (T + H + 4) / 4;  // <= NS * kThreads (checked by the launcher)
    const int64_t row = row0 + blockIdx.x / segs_per_row;
    const int64_t tfirst = (int64_t)(blockIdx.x % segs_per_row) * tiles_per_seg;
    const int ntiles = (int)min<int64_t>((int64_t)tiles_per_seg, tiles_per_row - tfirst);
    const float *__restrict__ xrow = x + row * ldx;
    float *__restrict__ yrow = y + row * ldy;
    const int lane_base = threadIdx.x * R;

    f32x4 pre[NS];
    auto fetch = [&](int64_t tile) {
        const int64_t g0 = tile * T - H;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int v = threadIdx.x + s * kThreads;
            if (v < count4) pre[s] = load4(xrow, g0 + 4 * (int64_t)v, N, rev != 0);
        }
    };

    fetch(tfirst);
    for (int it = 0; it < ntiles; ++it) {
        float *buf = lds + (it & 1) * buf_floats;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int v = threadIdx.x + s * kThreads;
            if (v < count4) *reinterpret_cast<f32x4 *>(buf + 4 * v) = pre[s];
        }
        __syncthreads();
        if (it + 1 < ntiles) fetch(tfirst + it + 1);
        float acc[R];
        taps_from_lds<R, EXACT, true>(acc, buf, H, lane_base, K, b);
        fix_ramp_in<R, EXACT>(acc, xrow, N, rev != 0, b, K, (tfirst + it) * T + lane_base);
        store_outputs<R>(acc, yrow, (tfirst + it) * T + lane_base, N, rev != 0);
    }
}

// LDS-DMA variant (forward direction, 16-byte aligned rows, N % 4 == 0):
// tiles are copied from HBM straight into LDS with buffer_load_dwordx4 ... lds,
// NB-1 tiles ahead (NB LDS buffers), so no registers or VALU instructions
// carry the staged samples and each workgroup keeps two tiles of loads in
// flight.  The row is addressed through a buffer resource of N*4 bytes:
// chunks before the row start (negative offsets wrap past the record count)
// and past its end read as zero, and output stores past the end are dropped,
// so every wave issues the same number of memory instructions per tile and
// its waits can be counted exactly.
//
// Wave w copies the LDS chunk groups q = w, w+4, ... (64 chunks of 16 bytes,
// one instruction each) with 64*q < count4; lane l of group q fills tile
// chunk 64q + l (the linear LDS layout).
__device__ __forceinline__ void dma_chunk16(u32x4 rsrc, unsigned voffset, unsigned lds_byte)
{
    // M0 (the LDS destination of the wave) is written by the compiler just
    // before this block; s_nop provides the wait state an LDS DMA needs after
    // an M0 write.
    asm volatile("s_nop 0\n\tbuffer_load_dwordx4 %0, %1, 0 offen lds"
                 :
                 : "v"(voffset), "s"(rsrc), "{m0}"(lds_byte)
                 : "memory");
}

// Buffer resource of a row of N floats (raw: stride 0, range N*4 bytes).
__device__ __forceinline__ u32x4 row_rsrc(const float *row, int64_t N)
{
    const uint64_t base = reinterpret_cast<uint64_t>(row);
    return u32x4{(unsigned)base, (unsigned)(base >> 32) & 0xffffu, (unsigned)(N * 4), 0x00020000u};
}

__device__ __forceinline__ void store_dwordx4(__amdgpu_buffer_rsrc_t rsrc, unsigned voffset,
                                              f32x4 v)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), rsrc, voffset, 0, 0);
}

// s_waitcnt vmcnt(n) for a wave-uniform n, leaving the other counters
// unconstrained; n > 15 waits for everything (never too early).
__device__ __forceinline__ void wait_vmcnt(int n)
{
#define SCIR_VMW(k) \
    case k: __builtin_amdgcn_s_waitcnt((k) | (7 << 4) | (0xf << 8)); break;
    switch (n) {
        SCIR_VMW(1) SCIR_VMW(2) SCIR_VMW(3) SCIR_VMW(4) SCIR_VMW(5) SCIR_VMW(6) SCIR_VMW(7)
        SCIR_VMW(8) SCIR_VMW(9) SCIR_VMW(10) SCIR_VMW(11) SCIR_VMW(12) SCIR_VMW(13)
        SCIR_VMW(14) SCIR_VMW(15)
        default: __builtin_amdgcn_s_waitcnt(0 | (7 << 4) | (0xf << 8)); break;
    }
#undef SCIR_VMW
}

template <int R, bool EXACT, int NB>
__global__ __launch_bounds__(kThreads) void fir_direct_dma_kernel(
    const float *__restrict__ x, int64_t ldx, const float *__restrict__ b, int K, int64_t N,
    float *__restrict__ y, int64_t ldy, int64_t row0, int segs_per_row, int tiles_per_seg,
    int64_t tiles_per_row, int buf_chunks)
{
    extern __shared__ __attribute__((aligned(16))) float lds[];
    constexpr int T = kThreads * R;
    constexpr int S = R / 4;  // output stores per lane and tile
    const int H = (K + 6) & ~3;
    const int count4 = (T + H + 4) / 4;  // tile chunks, <= buf_chunks (checked by the launcher)
    const int64_t row = row0 + blockIdx.x / segs_per_row;
    const int64_t tfirst = (int64_t)(blockIdx.x % segs_per_row) * tiles_per_seg;
    const int ntiles = (int)min<int64_t>((int64_t)tiles_per_seg, tiles_per_row - tfirst);
    const u32x4 xr = row_rsrc(x + row * ldx, N);
    const __amdgpu_buffer_rsrc_t yr =
        __builtin_amdgcn_make_buffer_rsrc(y + row * ldy, 0, (int)(N * 4), 0x00020000);
    // The wave index is uniform: keep it (and what derives from it) scalar.
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int lane_base = threadIdx.x * R;
    // DMA instructions of this wave per tile.
    const int nd = (count4 + 255 - 64 * wave) / 256;
    const unsigned lds0 =
        (unsigned)reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) float *)lds);

    auto issue = [&](int64_t tile, int slot) {
        const int64_t g0 = tile * T - H;  // sample of tile chunk 0
        for (int d = 0; d < nd; ++d) {
            const int q = wave + 4 * d;
            const int c = 64 * q + lane;
            // Chunks past the tile read as zero (offset beyond the record count).
            const unsigned voff = c < count4 ? (unsigned)(4 * (g0 + 4 * (int64_t)c)) : 0x80000000u;
            dma_chunk16(xr, voff, lds0 + 16u * (unsigned)(slot * buf_chunks + 64 * q));
        }
    };

    for (int t = 0; t < NB - 1 && t < ntiles; ++t) issue(tfirst + t, t);
    for (int it = 0; it < ntiles; ++it) {
        // VMEM instructions this wave issued after tile it's copies: the
        // copies of the tiles it+1 .. it+NB-2 and the stores of up to NB-1
        // earlier tiles.  Memory instructions of a wave complete in order.
        // High wave priority from the wait to the issue of the next copies,
        // so they leave while the other waves of the CU run FMAs (measured:
        // +0.5-0.7 % at K = 33).
        __builtin_amdgcn_s_setprio(3);
        wait_vmcnt(nd * max(0, min(NB - 2, ntiles - 1 - it)) + S * min(it, NB - 1));
        __syncthreads();
        if (it + NB - 1 < ntiles) issue(tfirst + it + NB - 1, (it + NB - 1) % NB);
        __builtin_amdgcn_s_setprio(0);
        const float *buf = lds + (it % NB) * buf_chunks * 4;
        float acc[R];
        taps_from_lds<R, EXACT, true>(acc, buf, H, lane_base, K, b);
        const int64_t i0 = (tfirst + it) * T + lane_base;
        fix_ramp_in<R, EXACT>(acc, x + row * ldx, N, false, b, K, i0);
#pragma unroll
        for (int r = 0; r < R; r += 4)
            store_dwordx4(yr, (unsigned)(4 * (i0 + r)),
                          f32x4{acc[r], acc[r + 1], acc[r + 2], acc[r + 3]});
    }
}

// ---- F32 mode on the matrix cores (direct.loader = 3, K <= kMfmaF32MaxK = 513) ----
//
// v_mfma_f32_16x16x4_f32 issues 1024 FP32 multiply-adds per 32 cycles,
// twice the rate of non-packed FP32 FMAs on the VALU.  A tile of 256
// consecutive outputs y[s + 16j + i] (i, j < 16) is the Toeplitz product
//   D[i][j] = sum_c b[i - c] x[s + 16j + c],  c in [-16(NG-1), 16),
// NG = ceil((K-1)/16) + 1, b[k] = 0 outside [0, K).  Step (J, v), J < NG,
// v < 4, takes c = -16J + 4g + v for the MFMA reduction index g < 4:
//   A[i][g] = b[i + 16J - 4g - v]          taps, in registers for the wave
//   B[g][j] = x[s + 16(j - J) + 4g + v]    samples
// Lane (g, j) (g = lane/16, j = lane%16) loads x[s + 16j + 4g .. +3] once
// per tile (the wave's 64 16-byte loads are the tile), which is B of step
// (0, v) in component v.  Step (J, v) takes component v of lane (g, j - J):
// a DPP row shift by J within the lane's row of 16, lanes j < J reading the
// previous tile's registers (DPP row rotate); J = 16 is the previous tile
// itself, J = 17 .. 31 the same window one tile further back, J = 32 the
// tile before the previous one.  Every (tap, output) pair of the filter occurs once; the rest of
// the 16 x 16NG block multiplies zero taps.  D comes back in lane (g, j) as
// y[s + 16j + 4g .. +3]: one 16-byte store.
//
// Each wave streams a strip of consecutive tiles of one row, loading three
// tiles ahead, and writes with non-temporal stores (measured faster than
// cached stores, and than non-temporal loads); no LDS, no barriers.  Per output: 16NG multiply-adds on the
// matrix cores (48 at K = 33, 144 at K = 129) against K FMAs on the VALU,
// which is left the DPP moves.  FP32 sums of the exact terms (and zeros)
// with FP32 rounding, within the F32 mode's (K+1)*2^-23*sum|b||x|.  A
// non-finite output (an Inf or NaN times a zero tap of the block, a
// non-finite tap times a zero before the row start) is recomputed from the
// reference loop's terms, so non-finite outputs sit where the reference
// loop's do.
constexpr int kMfmaF32MaxK = 513;  // NG <= 33: history from the two previous tiles
constexpr int kMfmaTile = 256;

// Lane j of a row of 16 reading lane j - J of `hi` (j >= J) or lane
// j - J + 16 of `lo` (j < J), 1 <= J <= 15.
template <int J>
__device__ __forceinline__ float row_window(float hi, float lo)
{
    // row_ror:J -- lane j reads lane (j - J) mod 16 of lo; then row_shr:J --
    // lanes j >= J read lane j - J of hi, the others keep the rotated lo (no
    // source lane: the write is skipped).
    const int p = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, lo), 0x120 + J, 0xF, 0xF, false);
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(p, __builtin_bit_cast(int, hi), 0x110 + J, 0xF, 0xF,
                                           false));
}

// Sample x[s + 16(j - J) + 4g + v] for lane (g, j), from component v of the
// current tile (cur) and the two before it (prev, prev2).
template <int J>
__device__ __forceinline__ float history_sample(float cur, float prev, float prev2)
{
    if constexpr (J == 0)
        return cur;
    else if constexpr (J < 16)
        return row_window<J>(cur, prev);
    else if constexpr (J == 16)
        return prev;
    else if constexpr (J < 32)
        return row_window<J - 16>(prev, prev2);
    else
        return prev2;
}

template <int J>
__device__ __forceinline__ void toeplitz_step(f32x4 &acc, const float (&a)[4], const f32x4 &cur,
                                              const f32x4 &prev, const f32x4 &prev2)
{
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], history_sample<J>(cur.x, prev.x, prev2.x), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], history_sample<J>(cur.y, prev.y, prev2.y), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], history_sample<J>(cur.z, prev.z, prev2.z), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], history_sample<J>(cur.w, prev.w, prev2.w), acc, 0, 0, 0);
}

template <int NG, int... Js>
__device__ __forceinline__ void toeplitz_tile(f32x4 &acc, const float (&a)[NG][4], const f32x4 &cur,
                                              const f32x4 &prev, const f32x4 &prev2,
                                              std::integer_sequence<int, Js...>)
{
    (toeplitz_step<Js>(acc, a[Js], cur, prev, prev2), ...);
}

// Every output of lane chunk `off` in tiles t0 .. t0+nt-1 of a row, from the
// reference loop's terms (fused multiply-add), overwriting what the tile
// loop stored: the slow path of a lane that met a non-finite output, taken
// after the strip so it adds nothing to the tile loop's registers.
__device__ __forceinline__ void recompute_strip(const float *__restrict__ xrow,
                                                float *__restrict__ yrow, int64_t N, bool rev,
                                                const float *__restrict__ b, int K, int64_t t0,
                                                int nt, int off)
{
    for (int m = 0; m < nt; ++m)
        for (int r = 0; r < 4; ++r) {
            const int64_t i = (t0 + m) * kMfmaTile + off + r;
            if (i >= N) break;
            float acc = 0.0f;
            const int64_t kend = i < (int64_t)K - 1 ? i : (int64_t)K - 1;
            for (int64_t k = 0; k <= kend; ++k)
                acc = __builtin_fmaf(b[k], xrow[rev ? N - 1 - (i - k) : i - k], acc);
            yrow[rev ? N - 1 - i : i] = acc;
        }
}

template <int NG, bool REV>
__global__ __launch_bounds__(kThreads) void fir_direct_mfma_kernel(
    const float *__restrict__ x, int64_t ldx, const float *__restrict__ b, int K, int64_t N,
    float *__restrict__ y, int64_t ldy, int64_t row0, int64_t strips, int64_t strips_per_row,
    int strip_tiles)
{
    // The wave index is uniform: keep it (and the strip bounds) scalar.
    const int64_t strip = (int64_t)blockIdx.x * (kThreads / 64) +
                          __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (strip >= strips) return;  // whole waves: nothing below synchronises
    const int64_t row = row0 + strip / strips_per_row;
    const int64_t t0 = (strip % strips_per_row) * (int64_t)strip_tiles;
    const int64_t tiles_in_row = (N + kMfmaTile - 1) / kMfmaTile;
    const int nt = (int)min<int64_t>((int64_t)strip_tiles, tiles_in_row - t0);
    const int lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
    const float *__restrict__ xrow = x + row * ldx;
    float *__restrict__ yrow = y + row * ldy;
    constexpr bool rv = REV;  // a compile-time direction keeps the loads' results unwaited

    // A of step (J, v) in lane (g = lq, i = li): b[i + 16J - 4g - v].
    float a[NG][4];
#pragma unroll
    for (int J = 0; J < NG; ++J)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int k = li + 16 * J - 4 * lq - v;
            a[J][v] = (k >= 0 && k < K) ? b[k] : 0.0f;
        }

    const int off = 16 * li + 4 * lq;  // the lane's chunk of a tile
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    // Ring of D + 1 + H tiles (H = 1 history tile up to NG = 17, else 2):
    // tile t0 + m sits in q[(m + H) % RING].  Step m reads tiles m - H .. m
    // and loads tile m + D into the slot of tile m - H - 1; the loop is
    // unrolled by RING so no register is copied (a copy would wait for the
    // load in flight).
    constexpr int D = 3;
    constexpr int H = NG > 17 ? 2 : 1;
    constexpr int RING = D + 1 + H;
    f32x4 q[RING];
    // The first tile's history: of tile t0 - h only lanes j >= 16h - (NG-1)
    // are read by the DPP rotates (J <= NG-1), so only they load it (at
    // K = 33 the last 32 samples, not 256: the strip's one extra read).
    // Zeros before the row start.
#pragma unroll
    for (int h = H; h >= 1; --h)
        q[H - h] = li >= 16 * h - (NG - 1) ? load4(xrow, (t0 - h) * kMfmaTile + off, N, rv) : zero;
#pragma unroll
    for (int d = 0; d < D; ++d)
        q[H + d] = d < nt ? load4(xrow, (t0 + d) * kMfmaTile + off, N, rv) : zero;
    q[RING - 1] = zero;
    // Strips wholly inside the row load and store with plain 16-byte
    // accesses: a bounds-checked load is a branch, and at its join the
    // compiler waits for every load in flight.
    const bool inside = (t0 + nt) * kMfmaTile <= N;
    bool nonfinite = false;  // a non-finite output in this lane's chunks
    auto step = [&](auto whole, int m, f32x4 &prev2, f32x4 &prev, f32x4 &cur, f32x4 &slot) {
        constexpr bool W = decltype(whole)::value;
        if (W) {
            // Unconditional (a branch around a load is a join that waits):
            // past the strip's end the last tile is loaded again, unused.
            const int64_t g = (t0 + min(m + D, nt - 1)) * kMfmaTile + off;
            const float *p = rv ? xrow + (N - 4 - g) : xrow + g;
            const f32x4 t = *reinterpret_cast<const f32x4u *>(p);
            slot = rv ? f32x4{t.w, t.z, t.y, t.x} : t;
        } else if (m + D < nt) {
            slot = load4(xrow, (t0 + m + D) * kMfmaTile + off, N, rv);
        }
        f32x4 acc = zero;
        toeplitz_tile<NG>(acc, a, cur, prev, prev2, std::make_integer_sequence<int, NG>{});
        float o[4] = {acc.x, acc.y, acc.z, acc.w};
        const int64_t i0 = (t0 + m) * kMfmaTile + off;
        // Past the row end o is 0 times zeros or a non-finite tap: ignored.
        const bool fin = __builtin_isfinite(o[0]) & __builtin_isfinite(o[1]) &
                         __builtin_isfinite(o[2]) & __builtin_isfinite(o[3]);
        nonfinite |= !fin & (i0 < N);
        if (W) {
            f32x4u *p = reinterpret_cast<f32x4u *>(rv ? yrow + (N - 4 - i0) : yrow + i0);
            const f32x4 v = rv ? f32x4{o[3], o[2], o[1], o[0]} : f32x4{o[0], o[1], o[2], o[3]};
            __builtin_nontemporal_store(v, p);
        } else {
            store_outputs<4>(o, yrow, i0, N, rv);
        }
    };
    auto run = [&](auto whole) {
        for (int m = 0; m < nt; m += RING) {
#pragma unroll
            for (int u = 0; u < RING; ++u) {
                if (m + u >= nt) break;
                step(whole, m + u, q[(u + H + RING - 2) % RING], q[(u + H - 1) % RING],
                     q[(u + H) % RING], q[(u + H + D) % RING]);
            }
        }
    };
    if (inside)
        run(std::true_type{});
    else
        run(std::false_type{});
    if (nonfinite) recompute_strip(xrow, yrow, N, rv, b, K, t0, nt, off);
}

template <int NG, bool REV>
int launch_direct_mfma(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                       float *y, int64_t ldy, int strip_tiles, hipStream_t stream)
{
    constexpr int64_t kWavesPerWg = kThreads / 64;
    const int64_t tiles = (N + kMfmaTile - 1) / kMfmaTile;
    const int64_t st = std::max<int64_t>(1, std::min<int64_t>(strip_tiles, tiles));
    const int64_t spr = (tiles + st - 1) / st;
    // One launch covers as many rows as fit in a 2^31-1 workgroup grid.
    const int64_t rows_per_launch = std::max<int64_t>(1, 0x7fffffffLL * kWavesPerWg / spr);
    if (spr > 0x7fffffffLL * kWavesPerWg)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row too long (N=%lld)", (long long)N);
    for (int64_t r0 = 0; r0 < B; r0 += rows_per_launch) {
        const int64_t rows = std::min(rows_per_launch, B - r0);
        const int64_t strips = rows * spr;
        hipLaunchKernelGGL((fir_direct_mfma_kernel<NG, REV>),
                           dim3((unsigned)((strips + kWavesPerWg - 1) / kWavesPerWg)),
                           dim3(kThreads), 0, stream, x, ldx, b, K, N, y, ldy, r0, strips, spr,
                           (int)st);
        SCIR_TRY(check_launch("fir_direct_mfma_kernel"));
    }
    return SCIR_HIP_OK;
}

int launch_direct_mfma_k(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                         float *y, int64_t ldy, int strip_tiles, bool rev, hipStream_t stream)
{
    const int ng = (K + 14) / 16 + 1;  // ceil((K-1)/16) + 1
    switch (ng) {
#define SCIR_NG(G) \
    case G:                                                                               \
        return rev ? launch_direct_mfma<G, true>(x, ldx, b, K, B, N, y, ldy, strip_tiles, stream) \
                   : launch_direct_mfma<G, false>(x, ldx, b, K, B, N, y, ldy, strip_tiles, stream);
        SCIR_NG(1) SCIR_NG(2) SCIR_NG(3) SCIR_NG(4) SCIR_NG(5) SCIR_NG(6) SCIR_NG(7) SCIR_NG(8)
        SCIR_NG(9) SCIR_NG(10) SCIR_NG(11) SCIR_NG(12) SCIR_NG(13) SCIR_NG(14) SCIR_NG(15)
        SCIR_NG(16) SCIR_NG(17) SCIR_NG(18) SCIR_NG(19) SCIR_NG(20) SCIR_NG(21) SCIR_NG(22)
        SCIR_NG(23) SCIR_NG(24) SCIR_NG(25) SCIR_NG(26) SCIR_NG(27) SCIR_NG(28) SCIR_NG(29)
        SCIR_NG(30) SCIR_NG(31) SCIR_NG(32) SCIR_NG(33)
#undef SCIR_NG
    default:
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "FP32 MFMA direct kernel needs K <= %d",
                    kMfmaF32MaxK);
    }
}

constexpr int kPipeMaxK = 1017;  // (K + 6) & ~3 <= 1020, so T + H + 4 <= T + 1024

int align_flags(const float *p, int64_t ld) {
    return ((reinterpret_cast<uintptr_t>(p) & 15) == 0) && (ld % 4 == 0);
}

template <int R, bool EXACT>
int launch_direct(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                  float *y, int64_t ldy, bool rev, hipStream_t stream)
{
    constexpr int T = kThreads * R;
    const int kmax = std::min(K, kSuper);
    const int H = (kmax + 6) & ~3;
    const int staged = T + H + 4;
    const size_t lds = (size_t)(staged + (staged / 16) * 4 + 4) * sizeof(float);
    const int64_t tiles = (N + T - 1) / T;
    if (tiles > 0x7fffffff)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row too long (N=%lld)", (long long)N);
    // One launch covers as many rows as fit in a 2^31-1 workgroup grid.
    const int64_t rows_per_launch = std::max<int64_t>(1, 0x7fffffff / tiles);
    for (int64_t r0 = 0; r0 < B; r0 += rows_per_launch) {
        const int64_t rows = std::min(rows_per_launch, B - r0);
        hipLaunchKernelGGL((fir_direct_kernel<R, EXACT>), dim3((unsigned)(rows * tiles)),
                           dim3(kThreads), lds, stream, x, ldx, b, K, N, y, ldy, r0, (int)tiles,
                           (int)rev);
        SCIR_TRY(check_launch("fir_direct_kernel"));
    }
    return SCIR_HIP_OK;
}

template <int R, bool EXACT>
int launch_direct_pipe(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                       float *y, int64_t ldy, int seg_tiles, bool rev, hipStream_t stream)
{
    constexpr int T = kThreads * R;
    constexpr int NS = (T + 1024 + kThreads * 4 - 1) / (kThreads * 4);
    const int H = (K + 6) & ~3;
    const int staged = T + H + 4;
    if (K > kPipeMaxK || (staged / 4) > NS * kThreads)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "pipelined direct kernel needs K <= %d",
                    kPipeMaxK);
    const int buf_floats = ((staged + (staged / 16) * 4 + 4) + 3) & ~3;
    const size_t lds = 2 * (size_t)buf_floats * sizeof(float);
    const int64_t tiles = (N + T - 1) / T;
    const int64_t tps = std::max<int64_t>(1, std::min<int64_t>(seg_tiles, tiles));
    const int64_t segs = (tiles + tps - 1) / tps;
    if (segs > 0x7fffffff)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row too long (N=%lld)", (long long)N);
    const int64_t rows_per_launch = std::max<int64_t>(1, 0x7fffffff / segs);
    for (int64_t r0 = 0; r0 < B; r0 += rows_per_launch) {
        const int64_t rows = std::min(rows_per_launch, B - r0);
        hipLaunchKernelGGL((fir_direct_pipe_kernel<R, EXACT, NS>), dim3((unsigned)(rows * segs)),
                           dim3(kThreads), lds, stream, x, ldx, b, K, N, y, ldy, r0, (int)segs,
                           (int)tps, tiles, buf_floats, (int)rev);
        SCIR_TRY(check_launch("fir_direct_pipe_kernel"));
    }
    return SCIR_HIP_OK;
}

template <int R, bool EXACT, int NB>
int launch_direct_dma(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                      float *y, int64_t ldy, int seg_tiles, hipStream_t stream)
{
    constexpr int T = kThreads * R;
    const int H = (K + 6) & ~3;
    const int count4 = (T + H + 4) / 4;
    const int buf_chunks = (count4 + 63) / 64 * 64;
    if (K > kPipeMaxK || N % 4 != 0 || N >= ((int64_t)1 << 29))
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "LDS-DMA direct kernel: unsupported shape");
    const size_t lds = NB * (size_t)buf_chunks * 16;
    const int64_t tiles = (N + T - 1) / T;
    const int64_t tps = std::max<int64_t>(1, std::min<int64_t>(seg_tiles, tiles));
    const int64_t segs = (tiles + tps - 1) / tps;
    const int64_t rows_per_launch = std::max<int64_t>(1, 0x7fffffff / segs);
    for (int64_t r0 = 0; r0 < B; r0 += rows_per_launch) {
        const int64_t rows = std::min(rows_per_launch, B - r0);
        hipLaunchKernelGGL((fir_direct_dma_kernel<R, EXACT, NB>), dim3((unsigned)(rows * segs)),
                           dim3(kThreads), lds, stream, x, ldx, b, K, N, y, ldy, r0, (int)segs,
                           (int)tps, tiles, buf_chunks);
        SCIR_TRY(check_launch("fir_direct_dma_kernel"));
    }
    return SCIR_HIP_OK;
}

}  // namespace

int fir_direct_dev(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                   float *y, int64_t ldy, bool exact, bool rev, hipStream_t stream, DeviceState *)
{
    if (B == 0 || N == 0) return SCIR_HIP_OK;
    if (K == 0) return fill_zero_2d(y, ldy, B, N, stream);
    const TuneChoice tc = tune_lookup(K, N);
    const bool r8 = (tc.direct_r == 8) || N <= 2048;
    // F32 mode on the matrix cores (strips of direct_mfma_strip 256-output
    // tiles); F32_EXACT and longer filters take the LDS-DMA loader's rules.
    if (tc.direct_loader == 3 && !exact && K <= kMfmaF32MaxK) {
        // Default strips, measured on [4096, 65536]: 4 tiles at K = 1 (319 us
        // against 352 with 8), 8 at K = 17-33, 16 at K = 48, 32 from K = 64;
        // on [2048, 262144] 64 above K = 257 (K = 512: 5559 us, 5640 with 32).
        const int strip = tc.direct_mfma_strip > 0 ? tc.direct_mfma_strip
                          : K <= 8                 ? 4
                          : K <= 40                ? 8
                          : K <= 56                ? 16
                          : K <= 257               ? 32
                                                   : 64;
        return launch_direct_mfma_k(x, ldx, b, K, B, N, y, ldy, strip, rev, stream);
    }
    // LDS-DMA loader: forward rows, 16-byte aligned, N % 4 == 0, N < 2^29.
    if (tc.direct_loader >= 2 && K <= kPipeMaxK && !rev && N % 4 == 0 &&
        N < ((int64_t)1 << 29) && align_flags(x, ldx) && align_flags(y, ldy)) {
        const int st = tc.direct_seg_tiles;
        const int nb = tc.direct_dma_bufs;
#define SCIR_DMA(RR, EX, NBV) launch_direct_dma<RR, EX, NBV>(x, ldx, b, K, B, N, y, ldy, st, stream)
        if (exact) {
            if (!r8) return nb == 2 ? SCIR_DMA(16, true, 2) : SCIR_DMA(16, true, 3);
            return nb == 2 ? SCIR_DMA(8, true, 2) : nb == 4 ? SCIR_DMA(8, true, 4) : SCIR_DMA(8, true, 3);
        }
        if (!r8) return nb == 2 ? SCIR_DMA(16, false, 2) : SCIR_DMA(16, false, 3);
        return nb == 2 ? SCIR_DMA(8, false, 2) : nb == 4 ? SCIR_DMA(8, false, 4) : SCIR_DMA(8, false, 3);
#undef SCIR_DMA
    }
    if (tc.direct_loader >= 1 && K <= kPipeMaxK) {
        const int st = tc.direct_seg_tiles;
        if (exact)
            return r8 ? launch_direct_pipe<8, true>(x, ldx, b, K, B, N, y, ldy, st, rev, stream)
                      : launch_direct_pipe<16, true>(x, ldx, b, K, B, N, y, ldy, st, rev, stream);
        return r8 ? launch_direct_pipe<8, false>(x, ldx, b, K, B, N, y, ldy, st, rev, stream)
                  : launch_direct_pipe<16, false>(x, ldx, b, K, B, N, y, ldy, st, rev, stream);
    }
    if (exact)
        return r8 ? launch_direct<8, true>(x, ldx, b, K, B, N, y, ldy, rev, stream)
                  : launch_direct<16, true>(x, ldx, b, K, B, N, y, ldy, rev, stream);
    return r8 ? launch_direct<8, false>(x, ldx, b, K, B, N, y, ldy, rev, stream)
              : launch_direct<16, false>(x, ldx, b, K, B, N, y, ldy, rev, stream);
}

// Dense dispatch shared by the host tier, the device tier and the dL/dx
// adjoint (rev = true: the same filter run back to front).
int fir_dense_dispatch(const float *x, int64_t ldx, const float *b, int32_t K, int64_t B,
                       int64_t N, float *y, int64_t ldy, int32_t precision, bool rev,
                       hipStream_t stream, DeviceState *st)
{
    switch (precision) {
    case SCIR_HIP_PREC_F32:
        return fir_direct_dev(x, ldx, b, K, B, N, y, ldy, false, rev, stream, st);
    case SCIR_HIP_PREC_F32_EXACT:
        return fir_direct_dev(x, ldx, b, K, B, N, y, ldy, true, rev, stream, st);
    case SCIR_HIP_PREC_BF16X3:
    case SCIR_HIP_PREC_FP16:
        return fir_mfma_dev(x, ldx, b, K, B, N, y, ldy, precision, rev, stream, st);
    case SCIR_HIP_PREC_FFT32:
        return fir_fft_dev(x, ldx, b, K, B, N, y, ldy, rev, stream, st);
    case SCIR_HIP_PREC_AUTO: {
        const TuneChoice tc = tune_lookup(K, N);
        if (tc.auto_fft && K >= tc.auto_fft_min_k && x != y)
            return fir_fft_dev(x, ldx, b, K, B, N, y, ldy, rev, stream, st);
        // BF16x3 MFMA from the cut-over up, while its LDS holds the taps.
        if (K >= tc.auto_mfma_min_k && fir_mfma_takes(SCIR_HIP_PREC_BF16X3, K))
            return fir_mfma_dev(x, ldx, b, K, B, N, y, ldy, SCIR_HIP_PREC_BF16X3, rev, stream, st);
        return fir_direct_dev(x, ldx, b, K, B, N, y, ldy, false, rev, stream, st);
    }
    default:
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "unknown precision %d", precision);
    }
}

namespace {

// Shared state of the three host threads of one pipelined host-tier call.
// Reporting a failure allocates nothing (the message goes into a fixed
// buffer), so it can run in any thread and while an exception unwinds.
struct HostPipe {
    std::mutex m;
    std::condition_variable cv;
    int64_t uploaded = 0;  // chunks whose upload event is recorded
    int64_t computed = 0;  // chunks whose kernel event is recorded
    std::atomic<bool> failed{false};
    int status = SCIR_HIP_OK;
    char msg[256] = {0};

    void fail_with(int st, const char *what, hipError_t e = hipSuccess) noexcept
    {
        std::lock_guard<std::mutex> g(m);
        if (!failed) {
            failed = true;
            status = st;
            if (e != hipSuccess)
                std::snprintf(msg, sizeof msg, "%s failed: %s", what, hipGetErrorString(e));
            else
                std::snprintf(msg, sizeof msg, "%s", what);
        }
        cv.notify_all();
    }
    // Waits until `counter` exceeds c; false when the call failed meanwhile.
    bool wait_for(const int64_t &counter, int64_t c)
    {
        std::unique_lock<std::mutex> g(m);
        cv.wait(g, [&] { return counter > c || failed; });
        return !failed;
    }
    void advance(int64_t &counter, int64_t v)
    {
        std::lock_guard<std::mutex> g(m);
        counter = v;
        cv.notify_all();
    }
};

// A team of threads for the staging copies of one direction: run(job) calls
// job(member, size) on every member, the calling thread being member 0, and
// returns when all have finished.  The job is passed by address, so a run
// allocates nothing and cannot throw.
class CopyTeam {
public:
    explicit CopyTeam(int size) : size_(size)
    {
        try {
            for (int i = 1; i < size; ++i) helpers_.emplace_back([this, i] { serve(i); });
        } catch (...) {
            stop();
            throw;
        }
    }
    ~CopyTeam() { stop(); }
    CopyTeam(const CopyTeam &) = delete;
    CopyTeam &operator=(const CopyTeam &) = delete;

    template <class F> void run(F &job)
    {
        {
            std::lock_guard<std::mutex> g(m_);
            fn_ = [](void *ctx, int member, int size) { (*static_cast<F *>(ctx))(member, size); };
            ctx_ = &job;
            pending_ = size_ - 1;
            ++generation_;
        }
        go_.notify_all();
        job(0, size_);
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [&] { return pending_ == 0; });
    }

private:
    void serve(int member)
    {
        uint64_t seen = 0;
        for (;;) {
            void (*fn)(void *, int, int);
            void *ctx;
            {
                std::unique_lock<std::mutex> g(m_);
                go_.wait(g, [&] { return stopping_ || generation_ != seen; });
                if (stopping_) return;
                seen = generation_;
                fn = fn_;
                ctx = ctx_;
            }
            fn(ctx, member, size_);
            std::lock_guard<std::mutex> g(m_);
            if (--pending_ == 0) done_.notify_all();
        }
    }
    void stop()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stopping_ = true;
        }
        go_.notify_all();
        for (std::thread &t : helpers_) t.join();
        helpers_.clear();
    }

    const int size_;
    std::vector<std::thread> helpers_;
    std::mutex m_;
    std::condition_variable go_, done_;
    void (*fn_)(void *, int, int) = nullptr;
    void *ctx_ = nullptr;
    uint64_t generation_ = 0;
    int pending_ = 0;
    bool stopping_ = false;
};

// Member m of n copies its share of a rows x N block of floats from pitch
// lsrc to pitch ldst: the block's samples in row-major order, cut into n
// ranges that start on 64-byte boundaries of that order.
void copy_share(float *dst, int64_t ldst, const float *src, int64_t lsrc, int64_t N,
                int64_t rows, int m, int n)
{
    const int64_t total = rows * N;
    int64_t lo = (total / n * m + total % n * m / n) & ~int64_t(15);
    const int64_t hi =
        m + 1 == n ? total : (total / n * (m + 1) + total % n * (m + 1) / n) & ~int64_t(15);
    while (lo < hi) {
        const int64_t r = lo / N, c = lo - r * N, len = std::min(hi - lo, N - c);
        std::memcpy(dst + r * ldst + c, src + r * lsrc + c, (size_t)len * sizeof(float));
        lo += len;
    }
}

// Four pinned host blocks of at least `bytes` each (two staging slots per
// direction), kept in the device state until shutdown.
int ensure_pinned(DeviceState *st, size_t bytes)
{
    if (st->pinned_bytes >= bytes) return SCIR_HIP_OK;
    for (void *&p : st->pinned) {
        if (p) SCIR_HIP_TRY(hipHostFree(p));
        p = nullptr;
    }
    st->pinned_bytes = 0;
    if (alloc_failure_injected())
        return fail(SCIR_HIP_ERR_OUT_OF_MEMORY, "host-tier staging: pinned allocation of %zu "
                    "bytes failed (SCIR_HIP_FAIL_ALLOC is set)", bytes);
    for (void *&p : st->pinned) {
        const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            for (void *&q : st->pinned) {
                if (q) (void)hipHostFree(q);
                q = nullptr;
            }
            return fail(e == hipErrorOutOfMemory ? SCIR_HIP_ERR_OUT_OF_MEMORY : SCIR_HIP_ERR_BACKEND,
                        "host-tier staging: hipHostMalloc(%zu) failed: %s", bytes,
                        hipGetErrorString(e));
        }
    }
    st->pinned_bytes = bytes;
    return SCIR_HIP_OK;
}

}  // namespace

// The host tier in row chunks staged through pinned memory, so that both
// PCIe directions run at the DMA engines' rate and overlap each other and the
// kernels.  Chunk c goes through five stages:
//   1. the input thread's team copies its rows from x into pinned slot c % 2
//      (after the upload of chunk c - 2 out of that slot has finished);
//   2. one copy moves the slot to the device (upload stream, event up[c]);
//   3. the calling thread enqueues the chunk's kernels behind up[c] (kernel
//      stream, event k[c]);
//   4. the output thread copies the chunk's result into pinned slot c % 2 of
//      the output side (download stream, behind k[c], event down[c]);
//   5. after down[c], the output thread's team copies the rows into y.
// The output thread enqueues the download of chunk c + 1 before it waits for
// chunk c, so the DMA engine and the staging copies overlap in both
// directions.  Device buffers are the full-size host-tier buffers.
int host_fir_pipelined(const float *x, int64_t ldx, const float *db, int32_t K, int64_t B,
                       int64_t N, float *y, int64_t ldy, int32_t precision, float *dx, float *dy,
                       int64_t ldd, int64_t rows_per_chunk, DeviceState *st)
{
    for (hipStream_t &cs : st->host_copy_stream)
        if (!cs) SCIR_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipStream_t sk = st->host_stream, sup = st->host_copy_stream[0],
                sdown = st->host_copy_stream[1];
    const int64_t nch = (B + rows_per_chunk - 1) / rows_per_chunk;
    const size_t slot_bytes = (size_t)rows_per_chunk * (size_t)ldd * sizeof(float);
    SCIR_TRY(ensure_pinned(st, slot_bytes));
    float *const pin_in[2] = {static_cast<float *>(st->pinned[0]),
                              static_cast<float *>(st->pinned[1])};
    float *const pin_out[2] = {static_cast<float *>(st->pinned[2]),
                               static_cast<float *>(st->pinned[3])};
    host_alloc_point();
    // [0, nch): uploads; [nch, 2 nch): kernels; [2 nch, 3 nch): downloads.
    std::vector<hipEvent_t> ev(3 * (size_t)nch, nullptr);
    struct EventsGuard {
        std::vector<hipEvent_t> &v;
        ~EventsGuard()
        {
            for (hipEvent_t e : v)
                if (e) (void)hipEventDestroy(e);
        }
    } guard{ev};
    for (hipEvent_t &e : ev) SCIR_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const int threads = host_copy_threads();
    CopyTeam team_in(threads), team_out(threads);
    HostPipe pipe;
    const int dev = st->device;
    auto rows_of = [&](int64_t c) { return std::min(rows_per_chunk, B - c * rows_per_chunk); };
    auto in = [&]() noexcept {
        (void)hipSetDevice(dev);  // the current device is per thread
        for (int64_t c = 0; c < nch; ++c) {
            if (pipe.failed) return;
            hipError_t e = c >= 2 ? hipEventSynchronize(ev[c - 2]) : hipSuccess;
            if (e != hipSuccess) return pipe.fail_with(SCIR_HIP_ERR_BACKEND, "upload", e);
            const int64_t r0 = c * rows_per_chunk, rows = rows_of(c);
            float *slot = pin_in[c & 1];
            const float *src = x + r0 * ldx;
            auto stage = [&](int m, int n) { copy_share(slot, ldd, src, ldx, N, rows, m, n); };
            team_in.run(stage);
            e = hipMemcpyAsync(dx + r0 * ldd, slot, (size_t)rows * ldd * sizeof(float),
                               hipMemcpyHostToDevice, sup);
            if (e == hipSuccess) e = hipEventRecord(ev[c], sup);
            if (e != hipSuccess) return pipe.fail_with(SCIR_HIP_ERR_BACKEND, "upload", e);
            pipe.advance(pipe.uploaded, c + 1);
        }
    };
    auto out = [&]() noexcept {
        (void)hipSetDevice(dev);
        auto enqueue = [&](int64_t c) {
            if (!pipe.wait_for(pipe.computed, c)) return false;
            const int64_t r0 = c * rows_per_chunk;
            hipError_t e = hipStreamWaitEvent(sdown, ev[nch + c], 0);
            if (e == hipSuccess)
                e = hipMemcpyAsync(pin_out[c & 1], dy + r0 * ldd,
                                   (size_t)rows_of(c) * ldd * sizeof(float),
                                   hipMemcpyDeviceToHost, sdown);
            if (e == hipSuccess) e = hipEventRecord(ev[2 * nch + c], sdown);
            if (e != hipSuccess) {
                pipe.fail_with(SCIR_HIP_ERR_BACKEND, "download", e);
                return false;
            }
            return true;
        };
        if (!enqueue(0)) return;
        for (int64_t c = 0; c < nch; ++c) {
            // Slot (c + 1) % 2 last held chunk c - 1, already copied out.
            if (c + 1 < nch && !enqueue(c + 1)) return;
            const hipError_t e = hipEventSynchronize(ev[2 * nch + c]);
            if (e != hipSuccess) return pipe.fail_with(SCIR_HIP_ERR_BACKEND, "download", e);
            if (pipe.failed) return;
            const int64_t r0 = c * rows_per_chunk, rows = rows_of(c);
            const float *slot = pin_out[c & 1];
            float *dst = y + r0 * ldy;
            auto unstage = [&](int m, int n) { copy_share(dst, ldy, slot, ldd, N, rows, m, n); };
            team_out.run(unstage);
        }
    };
    std::thread tin, tout;
    // Whatever leaves this function (an error, or an exception from the
    // kernel dispatch below, such as a failed host allocation), both
    // threads are stopped and joined first, and the streams drained.
    struct JoinOnExit {
        HostPipe &pipe;
        std::thread &a, &b;
        hipStream_t s0, s1, s2;
        ~JoinOnExit()
        {
            if (!a.joinable() && !b.joinable()) return;
            pipe.fail_with(SCIR_HIP_ERR_BACKEND, "host-tier pipeline stopped");
            if (a.joinable()) a.join();
            if (b.joinable()) b.join();
            (void)hipStreamSynchronize(s0);
            (void)hipStreamSynchronize(s1);
            (void)hipStreamSynchronize(s2);
        }
    } joiner{pipe, tin, tout, sup, sk, sdown};
    tin = std::thread(in);
    tout = std::thread(out);
    // Kernels of chunk c once its upload is enqueued (the stream waits for it).
    for (int64_t c = 0; c < nch; ++c) {
        if (!pipe.wait_for(pipe.uploaded, c)) break;
        const int64_t r0 = c * rows_per_chunk;
        hipError_t e = hipStreamWaitEvent(sk, ev[c], 0);
        if (e != hipSuccess) {
            pipe.fail_with(SCIR_HIP_ERR_BACKEND, "hipStreamWaitEvent", e);
            break;
        }
        const int st2 = fir_dense_dispatch(dx + r0 * ldd, ldd, db, K, rows_of(c), N,
                                           dy + r0 * ldd, ldd, precision, false, sk, st);
        if (st2 != SCIR_HIP_OK) {
            pipe.fail_with(st2, scir_hip_last_error());
            break;
        }
        e = hipEventRecord(ev[nch + c], sk);
        if (e != hipSuccess) {
            pipe.fail_with(SCIR_HIP_ERR_BACKEND, "hipEventRecord", e);
            break;
        }
        pipe.advance(pipe.computed, c + 1);
    }
    tin.join();
    tout.join();
    // Leave no copy or kernel of this call running on any of the streams (the
    // pinned slots are reused by the next call).
    const hipError_t e1 = hipStreamSynchronize(sup), e2 = hipStreamSynchronize(sk),
                     e3 = hipStreamSynchronize(sdown);
    if (pipe.failed) return fail(pipe.status, "%s", pipe.msg);
    if (e1 != hipSuccess) return fail(SCIR_HIP_ERR_BACKEND, "upload failed: %s", hipGetErrorString(e1));
    if (e2 != hipSuccess) return fail(SCIR_HIP_ERR_BACKEND, "kernels failed: %s", hipGetErrorString(e2));
    if (e3 != hipSuccess) return fail(SCIR_HIP_ERR_BACKEND, "download failed: %s", hipGetErrorString(e3));
    return SCIR_HIP_OK;
}

}  // namespace scir

using namespace scir;

extern "C" {

int scir_hip_fir_f32_dev(const float *x, int64_t ldx, const float *b, int32_t K, int64_t B,
                         int64_t N, float *y, int64_t ldy, int32_t precision, void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        SCIR_TRY(check_dense_args(x, ldx, b, K, B, N, y, ldy));
        if (B == 0 || N == 0) return SCIR_HIP_OK;
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        std::lock_guard<std::mutex> g(st->mu);
        return fir_dense_dispatch(x, ldx, b, K, B, N, y, ldy, precision, false, as_stream(stream),
                                  st);
    });
}

int scir_hip_fir_f32_host(const float *x, int64_t ldx, const float *b, int32_t K, int64_t B,
                          int64_t N, float *y, int64_t ldy, int32_t precision)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        SCIR_TRY(check_dense_args(x, ldx, b, K, B, N, y, ldy));
        if (B == 0 || N == 0) return SCIR_HIP_OK;
        if (precision < SCIR_HIP_PREC_F32 || precision > SCIR_HIP_PREC_FFT32)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "unknown precision %d", precision);
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        std::lock_guard<std::mutex> g(st->mu);
        if (!st->host_stream) SCIR_HIP_TRY(hipStreamCreateWithFlags(&st->host_stream, hipStreamNonBlocking));
        hipStream_t s = st->host_stream;
        const size_t row_bytes = (size_t)N * sizeof(float);
        // Device rows are padded to a multiple of 4 floats so the kernels can use
        // 16-byte accesses.
        const int64_t ldd = (N + 3) & ~int64_t(3);
        const size_t mat_bytes = (size_t)B * (size_t)ldd * sizeof(float);
        SCIR_TRY(ensure_buffer(st->host_x, mat_bytes, s));
        SCIR_TRY(ensure_buffer(st->host_y, mat_bytes, s));
        SCIR_TRY(ensure_buffer(st->host_b, (size_t)std::max(K, 1) * sizeof(float), s));
        float *dx = static_cast<float *>(st->host_x.ptr);
        float *dy = static_cast<float *>(st->host_y.ptr);
        float *db = static_cast<float *>(st->host_b.ptr);
        // Large batches: chunks of rows through the three-stream pipeline.
        const int64_t rows_per_chunk =
            std::max<int64_t>(1, (int64_t)(host_chunk_bytes() / ((size_t)ldd * sizeof(float))));
        if (B >= 3 * rows_per_chunk) {
            if (K > 0)
                SCIR_HIP_TRY(hipMemcpy(db, b, (size_t)K * sizeof(float), hipMemcpyHostToDevice));
            return host_fir_pipelined(x, ldx, db, K, B, N, y, ldy, precision, dx, dy, ldd,
                                      rows_per_chunk, st);
        }
        // One linear copy when host and device rows are both packed (a pitched
        // copy of many rows runs as many small transfers).
        if (ldx == N && ldd == N)
            SCIR_HIP_TRY(hipMemcpyAsync(dx, x, mat_bytes, hipMemcpyHostToDevice, s));
        else
            SCIR_HIP_TRY(hipMemcpy2DAsync(dx, ldd * sizeof(float), x, ldx * sizeof(float), row_bytes,
                                          B, hipMemcpyHostToDevice, s));
        if (K > 0)
            SCIR_HIP_TRY(hipMemcpyAsync(db, b, (size_t)K * sizeof(float), hipMemcpyHostToDevice, s));
        SCIR_TRY(fir_dense_dispatch(dx, ldd, db, K, B, N, dy, ldd, precision, false, s, st));
        if (ldy == N && ldd == N)
            SCIR_HIP_TRY(hipMemcpyAsync(y, dy, mat_bytes, hipMemcpyDeviceToHost, s));
        else
            SCIR_HIP_TRY(hipMemcpy2DAsync(y, ldy * sizeof(float), dy, ldd * sizeof(float), row_bytes,
                                          B, hipMemcpyDeviceToHost, s));
        SCIR_HIP_TRY(hipStreamSynchronize(s));
        return SCIR_HIP_OK;
    });
}

}  // extern "C"
