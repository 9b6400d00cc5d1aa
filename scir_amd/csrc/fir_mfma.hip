// Long-tap batched FIR as a Toeplitz-block contraction on gfx950 MFMA
// (v_mfma_f32_32x32x16_{bf16,f16}), FP32 accumulation.
//
// Math (lfilter order, y[n] = sum_k b[k] x[n-k]).  Cut each row's output into
// blocks of 32 samples; block p is y[32p .. 32p+31].  With half-steps
// m = -1 .. m_hi (m_hi = ceil((K-1)/16)):
//   A_m[r][c] = b[16m + r - c]            (32 x 16, zero outside [0, K))
//   B_m[c][p] = x[32p - 16m + c]          (16 x 32 block columns)
//   y[32p + r] = sum_m (A_m B_m)[r][p]
// Each (k, output) pair appears exactly once (m = ceil((k-r)/16), c = 16m-k+r).
// A_m depends only on the taps.  Each workgroup keeps the split taps in LDS
// as a compact reversed table, T[j] = b[C - j]: lane (i, g) of A_m holds the
// 8 consecutive entries starting at j0 = C - 16m - i + 8g, i.e. one 16-byte
// read per split and half-step.  The table is stored twice, the second copy
// shifted by one entry, so every lane's read is 4-byte aligned whatever the
// parity of j0; the two copies are placed 128 bytes apart modulo the 256-byte
// bank cycle so the lanes reading either copy do not collide.
// B_m is a shifted view of the same staged x: no im2col copy is made; each
// sample is split once into LDS and read by the ~2*(K+31)/32 half-steps that
// need it.
//
// Precision modes (template S = number of split terms):
//   S = 3 (BF16x3): v = hi + mid + lo, each bf16 (hi = rn(v), mid = rn(v-hi),
//     lo = rn(v-hi-mid); both subtractions are exact), for x and for b.  Six
//     products are kept: hh (own accumulator), hm, mh, mm, hl, lh (a second
//     accumulator); the dropped ml, lm, ll are below 2^-24 relative.
//   S = 1 (FP16): x and b rounded to fp16, one product.
//
// Work decomposition: a workgroup (4 waves) owns a row segment of
// 4*TW*32 blocks; each wave owns TW 32x32 output tiles (TW*1024 outputs).
// The segment's inputs plus 16*m_hi samples of history are split into LDS
// once.  LDS layout per split: unit u = 8 consecutive samples (16 B) at byte
// 16*(u + u/4); a wave's B read has a 4-unit lane stride, which this padding
// makes bank-conflict free.
//
// MFMA 32x32x16 operand layout (lane l = 32*g + i, g = l/32, i = l%32):
//   A: lane holds A[i][8g + e], e = 0..7;  B: lane holds B[8g + e][i];
//   D: lane holds D[8*(q) + 4g + j][i] in acc[4q + j], q, j = 0..3.
#include "common.hpp"

#include <algorithm>

namespace scir {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kMaxLds = 160 * 1024;
constexpr int kStageBatch = 4;  // LDS units (8 samples each) loaded per lane before splitting

inline int mfma_m_hi(int K) { return (K + 14) / 16; }  // ceil((K-1)/16) for K >= 1
inline int mfma_steps(int K) { return mfma_m_hi(K) + 2; }

__device__ __forceinline__ int unit_addr(int u) { return (u + (u >> 2)) * 16; }
inline size_t unit_bytes(int units) { return (size_t)(units + (units >> 2) + 1) * 16; }

__device__ __forceinline__ unsigned short bf16_bits(__bf16 v) { return __builtin_bit_cast(unsigned short, v); }
__device__ __forceinline__ unsigned short f16_bits(_Float16 v) { return __builtin_bit_cast(unsigned short, v); }

// Split term s (0 = hi, 1 = mid, 2 = lo) of v, as bf16 bits.
__device__ __forceinline__ void split3(float v, unsigned short &h, unsigned short &m,
                                       unsigned short &l)
{
    __bf16 bh = (__bf16)v;
    float r1 = v - (float)bh;
    __bf16 bm = (__bf16)r1;
    float r2 = r1 - (float)bm;
    __bf16 bl = (__bf16)r2;
    h = bf16_bits(bh);
    m = bf16_bits(bm);
    l = bf16_bits(bl);
}

__device__ __forceinline__ unsigned int pack2(unsigned short lo, unsigned short hi)
{
    return (unsigned int)lo | ((unsigned int)hi << 16);
}

template <int S>
__device__ __forceinline__ f32x16 mfma(u32x4 a, u32x4 b, f32x16 c)
{
    if (S == 3)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

// One half-step for TW tiles: acc0 += hi(b)*hi(x); acc1 += the five
// correction products (BF16x3), or acc0 += b*x (FP16).
template <int S, int TW>
__device__ __forceinline__ void step_mfma(f32x16 (&acc0)[TW], f32x16 (&acc1)[TW],
                                          const u32x4 (&a)[S], const u32x4 (&bv)[TW][S])
{
#pragma unroll
    for (int t = 0; t < TW; ++t) acc0[t] = mfma<S>(a[0], bv[t][0], acc0[t]);
    if (S == 3) {
        constexpr int M = S > 1 ? 1 : 0, L = S > 2 ? 2 : 0;
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            acc1[t] = mfma<S>(a[0], bv[t][M], acc1[t]);  // hi(b) * mid(x)
            acc1[t] = mfma<S>(a[M], bv[t][0], acc1[t]);  // mid(b) * hi(x)
            acc1[t] = mfma<S>(a[M], bv[t][M], acc1[t]);  // mid(b) * mid(x)
            acc1[t] = mfma<S>(a[0], bv[t][L], acc1[t]);  // hi(b) * lo(x)
            acc1[t] = mfma<S>(a[L], bv[t][0], acc1[t]);  // lo(b) * hi(x)
        }
    }
}

// 16 bytes from a 4-byte aligned LDS address.
__device__ __forceinline__ u32x4 lds_read16_a4(const unsigned char *p)
{
    const unsigned int *q = reinterpret_cast<const unsigned int *>(p);
    return u32x4{q[0], q[1], q[2], q[3]};
}

// A operands of half-steps mi and mi+1: `abase[s]` is this lane's byte
// address of half-step 0 in split s's table; each half-step moves 16 entries
// (32 bytes) down the reversed table.
template <int S>
__device__ __forceinline__ void load_a(u32x4 (&a)[2][S], const unsigned char *lds,
                                       const int (&abase)[S], int mi)
{
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int s = 0; s < S; ++s) a[h][s] = lds_read16_a4(lds + abase[s] - 32 * (mi + h));
}

// B operands of half-step mi for TW tiles (unit clamped to 0 for look-ahead
// steps past the last one, whose values are never used).
template <int S, int TW>
__device__ __forceinline__ void load_b(u32x4 (&bv)[TW][S], const unsigned char *lds,
                                       int split_stride, const int (&ubase)[TW], int mi)
{
#pragma unroll
    for (int t = 0; t < TW; ++t) {
        const int u = max(ubase[t] - 2 * mi, 0);
#pragma unroll
        for (int s = 0; s < S; ++s)
            bv[t][s] = *reinterpret_cast<const u32x4 *>(lds + s * split_stride + unit_addr(u));
    }
}

// Workgroup id -> logical segment id so that the workgroups one XCD receives
// (hardware dispatches id b to XCD b % 8) cover a contiguous range of
// segments: neighbouring segments share their input history in that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nb)
{
    const int64_t xcd = bid & 7, within = bid >> 3;
    const int64_t q = nb >> 3, rem = nb & 7;
    return xcd * q + (xcd < rem ? xcd : rem) + within;
}

template <int S, int TW>
__global__ __launch_bounds__(kThreads) void fir_mfma_kernel(
    const float *__restrict__ x, int64_t ldx, int64_t N, float *__restrict__ y, int64_t ldy,
    const float *__restrict__ b, int K, int msteps, int m_hi, int64_t row0, int segs_per_row,
    int64_t nblocks, int split_stride, int tab_offset, int copy_bytes, int vec_x, int vec_y,
    int rev)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    constexpr int SEG_BLOCKS = kWaves * TW * 32;
    const int64_t lid = xcd_remap(blockIdx.x, nblocks);
    const int64_t row = row0 + lid / segs_per_row;
    const int64_t P0 = (lid % segs_per_row) * SEG_BLOCKS;  // first output block of the segment
    const float *__restrict__ xrow = x + row * ldx;
    const int64_t xb = 32 * P0 - 16 * (int64_t)m_hi;        // sample index of LDS unit 0
    const int units = SEG_BLOCKS * 4 + 2 * m_hi;

    // ---- stage: x[xb .. xb + 8*units) -> S split arrays in LDS ----
    // Each lane issues the loads of kStageBatch units before splitting any of
    // them, so their memory latencies overlap.
    for (int u0 = threadIdx.x; u0 < units; u0 += kStageBatch * kThreads) {
        f32x4 raw[kStageBatch][2];
#pragma unroll
        for (int j = 0; j < kStageBatch; ++j) {
            const int u = u0 + j * kThreads;
            const int64_t g0 = xb + 8 * (int64_t)u;
            if (u >= units) {
                raw[j][0] = raw[j][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            } else if (vec_x && g0 >= 0 && g0 + 7 < N) {
                if (!rev) {
                    raw[j][0] =
                        __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(xrow + g0));
                    raw[j][1] =
                        __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(xrow + g0 + 4));
                } else {  // logical g0..g0+7 == physical N-1-g0 .. N-8-g0
                    const f32x4 a = __builtin_nontemporal_load(
                        reinterpret_cast<const f32x4 *>(xrow + (N - 4 - g0)));
                    const f32x4 c = __builtin_nontemporal_load(
                        reinterpret_cast<const f32x4 *>(xrow + (N - 8 - g0)));
                    raw[j][0] = f32x4{a.w, a.z, a.y, a.x};
                    raw[j][1] = f32x4{c.w, c.z, c.y, c.x};
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int64_t i = g0 + e;
                    raw[j][e >> 2][e & 3] = (i >= 0 && i < N) ? xrow[rev ? N - 1 - i : i] : 0.0f;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kStageBatch; ++j) {
            const int u = u0 + j * kThreads;
            if (u >= units) break;
            unsigned short t[S][8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float v = raw[j][e >> 2][e & 3];
                if (S == 3)
                    split3(v, t[0][e], t[S > 1 ? 1 : 0][e], t[S > 2 ? 2 : 0][e]);
                else
                    t[0][e] = f16_bits((_Float16)v);
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                u32x4 w = {pack2(t[s][0], t[s][1]), pack2(t[s][2], t[s][3]),
                           pack2(t[s][4], t[s][5]), pack2(t[s][6], t[s][7])};
                *reinterpret_cast<u32x4 *>(lds + s * split_stride + unit_addr(u)) = w;
            }
        }
    }
    // ---- tap tables: T_s[j] = split_s(b[C - j]), j in [0, L), two copies ----
    // (C = 16*m_hi + 79 keeps every look-ahead read at j >= 0; L = C + 32.)
    const int C = 16 * m_hi + 79, L = C + 32;
    for (int q = threadIdx.x; q < 2 * (L + 1); q += kThreads) {
        const int c = q / (L + 1), pos = q - c * (L + 1), j = pos - c;
        const int k = C - j;
        const float v = (j >= 0 && j < L && k >= 0 && k < K) ? b[k] : 0.0f;
        unsigned short t[3];
        if (S == 3)
            split3(v, t[0], t[1], t[2]);
        else
            t[0] = f16_bits((_Float16)v);
#pragma unroll
        for (int s2 = 0; s2 < S; ++s2)
            *reinterpret_cast<unsigned short *>(lds + tab_offset + (2 * s2 + c) * copy_bytes +
                                                2 * pos) = t[s2];
    }
    __syncthreads();

    // ---- contraction ----
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p = lane & 31, g = lane >> 5;
    f32x16 acc0[TW], acc1[TW];
#pragma unroll
    for (int t = 0; t < TW; ++t) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            acc0[t][i] = 0.0f;
            acc1[t][i] = 0.0f;
        }
    }
    // unit of B_m for tile t: 4*(tile block offset + p) + 2*(m_hi - m) + g
    int ubase[TW];
#pragma unroll
    for (int t = 0; t < TW; ++t) ubase[t] = 4 * ((wave * TW + t) * 32 + p) + g + 2 * m_hi + 2;

    // Software pipeline over half-steps, 4 per trip with no branches: A
    // operands are read two half-steps ahead and B operands one half-step
    // ahead, alternating between two register sets so no copies are needed.
    // The tap tables extend far enough for the look-ahead A reads (zeros) and
    // look-ahead B units are clamped to 0, so every read stays in bounds; the
    // few remaining half-steps run after the loop.
    // Lane's A read for half-step mi starts at entry j0 = C + 16 - p + 8g - 16 mi
    // of the copy whose shift c makes j0 + c even (4-byte aligned).
    int abase[S];
    {
        const int c = (C - p) & 1;
#pragma unroll
        for (int s2 = 0; s2 < S; ++s2)
            abase[s2] = tab_offset + (2 * s2 + c) * copy_bytes + 2 * (C + 16 - p + 8 * g + c);
    }
    u32x4 aA[2][S], aB[2][S], bA[TW][S], bB[TW][S];
    load_a<S>(aA, lds, abase, 0);
    load_b<S, TW>(bA, lds, split_stride, ubase, 0);
    int mi = 0;
    // sched_barrier(0) fences keep each look-ahead load group ahead of the
    // MFMAs it must overlap (the scheduler would otherwise sink the loads next
    // to their first use and wait on them).
    for (; mi + 4 <= msteps; mi += 4) {
        load_a<S>(aB, lds, abase, mi + 2);
        load_b<S, TW>(bB, lds, split_stride, ubase, mi + 1);
        __builtin_amdgcn_sched_barrier(0);
        step_mfma<S, TW>(acc0, acc1, aA[0], bA);
        __builtin_amdgcn_sched_barrier(0);
        load_b<S, TW>(bA, lds, split_stride, ubase, mi + 2);
        __builtin_amdgcn_sched_barrier(0);
        step_mfma<S, TW>(acc0, acc1, aA[1], bB);
        __builtin_amdgcn_sched_barrier(0);
        load_a<S>(aA, lds, abase, mi + 4);
        load_b<S, TW>(bB, lds, split_stride, ubase, mi + 3);
        __builtin_amdgcn_sched_barrier(0);
        step_mfma<S, TW>(acc0, acc1, aB[0], bA);
        __builtin_amdgcn_sched_barrier(0);
        load_b<S, TW>(bA, lds, split_stride, ubase, mi + 4);
        __builtin_amdgcn_sched_barrier(0);
        step_mfma<S, TW>(acc0, acc1, aB[1], bB);
        __builtin_amdgcn_sched_barrier(0);
    }
    // aA holds half-steps mi, mi+1 and bA half-step mi.
    if (mi < msteps) {
        load_b<S, TW>(bB, lds, split_stride, ubase, mi + 1);
        step_mfma<S, TW>(acc0, acc1, aA[0], bA);
        if (mi + 1 < msteps) {
            load_a<S>(aB, lds, abase, mi + 2);
            load_b<S, TW>(bA, lds, split_stride, ubase, mi + 2);
            step_mfma<S, TW>(acc0, acc1, aA[1], bB);
            if (mi + 2 < msteps) step_mfma<S, TW>(acc0, acc1, aB[0], bA);
        }
    }

    // ---- epilogue: y[32*(block) + 8q + 4g + j] = acc[4q + j] ----
    float *__restrict__ yrow = y + row * ldy;
#pragma unroll
    for (int t = 0; t < TW; ++t) {
        const int64_t n0 = 32 * (P0 + (wave * TW + t) * 32 + p) + 4 * g;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            f32x4 v;
            v.x = acc0[t][4 * q + 0] + acc1[t][4 * q + 0];
            v.y = acc0[t][4 * q + 1] + acc1[t][4 * q + 1];
            v.z = acc0[t][4 * q + 2] + acc1[t][4 * q + 2];
            v.w = acc0[t][4 * q + 3] + acc1[t][4 * q + 3];
            const int64_t n = n0 + 8 * q;
            if (vec_y && n + 3 < N) {
                if (!rev)
                    *reinterpret_cast<f32x4 *>(yrow + n) = v;
                else
                    *reinterpret_cast<f32x4 *>(yrow + (N - 4 - n)) = f32x4{v.w, v.z, v.y, v.x};
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (n + j < N) yrow[rev ? N - 1 - (n + j) : n + j] = v[j];
            }
        }
    }
}

int align_ok(const float *p, int64_t ld)
{
    return ((reinterpret_cast<uintptr_t>(p) & 15) == 0) && (ld % 4 == 0);
}

// Bytes of one copy of one split's tap table (L + 1 entries), rounded so
// consecutive copies start 128 bytes apart modulo the 256-byte LDS bank cycle.
inline int tab_copy_bytes(int m_hi)
{
    const int L = 16 * m_hi + 79 + 32;
    int bytes = (2 * (L + 1) + 15) & ~15;
    while (bytes % 256 != 128) bytes += 16;
    return bytes;
}

template <int S, int TW>
int launch_mfma(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                float *y, int64_t ldy, bool rev, hipStream_t stream)
{
    constexpr int SEG_BLOCKS = kWaves * TW * 32;
    const int m_hi = mfma_m_hi(K), msteps = mfma_steps(K);
    const int units = SEG_BLOCKS * 4 + 2 * m_hi;
    const int split_stride = (int)unit_bytes(units);
    const int tab_offset = S * split_stride;
    const int copy_bytes = tab_copy_bytes(m_hi);
    const size_t lds = (size_t)tab_offset + (size_t)2 * S * copy_bytes;
    if (lds > (size_t)kMaxLds)
        return fail(SCIR_HIP_ERR_UNSUPPORTED, "K=%d needs %zu bytes of LDS for the MFMA path", K,
                    lds);
    static bool attr_set = false;  // guarded by the device mutex held by the caller
    if (!attr_set) {
        SCIR_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&fir_mfma_kernel<S, TW>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds));
        attr_set = true;
    }
    const int64_t seg_len = 32 * (int64_t)SEG_BLOCKS;
    const int64_t segs = (N + seg_len - 1) / seg_len;
    if (segs > 0x7fffffff)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row too long (N=%lld)", (long long)N);
    const int nvec = !rev || (N % 4 == 0);
    const int vec_x = align_ok(x, ldx) && nvec, vec_y = align_ok(y, ldy) && nvec;
    const int64_t rows_per_launch = std::max<int64_t>(1, 0x7fffffff / segs);
    for (int64_t r0 = 0; r0 < B; r0 += rows_per_launch) {
        const int64_t rows = std::min(rows_per_launch, B - r0);
        const int64_t nb = rows * segs;
        hipLaunchKernelGGL((fir_mfma_kernel<S, TW>), dim3((unsigned)nb), dim3(kThreads), lds,
                           stream, x, ldx, N, y, ldy, b, K, msteps, m_hi, r0, (int)segs, nb,
                           split_stride, tab_offset, copy_bytes, vec_x, vec_y, (int)rev);
        SCIR_TRY(check_launch("fir_mfma_kernel"));
    }
    return SCIR_HIP_OK;
}

template <int S>
int run_mfma(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N, float *y,
             int64_t ldy, bool rev, hipStream_t stream)
{
    const TuneChoice tc = tune_lookup(K, N);
    int tw = S == 3 ? tc.mfma_tiles : tc.mfma_tiles_fp16;
    // Short rows: smaller segments keep the grid full.
    if (N <= 4096) tw = 1;
    switch (tw) {
    case 1: return launch_mfma<S, 1>(x, ldx, b, K, B, N, y, ldy, rev, stream);
    case 4: return launch_mfma<S, 4>(x, ldx, b, K, B, N, y, ldy, rev, stream);
    default: return launch_mfma<S, 2>(x, ldx, b, K, B, N, y, ldy, rev, stream);
    }
}

}  // namespace

int fir_mfma_dev(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                 float *y, int64_t ldy, int precision, bool rev, hipStream_t stream,
                 DeviceState *st)
{
    if (B == 0 || N == 0) return SCIR_HIP_OK;
    if (K == 0) return fill_zero_2d(y, ldy, B, N, stream);
    (void)st;
    if (precision == SCIR_HIP_PREC_BF16X3)
        return run_mfma<3>(x, ldx, b, K, B, N, y, ldy, rev, stream);
    if (precision == SCIR_HIP_PREC_FP16)
        return run_mfma<1>(x, ldx, b, K, B, N, y, ldy, rev, stream);
    return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "precision %d is not an MFMA mode", precision);
}

}  // namespace scir
