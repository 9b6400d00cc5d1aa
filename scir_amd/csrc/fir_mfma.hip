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
//
// Operand reuse.  A wave owns TW output tiles of 32 blocks each, interleaved:
// tile t covers blocks P + t + TW*i (i = 0..31, the MFMA column).  Then tile
// t's B operand at half-step m equals tile 0's at half-step m - 2t, so the
// loop runs over q and every iteration reads ONE B fragment D(q) (tile 0 at
// half-step q) and applies it to tile t with A_{q+2t}.  The A fragments live
// in a ring of 2*TW register sets: each is read once and used by TW tiles.
// Per iteration and split term: one A and one B read for TW MFMAs.
//
// A comes from the taps only.  Each workgroup keeps the split taps in LDS as a
// compact reversed table, T[j] = b[C - j]: lane (i, g) of A_m holds the 8
// consecutive entries starting at j0 = C - 16m - i + 8g (one 16-byte read).
// The table is stored twice, the second copy shifted by one entry, so every
// lane's read is 4-byte aligned; the copies sit 128 bytes apart modulo the
// 256-byte bank cycle.  Entries outside the taps are zero.
//
// B comes from x: the workgroup's segment of 4*TW*32 blocks plus 16*m_hi
// samples of history is split once into LDS.  LDS unit u (8 samples, 16 B)
// of a split array sits at byte 16*(u + u/(4*TW)): lanes read units 4*TW
// apart, and this padding spreads 16 such lanes over distinct bank groups.
//
// Precision modes (template S = number of split terms):
//   S = 3 (BF16x3): v = hi + mid + lo, each bf16 (hi = rn(v), mid = rn(v-hi),
//     lo = rn(v-hi-mid); both subtractions are exact), for x and for b.  Six
//     products are kept: hh (own accumulator), hm, mh, mm, hl, lh (a second
//     accumulator); the dropped ml, lm, ll are below 2^-24 relative.
//   S = 1 (FP16): x and b rounded to fp16, one product.
// Non-finite values: a zero tap of the Toeplitz block still multiplies its
// sample (and a zero sample its tap), so an Inf or NaN spreads to every output
// of the 32-output blocks it enters, and a finite value beyond the bf16/fp16
// range splits into Inf.  The epilogue therefore recomputes every non-finite
// output from the reference loop's terms (direct_output, FP32 fused
// multiply-add): outputs are non-finite exactly where the reference loop's
// are, at the cost of one class check per output when none is.
//
// MFMA 32x32x16 operand layout (lane l = 32*g + i, g = l/32, i = l%32):
//   A: lane holds A[i][8g + e], e = 0..7;  B: lane holds B[8g + e][i];
//   D: lane holds D[8q + 4g + j][i] in acc[4q + j], q, j = 0..3.
#include "direct_tiles.hpp"

#include <algorithm>

namespace scir {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef float f32x4 __attribute__((ext_vector_type(4)));
// Four floats at any 4-byte aligned global address: the kernels load rows and
// store outputs with 16-byte accesses whatever the row's offset or pitch (the
// hardware runs in unaligned access mode; an access that crosses a cache line
// costs two requests, far less than four 4-byte accesses).
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kMaxLds = 160 * 1024;
// LDS units (8 samples each) per lane loaded together when a segment's history
// is longer than the prefetched part (2*m_hi > 256 units, K > 2049).
constexpr int kRestBatch = 2;

__host__ __device__ inline int mfma_m_hi(int K) { return (K + 14) / 16; }  // ceil((K-1)/16), K >= 1
inline int mfma_steps(int K) { return mfma_m_hi(K) + 2; }

// Byte address of LDS unit u in a split array whose lanes read 4*TW apart.
template <int TW>
__device__ __forceinline__ int unit_addr(int u)
{
    return (u + u / (4 * TW)) * 16;
}
inline size_t unit_bytes(int units, int tw) { return (size_t)(units + units / (4 * tw) + 1) * 16; }

// Tap table: entries T[j] for j in [0, L), with C chosen so that every A read
// of the loop (half-step indices -2*TW+2 .. msteps+2*TW-1) lands inside it.
inline int tab_C(int m_hi, int tw) { return 16 * (m_hi + 2 * tw + 1) + 32; }
inline int tab_L(int m_hi, int tw) { return tab_C(m_hi, tw) + 32 * tw + 32; }
inline int tab_copy_bytes(int m_hi, int tw)
{
    int bytes = (2 * (tab_L(m_hi, tw) + 1) + 15) & ~15;
    while (bytes % 256 != 128) bytes += 16;
    return bytes;
}

__device__ __forceinline__ unsigned short bf16_bits(__bf16 v) { return __builtin_bit_cast(unsigned short, v); }
__device__ __forceinline__ unsigned short f16_bits(_Float16 v) { return __builtin_bit_cast(unsigned short, v); }

// hi, mid, lo bf16 terms of v (v == hi + mid + lo exactly for normal v).
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

// The same split for two values at once, as packed words (v0 in the low half,
// v1 in the high half): one v_cvt_pk_bf16_f32 per term for both values, and
// the words come out in the layout the split arrays store.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned int bf16x2_bits(f32x2 v)
{
    return __builtin_bit_cast(unsigned int, __builtin_convertvector(v, bf16x2));
}

__device__ __forceinline__ f32x2 bf16x2_widen(unsigned int w)
{
    return f32x2{__builtin_bit_cast(float, w << 16), __builtin_bit_cast(float, w & 0xffff0000u)};
}

__device__ __forceinline__ void split3_pair(float v0, float v1, unsigned int &h,
                                            unsigned int &m, unsigned int &l)
{
    const f32x2 v = {v0, v1};
    h = bf16x2_bits(v);
    const f32x2 r1 = v - bf16x2_widen(h);
    m = bf16x2_bits(r1);
    const f32x2 r2 = r1 - bf16x2_widen(m);
    l = bf16x2_bits(r2);
}

__device__ __forceinline__ unsigned int pack2(unsigned short lo, unsigned short hi)
{
    return (unsigned int)lo | ((unsigned int)hi << 16);
}

// Tiny magnitudes (BF16x3).  The split terms of v (hi ~ v, mid ~ 2^-8 v,
// lo ~ 2^-16 v) are exact bf16 values, and the kept cross products exact FP32
// values, only while they stay in the normal ranges: for samples and taps
// whose nonzero magnitudes are at least 2^-110, with products of at least
// 2^-100.  A segment whose staged samples or taps go below that (subnormal
// and near-subnormal data) is recomputed from the reference loop's terms.
// Magnitudes are tracked by binary exponent (v = m 2^e, m in [0.5, 1); a
// zero, Inf or NaN gives 0, so it never looks tiny): 2^(e-1) <= |v| < 2^e.
__device__ __forceinline__ int min_exponent(int m, float v)
{
    return min(m, __builtin_amdgcn_frexp_expf(v));
}

// Minimum over the wave: DPP row shifts, then row broadcasts; all VALU.
__device__ __forceinline__ int wave_min(int v)
{
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x111, 0xF, 0xF, false));  // row_shr:1
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x112, 0xF, 0xF, false));  // row_shr:2
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x114, 0xF, 0xF, false));  // row_shr:4
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x118, 0xF, 0xF, false));  // row_shr:8
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}

// Smallest exponents of the samples (e_x) and taps (e_b) against the limits
// above: 2^(e-1) < 2^-110, or 2^(e_x-1) 2^(e_b-1) < 2^-100.
__device__ __forceinline__ bool tiny_magnitudes(int e_x, int e_b)
{
    return e_x <= -110 || e_b <= -110 || e_x + e_b <= -99;
}

// LDS after the tap tables (after the item slots, ragged): the smallest
// sample exponent of the staged segment per wave, then the taps' per wave.
// Each wave writes its own word after a barrier that every lane passes only
// after reading the previous values, so the words need no reset.
constexpr int kMinsBytes = 32;

// Whether the segment just staged has tiny magnitudes (after the barrier
// that follows staging).
__device__ __forceinline__ bool segment_is_tiny(const int *mins)
{
    const int ex = min(min(mins[0], mins[1]), min(mins[2], mins[3]));
    const int eb = min(min(mins[4], mins[5]), min(mins[6], mins[7]));
    return tiny_magnitudes(ex, eb);
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

// acc0 += hi(b)*hi(x); acc1 += the five correction products (BF16x3), or
// acc0 += b*x (FP16).
template <int S>
__device__ __forceinline__ void tile_mfma(f32x16 &acc0, f32x16 &acc1, const u32x4 (&a)[S],
                                          const u32x4 (&d)[S])
{
    acc0 = mfma<S>(a[0], d[0], acc0);
    if (S == 3) {
        constexpr int M = S > 1 ? 1 : 0, L = S > 2 ? 2 : 0;
        acc1 = mfma<S>(a[0], d[M], acc1);  // hi(b) * mid(x)
        acc1 = mfma<S>(a[M], d[0], acc1);  // mid(b) * hi(x)
        acc1 = mfma<S>(a[M], d[M], acc1);  // mid(b) * mid(x)
        acc1 = mfma<S>(a[0], d[L], acc1);  // hi(b) * lo(x)
        acc1 = mfma<S>(a[L], d[0], acc1);  // lo(b) * hi(x)
    }
}

// 16 bytes from a 4-byte aligned LDS address.
__device__ __forceinline__ u32x4 lds_read16_a4(const unsigned char *p)
{
    const unsigned int *q = reinterpret_cast<const unsigned int *>(p);
    return u32x4{q[0], q[1], q[2], q[3]};
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

// Loop state of one wave: accumulators, A ring, double-buffered B fragment.
template <int S, int TW>
struct Contraction {
    static constexpr int RS = 2 * TW;  // A ring slots
    f32x16 acc0[TW], acc1[TW];
    u32x4 ring[RS][S];
    u32x4 dA[S], dB[S];
    const unsigned char *lds;
    int split_stride;
    int ubase;      // B unit of iteration 0
    int abase[S];   // A table byte address of half-step index 0
    int msteps, Q;  // half-steps per tile, loop iterations

    __device__ __forceinline__ void load_a(u32x4 (&a)[S], int sa) const
    {
#pragma unroll
        for (int s = 0; s < S; ++s) a[s] = lds_read16_a4(lds + abase[s] - 32 * sa);
    }
    __device__ __forceinline__ void load_d(u32x4 (&d)[S], int n) const
    {
        const int u = max(ubase - 2 * n, 0);  // past the last iteration: any in-range unit
#pragma unroll
        for (int s = 0; s < S; ++s)
            d[s] = *reinterpret_cast<const u32x4 *>(lds + s * split_stride + unit_addr<TW>(u));
    }

    // Iteration n (n % RS == KK): tile t uses A of half-step index
    // sa = n - 2*TW + 2 + 2t (ring slot (KK + 2 + 2t) % RS) and D(n).  After
    // tile 0, whose A is then dead, the slot is refilled with half-step n + 2.
    // With PRED, a tile whose sa is outside [0, msteps) is skipped.
    template <int KK, bool PRED>
    __device__ __forceinline__ void step(int n)
    {
        u32x4(&dcur)[S] = (KK & 1) ? dB : dA;
        u32x4(&dnxt)[S] = (KK & 1) ? dA : dB;
        load_d(dnxt, n + 1);
        __builtin_amdgcn_sched_barrier(0);
        {
            constexpr int slot = (KK + 2) % RS;
            const int sa = n - 2 * TW + 2;
            if (!PRED || (sa >= 0 && sa < msteps)) tile_mfma<S>(acc0[0], acc1[0], ring[slot], dcur);
            __builtin_amdgcn_sched_barrier(0);
            load_a(ring[slot], n + 2);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int t = 1; t < TW; ++t) {
            const int slot = (KK + 2 + 2 * t) % RS;
            const int sa = n - 2 * TW + 2 + 2 * t;
            if (!PRED || (sa >= 0 && sa < msteps)) tile_mfma<S>(acc0[t], acc1[t], ring[slot], dcur);
        }
        __builtin_amdgcn_sched_barrier(0);
    }

    template <int KK>
    __device__ __forceinline__ void tail(int n)
    {
        if constexpr (KK < RS) {
            if (n < Q) {
                step<KK, true>(n);
                tail<KK + 1>(n + 1);
            }
        }
    }

    template <int KK, bool PRED>
    __device__ __forceinline__ void body(int n0)
    {
        if constexpr (KK < RS) {
            step<KK, PRED>(n0 + KK);
            body<KK + 1, PRED>(n0);
        }
    }

    __device__ __forceinline__ void run()
    {
#pragma unroll
        for (int t = 0; t < TW; ++t)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc0[t][i] = acc1[t][i] = 0.0f;
        // The ring at iteration 0 holds half-step indices -2TW+2 .. 1 in slot
        // (index mod RS); negative indices read zero table entries.
#pragma unroll
        for (int k = 0; k < RS; ++k) {
            const int sa = -2 * TW + 2 + k;
            load_a(ring[(k + 2) % RS], sa);  // (sa mod RS) == (k + 2) mod RS
        }
        load_d(dA, 0);
        // The first and last 2TW-2 iterations have tiles outside their
        // half-steps: they run predicated; the rest as unrolled RS-iteration
        // bodies with no branches.
        const int ramp = 2 * TW - 2;
        int n0 = 0;
        if (ramp > 0 && Q >= RS) {
            body<0, true>(0);
            n0 = RS;
        }
        for (; n0 + RS <= Q - ramp; n0 += RS) body<0, false>(n0);
        // At most RS - 1 + ramp < 2*RS iterations remain.
        tail<0>(n0);
        tail<0>(n0 + RS);
    }
};

// Samples of LDS unit u: 8 samples from xb + 8u of the row (logical order),
// zero outside [0, N) and for u >= units.  With REV logical sample i is
// physical sample N-1-i, and r holds the unit in physical order (slot p =
// logical sample g0 + 7 - p), which unit_sample reads back to front.  Every
// load lands in its slot of r with no register move after it: a move (or a
// select) would wait for the load, and the prefetch of the next segment must
// stay in flight through the contraction.
template <bool REV>
__device__ __forceinline__ void load_unit(const float *__restrict__ xrow, int64_t N, int64_t xb,
                                          int u, int units, f32x4 (&r)[2])
{
    const int64_t g0 = xb + 8 * (int64_t)u;
    if (u >= units) {
        r[0] = r[1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    } else if (g0 >= 0 && g0 + 7 < N) {
        const float *p = REV ? xrow + (N - 8 - g0) : xrow + g0;
        r[0] = __builtin_nontemporal_load(reinterpret_cast<const f32x4u *>(p));
        r[1] = __builtin_nontemporal_load(reinterpret_cast<const f32x4u *>(p + 4));
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int64_t i = g0 + e;
            const int slot = REV ? 7 - e : e;
            r[slot >> 2][slot & 3] = 0.0f;
            if (i >= 0 && i < N) r[slot >> 2][slot & 3] = xrow[REV ? N - 1 - i : i];
        }
    }
}

// Logical sample e (0..7) of a unit loaded by load_unit<REV>.
template <bool REV>
__device__ __forceinline__ float unit_sample(const f32x4 (&r)[2], int e)
{
    const int slot = REV ? 7 - e : e;
    return r[slot >> 2][slot & 3];
}

// Split the 8 samples of unit u into the S split arrays of the LDS.
template <int S, int TW, bool REV>
__device__ __forceinline__ void store_unit(unsigned char *lds, int split_stride, int u,
                                           const f32x4 (&r)[2], int &min_e)
{
    if (S == 3) {
        // Smallest binary exponent of the unit's samples: one instruction per
        // sample and a three-input minimum per two.
#pragma unroll
        for (int e = 0; e < 8; e += 2)
            min_e = min(min_e, min(__builtin_amdgcn_frexp_expf(r[e >> 2][e & 3]),
                                   __builtin_amdgcn_frexp_expf(r[e >> 2][(e & 3) + 1])));
        // Pairs in the order the loads left them (slots 2p, 2p + 1).  Back to
        // front, slot pair p holds logical samples 7 - 2p and 6 - 2p, so word
        // 3 - p takes it with its halves swapped.  (Pairs formed from reversed
        // slots made the compiler keep the prefetch registers in scratch.)
        unsigned int w[3][4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            unsigned int h, m, l;
            split3_pair(r[p >> 1][2 * (p & 1)], r[p >> 1][2 * (p & 1) + 1], h, m, l);
            const int q = REV ? 3 - p : p;
            w[0][q] = REV ? __builtin_amdgcn_alignbit(h, h, 16) : h;
            w[1][q] = REV ? __builtin_amdgcn_alignbit(m, m, 16) : m;
            w[2][q] = REV ? __builtin_amdgcn_alignbit(l, l, 16) : l;
        }
#pragma unroll
        for (int s = 0; s < 3; ++s)
            *reinterpret_cast<u32x4 *>(lds + s * split_stride + unit_addr<TW>(u)) =
                u32x4{w[s][0], w[s][1], w[s][2], w[s][3]};
    } else {
        unsigned short t[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) t[e] = f16_bits((_Float16)unit_sample<REV>(r, e));
#pragma unroll
        for (int s = 0; s < S; ++s) {
            u32x4 w = {pack2(t[0], t[1]), pack2(t[2], t[3]), pack2(t[4], t[5]), pack2(t[6], t[7])};
            *reinterpret_cast<u32x4 *>(lds + s * split_stride + unit_addr<TW>(u)) = w;
        }
    }
}

// Tap tables for taps b[0, K) (lfilter order): T_s[j] = split_s(b[C - j]),
// j in [0, L), in two copies one entry apart (copy c holds T_s[pos - c] at
// entry pos, pos in [0, L]), so every lane's 16-byte A read starts 4-byte
// aligned in one of them.  Zero outside the taps.  Each tap is loaded once for
// both copies, four loads in flight per thread before any is used.
template <int S, int TW>
__device__ __forceinline__ void build_tap_table(unsigned char *lds, int tab_offset,
                                                int copy_bytes, const float *__restrict__ b,
                                                int K, int m_hi, int &min_b)
{
    const int C = 16 * (m_hi + 2 * TW + 1) + 32, L = C + 32 * TW + 32;
    constexpr int R = 4;
    // Entry j in [-1, L]: copy 0 at pos j (j >= 0), copy 1 at pos j + 1 (j < L).
    for (int base = -1; base <= L; base += R * kThreads) {
        float v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int jj = base + (int)threadIdx.x + r * kThreads, k = C - jj;
            v[r] = (jj >= 0 && jj < L && k >= 0 && k < K) ? b[k] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int jj = base + (int)threadIdx.x + r * kThreads;
            if (jj > L) break;
            if (S == 3) min_b = min_exponent(min_b, v[r]);
            unsigned short t[3];
            if (S == 3)
                split3(v[r], t[0], t[1], t[2]);
            else
                t[0] = f16_bits((_Float16)v[r]);
#pragma unroll
            for (int s2 = 0; s2 < S; ++s2) {
                unsigned char *tab = lds + tab_offset + 2 * s2 * copy_bytes;
                if (jj >= 0) *reinterpret_cast<unsigned short *>(tab + 2 * jj) = t[s2];
                if (jj < L)
                    *reinterpret_cast<unsigned short *>(tab + copy_bytes + 2 * (jj + 1)) = t[s2];
            }
        }
    }
    // Every tap load above is consumed, but the compiler cannot tell after the
    // early exit; without this vmcnt(0) it would carry them as pending into
    // the segment loop and wait for each segment's prefetch before the
    // contraction.  (The first segment's samples, loaded before the table,
    // are staged right after this anyway.)
    __builtin_amdgcn_s_waitcnt(0x0f70);
}

// Stage LDS units 0 .. units-1 of a segment (unit 0 = sample xb of the row):
// the first PF units of each lane come from `raw` (loaded ahead), the rest
// are loaded here.
template <int S, int TW, int PF, bool REV>
__device__ __forceinline__ void stage_segment(unsigned char *lds, int split_stride,
                                              const f32x4 (&raw)[PF][2],
                                              const float *__restrict__ xrow, int64_t N,
                                              int64_t xb, int units, int tid, int &min_e)
{
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        const int u = tid + j * kThreads;
        if (u < units) store_unit<S, TW, REV>(lds, split_stride, u, raw[j], min_e);
    }
    if (units <= PF * kThreads) return;  // K <= 2049
    for (int u0 = tid + PF * kThreads; u0 < units; u0 += kRestBatch * kThreads) {
        f32x4 rest[kRestBatch][2];
#pragma unroll
        for (int j = 0; j < kRestBatch; ++j)
            load_unit<REV>(xrow, N, xb, u0 + j * kThreads, units, rest[j]);
#pragma unroll
        for (int j = 0; j < kRestBatch; ++j) {
            const int u = u0 + j * kThreads;
            if (u >= units) break;
            store_unit<S, TW, REV>(lds, split_stride, u, rest[j], min_e);
        }
    }
    // Every load above has been consumed (a unit past `units` is zeros, not
    // loaded), but the compiler cannot tell, and would wait for the next
    // segment's prefetch before the contraction instead: vmcnt(0) here.
    __builtin_amdgcn_s_waitcnt(0x0f70);
}

// Contraction of one staged segment by wave `wave`; lane (i, g) = (lane % 32,
// lane / 32).  The caller's barrier has made the staged units and the tap
// tables visible.
template <int S, int TW>
__device__ __forceinline__ void contract_segment(Contraction<S, TW> &ct, const unsigned char *lds,
                                                 int split_stride, int tab_offset,
                                                 int copy_bytes, int msteps, int m_hi, int wave,
                                                 int i, int g)
{
    const int C = 16 * (m_hi + 2 * TW + 1) + 32;
    ct.lds = lds;
    ct.split_stride = split_stride;
    ct.msteps = msteps;
    ct.Q = msteps + 2 * TW - 2;
    // D(n) = tile 0's B at half-step q = n - 2TW + 1: unit
    // 4*(wave offset + TW*i) + 2*(m_hi - q) + g.
    ct.ubase = 4 * (wave * 32 * TW + TW * i) + 2 * (m_hi + 2 * TW - 1) + g;
    {
        // A of half-step index sa (m = sa - 1) starts at entry
        // j0 = C - 16*(sa - 1) - i + 8g, in the copy whose shift c makes j0 + c even.
        const int c = (C - i) & 1;
#pragma unroll
        for (int s2 = 0; s2 < S; ++s2)
            ct.abase[s2] = tab_offset + (2 * s2 + c) * copy_bytes + 2 * (C + 16 - i + 8 * g + c);
    }
    ct.run();
}

// Epilogue: tile t, lane (i, g) holds block P0 + wave*32*TW + t + TW*i,
// samples 8q + 4g + j; outputs past N are dropped.  A non-finite output, and
// every output of a segment with tiny magnitudes, is recomputed from the
// reference loop's terms (x row xrow, taps b[0, K)).
template <int S, int TW, bool REV>
__device__ __forceinline__ void store_segment(const Contraction<S, TW> &ct,
                                              float *__restrict__ yrow, int64_t N, int64_t P0,
                                              int wave, int i, int g,
                                              const float *__restrict__ xrow,
                                              const float *__restrict__ b, int K, bool tiny)
{
    // Every output of the wave inside the row (a wave-uniform test): the
    // stores go out unconditionally, one class test per output sets a flag,
    // and only a flagged (or tiny) segment runs the recomputation pass over
    // the stored values (rolled loops: indexing the accumulators there would
    // put them in scratch).
    if (32 * (P0 + (int64_t)(wave + 1) * 32 * TW) <= N) {
        bool bad = false;
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            const int64_t n0 = 32 * (P0 + wave * 32 * TW + t + TW * i) + 4 * g;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                f32x4 v;
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = ct.acc0[t][4 * q + j] + ct.acc1[t][4 * q + j];
                bad |= !__builtin_isfinite(v.x) | !__builtin_isfinite(v.y) |
                       !__builtin_isfinite(v.z) | !__builtin_isfinite(v.w);
                const int64_t n = n0 + 8 * q;
                if (!REV)
                    *reinterpret_cast<f32x4u *>(yrow + n) = v;
                else
                    *reinterpret_cast<f32x4u *>(yrow + (N - 4 - n)) = f32x4{v.w, v.z, v.y, v.x};
            }
        }
        if (tiny || bad) {
#pragma unroll 1
            for (int t = 0; t < TW; ++t) {
                const int64_t n0 = 32 * (P0 + wave * 32 * TW + t + TW * i) + 4 * g;
#pragma unroll 1
                for (int q = 0; q < 4; ++q) {
                    const int64_t n = n0 + 8 * q;
#pragma unroll 1
                    for (int j = 0; j < 4; ++j) {
                        // The thread's own store above, read back.
                        float &y = yrow[REV ? N - 1 - (n + j) : n + j];
                        if (tiny || !__builtin_isfinite(y))
                            y = direct_output<false>(xrow, N, REV, b, K, n + j);
                    }
                }
            }
        }
        return;
    }
#pragma unroll
    for (int t = 0; t < TW; ++t) {
        const int64_t n0 = 32 * (P0 + wave * 32 * TW + t + TW * i) + 4 * g;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            f32x4 v;
            v.x = ct.acc0[t][4 * q + 0] + ct.acc1[t][4 * q + 0];
            v.y = ct.acc0[t][4 * q + 1] + ct.acc1[t][4 * q + 1];
            v.z = ct.acc0[t][4 * q + 2] + ct.acc1[t][4 * q + 2];
            v.w = ct.acc0[t][4 * q + 3] + ct.acc1[t][4 * q + 3];
            const int64_t n = n0 + 8 * q;
            if (tiny || !__builtin_isfinite(v.x) || !__builtin_isfinite(v.y) ||
                !__builtin_isfinite(v.z) || !__builtin_isfinite(v.w)) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (n + j < N && (tiny || !__builtin_isfinite(v[j])))
                        v[j] = direct_output<false>(xrow, N, REV, b, K, n + j);
            }
            if (n + 3 < N) {
                if (!REV)
                    *reinterpret_cast<f32x4u *>(yrow + n) = v;
                else
                    *reinterpret_cast<f32x4u *>(yrow + (N - 4 - n)) = f32x4{v.w, v.z, v.y, v.x};
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (n + j < N) yrow[REV ? N - 1 - (n + j) : n + j] = v[j];
            }
        }
    }
}

// One workgroup (4 waves) computes segs_per_wg consecutive segments of
// kWaves*TW*32 output blocks each (segment id = row * segs_per_row + index in
// the row).  A segment needs LDS units 0 .. units-1: 2*m_hi units of history
// and 4*SEG_BLOCKS of its own samples.  The first PF = 2*TW + 1 units of
// each lane (unit lane + 256*j) are loaded into registers one segment ahead:
// they are in flight while the waves run the previous segment's MFMAs.
// Units past PF*256 (only when 2*m_hi > 256) are loaded when the segment is
// staged.
template <int S, int TW, bool REV>
__global__ __launch_bounds__(kThreads) void fir_mfma_kernel(
    const float *__restrict__ x, int64_t ldx, int64_t N, float *__restrict__ y, int64_t ldy,
    const float *__restrict__ b, int K, int msteps, int m_hi, int64_t row0, int segs_per_row,
    int64_t nsegs, int segs_per_wg, int64_t nwg, int split_stride, int tab_offset,
    int copy_bytes)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    constexpr int SEG_BLOCKS = kWaves * TW * 32;
    constexpr int PF = 2 * TW + 1;
    const int64_t lid = xcd_remap(blockIdx.x, nwg);
    const int64_t seg_first = lid * segs_per_wg;
    const int64_t seg_end = min(nsegs, seg_first + segs_per_wg);
    const int units = SEG_BLOCKS * 4 + 2 * m_hi;

    // Row pointer and sample index of LDS unit 0 of segment sg.
    auto seg_base = [&](int64_t sg, const float *&xrow, int64_t &xb, int64_t &P0, int64_t &row) {
        row = row0 + sg / segs_per_row;
        P0 = (sg % segs_per_row) * SEG_BLOCKS;
        xrow = x + row * ldx;
        xb = 32 * P0 - 16 * (int64_t)m_hi;
    };

    f32x4 raw[PF][2];
    auto prefetch = [&](int64_t sg, int tid) {
        const float *xrow;
        int64_t xb, P0, row;
        seg_base(sg, xrow, xb, P0, row);
#pragma unroll
        for (int j = 0; j < PF; ++j)
            load_unit<REV>(xrow, N, xb, tid + j * kThreads, units, raw[j]);
    };
    prefetch(seg_first, (int)threadIdx.x);

    // Tap tables, built while the first segment's loads are in flight.
    int *mins = reinterpret_cast<int *>(lds + tab_offset + 2 * S * copy_bytes);
    {
        int min_b = 0;
        build_tap_table<S, TW>(lds, tab_offset, copy_bytes, b, K, m_hi, min_b);
        if (S == 3) {
            min_b = wave_min(min_b);
            if ((threadIdx.x & 63) == 0) mins[4 + (threadIdx.x >> 6)] = min_b;
        }
    }

    for (int64_t sg = seg_first; sg < seg_end; ++sg) {
        // Lane-dependent values are recomputed in every iteration rather than
        // hoisted out of the loop, where they would hold registers through
        // the contraction.
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int wave = tid >> 6, lane = tid & 63;
        const int i = lane & 31, g = lane >> 5;
        const float *xrow;
        int64_t xb, P0, row;
        seg_base(sg, xrow, xb, P0, row);

        // ---- stage: x[xb .. xb + 8*units) -> S split arrays in LDS ----
        if (sg != seg_first) __syncthreads();  // the previous segment's MFMAs are done
        int min_e = 0;
        stage_segment<S, TW, PF, REV>(lds, split_stride, raw, xrow, N, xb, units, tid, min_e);
        if (S == 3) {
            const int m = wave_min(min_e);
            if (lane == 0) mins[wave] = m;
        }
        __syncthreads();
        if (sg + 1 < seg_end) prefetch(sg + 1, tid);

        // ---- contraction and epilogue ----
        // Wave w's tiles cover outputs 32*(P0 + 32*TW*w) .. +1024*TW-1 of the
        // row: a wave past the row end has no output and skips both (the
        // contraction has no barrier), leaving the SIMD's matrix pipe to the
        // other wave on it.
        if (32 * (P0 + (int64_t)__builtin_amdgcn_readfirstlane(wave) * 32 * TW) < N) {
            Contraction<S, TW> ct;
            contract_segment<S, TW>(ct, lds, split_stride, tab_offset, copy_bytes, msteps, m_hi,
                                    wave, i, g);
            // Read after the contraction (the words change only after the
            // next segment's first barrier): nothing extra stays live through it.
            const bool tiny = S == 3 && segment_is_tiny(mins);
            store_segment<S, TW, REV>(ct, y + row * ldy, N, P0, wave, i, g, xrow, b, K, tiny);
        }
    }
}

// Ragged (CSR) batches: a persistent grid pulls work items (a run of
// consecutive segments of one signal, longest first) through an atomic
// counter until the queue is drained; every workgroup leaves the loop on its
// first fetch past the end.  Each segment is staged, contracted and stored as
// in the dense kernel, with the item's own taps and length; the tap tables are
// rebuilt only when the item's taps differ from the ones in LDS.  Items are
// claimed one ahead and broadcast through the 16 bytes of LDS at
// `slot_offset`; the next item's first segment loads during the current
// item's last contraction.
template <int S, int TW>
__global__ __launch_bounds__(kThreads) void fir_mfma_ragged_kernel(
    const float *__restrict__ x, const float *__restrict__ b, float *__restrict__ y,
    const MfmaRaggedItem *__restrict__ items, int n_items, unsigned int *__restrict__ counter,
    int split_stride, int tab_offset, int copy_bytes, int slot_offset)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    constexpr int SEG_BLOCKS = kWaves * TW * 32;
    constexpr int PF = 2 * TW + 1;
    // Two item slots, used in turn: the index claimed at the start of item n
    // is read by every lane right after that item's first barrier, and the
    // slot is written again two items later, after the next item's barrier.
    int *s_item = reinterpret_cast<int *>(lds + slot_offset);
    int *mins = reinterpret_cast<int *>(lds + slot_offset + 16);
    int64_t tab_b0 = -1;
    int tab_K = -1;
    f32x4 raw[PF][2];
    // The first segment of item q, into raw (also for the next item, during
    // the current item's last contraction).
    auto prefetch = [&](const MfmaRaggedItem &q, int64_t sg, int tid) {
        const int mh = mfma_m_hi(q.K);
        const int64_t xb = 32 * sg * SEG_BLOCKS - 16 * (int64_t)mh;
#pragma unroll
        for (int j = 0; j < PF; ++j)
            load_unit<false>(x + q.x0, q.len, xb, tid + j * kThreads, SEG_BLOCKS * 4 + 2 * mh,
                             raw[j]);
    };
    if (threadIdx.x == 0) s_item[0] = (int)atomicAdd(counter, 1u);
    __syncthreads();
    int nxt = s_item[0];
    MfmaRaggedItem wn{};
    if (nxt < n_items) {
        wn = items[nxt];
        prefetch(wn, wn.seg0, (int)threadIdx.x);
    }
    for (int parity = 1;; parity ^= 1) {
        const int it = nxt;
        if (it >= n_items) break;  // uniform: read from LDS after a barrier
        const MfmaRaggedItem w = wn;
        // Claim the item after this one now, so its first segment can load
        // during this item's last contraction.
        if (threadIdx.x == 0) s_item[parity] = (int)atomicAdd(counter, 1u);
        __syncthreads();  // every wave is done with the previous item's LDS; slot written
        nxt = s_item[parity];
        const int m_hi = mfma_m_hi(w.K), msteps = m_hi + 2;
        const int units = SEG_BLOCKS * 4 + 2 * m_hi;
        const float *__restrict__ xrow = x + w.x0;
        float *__restrict__ yrow = y + w.x0;
        const int64_t seg_end = (int64_t)w.seg0 + w.nseg;
        if (w.b0 != tab_b0 || w.K != tab_K) {
            int min_b = 0;
            build_tap_table<S, TW>(lds, tab_offset, copy_bytes, b + w.b0, w.K, m_hi, min_b);
            if (S == 3) {
                min_b = wave_min(min_b);
                if ((threadIdx.x & 63) == 0) mins[4 + (threadIdx.x >> 6)] = min_b;
            }
            tab_b0 = w.b0;
            tab_K = w.K;
        }
        for (int64_t sg = w.seg0; sg < seg_end; ++sg) {
            int tid = (int)threadIdx.x;
            asm volatile("" : "+v"(tid));
            const int wave = tid >> 6, lane = tid & 63;
            const int i = lane & 31, g = lane >> 5;
            const int64_t P0 = sg * SEG_BLOCKS;
            const int64_t xb = 32 * P0 - 16 * (int64_t)m_hi;
            if (sg != w.seg0) __syncthreads();  // the previous segment's MFMAs are done
            int min_e = 0;
            stage_segment<S, TW, PF, false>(lds, split_stride, raw, xrow, w.len, xb, units, tid,
                                            min_e);
            if (S == 3) {
                const int m = wave_min(min_e);
                if (lane == 0) mins[wave] = m;
            }
            __syncthreads();
            if (sg + 1 < seg_end) {
                prefetch(w, sg + 1, tid);
            } else if (nxt < n_items) {
                wn = items[nxt];
                prefetch(wn, wn.seg0, tid);
            }
            // A wave whose outputs all lie past the signal's end skips its
            // contraction and epilogue (as in the dense kernel).
            if (32 * (P0 + (int64_t)__builtin_amdgcn_readfirstlane(wave) * 32 * TW) < w.len) {
                Contraction<S, TW> ct;
                contract_segment<S, TW>(ct, lds, split_stride, tab_offset, copy_bytes, msteps,
                                        m_hi, wave, i, g);
                const bool tiny = S == 3 && segment_is_tiny(mins);
                store_segment<S, TW, false>(ct, yrow, w.len, P0, wave, i, g, xrow, b + w.b0,
                                            w.K, tiny);
            }
        }
    }
}

constexpr int kRaggedTW = 2;
static_assert(kMfmaRaggedSegLen == 32 * kWaves * kRaggedTW * 32, "segment length");

template <int S, int TW>
size_t mfma_lds_bytes(int K)
{
    const int m_hi = mfma_m_hi(K);
    const int units = kWaves * TW * 32 * 4 + 2 * m_hi;
    return (size_t)S * unit_bytes(units, TW) + (size_t)2 * S * tab_copy_bytes(m_hi, TW) +
           kMinsBytes;
}

template <int S, int TW, bool REV>
int launch_mfma_dir(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                    float *y, int64_t ldy, int segs_per_wg, hipStream_t stream)
{
    constexpr int SEG_BLOCKS = kWaves * TW * 32;
    const int m_hi = mfma_m_hi(K), msteps = mfma_steps(K);
    const int units = SEG_BLOCKS * 4 + 2 * m_hi;
    const int split_stride = (int)unit_bytes(units, TW);
    const int tab_offset = S * split_stride;
    const int copy_bytes = tab_copy_bytes(m_hi, TW);
    const size_t lds = mfma_lds_bytes<S, TW>(K);
    SCIR_TRY(
        allow_dynamic_lds(reinterpret_cast<const void *>(&fir_mfma_kernel<S, TW, REV>), kMaxLds));
    const int64_t seg_len = 32 * (int64_t)SEG_BLOCKS;
    const int64_t segs = (N + seg_len - 1) / seg_len;
    if (segs > 0x7fffffff)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "row too long (N=%lld)", (long long)N);
    const int64_t rows_per_launch = std::max<int64_t>(1, 0x7fffffff / segs);
    for (int64_t r0 = 0; r0 < B; r0 += rows_per_launch) {
        const int64_t rows = std::min(rows_per_launch, B - r0);
        const int64_t nsegs = rows * segs;
        const int64_t nwg = (nsegs + segs_per_wg - 1) / segs_per_wg;
        hipLaunchKernelGGL((fir_mfma_kernel<S, TW, REV>), dim3((unsigned)nwg), dim3(kThreads),
                           lds, stream, x, ldx, N, y, ldy, b, K, msteps, m_hi, r0, (int)segs,
                           nsegs, segs_per_wg, nwg, split_stride, tab_offset, copy_bytes);
        SCIR_TRY(check_launch("fir_mfma_kernel"));
    }
    return SCIR_HIP_OK;
}

// The direction is a template argument: a run-time flag puts a join after
// every prefetch load, at which the compiler waits for the loads in flight.
template <int S, int TW>
int launch_mfma(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                float *y, int64_t ldy, bool rev, int segs_per_wg, hipStream_t stream)
{
    return rev ? launch_mfma_dir<S, TW, true>(x, ldx, b, K, B, N, y, ldy, segs_per_wg, stream)
               : launch_mfma_dir<S, TW, false>(x, ldx, b, K, B, N, y, ldy, segs_per_wg, stream);
}

template <int S>
int run_mfma(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N, float *y,
             int64_t ldy, bool rev, hipStream_t stream, const DeviceState *st)
{
    const TuneChoice tc = tune_lookup(K, N);
    int tw = S == 3 ? tc.mfma_tiles : tc.mfma_tiles_fp16;
    // Short rows: smaller segments keep the grid full.
    if (N <= 4096) tw = 1;
    // Long filters: fewer tiles per wave when the LDS would overflow.
    if (tw >= 4 && mfma_lds_bytes<S, 4>(K) > (size_t)kMaxLds) tw = 2;
    if (tw >= 2 && mfma_lds_bytes<S, 2>(K) > (size_t)kMaxLds) tw = 1;
    if (mfma_lds_bytes<S, 1>(K) > (size_t)kMaxLds)
        return fail(SCIR_HIP_ERR_UNSUPPORTED, "K=%d needs more LDS than the MFMA path has", K);
    // Segments per workgroup: as tuned, but (unless forced) never so many
    // that fewer than four workgroups per CU remain.
    const int64_t seg_len = 32 * (int64_t)kWaves * tw * 32;
    const int64_t total = B * ((N + seg_len - 1) / seg_len);
    const int64_t cap = tc.mfma_segs_forced
                            ? total
                            : std::max<int64_t>(1, total / (4 * (int64_t)std::max(1, st->cu_count)));
    const int spw = (int)std::max<int64_t>(1, std::min<int64_t>(tc.mfma_segs, cap));
    switch (tw) {
    case 1: return launch_mfma<S, 1>(x, ldx, b, K, B, N, y, ldy, rev, spw, stream);
    case 4: return launch_mfma<S, 4>(x, ldx, b, K, B, N, y, ldy, rev, spw, stream);
    default: return launch_mfma<S, 2>(x, ldx, b, K, B, N, y, ldy, rev, spw, stream);
    }
}

}  // namespace

bool fir_mfma_takes(int precision, int K)
{
    if (K <= 0) return true;
    const size_t need = precision == SCIR_HIP_PREC_BF16X3 ? mfma_lds_bytes<3, 1>(K)
                                                          : mfma_lds_bytes<1, 1>(K);
    return need <= (size_t)kMaxLds;
}

int fir_mfma_dev(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N,
                 float *y, int64_t ldy, int precision, bool rev, hipStream_t stream,
                 DeviceState *st)
{
    if (B == 0 || N == 0) return SCIR_HIP_OK;
    if (K == 0) return fill_zero_2d(y, ldy, B, N, stream);
    if (precision == SCIR_HIP_PREC_BF16X3)
        return run_mfma<3>(x, ldx, b, K, B, N, y, ldy, rev, stream, st);
    if (precision == SCIR_HIP_PREC_FP16)
        return run_mfma<1>(x, ldx, b, K, B, N, y, ldy, rev, stream, st);
    return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "precision %d is not an MFMA mode", precision);
}

namespace {

// Ragged LDS layout: S split arrays, the tap tables, then the item slot.
template <int S>
struct RaggedLayout {
    int split_stride, tab_offset, copy_bytes, slot_offset;
    size_t bytes;
    explicit RaggedLayout(int kmax)
    {
        const int m_hi = mfma_m_hi(kmax);
        const int units = kWaves * kRaggedTW * 32 * 4 + 2 * m_hi;
        split_stride = (int)unit_bytes(units, kRaggedTW);
        tab_offset = S * split_stride;
        copy_bytes = tab_copy_bytes(m_hi, kRaggedTW);
        slot_offset = tab_offset + 2 * S * copy_bytes;
        bytes = (size_t)slot_offset + 16 + kMinsBytes;
    }
};

template <int S>
int ragged_launch(const float *x, const float *b, float *y, const MfmaRaggedItem *items,
                  int n_items, unsigned int *counter, int kmax, size_t lds, int grid,
                  hipStream_t stream)
{
    const RaggedLayout<S> lay(kmax);
    if (lay.bytes != lds || lds > (size_t)kMaxLds)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "ragged MFMA layout mismatch");
    SCIR_TRY(allow_dynamic_lds(
        reinterpret_cast<const void *>(&fir_mfma_ragged_kernel<S, kRaggedTW>), kMaxLds));
    hipLaunchKernelGGL((fir_mfma_ragged_kernel<S, kRaggedTW>), dim3((unsigned)grid),
                       dim3(kThreads), lds, stream, x, b, y, items, n_items, counter,
                       lay.split_stride, lay.tab_offset, lay.copy_bytes, lay.slot_offset);
    return check_launch("fir_mfma_ragged_kernel");
}

}  // namespace

int fir_mfma_ragged_lds(int precision, int kmax, size_t *lds)
{
    const size_t bytes = precision == SCIR_HIP_PREC_BF16X3 ? RaggedLayout<3>(kmax).bytes
                                                           : RaggedLayout<1>(kmax).bytes;
    if (bytes > (size_t)kMaxLds)
        return fail(SCIR_HIP_ERR_UNSUPPORTED,
                    "ragged MFMA batches take taps up to the LDS limit; K=%d needs %zu bytes",
                    kmax, bytes);
    *lds = bytes;
    return SCIR_HIP_OK;
}

int fir_mfma_ragged_occupancy(int precision, size_t lds, int *per_cu)
{
    int n = 0;
    const void *fn = precision == SCIR_HIP_PREC_BF16X3
                         ? reinterpret_cast<const void *>(&fir_mfma_ragged_kernel<3, kRaggedTW>)
                         : reinterpret_cast<const void *>(&fir_mfma_ragged_kernel<1, kRaggedTW>);
    SCIR_TRY(allow_dynamic_lds(fn, kMaxLds));
    SCIR_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kThreads, lds));
    *per_cu = std::max(1, n);
    return SCIR_HIP_OK;
}

int fir_mfma_ragged_launch(const float *x, const float *b, float *y, const MfmaRaggedItem *items,
                           int n_items, unsigned int *counter, int precision, int kmax,
                           size_t lds, int grid, hipStream_t stream)
{
    if (precision == SCIR_HIP_PREC_BF16X3)
        return ragged_launch<3>(x, b, y, items, n_items, counter, kmax, lds, grid, stream);
    if (precision == SCIR_HIP_PREC_FP16)
        return ragged_launch<1>(x, b, y, items, n_items, counter, kmax, lds, grid, stream);
    return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "precision %d is not an MFMA mode", precision);
}

}  // namespace scir
