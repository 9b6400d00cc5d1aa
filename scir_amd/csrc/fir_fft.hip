// FP32 FFT overlap-save FIR for long filters (SCIR_HIP_PREC_FFT32).
//
// The direct and Toeplitz-MFMA kernels cost O(K) per output; a segment of
// NF samples convolved through the FFT costs O(log NF) per output.  Each
// workgroup filters two segments of NF - K + 1 outputs at once: their inputs
// (the K - 1 samples before each plus the segment) are the real and
// imaginary parts of one complex NF-point signal z; since the filter is
// real, IFFT(FFT(z) * H) carries the first segment's circular convolution in
// its real part and the second's in its imaginary part, and entries K-1 ..
// NF-1 of a circular convolution are the linear convolution's outputs.
// H = FFT(b zero-padded to NF) / NF comes from fir_fft_spectrum_kernel,
// launched first on the same stream.
//
// The transforms run in LDS: radix-16/32 Stockham passes (natural order in
// and out), each thread loading R points with stride NF/R, multiplying by
// the pass twiddles, taking an R-point DFT in registers (radix-2 stages
// with compile-time twiddles) and storing with stride NS.  LDS indices are
// padded by one complex every 32 (i + i/32), which keeps every pass's loads
// and stores free of bank conflicts.  Pass twiddles are products of two
// small tables (64 fine and NF/64 coarse angles) computed on the host in
// double and rounded once.
//
// Error: the FFT's rounding is normwise, about log2(NF) * 2^-24 * |x|_2 |b|_1
// per segment, not relative to each output; the direct and MFMA kernels keep
// the elementwise (K+1) 2^-24 sum|b||x| bound.
//
// Non-finite values: an Inf or NaN anywhere in a segment's window (or an
// overflow inside the transform) makes every output of the segment NaN.  The
// epilogue recomputes each non-finite output from the reference loop's terms
// (direct_output with the whole filter), so outputs are non-finite exactly
// where the reference loop's are.  With a partitioned filter the check is
// made by the block that writes the output last.
#include "direct_tiles.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace scir {
namespace {

// ---- compile-time trigonometry for the in-register DFTs ----------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ct_sin_taylor(double x)
{
    double term = x, sum = x;
    for (int n = 1; n < 30; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
// cos / sin of 2*pi*m/R
constexpr double ct_cos_frac(int m, int R)
{
    double a = 2.0 * kPi * (double)m / (double)R;
    if (a > kPi) a -= 2.0 * kPi;
    return ct_sin_taylor(a + kPi / 2.0 > kPi ? a + kPi / 2.0 - 2.0 * kPi : a + kPi / 2.0);
}
constexpr double ct_sin_frac(int m, int R)
{
    double a = 2.0 * kPi * (double)m / (double)R;
    if (a > kPi) a -= 2.0 * kPi;
    return ct_sin_taylor(a);
}

template <int R>
struct SmallTw {
    float c[R], s[R];
    constexpr SmallTw() : c(), s()
    {
        for (int m = 0; m < R; ++m) {
            c[m] = (float)ct_cos_frac(m, R);
            s[m] = (float)ct_sin_frac(m, R);
        }
    }
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// v <- DFT_R(v) (forward: exp(-2 pi i f n / R); INV: exp(+...)), natural
// order in and out.  Radix-2 decimation in frequency, then the bit-reversal
// permutation (a renaming of registers).
template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&v)[R])
{
    constexpr SmallTw<R> tw{};
#pragma unroll
    for (int len = R; len >= 2; len >>= 1) {
        const int half = len / 2;
#pragma unroll
        for (int s0 = 0; s0 < R; s0 += len) {
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const float2 a = v[s0 + j], b = v[s0 + j + half];
                v[s0 + j] = cadd(a, b);
                const float2 t = csub(a, b);
                const int m = j * (R / len);  // W_len^j = W_R^m
                if (m == 0) {
                    v[s0 + j + half] = t;
                } else if (4 * m == R) {  // -i (forward) or +i (inverse)
                    v[s0 + j + half] = INV ? make_float2(-t.y, t.x) : make_float2(t.y, -t.x);
                } else {
                    const float2 w = make_float2(tw.c[m], INV ? tw.s[m] : -tw.s[m]);
                    v[s0 + j + half] = cmul(t, w);
                }
            }
        }
    }
    float2 o[R];
#pragma unroll
    for (int f = 0; f < R; ++f) {
        int r = 0;
#pragma unroll
        for (int b = 1, q = f; b < R; b <<= 1, q >>= 1) r = (r << 1) | (q & 1);
        o[f] = v[r];
    }
#pragma unroll
    for (int f = 0; f < R; ++f) v[f] = o[f];
}

__device__ __forceinline__ int pad(int i) { return i + (i >> 5); }

// pad(base + r * NS) - pad(base) for the store pattern of a pass with radix R
// after NS points (base = (j / NS) * NS * R + j % NS): a compile-time offset,
// so the LDS instructions carry it and no per-point address arithmetic is
// left.  NS a multiple of 32: r * NS adds r * NS / 32 pads.  NS == 16 (R
// even): base % 32 < 16, so 16r adds r / 2 pads.  NS == 1, R == 16: base is
// 16j and r < 16 never crosses a 32-point line.
template <int NS, int R>
__device__ __forceinline__ constexpr int store_off(int r)
{
    static_assert(NS % 32 == 0 || (NS == 16 && R % 2 == 0) || (NS == 1 && R == 16),
                  "pass shape without a constant store offset");
    return NS % 32 == 0 ? r * (NS + NS / 32) : NS == 16 ? 16 * r + (r >> 1) : r;
}

// Twiddle tables (cos, sin of positive angles; the forward transform uses
// the conjugates), in this order:
//   tq[i], i < 64:        2*pi*i/NF        (fine)
//   tq[64 + a], a < NF/64: 2*pi*64a/NF      (coarse; fine * coarse gives any
//                                            angle 2*pi*m/NF within ~1.5 ulp)
//   tq[64 + NF/64 + m]:   2*pi*m/(R0*R1)   (the second pass's angles, direct)
constexpr int kFine = 64;
template <int NF>
constexpr int tw_fine_coarse() { return kFine + NF / kFine; }

template <int NF, bool INV>
__device__ __forceinline__ float2 twiddle(const float2 *tq, int m)
{
    const float2 p = cmul(tq[m & (kFine - 1)], tq[kFine + m / kFine]);
    return make_float2(p.x, INV ? p.y : -p.y);
}

// One Stockham pass: radix R, NS = product of the earlier radices.
// MULH multiplies the loaded points by the spectrum H first (the pointwise
// product, fused into the inverse transform's first pass).
template <int NF, int T, int R, int NS, bool INV, bool MULH, bool DIRECT = false>
__device__ __forceinline__ void fft_pass(float2 *d, const float2 *tq,
                                         const float2 *__restrict__ H, int tid)
{
    constexpr int B = NF / R;
    constexpr int PER = (B + T - 1) / T;
    static_assert(B % 32 == 0, "loads use pad(j + r*B) == pad(j) + r*(B + B/32)");
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = tid + p * T;
        if (B % T == 0 || j < B) {
            const int jp = pad(j);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                v[p][r] = d[jp + r * (B + B / 32)];
                if (MULH) {
                    // H[j + r*B] through a buffer resource: the per-point offset
                    // r*B*8 is a scalar constant, so no 64-bit address arithmetic
                    const __amdgpu_buffer_rsrc_t hr = __builtin_amdgcn_make_buffer_rsrc(
                        const_cast<float2 *>(H), 0, NF * (int)sizeof(float2), 0x00020000);
                    const float2 h = __builtin_bit_cast(
                        float2, __builtin_amdgcn_raw_buffer_load_b64(hr, 8 * j, 8 * r * B, 0));
                    v[p][r] = cmul(v[p][r], h);
                }
            }
            if (NS > 1) {
                const int k = j % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    float2 w;
                    if (DIRECT) {  // tq points at this pass's table of NS*R angles
                        const float2 t = tq[k * r];
                        w = make_float2(t.x, INV ? t.y : -t.y);
                    } else {
                        w = twiddle<NF, INV>(tq, k * r * (NF / (NS * R)));
                    }
                    v[p][r] = cmul(v[p][r], w);
                }
            }
            dft<R, INV>(v[p]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = tid + p * T;
        if (B % T == 0 || j < B) {
            const int k = j % NS;
            const int bp = pad((j / NS) * NS * R + k);
#pragma unroll
            for (int r = 0; r < R; ++r) d[bp + store_off<NS, R>(r)] = v[p][r];
        }
    }
    __syncthreads();
}

template <int LOGN> struct Plan;
template <> struct Plan<12> { static constexpr int T = 256, R0 = 16, R1 = 16, R2 = 16; };
template <> struct Plan<13> { static constexpr int T = 256, R0 = 16, R1 = 16, R2 = 32; };
template <> struct Plan<14> { static constexpr int T = 512, R0 = 16, R1 = 32, R2 = 32; };

template <int LOGN, bool INV, bool MULH>
__device__ __forceinline__ void fft(float2 *d, const float2 *tq, const float2 *__restrict__ H,
                                    int tid)
{
    using P = Plan<LOGN>;
    constexpr int NF = 1 << LOGN;
    fft_pass<NF, P::T, P::R0, 1, INV, MULH>(d, tq, H, tid);
    fft_pass<NF, P::T, P::R1, P::R0, INV, false, true>(d, tq + tw_fine_coarse<NF>(), H, tid);
    fft_pass<NF, P::T, P::R2, P::R0 * P::R1, INV, false>(d, tq, H, tid);
}

template <int LOGN>
constexpr int tw_entries()
{
    return tw_fine_coarse<1 << LOGN>() + Plan<LOGN>::R0 * Plan<LOGN>::R1;
}

template <int LOGN>
constexpr size_t lds_bytes()
{
    constexpr int NF = 1 << LOGN;
    return (size_t)(NF + NF / 32 + tw_entries<LOGN>()) * sizeof(float2);
}

// H[i] = FFT(b zero-padded)[i] / NF, one workgroup.
template <int LOGN>
__global__ __launch_bounds__(Plan<LOGN>::T) void fir_fft_spectrum_kernel(
    const float *__restrict__ b, int K, const float2 *__restrict__ tq_g, float2 *__restrict__ H)
{
    constexpr int NF = 1 << LOGN, T = Plan<LOGN>::T;
    extern __shared__ float2 lds_f2[];
    float2 *d = lds_f2;
    float2 *tq = lds_f2 + NF + NF / 32;
    const int tid = threadIdx.x;
    for (int i = tid; i < tw_entries<LOGN>(); i += T) tq[i] = tq_g[i];
    for (int i = tid; i < NF; i += T) d[pad(i)] = make_float2(i < K ? b[i] : 0.0f, 0.0f);
    __syncthreads();
    fft<LOGN, false, false>(d, tq, nullptr, tid);
    const float inv = 1.0f / (float)NF;  // exact: NF is a power of two
    for (int i = tid; i < NF; i += T) {
        const float2 v = d[pad(i)];
        H[i] = make_float2(v.x * inv, v.y * inv);
    }
}

// Segment s of the (row, segment) list: row s / spr, outputs [o, o + V),
// V = NF - K + 1, o = (s % spr) * V.  REV reads and writes every row back to
// front (the anti-causal correlation of the dL/dx adjoint).  With `accumulate`
// the outputs are added to y instead of stored (the later tap blocks of a
// partitioned filter).  The block holds taps bfull[j0, j0 + K) of a filter of
// Kfull taps: its output t is output t + j0 of the whole filter, its input t
// is the row's input t, and it writes outputs t < final_len last (the later
// blocks start at t = kFftBlockTaps).
template <int LOGN, bool REV>
__global__ __launch_bounds__(Plan<LOGN>::T) void fir_fft_kernel(
    const float *__restrict__ x, int64_t ldx, float *__restrict__ y, int64_t ldy, int64_t N,
    int K, const float2 *__restrict__ H, const float2 *__restrict__ tq_g, int64_t spr,
    int64_t nseg, int accumulate, const float *__restrict__ bfull, int Kfull, int64_t j0,
    int64_t final_len)
{
    constexpr int NF = 1 << LOGN, T = Plan<LOGN>::T;
    extern __shared__ float2 lds_f2[];
    float2 *d = lds_f2;
    float2 *tq = lds_f2 + NF + NF / 32;
    const int tid = threadIdx.x;
    const int V = NF - K + 1;
    const int64_t s0 = 2 * (int64_t)blockIdx.x, s1 = s0 + 1;
    const bool has1 = s1 < nseg;
    const int64_t ra = s0 / spr, oa = (s0 - ra * spr) * V;
    const int64_t rb = has1 ? s1 / spr : ra, ob = has1 ? (s1 - rb * spr) * V : 0;
    const float *xa = x + ra * ldx, *xb = x + rb * ldx;
    for (int i = tid; i < tw_entries<LOGN>(); i += T) tq[i] = tq_g[i];
    const int64_t wa = oa - (K - 1), wb = ob - (K - 1);  // first input of each window
    if (has1 && wa >= 0 && wa + NF <= N && wb >= 0 && wb + NF <= N) {
        // both windows inside their rows: no bounds checks
        const float *pa = REV ? xa + (N - 1 - wa) : xa + wa;
        const float *pb = REV ? xb + (N - 1 - wb) : xb + wb;
        // pad(tid + m*T) == pad(tid) + m*(T + T/32): T is a multiple of 32
        static_assert(T % 32 == 0, "constant LDS offsets need T % 32 == 0");
        const int tp = pad(tid);
#pragma unroll 4
        for (int m = 0; m < NF / T; ++m) {
            const int i = tid + m * T;
            d[tp + m * (T + T / 32)] = make_float2(pa[REV ? -i : i], pb[REV ? -i : i]);
        }
    } else {
        for (int i = tid; i < NF; i += T) {
            const int64_t ta = wa + i, tb = wb + i;
            float va = 0.0f, vb = 0.0f;
            if (ta >= 0 && ta < N) va = xa[REV ? N - 1 - ta : ta];
            if (has1 && tb >= 0 && tb < N) vb = xb[REV ? N - 1 - tb : tb];
            d[pad(i)] = make_float2(va, vb);
        }
    }
    __syncthreads();
    fft<LOGN, false, false>(d, tq, nullptr, tid);
    fft<LOGN, true, true>(d, tq, H, tid);
    float *ya = y + ra * ldy, *yb = y + rb * ldy;
    // The rows of the whole filter (for recomputing non-finite outputs): in
    // REV mode the block's rows start j0 samples into the physical row.
    const float *fa = REV ? xa - j0 : xa, *fb = REV ? xb - j0 : xb;
    const int64_t Nf = N + j0;
    auto finish = [&](float v, const float *frow, int64_t t) {
        if (!__builtin_isfinite(v) && t < final_len)
            v = direct_output<false>(frow, Nf, REV, bfull, Kfull, t + j0);
        return v;
    };
    if (has1 && oa + V <= N && ob + V <= N) {
        float *qa = REV ? ya + (N - 1 - oa) : ya + oa;
        float *qb = REV ? yb + (N - 1 - ob) : yb + ob;
        // Outputs are stored as computed; a flag notes whether any came out
        // non-finite (one class test each, no branch in the loop), and only
        // then are the thread's outputs read back and fixed.
        bool bad = false;
        if (accumulate) {
#pragma unroll 4
            for (int v = tid; v < V; v += T) {
                const float2 o = d[pad(K - 1 + v)];
                float &pa = qa[REV ? -v : v], &pb = qb[REV ? -v : v];
                const float a = pa + o.x, b = pb + o.y;
                pa = a;
                pb = b;
                bad |= !__builtin_isfinite(a) | !__builtin_isfinite(b);
            }
        } else {
#pragma unroll 4
            for (int v = tid; v < V; v += T) {
                const float2 o = d[pad(K - 1 + v)];
                qa[REV ? -v : v] = o.x;
                qb[REV ? -v : v] = o.y;
                bad |= !__builtin_isfinite(o.x) | !__builtin_isfinite(o.y);
            }
        }
        if (bad) {
            for (int v = tid; v < V; v += T) {
                float &pa = qa[REV ? -v : v], &pb = qb[REV ? -v : v];
                pa = finish(pa, fa, oa + v);
                pb = finish(pb, fb, ob + v);
            }
        }
    } else {
        for (int v = tid; v < V; v += T) {
            const float2 o = d[pad(K - 1 + v)];
            const int64_t ta = oa + v, tb = ob + v;
            if (ta < N) {
                float &q = ya[REV ? N - 1 - ta : ta];
                q = finish(accumulate ? q + o.x : o.x, fa, ta);
            }
            if (has1 && tb < N) {
                float &q = yb[REV ? N - 1 - tb : tb];
                q = finish(accumulate ? q + o.y : o.y, fb, tb);
            }
        }
    }
}

// Twiddle tables, one per size, built on first use per device (kept in the
// device state, freed by scir_hip_shutdown).
int twiddle_table(DeviceState *st, int logn, hipStream_t s, const float2 **out)
{
    const int NF = 1 << logn;
    DeviceBuffer &buf = st->fft_twiddles[logn - 12];
    if (!buf.ptr) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (s != nullptr && hipStreamIsCapturing(s, &cs) == hipSuccess &&
            cs != hipStreamCaptureStatusNone)
            return fail(SCIR_HIP_ERR_CAPTURE,
                        "the FFT twiddle table is built on first use; run once outside capture");
        const int r01 = 16 * (logn == 14 ? 32 : 16);  // Plan<logn>::R0 * R1
        std::vector<float2> h(kFine + NF / kFine + r01);
        for (int i = 0; i < kFine; ++i) {
            const double a = 2.0 * kPi * (double)i / (double)NF;
            h[i] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        for (int i = 0; i < NF / kFine; ++i) {
            const double a = 2.0 * kPi * (double)(i * kFine) / (double)NF;
            h[kFine + i] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        for (int m = 0; m < r01; ++m) {
            const double a = 2.0 * kPi * (double)m / (double)r01;
            h[kFine + NF / kFine + m] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        void *p = nullptr;
        SCIR_TRY(device_alloc(&p, h.size() * sizeof(float2), "FFT twiddle table"));
        const hipError_t e = hipMemcpy(p, h.data(), h.size() * sizeof(float2),
                                       hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(p);
            return fail(SCIR_HIP_ERR_BACKEND, "FFT twiddle upload failed: %s",
                        hipGetErrorString(e));
        }
        buf.ptr = p;
        buf.bytes = h.size() * sizeof(float2);
    }
    *out = static_cast<const float2 *>(buf.ptr);
    return SCIR_HIP_OK;
}

template <int LOGN>
int launch_fft(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N, float *y,
               int64_t ldy, bool rev, bool accumulate, const float *bfull, int Kfull, int64_t j0,
               int64_t final_len, hipStream_t s, DeviceState *st)
{
    constexpr int NF = 1 << LOGN, T = Plan<LOGN>::T;
    const float2 *tq = nullptr;
    SCIR_TRY(twiddle_table(st, LOGN, s, &tq));
    StreamScratch &sc = scratch_for(st, s);
    SCIR_TRY(ensure_scratch(st, sc.fft_spectrum, (size_t)NF * sizeof(float2), s));
    float2 *H = static_cast<float2 *>(sc.fft_spectrum.ptr);
    constexpr size_t lds = lds_bytes<LOGN>();
    SCIR_TRY(allow_dynamic_lds(reinterpret_cast<const void *>(&fir_fft_kernel<LOGN, false>),
                               (int)lds));
    SCIR_TRY(allow_dynamic_lds(reinterpret_cast<const void *>(&fir_fft_kernel<LOGN, true>),
                               (int)lds));
    SCIR_TRY(allow_dynamic_lds(reinterpret_cast<const void *>(&fir_fft_spectrum_kernel<LOGN>),
                               (int)lds));
    hipLaunchKernelGGL((fir_fft_spectrum_kernel<LOGN>), dim3(1), dim3(T), lds, s, b, K, tq, H);
    SCIR_TRY(check_launch("fir_fft_spectrum_kernel"));
    const int64_t V = NF - K + 1;
    const int64_t spr = (N + V - 1) / V;
    const int64_t nseg = B * spr;
    const int64_t blocks = (nseg + 1) / 2;
    if (blocks > 0x7fffffff) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "batch too large");
    const int acc = accumulate ? 1 : 0;
    if (rev)
        hipLaunchKernelGGL((fir_fft_kernel<LOGN, true>), dim3((unsigned)blocks), dim3(T), lds, s, x,
                           ldx, y, ldy, N, K, H, tq, spr, nseg, acc, bfull, Kfull, j0, final_len);
    else
        hipLaunchKernelGGL((fir_fft_kernel<LOGN, false>), dim3((unsigned)blocks), dim3(T), lds, s,
                           x, ldx, y, ldy, N, K, H, tq, spr, nseg, acc, bfull, Kfull, j0,
                           final_len);
    return check_launch("fir_fft_kernel");
}

// Taps per block of the FFT path: the largest transform (2^14) keeps at
// least half of each segment as outputs.  Longer filters are cut into blocks
// of kFftBlockTaps taps (uniformly partitioned overlap-save): block j, taps
// b[jP, jP + P), filters the input delayed by jP samples, and its outputs are
// added to those of the blocks before it.
constexpr int kFftBlockTaps = 8192;

// FFT size for K taps: the smallest of 2^12, 2^13, 2^14 with NF >= 4K
// (outputs per segment >= 3/4 NF), else 2^14.
int fft_log_size(int K)
{
    if (4 * K <= 4096) return 12;
    if (4 * K <= 8192) return 13;
    return 14;
}

int fft_block(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N, float *y,
              int64_t ldy, bool rev, bool accumulate, const float *bfull, int Kfull, int64_t j0,
              int64_t final_len, hipStream_t stream, DeviceState *st)
{
#define SCIR_FFT_BLOCK(L)                                                                      \
    launch_fft<L>(x, ldx, b, K, B, N, y, ldy, rev, accumulate, bfull, Kfull, j0, final_len, \
                  stream, st)
    switch (fft_log_size(K)) {
    case 12: return SCIR_FFT_BLOCK(12);
    case 13: return SCIR_FFT_BLOCK(13);
    default: return SCIR_FFT_BLOCK(14);
    }
#undef SCIR_FFT_BLOCK
}

}  // namespace

int fir_fft_dev(const float *x, int64_t ldx, const float *b, int K, int64_t B, int64_t N, float *y,
                int64_t ldy, bool rev, hipStream_t stream, DeviceState *st)
{
    if (K == 0)  // no taps: zeros, as every path gives
        return fir_direct_dev(x, ldx, b, K, B, N, y, ldy, false, rev, stream, st);
    if (K < 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative tap count");
    if (x == y)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "the FFT path does not filter in place");
    if (B == 0 || N == 0) return SCIR_HIP_OK;
    for (int64_t j0 = 0; j0 < K && j0 < N; j0 += kFftBlockTaps) {
        const int Kj = (int)std::min<int64_t>(kFftBlockTaps, K - j0);
        // Forward: y[j0 + i] += sum_k b[j0 + k] x[i - k] for i < N - j0.  Back
        // to front the same sums read x from sample j0 on and write y from 0.
        const float *xs = rev ? x + j0 : x;
        float *ys = rev ? y : y + j0;
        // Block j writes outputs j0 .. j0 + kFftBlockTaps - 1 last (all the
        // rest when it is the final block).
        const bool last = j0 + kFftBlockTaps >= K || j0 + kFftBlockTaps >= N;
        const int64_t final_len = last ? N - j0 : (int64_t)kFftBlockTaps;
        SCIR_TRY(fft_block(xs, ldx, b + j0, Kj, B, N - j0, ys, ldy, rev, j0 > 0, b, K, j0,
                           final_len, stream, st));
    }
    return SCIR_HIP_OK;
}

}  // namespace scir
