/*
 * CPU oracle for SciR's batched FIR path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py load this file's shared
 * library: as the checker (error figures against FP64) or as the timed
 * single-thread CPU baseline.  The product path (scir_amd/, include/scir_hip.h)
 * never links or calls it.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math -shared -fPIC
 * (oracle/Makefile).  -ffp-contract=off matters: rustc never contracts a*b+c
 * into an FMA, gcc's default in GNU mode does, and the f32 restatement below
 * must reproduce the reference loop bit for bit.  x86-64 uses SSE scalar
 * arithmetic for float/double, so every operation rounds to its own type.
 *
 * Conventions
 *   "reference order" taps: y[i] = sum_{t=0}^{min(i,K-1)} taps[K-1-t] * x[i-t]
 *       (crates/scir-gpu/src/lib.rs:1134-1152), i.e. lfilter(taps[::-1], [1], x).
 *   "lfilter order" b:      y[n] = sum_{k=0}^{K-1} b[k] * x[n-k], x[<0] = 0,
 *       i.e. scipy.signal.lfilter(b, [1.0], x) with zero initial state.
 *   b[k] == taps[K-1-k].
 *
 * How each function is pinned (tests/test_oracle.py):
 *   oracle_fir_ref_f32   reference KAT (crates/scir-gpu/src/lib.rs:1249-1261),
 *                        bitwise against an independent numpy float32
 *                        emulation, impulse/step closed forms, scipy lfilter
 *                        golden vectors within the K*eps bound.
 *   oracle_fir_ref_f64   scipy lfilter golden vectors (1e-12), brute force.
 *   oracle_lfilter_f64   scipy lfilter golden vectors, == ref_f64 on reversed taps.
 *   oracle_fir_absbound  closed form on |x|, |b| (a sum of absolute values).
 *   oracle_resample_poly_f64  scipy resample_poly golden output for the
 *                        reference's fixture input within the reference test's
 *                        own tolerance, brute-force polyphase sums, impulse
 *                        responses, the reference's output length.
 *   oracle_sosfilt_f64 / oracle_filtfilt_f64  the reference's own IIR fixtures
 *                        (crates/scir-iir-filters/test_data, 1e-12), scipy
 *                        sosfilt on random data, the closed-form impulse
 *                        response of a resonator section, filtfilt as the
 *                        reference fixture script's composition.
 *   oracle_fir_vjp_f64   central finite differences of sum(w*y) (exact for a
 *                        bilinear loss up to rounding) and torch float64
 *                        autograd through conv1d.
 */
#include <stdint.h>
#include <stddef.h>

/* crates/scir-gpu/src/lib.rs:1134-1152 (fir1d_batched_f32).
 * Same loop order as the reference: for each row, for each i, acc starts at
 * 0.0f, then for t_idx = 0.. over xi = i down to start, acc += taps[K-1-t_idx]
 * * x[xi] with a separately rounded multiply and add. */
void oracle_fir_ref_f32(const float *x, int64_t ldx, const float *taps, int64_t K,
                        int64_t B, int64_t N, float *y, int64_t ldy)
{
    for (int64_t bi = 0; bi < B; ++bi) {
        const float *xin = x + bi * ldx;
        float *yout = y + bi * ldy;
        for (int64_t i = 0; i < N; ++i) {
            float acc = 0.0f;
            int64_t start = (i + 1 > K) ? (i + 1 - K) : 0;
            int64_t t_idx = 0;
            for (int64_t xi = i; xi >= start; --xi, ++t_idx) {
                float prod = taps[K - 1 - t_idx] * xin[xi];
                acc = acc + prod;
            }
            yout[i] = acc;
        }
    }
}

/* crates/scir-gpu/src/lib.rs:1166-1184 (fir1d_batched_f64): same formula and
 * order in double precision. */
void oracle_fir_ref_f64(const double *x, int64_t ldx, const double *taps, int64_t K,
                        int64_t B, int64_t N, double *y, int64_t ldy)
{
    for (int64_t bi = 0; bi < B; ++bi) {
        const double *xin = x + bi * ldx;
        double *yout = y + bi * ldy;
        for (int64_t i = 0; i < N; ++i) {
            double acc = 0.0;
            int64_t start = (i + 1 > K) ? (i + 1 - K) : 0;
            int64_t t_idx = 0;
            for (int64_t xi = i; xi >= start; --xi, ++t_idx)
                acc += taps[K - 1 - t_idx] * xin[xi];
            yout[i] = acc;
        }
    }
}

/* FP64 oracle of the lfilter convention on float32 inputs:
 * y[n] = sum_k b[k] * x[n-k] evaluated in double from the float values. */
void oracle_lfilter_f64(const float *x, int64_t ldx, const float *b, int64_t K,
                        int64_t B, int64_t N, double *y, int64_t ldy)
{
    for (int64_t bi = 0; bi < B; ++bi) {
        const float *xin = x + bi * ldx;
        double *yout = y + bi * ldy;
        for (int64_t n = 0; n < N; ++n) {
            double acc = 0.0;
            int64_t kmax = (n < K - 1) ? n : K - 1;
            for (int64_t k = 0; k <= kmax; ++k)
                acc += (double)b[k] * (double)xin[n - k];
            yout[n] = acc;
        }
    }
}

/* sum_k |b[k]| * |x[n-k]| in double: the scale of the error bound
 * |y32 - y_exact| <= (K+1) * 2^-23 * sum_k |b_k||x_{n-k}|. */
void oracle_fir_absbound(const float *x, int64_t ldx, const float *b, int64_t K,
                         int64_t B, int64_t N, double *s, int64_t lds)
{
    for (int64_t bi = 0; bi < B; ++bi) {
        const float *xin = x + bi * ldx;
        double *sout = s + bi * lds;
        for (int64_t n = 0; n < N; ++n) {
            double acc = 0.0;
            int64_t kmax = (n < K - 1) ? n : K - 1;
            for (int64_t k = 0; k <= kmax; ++k) {
                double bv = (double)b[k], xv = (double)xin[n - k];
                acc += (bv < 0 ? -bv : bv) * (xv < 0 ? -xv : xv);
            }
            sout[n] = acc;
        }
    }
}

/* Reference gradient of L = sum_{r,n} g[r,n] * y[r,n], y = lfilter(b,[1],x)
 * per row, in double:
 *   dx[r,m] = sum_k b[k] * g[r,m+k]            (m+k < N)
 *   db[k]   = sum_{r,n} g[r,n] * x[r,n-k]      (n-k >= 0)
 * db must hold K doubles and is overwritten. */
void oracle_fir_vjp_f64(const float *g, int64_t ldg, const float *x, int64_t ldx,
                        const float *b, int64_t K, int64_t B, int64_t N,
                        double *dx, int64_t lddx, double *db)
{
    for (int64_t k = 0; k < K; ++k) db[k] = 0.0;
    for (int64_t r = 0; r < B; ++r) {
        const float *gr = g + r * ldg;
        const float *xr = x + r * ldx;
        double *dxr = dx + r * lddx;
        for (int64_t m = 0; m < N; ++m) {
            double acc = 0.0;
            for (int64_t k = 0; k < K && m + k < N; ++k)
                acc += (double)b[k] * (double)gr[m + k];
            dxr[m] = acc;
        }
        for (int64_t k = 0; k < K; ++k) {
            double acc = 0.0;
            for (int64_t n = k; n < N; ++n)
                acc += (double)gr[n] * (double)xr[n - k];
            db[k] += acc;
        }
    }
}

/* Ragged (CSR) form of oracle_lfilter_f64: signal s occupies
 * x[offs[s] .. offs[s+1]) and uses taps b[boffs[s] .. boffs[s+1]). */
void oracle_lfilter_ragged_f64(const float *x, const int64_t *offs, const float *b,
                               const int64_t *boffs, int64_t S, double *y)
{
    for (int64_t s = 0; s < S; ++s) {
        int64_t n0 = offs[s], len = offs[s + 1] - offs[s];
        int64_t k0 = boffs[s], K = boffs[s + 1] - boffs[s];
        oracle_lfilter_f64(x + n0, len, b + k0, K, 1, len, y + n0, len);
    }
}

/* crates/scir-signal/src/lib.rs:293-302 (convolve) and :313-362
 * (resample_poly), generalised from the fixed up = 2, down = 3 and 31-tap H
 * to any up, down >= 1 and taps h[K], K >= 1, with the same steps and
 * operation order: x is upsampled by inserting up-1 zeros after every
 * sample; the full linear convolution with h is accumulated as
 * conv[i + j] += xu[i] * h[j] for i ascending, then j ascending (separately
 * rounded multiply and add); outputs are conv[off + down * m] for
 * off = (K - 1) / 2 while off + down * m < len(conv) - off.
 * Returns the number of outputs written to y (y must hold
 * oracle_resample_len(n, up, down, K) doubles); `work` must hold
 * n * up + K - 1 doubles. */
int64_t oracle_resample_len(int64_t n, int64_t up, int64_t down, int64_t K)
{
    const int64_t len = n * up + K - 1, off = (K - 1) / 2, end = len - off;
    if (end <= off) return 0;
    return (end - off + down - 1) / down;
}

int64_t oracle_resample_poly_f64(const double *x, int64_t n, int64_t up, int64_t down,
                                 const double *h, int64_t K, double *work, double *y)
{
    const int64_t nu = n * up, len = nu + K - 1, off = (K - 1) / 2, end = len - off;
    for (int64_t i = 0; i < len; ++i) work[i] = 0.0;
    for (int64_t i = 0; i < nu; ++i) {
        const double xi = (i % up == 0) ? x[i / up] : 0.0;
        for (int64_t j = 0; j < K; ++j) {
            const double p = xi * h[j];
            work[i + j] = work[i + j] + p;
        }
    }
    int64_t m = 0;
    for (int64_t idx = off; idx < end; idx += down) y[m++] = work[idx];
    return m;
}

/* crates/scir-iir-filters/src/filter.rs:47-62 (DirectForm2Transposed::filter)
 * as driven by crates/scir-signal/src/lib.rs:260-267 (sosfilt): zero state
 * s1[i] = s2[i] = 0, then for each sample v = x[t] and each section i,
 *   p = v;  v = b0*p + s1[i];  s1[i] = b1*p - a1*v + s2[i];  s2[i] = b2*p - a2*v
 * with the expressions evaluated left to right, every operation rounded.
 * sos holds S rows {b0, b1, b2, a0, a1, a2}; a0 is not read.  `state` must
 * hold 2*S doubles. */
void oracle_sosfilt_f64(const double *x, int64_t n, const double *sos, int64_t S, double *state,
                        double *y)
{
    double *s1 = state, *s2 = state + S;
    for (int64_t i = 0; i < S; ++i) s1[i] = s2[i] = 0.0;
    for (int64_t t = 0; t < n; ++t) {
        double output = x[t];
        for (int64_t i = 0; i < S; ++i) {
            const double *c = sos + 6 * i;
            const double prev_output = output;
            output = c[0] * prev_output + s1[i];
            s1[i] = c[1] * prev_output - c[4] * output + s2[i];
            s2[i] = c[2] * prev_output - c[5] * output;
        }
        y[t] = output;
    }
}

/* crates/scir-signal/src/lib.rs:278-291 (filtfilt): tmp = the cascade over
 * x from zero state; a fresh cascade over tmp in reverse order gives out;
 * the result is out reversed.  `tmp` must hold n doubles, `state` 2*S. */
void oracle_filtfilt_f64(const double *x, int64_t n, const double *sos, int64_t S, double *state,
                         double *tmp, double *y)
{
    oracle_sosfilt_f64(x, n, sos, S, state, tmp);
    for (int64_t t = 0; t < n / 2; ++t) {  /* tmp.iter().rev() */
        const double a = tmp[t];
        tmp[t] = tmp[n - 1 - t];
        tmp[n - 1 - t] = a;
    }
    oracle_sosfilt_f64(tmp, n, sos, S, state, y);
    for (int64_t t = 0; t < n / 2; ++t) {  /* out.reverse() */
        const double a = y[t];
        y[t] = y[n - 1 - t];
        y[n - 1 - t] = a;
    }
}
