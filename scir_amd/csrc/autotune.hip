// Tuning table and the autotuner.
//
// The table maps (K bucket, N bucket) to launch shapes:
//   direct_r / direct_loader / direct_seg_tiles   FP32 direct kernel (loader 0: one
//                                                 tile per workgroup, 1: register
//                                                 prefetch, 2: LDS DMA)
//   mfma_tiles / mfma_tiles_fp16                  MFMA kernel tiles per wave
//   mfma_segs                                     MFMA row segments per workgroup
// plus one global AUTO cut-over K (F32 direct below, BF16x3 MFMA from there),
// the K from which the FP32 FFT path beats BF16x3 MFMA (auto_fft_min_k; AUTO
// takes the FFT path from there only when the tuning key auto.fft is 1, since
// its error is normwise rather than elementwise), and the launch shapes of
// ragged (CSR) batches (persistent workgroups per CU and work per item, for
// the FP32 direct form and for MFMA).
//
// scir_hip_autotune() fills it by timing every candidate with HIP events on
// synthetic data (2^26 samples per problem: rows of the bucket's
// representative length, or a ragged batch with log-uniform lengths in
// [256, 2^20] and K uniform in [15, 1023]) on the current device, keeps the
// fastest and writes the table in the format of tune_table.cpp;
// scir_hip_autotune_load() reads it back (a table written for another arch is
// ignored).  Overrides set with scir_hip_tune_set() win over the table.
#include "common.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

namespace scir {
namespace {

using Entry = TuneEntry;

std::mutex g_tune_mu;
std::vector<Entry> g_table;
int g_auto_min_k = -1;  // from the table; -1 = built-in default
int g_auto_fft_min_k = -1;  // from the table; -1 = built-in default
RaggedTune g_ragged;        // from the table (defaults until one is loaded)
int g_over_auto_fft = -1, g_over_auto_fft_min_k = -1;
int g_over_vjp_db_mfma = -1;
int g_over_resample_fixed = -1;
// Overrides set through scir_hip_tune_set (-1 = unset).
int g_over_direct_r = -1, g_over_direct_loader = -1, g_over_direct_seg = -1,
    g_over_mfma_tiles = -1, g_over_auto_min_k = -1, g_over_mfma_segs = -1,
    g_over_dma_bufs = -1;
int g_over_rg_direct_wg = -1, g_over_rg_direct_kmacs = -1, g_over_rg_mfma_segs = -1,
    g_over_rg_mfma_wg = -1;
int g_over_host_chunk_kib = -1;
int g_over_host_copy_threads = -1;
int g_over_direct_mfma_strip = -1;

// Buckets searched by the autotuner.
const int kKBuckets[][2] = {{1, 16}, {17, 64}, {65, 256}, {257, 1017}, {1018, 4096}};
const int64_t kNBuckets[][2] = {{0, 8192}, {8193, (int64_t)1 << 62}};

}  // namespace

TuneChoice tune_lookup(int K, int64_t N)
{
    TuneChoice c;
    // Built-in defaults (measured on MI355X): long filters share each tap
    // fragment across four tiles per wave.
    if (K > 1017) c.mfma_tiles = 4;
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (const Entry &e : g_table)
        if (K >= e.k_lo && K <= e.k_hi && N >= e.n_lo && N <= e.n_hi) {
            c = e.c;
            break;
        }
    if (g_auto_min_k >= 0) c.auto_mfma_min_k = g_auto_min_k;
    if (g_auto_fft_min_k >= 0) c.auto_fft_min_k = g_auto_fft_min_k;
    if (g_over_auto_fft >= 0) c.auto_fft = g_over_auto_fft != 0;
    if (g_over_auto_fft_min_k >= 0) c.auto_fft_min_k = g_over_auto_fft_min_k;
    if (g_over_vjp_db_mfma >= 0) c.vjp_db_mfma = g_over_vjp_db_mfma != 0;
    if (g_over_resample_fixed >= 0) c.resample_fixed = g_over_resample_fixed != 0;
    if (g_over_direct_r >= 0) c.direct_r = g_over_direct_r;
    if (g_over_direct_loader >= 0) c.direct_loader = g_over_direct_loader;
    if (g_over_direct_seg >= 0) c.direct_seg_tiles = g_over_direct_seg;
    if (g_over_dma_bufs >= 0) c.direct_dma_bufs = g_over_dma_bufs;
    if (g_over_direct_mfma_strip >= 0) c.direct_mfma_strip = g_over_direct_mfma_strip;
    if (g_over_mfma_tiles >= 0) c.mfma_tiles = c.mfma_tiles_fp16 = g_over_mfma_tiles;
    if (g_over_auto_min_k >= 0) c.auto_mfma_min_k = g_over_auto_min_k;
    if (g_over_mfma_segs >= 0) {
        c.mfma_segs = g_over_mfma_segs;
        c.mfma_segs_forced = true;
    }
    return c;
}

size_t host_chunk_bytes()
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    return g_over_host_chunk_kib > 0 ? (size_t)g_over_host_chunk_kib << 10 : (size_t)16 << 20;
}

int host_copy_threads()
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    return g_over_host_copy_threads > 0 ? g_over_host_copy_threads : 4;
}

RaggedTune ragged_tune_lookup()
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    RaggedTune r = g_ragged;
    if (g_over_rg_direct_wg >= 0) r.direct_wg_per_cu = g_over_rg_direct_wg;
    if (g_over_rg_direct_kmacs >= 0) r.direct_min_item_kmacs = g_over_rg_direct_kmacs;
    if (g_over_rg_mfma_segs >= 0) r.mfma_segs_per_item = g_over_rg_mfma_segs;
    if (g_over_rg_mfma_wg >= 0) r.mfma_wg_per_cu = g_over_rg_mfma_wg;
    if (g_over_auto_min_k >= 0) r.auto_mfma_min_k = g_over_auto_min_k;
    return r;
}

namespace {

// Median of `reps` HIP-event timings of `fn` (which enqueues on `s`), after
// one untimed call.
template <class F>
int time_ms(F fn, hipStream_t s, int reps, float *ms_out)
{
    hipEvent_t a, b;
    SCIR_HIP_TRY(hipEventCreate(&a));
    SCIR_HIP_TRY(hipEventCreate(&b));
    std::vector<float> t;
    int st = fn();
    for (int r = 0; st == SCIR_HIP_OK && r < reps; ++r) {
        (void)hipEventRecord(a, s);
        st = fn();
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        t.push_back(ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st != SCIR_HIP_OK) return st;
    std::sort(t.begin(), t.end());
    *ms_out = t[t.size() / 2];
    return SCIR_HIP_OK;
}

}  // namespace
}  // namespace scir

using namespace scir;

extern "C" int scir_hip_tune_set(const char *key, int32_t value)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!key) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "key is null");
        std::lock_guard<std::mutex> g(g_tune_mu);
        int v = value < 0 ? -1 : value;
        if (!std::strcmp(key, "direct.r")) {
            if (v >= 0 && v != 8 && v != 16)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.r must be 8 or 16");
            g_over_direct_r = v;
        } else if (!std::strcmp(key, "direct.loader")) {
            if (v > 3) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.loader must be 0 .. 3");
            g_over_direct_loader = v;
        } else if (!std::strcmp(key, "direct.dma_bufs")) {
            if (v >= 0 && (v < 2 || v > 4))
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.dma_bufs must be 2, 3 or 4");
            g_over_dma_bufs = v;
        } else if (!std::strcmp(key, "direct.mfma_strip")) {
            if (v == 0 || v > 4096)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.mfma_strip: 1 .. 4096");
            g_over_direct_mfma_strip = v;
        } else if (!std::strcmp(key, "direct.seg_tiles")) {
            if (v == 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.seg_tiles must be >= 1");
            g_over_direct_seg = v;
        } else if (!std::strcmp(key, "mfma.tiles")) {
            if (v >= 0 && v != 1 && v != 2 && v != 4)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "mfma.tiles must be 1, 2 or 4");
            g_over_mfma_tiles = v;
        } else if (!std::strcmp(key, "auto.mfma_min_k")) {
            g_over_auto_min_k = v;
        } else if (!std::strcmp(key, "auto.fft")) {
            if (v > 1) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "auto.fft must be 0 or 1");
            g_over_auto_fft = v;
        } else if (!std::strcmp(key, "auto.fft_min_k")) {
            g_over_auto_fft_min_k = v;
        } else if (!std::strcmp(key, "vjp.db_mfma")) {
            if (v > 1) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "vjp.db_mfma must be 0 or 1");
            g_over_vjp_db_mfma = v;
        } else if (!std::strcmp(key, "resample.fixed")) {
            if (v > 1) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "resample.fixed must be 0 or 1");
            g_over_resample_fixed = v;
        } else if (!std::strcmp(key, "mfma.segs")) {
            if (v == 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "mfma.segs must be >= 1");
            g_over_mfma_segs = v;
        } else if (!std::strcmp(key, "ragged.direct_wg_per_cu")) {
            if (v == 0 || v > 32)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "ragged.direct_wg_per_cu: 1 .. 32");
            g_over_rg_direct_wg = v;
        } else if (!std::strcmp(key, "ragged.direct_min_item_kmacs")) {
            if (v == 0)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "ragged.direct_min_item_kmacs >= 1");
            g_over_rg_direct_kmacs = v;
        } else if (!std::strcmp(key, "ragged.mfma_segs_per_item")) {
            if (v == 0 || v > 1024)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "ragged.mfma_segs_per_item: 1 .. 1024");
            g_over_rg_mfma_segs = v;
        } else if (!std::strcmp(key, "host.chunk_kib")) {
            if (v == 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "host.chunk_kib must be >= 1");
            g_over_host_chunk_kib = v;
        } else if (!std::strcmp(key, "host.copy_threads")) {
            if (v == 0 || v > 64)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "host.copy_threads: 1 .. 64");
            g_over_host_copy_threads = v;
        } else if (!std::strcmp(key, "ragged.mfma_wg_per_cu")) {
            if (v > 8)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT,
                            "ragged.mfma_wg_per_cu: 0 (occupancy) .. 8");
            g_over_rg_mfma_wg = v;
        } else {
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "unknown tuning key '%s'", key);
        }
        return SCIR_HIP_OK;
    });
}

extern "C" int scir_hip_autotune_load(const char *cache_path)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!cache_path) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "cache_path is null");
        std::ifstream f(cache_path);
        if (!f) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "cannot open %s", cache_path);
        std::stringstream ss;
        ss << f.rdbuf();
        TuneTable t;
        SCIR_TRY(parse_tune_table(ss.str(), cache_path, &t));
        DeviceState *st = nullptr;
        if (current_device_state(&st) == SCIR_HIP_OK &&
            st->arch.compare(0, t.arch.size(), t.arch) != 0)
            return SCIR_HIP_OK;  // a table for another device: ignored
        clear_error();
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_table = t.entries;
        g_auto_min_k = (int)t.auto_mfma_min_k;
        g_auto_fft_min_k = (int)t.auto_fft_min_k;
        g_ragged = t.has_ragged ? t.ragged : RaggedTune{};
        return SCIR_HIP_OK;
    });
}

extern "C" int scir_hip_autotune(const char *cache_path, int32_t budget_ms)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        const auto t_start = std::chrono::steady_clock::now();
        const double budget = budget_ms > 0 ? budget_ms : 120000.0;
        auto over_budget = [&]() {
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                             t_start).count() > budget;
        };
        const int64_t total = (int64_t)1 << 26;  // samples per timing problem
        const int64_t long_n = (int64_t)1 << 18;
        hipStream_t s = nullptr;
        SCIR_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        float *x = nullptr, *y = nullptr, *b = nullptr;
        auto cleanup = [&]() {
            (void)hipStreamSynchronize(s);
            if (x) (void)hipFree(x);
            if (y) (void)hipFree(y);
            if (b) (void)hipFree(b);
            (void)hipStreamDestroy(s);
        };
        const int64_t n_taps = (int64_t)1 << 20;  // dense taps, and the ragged batch's
        int status = device_alloc(reinterpret_cast<void **>(&x), total * sizeof(float),
                                  "autotune input");
        if (status == SCIR_HIP_OK)
            status = device_alloc(reinterpret_cast<void **>(&y), total * sizeof(float),
                                  "autotune output");
        if (status == SCIR_HIP_OK)
            status = device_alloc(reinterpret_cast<void **>(&b), n_taps * sizeof(float),
                                  "autotune taps");
        if (status != SCIR_HIP_OK) {
            const std::string msg = scir_hip_last_error();
            cleanup();
            return fail(status, "%s", msg.c_str());
        }
        status = scir_hip_synth_normal_f32(x, total, 1, total, 0x5C12AAAAull, 0.0f, 1.0f, s);
        if (status == SCIR_HIP_OK) {
            // 1/(i+1) in the first 4096 taps (dense problems), N(0,1)/32 after
            std::vector<float> hb(4096);
            for (int i = 0; i < 4096; ++i) hb[i] = 1.0f / (float)(i + 1);
            status = scir_hip_synth_normal_f32(b + 4096, n_taps - 4096, 1, n_taps - 4096,
                                               0x5C12BBBBull, 0.0f, 1.0f, s);
            if (status == SCIR_HIP_OK &&
                hipMemcpy(b, hb.data(), hb.size() * sizeof(float), hipMemcpyHostToDevice) !=
                    hipSuccess)
                status = fail(SCIR_HIP_ERR_BACKEND, "autotune tap upload failed");
        }
        // Search with the table cleared; overrides drive each candidate.
        int saved[7];
        {
            std::lock_guard<std::mutex> g(g_tune_mu);
            saved[6] = g_over_direct_mfma_strip;
            saved[0] = g_over_direct_r;
            saved[1] = g_over_direct_loader;
            saved[2] = g_over_direct_seg;
            saved[3] = g_over_mfma_tiles;
            saved[4] = g_over_auto_min_k;
            saved[5] = g_over_mfma_segs;
            g_table.clear();
            g_auto_min_k = -1;
        }
        auto set = [&](int r, int loader, int seg, int tiles, int msegs, int strip = -1) {
            std::lock_guard<std::mutex> g(g_tune_mu);
            g_over_direct_mfma_strip = strip;
            g_over_direct_r = r;
            g_over_direct_loader = loader;
            g_over_direct_seg = seg;
            g_over_mfma_tiles = tiles;
            g_over_mfma_segs = msegs;
            g_over_auto_min_k = -1;
        };
        std::vector<Entry> found;
        for (const auto &kb : kKBuckets) {
            for (const auto &nb : kNBuckets) {
                if (status != SCIR_HIP_OK || over_budget()) break;
                const int K = (kb[0] + kb[1]) / 2;
                const int64_t N = nb[1] <= 8192 ? 4096 : long_n;
                const int64_t B = total / N;
                Entry e;
                e.k_lo = kb[0];
                e.k_hi = kb[1];
                e.n_lo = nb[0];
                e.n_hi = nb[1];
                float best = 1e30f;
                // Loader 3 (FP32 MFMA) has no outputs-per-lane choice; its
                // four candidates are strips of 8, 16, 32 and 64 tiles.
                const int rs[2] = {8, 16}, loaders[4] = {0, 1, 2, 3}, segs[4] = {1, 2, 4, 8};
                for (int r : rs)
                    for (int ld : loaders)
                        for (int sg : segs) {
                            if ((ld == 0 && sg != 1) || (ld == 3 && r != 8) ||
                                status != SCIR_HIP_OK)
                                continue;
                            if (ld == 3)
                                set(r, ld, -1, -1, -1, 8 * sg);
                            else
                                set(r, ld, sg, -1, -1);
                            float ms = 0;
                            status = time_ms([&]() {
                                return scir_hip_fir_f32_dev(x, N, b, K, B, N, y, N,
                                                            SCIR_HIP_PREC_F32, s);
                            }, s, 3, &ms);
                            if (status == SCIR_HIP_OK && ms < best) {
                                best = ms;
                                e.c.direct_r = r;
                                e.c.direct_loader = ld;
                                if (ld == 3)
                                    e.c.direct_mfma_strip = 8 * sg;
                                else
                                    e.c.direct_seg_tiles = sg;
                            }
                        }
                const int tiles[3] = {1, 2, 4};
                const int precs[2] = {SCIR_HIP_PREC_BF16X3, SCIR_HIP_PREC_FP16};
                for (int pi = 0; pi < 2; ++pi) {
                    float bt = 1e30f;
                    for (int tw : tiles) {
                        if (status != SCIR_HIP_OK) break;
                        set(-1, -1, -1, tw, -1);
                        float ms = 0;
                        status = time_ms([&]() {
                            return scir_hip_fir_f32_dev(x, N, b, K, B, N, y, N, precs[pi], s);
                        }, s, 3, &ms);
                        if (status == SCIR_HIP_OK && ms < bt) {
                            bt = ms;
                            (pi == 0 ? e.c.mfma_tiles : e.c.mfma_tiles_fp16) = tw;
                        }
                    }
                }
                // Row segments per MFMA workgroup, with the BF16x3 tile choice.
                {
                    float bt = 1e30f;
                    const int msegs[3] = {1, 4, 8};
                    for (int sg : msegs) {
                        if (status != SCIR_HIP_OK) break;
                        set(-1, -1, -1, e.c.mfma_tiles, sg);
                        float ms = 0;
                        status = time_ms([&]() {
                            return scir_hip_fir_f32_dev(x, N, b, K, B, N, y, N,
                                                        SCIR_HIP_PREC_BF16X3, s);
                        }, s, 3, &ms);
                        if (status == SCIR_HIP_OK && ms < bt) {
                            bt = ms;
                            e.c.mfma_segs = sg;
                        }
                    }
                }
                found.push_back(e);
            }
        }
        // AUTO cut-over on long rows: the smallest K at which BF16x3 beats F32.
        int cut = TuneChoice{}.auto_mfma_min_k;
        if (status == SCIR_HIP_OK && !over_budget()) {
            {
                std::lock_guard<std::mutex> g(g_tune_mu);
                g_table = found;
            }
            set(-1, -1, -1, -1, -1);
            const int ks[] = {48, 64, 96, 128, 160, 192, 256, 384, 512};
            cut = -1;
            for (int K : ks) {
                if (status != SCIR_HIP_OK) break;
                const int64_t N = long_n, B = total / N;
                float m32 = 0, m3 = 0;
                status = time_ms([&]() {
                    return scir_hip_fir_f32_dev(x, N, b, K, B, N, y, N, SCIR_HIP_PREC_F32, s);
                }, s, 3, &m32);
                if (status == SCIR_HIP_OK)
                    status = time_ms([&]() {
                        return scir_hip_fir_f32_dev(x, N, b, K, B, N, y, N, SCIR_HIP_PREC_BF16X3, s);
                    }, s, 3, &m3);
                if (status == SCIR_HIP_OK && m3 < m32) {
                    cut = K;
                    break;
                }
            }
            if (cut < 0) cut = 1 << 20;
        }
        // FFT cut-over: the smallest K at which FP32 FFT beats BF16x3 MFMA.
        int fft_cut = TuneChoice{}.auto_fft_min_k;
        if (status == SCIR_HIP_OK && !over_budget()) {
            const int ks[] = {128, 192, 256, 384, 512, 768, 1024, 2048};
            fft_cut = -1;
            for (int K : ks) {
                if (status != SCIR_HIP_OK) break;
                const int64_t N = long_n, B = total / N;
                float mf = 0, m3 = 0;
                status = time_ms([&]() {
                    return scir_hip_fir_f32_dev(x, N, b, K, B, N, y, N, SCIR_HIP_PREC_FFT32, s);
                }, s, 3, &mf);
                if (status == SCIR_HIP_OK)
                    status = time_ms([&]() {
                        return scir_hip_fir_f32_dev(x, N, b, K, B, N, y, N, SCIR_HIP_PREC_BF16X3, s);
                    }, s, 3, &m3);
                if (status == SCIR_HIP_OK && mf < m3) {
                    fft_cut = K;
                    break;
                }
            }
            if (fft_cut < 0) fft_cut = 1 << 20;
        }
        // Ragged launch shapes: a synthetic CSR batch with log-uniform lengths
        // in [256, 2^20] and K uniform in [15, 1023] (about 2^26 samples),
        // every candidate timed on one persistent launch.
        RaggedTune rbest;
        if (status == SCIR_HIP_OK && !over_budget()) {
            std::vector<int64_t> offs(1, 0), boffs(1, 0);
            uint64_t lcg = 0x5C12CCCCull;
            auto uni = [&]() {
                lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
                return (double)(lcg >> 11) * (1.0 / 9007199254740992.0);
            };
            while (offs.back() < total) {
                int64_t len = (int64_t)std::llround(std::exp(std::log(256.0) +
                                                             uni() * std::log(4096.0)));
                len = std::min<int64_t>(len, total - offs.back());
                const int64_t K = 15 + (int64_t)(uni() * 1009.0);
                if (boffs.back() + K > n_taps) break;
                offs.push_back(offs.back() + len);
                boffs.push_back(boffs.back() + K);
            }
            const int64_t S = (int64_t)offs.size() - 1;
            {
                std::lock_guard<std::mutex> g(g_tune_mu);
                g_ragged = RaggedTune{};
            }
            auto time_plan = [&](int prec, float *ms) {
                scir_hip_ragged_plan plan = nullptr;
                int st2 = scir_hip_ragged_plan_create(offs.data(), boffs.data(), S, prec, &plan);
                if (st2 == SCIR_HIP_OK)
                    st2 = time_ms([&]() { return scir_hip_ragged_plan_execute(plan, x, b, y, s); },
                                  s, 3, ms);
                (void)hipStreamSynchronize(s);
                const std::string msg = st2 == SCIR_HIP_OK ? "" : scir_hip_last_error();
                scir_hip_ragged_plan_destroy(plan);
                return st2 == SCIR_HIP_OK ? st2 : fail(st2, "%s", msg.c_str());
            };
            auto set_ragged = [&](int dwg, int dk, int msegs, int mwg) {
                std::lock_guard<std::mutex> g(g_tune_mu);
                g_over_rg_direct_wg = dwg;
                g_over_rg_direct_kmacs = dk;
                g_over_rg_mfma_segs = msegs;
                g_over_rg_mfma_wg = mwg;
            };
            float best = 1e30f;
            const int dwgs[3] = {4, 8, 16}, dks[3] = {128, 512, 2048};
            for (int dwg : dwgs)
                for (int dk : dks) {
                    if (status != SCIR_HIP_OK || over_budget()) break;
                    set_ragged(dwg, dk, -1, -1);
                    float ms = 0;
                    status = time_plan(SCIR_HIP_PREC_F32, &ms);
                    if (status == SCIR_HIP_OK && ms < best) {
                        best = ms;
                        rbest.direct_wg_per_cu = dwg;
                        rbest.direct_min_item_kmacs = dk;
                    }
                }
            best = 1e30f;
            const int msegs[5] = {2, 4, 8, 16, 32}, mwgs[2] = {1, 0};
            for (int sg : msegs)
                for (int mw : mwgs) {
                    if (status != SCIR_HIP_OK || over_budget()) break;
                    set_ragged(-1, -1, sg, mw);
                    float ms = 0;
                    status = time_plan(SCIR_HIP_PREC_BF16X3, &ms);
                    if (status == SCIR_HIP_OK && ms < best) {
                        best = ms;
                        rbest.mfma_segs_per_item = sg;
                        rbest.mfma_wg_per_cu = mw;
                    }
                }
            set_ragged(-1, -1, -1, -1);
        }
        {
            std::lock_guard<std::mutex> g(g_tune_mu);
            g_over_direct_r = saved[0];
            g_over_direct_loader = saved[1];
            g_over_direct_seg = saved[2];
            g_over_mfma_tiles = saved[3];
            g_over_auto_min_k = saved[4];
            g_over_mfma_segs = saved[5];
            g_over_direct_mfma_strip = saved[6];
            if (status == SCIR_HIP_OK) {
                g_table = found;
                g_auto_min_k = cut;
                g_auto_fft_min_k = fft_cut;
                g_ragged = rbest;
            }
        }
        cleanup();
        if (status != SCIR_HIP_OK) return status;
        if (cache_path) {
            TuneTable t;
            {
                std::lock_guard<std::mutex> g(g_tune_mu);
                t.arch = st->arch.substr(0, 6);
                t.entries = g_table;
                t.auto_mfma_min_k = g_auto_min_k;
                t.auto_fft_min_k = g_auto_fft_min_k;
                t.has_ragged = true;
                t.ragged = g_ragged;
            }
            std::ofstream f(cache_path);
            if (!f) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "cannot write %s", cache_path);
            f << tune_table_json(t);
        }
        return SCIR_HIP_OK;
    });
}
