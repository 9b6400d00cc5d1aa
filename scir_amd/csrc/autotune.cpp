// Tuning table and the autotuner.
//
// The table maps (K bucket, N bucket) to launch shapes:
//   direct_r / direct_loader / direct_seg_tiles   FP32 direct kernel (loader 0: one
//                                                 tile per workgroup, 1: register
//                                                 prefetch, 2: LDS DMA)
//   mfma_tiles / mfma_tiles_fp16                  MFMA kernel tiles per wave
//   mfma_segs                                     MFMA row segments per workgroup
// plus one global AUTO cut-over K (F32 direct below, BF16x3 MFMA from there)
// and the K from which the FP32 FFT path beats BF16x3 MFMA (auto_fft_min_k;
// AUTO takes the FFT path from there only when the tuning key auto.fft is 1,
// since its error is normwise rather than elementwise).
//
// scir_hip_autotune() fills it by timing every candidate with HIP events on
// synthetic rows (2^26 samples per bucket, N = the bucket's representative
// row length) on the current device, keeps the fastest, and writes
//   {"arch": "gfx950", "auto_mfma_min_k": 160, "auto_fft_min_k": 512,
//    "entries": [{"k_lo": 1, "k_hi": 16, "n_lo": 0, "n_hi": 8192,
//                 "direct_r": 8, "direct_loader": 1, "direct_seg_tiles": 2,
//                 "mfma_tiles": 2, "mfma_tiles_fp16": 4, "mfma_segs": 16}, ...]}
// scir_hip_autotune_load() reads that format back (a table written for
// another arch is ignored).  Overrides set with scir_hip_tune_set() win over
// the table.
#include "common.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

namespace scir {
namespace {

struct Entry {
    int k_lo, k_hi;
    int64_t n_lo, n_hi;
    TuneChoice c;
};

std::mutex g_tune_mu;
std::vector<Entry> g_table;
int g_auto_min_k = -1;  // from the table; -1 = built-in default
int g_auto_fft_min_k = -1;  // from the table; -1 = built-in default
int g_over_auto_fft = -1, g_over_auto_fft_min_k = -1;
int g_over_vjp_db_mfma = -1;
// Overrides set through scir_hip_tune_set (-1 = unset).
int g_over_direct_r = -1, g_over_direct_loader = -1, g_over_direct_seg = -1,
    g_over_mfma_tiles = -1, g_over_auto_min_k = -1, g_over_mfma_segs = -1,
    g_over_dma_bufs = -1;

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
    if (g_over_direct_r >= 0) c.direct_r = g_over_direct_r;
    if (g_over_direct_loader >= 0) c.direct_loader = g_over_direct_loader;
    if (g_over_direct_seg >= 0) c.direct_seg_tiles = g_over_direct_seg;
    if (g_over_dma_bufs >= 0) c.direct_dma_bufs = g_over_dma_bufs;
    if (g_over_mfma_tiles >= 0) c.mfma_tiles = c.mfma_tiles_fp16 = g_over_mfma_tiles;
    if (g_over_auto_min_k >= 0) c.auto_mfma_min_k = g_over_auto_min_k;
    if (g_over_mfma_segs >= 0) {
        c.mfma_segs = g_over_mfma_segs;
        c.mfma_segs_forced = true;
    }
    return c;
}

namespace {

// Minimal reader for the flat format written below: finds `"key": <int>`
// after position `from` and before `to`.
bool find_int(const std::string &s, size_t from, size_t to, const char *key, int64_t *out)
{
    std::string pat = std::string("\"") + key + "\"";
    size_t p = s.find(pat, from);
    if (p == std::string::npos || p >= to) return false;
    p = s.find(':', p + pat.size());
    if (p == std::string::npos || p >= to) return false;
    ++p;
    while (p < to && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n')) ++p;
    char *end = nullptr;
    long long v = std::strtoll(s.c_str() + p, &end, 10);
    if (end == s.c_str() + p) return false;
    *out = v;
    return true;
}

bool find_str(const std::string &s, const char *key, std::string *out)
{
    std::string pat = std::string("\"") + key + "\"";
    size_t p = s.find(pat);
    if (p == std::string::npos) return false;
    p = s.find(':', p + pat.size());
    if (p == std::string::npos) return false;
    p = s.find('"', p);
    if (p == std::string::npos) return false;
    size_t q = s.find('"', p + 1);
    if (q == std::string::npos) return false;
    *out = s.substr(p + 1, q - p - 1);
    return true;
}

std::string table_json(const std::string &arch)
{
    std::ostringstream o;
    o << "{\"arch\": \"" << arch << "\", \"auto_mfma_min_k\": "
      << (g_auto_min_k >= 0 ? g_auto_min_k : TuneChoice{}.auto_mfma_min_k)
      << ", \"auto_fft_min_k\": "
      << (g_auto_fft_min_k >= 0 ? g_auto_fft_min_k : TuneChoice{}.auto_fft_min_k)
      << ",\n \"entries\": [";
    for (size_t i = 0; i < g_table.size(); ++i) {
        const Entry &e = g_table[i];
        o << (i ? ",\n  " : "\n  ") << "{\"k_lo\": " << e.k_lo << ", \"k_hi\": " << e.k_hi
          << ", \"n_lo\": " << e.n_lo << ", \"n_hi\": " << e.n_hi
          << ", \"direct_r\": " << e.c.direct_r << ", \"direct_loader\": " << e.c.direct_loader
          << ", \"direct_seg_tiles\": " << e.c.direct_seg_tiles
          << ", \"mfma_tiles\": " << e.c.mfma_tiles
          << ", \"mfma_tiles_fp16\": " << e.c.mfma_tiles_fp16
          << ", \"mfma_segs\": " << e.c.mfma_segs << "}";
    }
    o << "]}\n";
    return o.str();
}

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
    clear_error();
    if (!key) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "key is null");
    std::lock_guard<std::mutex> g(g_tune_mu);
    int v = value < 0 ? -1 : value;
    if (!std::strcmp(key, "direct.r")) {
        if (v >= 0 && v != 8 && v != 16)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.r must be 8 or 16");
        g_over_direct_r = v;
    } else if (!std::strcmp(key, "direct.loader")) {
        if (v > 2) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.loader must be 0, 1 or 2");
        g_over_direct_loader = v;
    } else if (!std::strcmp(key, "direct.dma_bufs")) {
        if (v >= 0 && (v < 2 || v > 4))
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.dma_bufs must be 2, 3 or 4");
        g_over_dma_bufs = v;
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
    } else if (!std::strcmp(key, "mfma.segs")) {
        if (v == 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "mfma.segs must be >= 1");
        g_over_mfma_segs = v;
    } else {
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "unknown tuning key '%s'", key);
    }
    return SCIR_HIP_OK;
}

extern "C" int scir_hip_autotune_load(const char *cache_path)
{
    clear_error();
    if (!cache_path) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "cache_path is null");
    std::ifstream f(cache_path);
    if (!f) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "cannot open %s", cache_path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    std::string arch;
    if (!find_str(s, "arch", &arch))
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "%s: no \"arch\"", cache_path);
    DeviceState *st = nullptr;
    if (current_device_state(&st) == SCIR_HIP_OK && st->arch.compare(0, arch.size(), arch) != 0)
        return SCIR_HIP_OK;  // a table for another device: ignored
    clear_error();
    std::vector<Entry> table;
    size_t p = s.find("\"entries\"");
    while (p != std::string::npos) {
        size_t a = s.find('{', p + 1);
        if (a == std::string::npos) break;
        size_t b = s.find('}', a);
        if (b == std::string::npos) break;
        Entry e;
        int64_t v[9];
        const char *keys[9] = {"k_lo", "k_hi", "n_lo", "n_hi", "direct_r", "direct_loader",
                               "direct_seg_tiles", "mfma_tiles", "mfma_tiles_fp16"};
        bool ok = true;
        for (int i = 0; i < 9; ++i) ok = ok && find_int(s, a, b, keys[i], &v[i]);
        if (!ok) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "%s: malformed entry", cache_path);
        e.k_lo = (int)v[0];
        e.k_hi = (int)v[1];
        e.n_lo = v[2];
        e.n_hi = v[3];
        e.c.direct_r = (v[4] == 16) ? 16 : 8;
        e.c.direct_loader = (v[5] >= 0 && v[5] <= 2) ? (int)v[5] : 1;
        e.c.direct_seg_tiles = (int)std::max<int64_t>(1, v[6]);
        e.c.mfma_tiles = (v[7] == 1 || v[7] == 4) ? (int)v[7] : 2;
        e.c.mfma_tiles_fp16 = (v[8] == 1 || v[8] == 2) ? (int)v[8] : 4;
        int64_t segs = 0;  // optional (tables written before the key existed)
        if (find_int(s, a, b, "mfma_segs", &segs) && segs >= 1) e.c.mfma_segs = (int)segs;
        table.push_back(e);
        p = b;
    }
    int64_t amk = -1, afk = -1;
    find_int(s, 0, s.size(), "auto_mfma_min_k", &amk);
    find_int(s, 0, s.size(), "auto_fft_min_k", &afk);  // optional
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_table = table;
    g_auto_min_k = (int)amk;
    g_auto_fft_min_k = (int)afk;
    return SCIR_HIP_OK;
}

extern "C" int scir_hip_autotune(const char *cache_path, int32_t budget_ms)
{
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
    if (hipMalloc(&x, total * sizeof(float)) != hipSuccess ||
        hipMalloc(&y, total * sizeof(float)) != hipSuccess ||
        hipMalloc(&b, 4096 * sizeof(float)) != hipSuccess) {
        cleanup();
        return fail(SCIR_HIP_ERR_OUT_OF_MEMORY, "autotune buffers");
    }
    int status = scir_hip_synth_normal_f32(x, total, 1, total, 0x5C12AAAAull, 0.0f, 1.0f, s);
    if (status == SCIR_HIP_OK) {
        std::vector<float> hb(4096);
        for (int i = 0; i < 4096; ++i) hb[i] = 1.0f / (float)(i + 1);
        if (hipMemcpy(b, hb.data(), hb.size() * sizeof(float), hipMemcpyHostToDevice) !=
            hipSuccess)
            status = fail(SCIR_HIP_ERR_BACKEND, "autotune tap upload failed");
    }
    // Search with the table cleared; overrides drive each candidate.
    int saved[6];
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        saved[0] = g_over_direct_r;
        saved[1] = g_over_direct_loader;
        saved[2] = g_over_direct_seg;
        saved[3] = g_over_mfma_tiles;
        saved[4] = g_over_auto_min_k;
        saved[5] = g_over_mfma_segs;
        g_table.clear();
        g_auto_min_k = -1;
    }
    auto set = [&](int r, int loader, int seg, int tiles, int msegs) {
        std::lock_guard<std::mutex> g(g_tune_mu);
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
            const int rs[2] = {8, 16}, loaders[3] = {0, 1, 2}, segs[4] = {1, 2, 4, 8};
            for (int r : rs)
                for (int ld : loaders)
                    for (int sg : segs) {
                        if ((ld == 0 && sg != 1) || status != SCIR_HIP_OK) continue;
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
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_over_direct_r = saved[0];
        g_over_direct_loader = saved[1];
        g_over_direct_seg = saved[2];
        g_over_mfma_tiles = saved[3];
        g_over_auto_min_k = saved[4];
        g_over_mfma_segs = saved[5];
        if (status == SCIR_HIP_OK) {
            g_table = found;
            g_auto_min_k = cut;
            g_auto_fft_min_k = fft_cut;
        }
    }
    cleanup();
    if (status != SCIR_HIP_OK) return status;
    if (cache_path) {
        std::string js;
        {
            std::lock_guard<std::mutex> g(g_tune_mu);
            js = table_json(st->arch.substr(0, 6));
        }
        std::ofstream f(cache_path);
        if (!f) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "cannot write %s", cache_path);
        f << js;
    }
    return SCIR_HIP_OK;
}
