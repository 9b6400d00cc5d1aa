// Host code of libscir_hip.so under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by tests/test_sanitize_cpu.py with g++,
// no device code): the CPU arms (FIR, sosfilt / filtfilt, resample_poly),
// the tuning-table reader and writer, the ragged work-item builder, the
// exception guard of the C ABI and allocation-failure injection, on random
// and edge-case inputs.  Any out-of-bounds access, leak or undefined
// behaviour aborts the run; value checks print FAIL lines.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../scir_amd/csrc/host_common.hpp"

namespace {

int failures = 0;
#define CHECK(c)                                                               \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);           \
            ++failures;                                                        \
        }                                                                      \
    } while (0)

bool same_bits(float a, float b)
{
    return (std::isnan(a) && std::isnan(b)) || std::memcmp(&a, &b, sizeof a) == 0;
}

// crates/scir-gpu/src/lib.rs:1141-1148 with b[t] == taps[K-1-t].
float ref_output(const float *xr, const float *b, int K, int64_t i)
{
    float acc = 0.0f;
    for (int64_t t = 0; t < K && t <= i; ++t) {
        volatile float p = b[t] * xr[i - t];  // no contraction
        acc = acc + p;
    }
    return acc;
}

void test_cpu_fir(std::mt19937_64 &g)
{
    std::uniform_int_distribution<int> dB(1, 4), dN(0, 300), dK(0, 40), dpad(0, 5), dthr(1, 4);
    std::normal_distribution<float> nd;
    for (int trial = 0; trial < 300; ++trial) {
        const int B = dB(g), N = dN(g), K = dK(g);
        const int64_t ldx = N + dpad(g), ldy = N + dpad(g);
        std::vector<float> x(B * ldx), b(std::max(K, 1)), y(B * ldy, 12345.0f);
        for (float &v : x) v = nd(g);
        for (float &v : b) v = nd(g);
        if (trial % 5 == 0 && N > 0) x[g() % x.size()] = std::numeric_limits<float>::infinity();
        if (trial % 7 == 0 && K > 0) b[g() % K] = std::numeric_limits<float>::quiet_NaN();
        CHECK(scir_hip_fir_f32_cpu(x.data(), ldx, b.data(), K, B, N, y.data(), ldy, dthr(g)) ==
              SCIR_HIP_OK);
        for (int r = 0; r < B; ++r) {
            for (int i = 0; i < N; ++i)
                CHECK(same_bits(y[r * ldy + i], ref_output(&x[r * ldx], b.data(), K, i)));
            for (int64_t i = N; i < ldy; ++i) CHECK(y[r * ldy + i] == 12345.0f);  // pad untouched
        }
    }
    float v = 0.0f;
    CHECK(scir_hip_fir_f32_cpu(&v, 0, &v, 1, 1, 1, &v, 1, 1) == SCIR_HIP_ERR_INVALID_ARGUMENT);
    CHECK(std::strstr(scir_hip_last_error(), "stride") != nullptr);
    CHECK(scir_hip_fir_f32_cpu(nullptr, 1, &v, 1, 1, 1, &v, 1, 1) ==
          SCIR_HIP_ERR_INVALID_ARGUMENT);
    CHECK(scir_hip_fir_f32_cpu(&v, 1, &v, -1, 1, 1, &v, 1, 1) == SCIR_HIP_ERR_INVALID_ARGUMENT);
}

void test_sos_and_resample(std::mt19937_64 &g)
{
    std::uniform_int_distribution<int> dR(1, 4), dN(0, 200), dS(0, SCIR_HIP_SOS_MAX_SECTIONS);
    std::normal_distribution<double> nd;
    for (int trial = 0; trial < 100; ++trial) {
        const int R = dR(g), N = dN(g), S = dS(g);
        std::vector<double> x(R * N + 1), y(R * N + 1), sos(6 * std::max(S, 1));
        for (double &v : x) v = nd(g);
        for (int s = 0; s < S; ++s) {
            double *c = &sos[6 * s];
            c[0] = nd(g) * 0.5; c[1] = nd(g) * 0.5; c[2] = nd(g) * 0.5;
            c[3] = 1.0; c[4] = -0.5 + 0.2 * nd(g); c[5] = 0.1 * nd(g);
        }
        CHECK(scir_hip_sosfilt_f64_cpu(x.data(), N, R, N, sos.data(), S, y.data(), N, 3) ==
              SCIR_HIP_OK);
        CHECK(scir_hip_filtfilt_f64_cpu(x.data(), N, R, N, sos.data(), S, y.data(), N, 2) ==
              SCIR_HIP_OK);
        // in place
        CHECK(scir_hip_sosfilt_f64_cpu(y.data(), N, R, N, sos.data(), S, y.data(), N, 1) ==
              SCIR_HIP_OK);
    }
    double d = 0;
    CHECK(scir_hip_sosfilt_f64_cpu(&d, 1, 1, 1, &d, SCIR_HIP_SOS_MAX_SECTIONS + 1, &d, 1, 1) ==
          SCIR_HIP_ERR_INVALID_ARGUMENT);
    std::uniform_int_distribution<int> dup(1, 5), dK(1, 40);
    for (int trial = 0; trial < 100; ++trial) {
        const int R = dR(g), N = dN(g), up = dup(g), down = dup(g), K = dK(g);
        int64_t M = -1;
        CHECK(scir_hip_resample_len(N, up, down, K, &M) == SCIR_HIP_OK);
        CHECK(M >= 0);
        std::vector<double> x(R * N + 1), h(K), y(R * std::max<int64_t>(M, 1));
        for (double &v : x) v = nd(g);
        for (double &v : h) v = nd(g);
        CHECK(scir_hip_resample_poly_f64_cpu(x.data(), N, R, N, up, down, h.data(), K, y.data(),
                                             std::max<int64_t>(M, 1), 3) == SCIR_HIP_OK);
        // y[m] = sum_k h[k] * xs[n - k], xs the zero-stuffed row, n = (K-1)/2 + down*m
        for (int r = 0; r < R; ++r)
            for (int64_t m = 0; m < M; ++m) {
                const int64_t n = (K - 1) / 2 + down * m;
                double acc = 0.0;
                for (int64_t k = 0; k < K; ++k) {
                    const int64_t j = n - k;
                    if (j >= 0 && j % up == 0 && j / up < N) acc += h[k] * x[r * N + j / up];
                }
                CHECK(std::fabs(acc - y[r * std::max<int64_t>(M, 1) + m]) <=
                      1e-12 * (1.0 + std::fabs(acc)) * K);
            }
    }
    int64_t M = 0;
    CHECK(scir_hip_resample_len(10, 0, 1, 3, &M) == SCIR_HIP_ERR_INVALID_ARGUMENT);
    CHECK(scir_hip_resample_len(10, 1, 1, 3, nullptr) == SCIR_HIP_ERR_INVALID_ARGUMENT);
}

const char *kTable =
    "{\"arch\": \"gfx950\", \"auto_mfma_min_k\": 96, \"auto_fft_min_k\": 300,\n"
    " \"ragged\": {\"direct_wg_per_cu\": 4, \"direct_min_item_kmacs\": 128,"
    " \"mfma_segs_per_item\": 16, \"mfma_wg_per_cu\": 1},\n"
    " \"entries\": [\n"
    "  {\"k_lo\": 1, \"k_hi\": 16, \"n_lo\": 0, \"n_hi\": 8192, \"direct_r\": 16,"
    " \"direct_loader\": 1, \"direct_seg_tiles\": 2, \"mfma_tiles\": 1, \"mfma_tiles_fp16\": 2,"
    " \"mfma_segs\": 4},\n"
    "  {\"k_lo\": 17, \"k_hi\": 4096, \"n_lo\": 8193, \"n_hi\": 4611686018427387904,"
    " \"direct_r\": 8, \"direct_loader\": 2, \"direct_seg_tiles\": 4, \"mfma_tiles\": 4,"
    " \"mfma_tiles_fp16\": 4, \"mfma_segs\": 16}]}\n";

void test_tune_table(std::mt19937_64 &g)
{
    scir::TuneTable t;
    CHECK(scir::parse_tune_table(kTable, "t", &t) == SCIR_HIP_OK);
    CHECK(t.arch == "gfx950" && t.entries.size() == 2 && t.auto_mfma_min_k == 96);
    CHECK(t.auto_fft_min_k == 300 && t.has_ragged && t.ragged.direct_wg_per_cu == 4);
    CHECK(t.ragged.mfma_segs_per_item == 16 && t.ragged.mfma_wg_per_cu == 1);
    CHECK(t.entries[0].c.direct_r == 16 && t.entries[1].c.mfma_tiles == 4);
    CHECK(t.entries[1].n_hi == (int64_t)1 << 62);
    // write, read back: the same table
    scir::TuneTable u;
    CHECK(scir::parse_tune_table(scir::tune_table_json(t), "round trip", &u) == SCIR_HIP_OK);
    CHECK(scir::tune_table_json(u) == scir::tune_table_json(t));
    // every prefix and random corruptions: OK or a message, never a fault
    const std::string full(kTable);
    for (size_t cut = 0; cut <= full.size(); ++cut) {
        scir::TuneTable p;
        const int st = scir::parse_tune_table(full.substr(0, cut), "prefix", &p);
        CHECK(st == SCIR_HIP_OK || st == SCIR_HIP_ERR_INVALID_ARGUMENT);
    }
    std::uniform_int_distribution<size_t> pos(0, full.size() - 1);
    const char junk[] = "{}[]:,\"0123456789- x";
    for (int trial = 0; trial < 2000; ++trial) {
        std::string s = full;
        for (int k = 0; k < 4; ++k) s[pos(g)] = junk[g() % (sizeof junk - 1)];
        scir::TuneTable p;
        const int st = scir::parse_tune_table(s, "fuzz", &p);
        CHECK(st == SCIR_HIP_OK || st == SCIR_HIP_ERR_INVALID_ARGUMENT);
        if (st == SCIR_HIP_OK)
            for (const auto &e : p.entries) CHECK(e.k_lo <= e.k_hi && e.n_lo <= e.n_hi);
    }
}

void test_ragged_items(std::mt19937_64 &g)
{
    std::uniform_int_distribution<int> dS(0, 60), dk(0, 3000), pick(0, 9);
    std::uniform_int_distribution<int64_t> dlen(0, 100000);
    for (int trial = 0; trial < 200; ++trial) {
        const int S = dS(g);
        std::vector<int64_t> offs(S + 1, 0), boffs(S + 1, 0);
        for (int s = 0; s < S; ++s) {
            const int64_t len = pick(g) < 2 ? 0 : pick(g) < 5 ? dlen(g) % 300 : dlen(g);
            offs[s + 1] = offs[s] + len;
            boffs[s + 1] = boffs[s] + (pick(g) == 0 ? 0 : dk(g));
        }
        CHECK(scir::validate_csr(offs.data(), boffs.data(), S) == SCIR_HIP_OK);
        scir::RaggedTune tune;
        tune.direct_min_item_kmacs = 1 << (g() % 12);
        tune.mfma_segs_per_item = 1 + (int)(g() % 40);
        for (int mfma = 0; mfma < 2; ++mfma) {
            std::vector<scir::RaggedItem> items;
            std::vector<scir::MfmaRaggedItem> mitems;
            int kmax = -1;
            CHECK(scir::build_ragged_items(offs.data(), boffs.data(), S, mfma != 0, tune, &items,
                                           &mitems, &kmax) == SCIR_HIP_OK);
            // every tile / segment of every non-empty signal exactly once
            std::vector<std::vector<int>> seen(S);
            int km = 0;
            for (int s = 0; s < S; ++s) {
                const int64_t len = offs[s + 1] - offs[s];
                const int64_t unit = mfma ? scir::kMfmaRaggedSegLen : scir::kRaggedDirectTile;
                seen[s].assign((size_t)((len + unit - 1) / unit), 0);
                if (len > 0) km = std::max(km, (int)(boffs[s + 1] - boffs[s]));
            }
            CHECK(kmax == km);
            auto signal_of = [&](int64_t x0) {
                int lo = 0, hi = S;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) / 2;
                    if (offs[mid] <= x0) lo = mid; else hi = mid;
                }
                while (lo < S && offs[lo + 1] == offs[lo]) ++lo;  // skip empty signals
                return lo;
            };
            if (!mfma) {
                int64_t prev = std::numeric_limits<int64_t>::max();
                for (const auto &it : items) {
                    const int s = signal_of(it.x0);
                    CHECK(s < S && it.len == offs[s + 1] - offs[s] && it.b0 == boffs[s]);
                    CHECK(it.t0 % scir::kRaggedDirectTile == 0 && it.ntiles >= 1);
                    const int64_t cost = (int64_t)it.ntiles * scir::kRaggedDirectTile *
                                         std::max(it.K, 1);
                    CHECK(cost <= prev);
                    prev = cost;
                    for (int k = 0; k < it.ntiles; ++k) {
                        const int64_t tile = it.t0 / scir::kRaggedDirectTile + k;
                        CHECK(tile < (int64_t)seen[s].size());
                        if (tile < (int64_t)seen[s].size()) ++seen[s][tile];
                    }
                }
            } else {
                for (const auto &it : mitems) {
                    const int s = signal_of(it.x0);
                    CHECK(s < S && it.len == offs[s + 1] - offs[s] && it.b0 == boffs[s]);
                    CHECK(it.nseg >= 1 && it.nseg <= tune.mfma_segs_per_item);
                    for (int k = 0; k < it.nseg; ++k) {
                        const int64_t seg = it.seg0 + k;
                        CHECK(seg < (int64_t)seen[s].size());
                        if (seg < (int64_t)seen[s].size()) ++seen[s][seg];
                    }
                }
            }
            for (int s = 0; s < S; ++s)
                for (int c : seen[s]) CHECK(c == 1);
        }
    }
    std::vector<int64_t> bad = {0, 10, 5}, boffs = {0, 1, 2};
    CHECK(scir::validate_csr(bad.data(), boffs.data(), 2) == SCIR_HIP_ERR_INVALID_ARGUMENT);
    std::vector<int64_t> off1 = {1, 2};
    CHECK(scir::validate_csr(off1.data(), boffs.data(), 1) == SCIR_HIP_ERR_INVALID_ARGUMENT);
    CHECK(scir::validate_csr(nullptr, boffs.data(), 1) == SCIR_HIP_ERR_INVALID_ARGUMENT);
}

void test_guard_and_injection()
{
    // exceptions stay inside the library and become statuses
    CHECK(scir::guarded("e1", []() -> int { throw std::bad_alloc(); }) ==
          SCIR_HIP_ERR_OUT_OF_MEMORY);
    CHECK(std::strstr(scir_hip_last_error(), "e1: host memory allocation failed") != nullptr);
    CHECK(scir::guarded("e2", []() -> int { throw std::runtime_error("boom"); }) ==
          SCIR_HIP_ERR_BACKEND);
    CHECK(std::strcmp(scir_hip_last_error(), "e2: boom") == 0);
    CHECK(scir::guarded("e3", []() -> int { throw 7; }) == SCIR_HIP_ERR_BACKEND);
    // injected allocation failure in a CPU arm (threads need host memory)
    std::vector<float> x(4 * 64, 1.0f), y(4 * 64), b(5, 1.0f);
    setenv("SCIR_HIP_FAIL_ALLOC", "1", 1);
    CHECK(scir_hip_fir_f32_cpu(x.data(), 64, b.data(), 5, 4, 64, y.data(), 64, 4) ==
          SCIR_HIP_ERR_OUT_OF_MEMORY);
    CHECK(std::strstr(scir_hip_last_error(), "scir_hip_fir_f32_cpu") != nullptr);
    std::vector<int64_t> offs = {0, 100}, boffs = {0, 3};
    std::vector<scir::RaggedItem> items;
    std::vector<scir::MfmaRaggedItem> mitems;
    int kmax = 0;
    CHECK(scir::guarded("plan", [&]() -> int {
              return scir::build_ragged_items(offs.data(), boffs.data(), 1, false,
                                              scir::RaggedTune{}, &items, &mitems, &kmax);
          }) == SCIR_HIP_ERR_OUT_OF_MEMORY);
    unsetenv("SCIR_HIP_FAIL_ALLOC");
    CHECK(scir_hip_fir_f32_cpu(x.data(), 64, b.data(), 5, 4, 64, y.data(), 64, 4) == SCIR_HIP_OK);
    CHECK(scir_hip_version() >= 10100);
}

}  // namespace

int main()
{
    std::mt19937_64 g(0x5C12DDDDull);
    test_cpu_fir(g);
    test_sos_and_resample(g);
    test_tune_table(g);
    test_ragged_items(g);
    test_guard_and_injection();
    std::printf("sanitize: %d failures\n", failures);
    return failures != 0;
}
