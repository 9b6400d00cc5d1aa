// Reader and writer of the tuning table (host code only).  The table is the
// flat JSON scir_hip_autotune writes:
//   {"arch": "gfx950", "auto_mfma_min_k": 160, "auto_fft_min_k": 512,
//    "ragged": {"direct_wg_per_cu": 8, "direct_min_item_kmacs": 512,
//               "mfma_segs_per_item": 8, "mfma_wg_per_cu": 0},
//    "entries": [{"k_lo": 1, "k_hi": 16, "n_lo": 0, "n_hi": 8192,
//                 "direct_r": 8, "direct_loader": 1, "direct_seg_tiles": 2,
//                 "direct_mfma_strip": 8,
//                 "mfma_tiles": 2, "mfma_tiles_fp16": 4, "mfma_segs": 16}, ...]}
// "auto_fft_min_k", "ragged", "mfma_segs" and "direct_mfma_strip" are
// optional (tables written before they existed); out-of-range values fall
// back to the defaults.
#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <sstream>

#include "host_common.hpp"

namespace scir {
namespace {

// Finds `"key": <int>` in s[from, to); false when absent or not an integer.
bool find_int(const std::string &s, size_t from, size_t to, const char *key, int64_t *out)
{
    const std::string pat = std::string("\"") + key + "\"";
    size_t p = s.find(pat, from);
    if (p == std::string::npos || p >= to) return false;
    p = s.find(':', p + pat.size());
    if (p == std::string::npos || p >= to) return false;
    ++p;
    while (p < to && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
    if (p >= to) return false;
    const std::string num = s.substr(p, std::min<size_t>(to - p, 24));
    char *end = nullptr;
    errno = 0;
    const long long v = std::strtoll(num.c_str(), &end, 10);
    if (end == num.c_str() || errno == ERANGE) return false;
    *out = v;
    return true;
}

bool find_str(const std::string &s, const char *key, std::string *out)
{
    const std::string pat = std::string("\"") + key + "\"";
    size_t p = s.find(pat);
    if (p == std::string::npos) return false;
    p = s.find(':', p + pat.size());
    if (p == std::string::npos) return false;
    p = s.find('"', p);
    if (p == std::string::npos) return false;
    const size_t q = s.find('"', p + 1);
    if (q == std::string::npos) return false;
    *out = s.substr(p + 1, q - p - 1);
    return true;
}

int clamp_int(int64_t v, int lo, int hi, int dflt) { return (v >= lo && v <= hi) ? (int)v : dflt; }

}  // namespace

int parse_tune_table(const std::string &s, const char *source, TuneTable *out)
{
    TuneTable t;
    if (!find_str(s, "arch", &t.arch))
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "%s: malformed table (no \"arch\")", source);
    size_t p = s.find("\"entries\"");
    const size_t end_entries = p == std::string::npos ? p : s.find(']', p);
    while (p != std::string::npos) {
        const size_t a = s.find('{', p + 1);
        if (a == std::string::npos || (end_entries != std::string::npos && a > end_entries)) break;
        const size_t b = s.find('}', a);
        if (b == std::string::npos)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "%s: malformed table (unterminated entry)",
                        source);
        int64_t v[9];
        const char *keys[9] = {"k_lo", "k_hi", "n_lo", "n_hi", "direct_r", "direct_loader",
                               "direct_seg_tiles", "mfma_tiles", "mfma_tiles_fp16"};
        for (int i = 0; i < 9; ++i)
            if (!find_int(s, a, b, keys[i], &v[i]))
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT,
                            "%s: malformed table (entry without \"%s\")", source, keys[i]);
        if (v[0] < 0 || v[1] < v[0] || v[1] > 0x7fffffff || v[2] < 0 || v[3] < v[2])
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "%s: malformed table (empty bucket)",
                        source);
        TuneEntry e;
        e.k_lo = (int)v[0];
        e.k_hi = (int)v[1];
        e.n_lo = v[2];
        e.n_hi = v[3];
        e.c.direct_r = v[4] == 16 ? 16 : 8;
        e.c.direct_loader = clamp_int(v[5], 0, 3, 1);
        e.c.direct_seg_tiles = clamp_int(v[6], 1, 1 << 20, 4);
        e.c.mfma_tiles = (v[7] == 1 || v[7] == 4) ? (int)v[7] : 2;
        e.c.mfma_tiles_fp16 = (v[8] == 1 || v[8] == 2) ? (int)v[8] : 4;
        int64_t segs = 0;
        if (find_int(s, a, b, "mfma_segs", &segs)) e.c.mfma_segs = clamp_int(segs, 1, 1 << 20, 16);
        int64_t strip = 0;
        if (find_int(s, a, b, "direct_mfma_strip", &strip))
            e.c.direct_mfma_strip = clamp_int(strip, 0, 4096, 0);
        t.entries.push_back(e);
        p = b;
    }
    find_int(s, 0, s.size(), "auto_mfma_min_k", &t.auto_mfma_min_k);
    find_int(s, 0, s.size(), "auto_fft_min_k", &t.auto_fft_min_k);
    const size_t r = s.find("\"ragged\"");
    if (r != std::string::npos) {
        const size_t a = s.find('{', r);
        const size_t b = a == std::string::npos ? a : s.find('}', a);
        if (b == std::string::npos)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "%s: malformed table (\"ragged\")", source);
        int64_t v = 0;
        t.has_ragged = true;
        if (find_int(s, a, b, "direct_wg_per_cu", &v))
            t.ragged.direct_wg_per_cu = clamp_int(v, 1, 32, t.ragged.direct_wg_per_cu);
        if (find_int(s, a, b, "direct_min_item_kmacs", &v))
            t.ragged.direct_min_item_kmacs =
                clamp_int(v, 1, 1 << 20, t.ragged.direct_min_item_kmacs);
        if (find_int(s, a, b, "mfma_segs_per_item", &v))
            t.ragged.mfma_segs_per_item = clamp_int(v, 1, 1024, t.ragged.mfma_segs_per_item);
        if (find_int(s, a, b, "mfma_wg_per_cu", &v))
            t.ragged.mfma_wg_per_cu = clamp_int(v, 0, 8, t.ragged.mfma_wg_per_cu);
    }
    *out = std::move(t);
    return SCIR_HIP_OK;
}

std::string tune_table_json(const TuneTable &t)
{
    std::ostringstream o;
    const TuneChoice d;
    o << "{\"arch\": \"" << t.arch << "\", \"auto_mfma_min_k\": "
      << (t.auto_mfma_min_k >= 0 ? t.auto_mfma_min_k : d.auto_mfma_min_k)
      << ", \"auto_fft_min_k\": " << (t.auto_fft_min_k >= 0 ? t.auto_fft_min_k : d.auto_fft_min_k)
      << ",\n \"ragged\": {\"direct_wg_per_cu\": " << t.ragged.direct_wg_per_cu
      << ", \"direct_min_item_kmacs\": " << t.ragged.direct_min_item_kmacs
      << ", \"mfma_segs_per_item\": " << t.ragged.mfma_segs_per_item
      << ", \"mfma_wg_per_cu\": " << t.ragged.mfma_wg_per_cu << "},\n \"entries\": [";
    for (size_t i = 0; i < t.entries.size(); ++i) {
        const TuneEntry &e = t.entries[i];
        o << (i ? ",\n  " : "\n  ") << "{\"k_lo\": " << e.k_lo << ", \"k_hi\": " << e.k_hi
          << ", \"n_lo\": " << e.n_lo << ", \"n_hi\": " << e.n_hi
          << ", \"direct_r\": " << e.c.direct_r << ", \"direct_loader\": " << e.c.direct_loader
          << ", \"direct_seg_tiles\": " << e.c.direct_seg_tiles
          << ", \"direct_mfma_strip\": " << e.c.direct_mfma_strip
          << ", \"mfma_tiles\": " << e.c.mfma_tiles
          << ", \"mfma_tiles_fp16\": " << e.c.mfma_tiles_fp16
          << ", \"mfma_segs\": " << e.c.mfma_segs << "}";
    }
    o << "]}\n";
    return o.str();
}

}  // namespace scir
