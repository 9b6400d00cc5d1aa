// Tuning table and overrides (the search itself is added with the autotuner).
#include "common.hpp"

#include <cstring>

namespace scir {
namespace {
std::mutex g_tune_mu;
// Overrides set through scir_hip_tune_set (-1 = unset).
int g_over_direct_r = -1, g_over_direct_loader = -1, g_over_direct_seg = -1,
    g_over_mfma_tiles = -1, g_over_auto_min_k = -1;
}  // namespace

TuneChoice tune_lookup(int, int64_t)
{
    TuneChoice c;
    std::lock_guard<std::mutex> g(g_tune_mu);
    if (g_over_direct_r >= 0) c.direct_r = g_over_direct_r;
    if (g_over_direct_loader >= 0) c.direct_loader = g_over_direct_loader;
    if (g_over_direct_seg >= 0) c.direct_seg_tiles = g_over_direct_seg;
    if (g_over_mfma_tiles >= 0) c.mfma_tiles = c.mfma_tiles_fp16 = g_over_mfma_tiles;
    if (g_over_auto_min_k >= 0) c.auto_mfma_min_k = g_over_auto_min_k;
    return c;
}

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
        if (v > 1) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.loader must be 0 or 1");
        g_over_direct_loader = v;
    } else if (!std::strcmp(key, "direct.seg_tiles")) {
        if (v == 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "direct.seg_tiles must be >= 1");
        g_over_direct_seg = v;
    } else if (!std::strcmp(key, "mfma.tiles")) {
        if (v >= 0 && v != 1 && v != 2 && v != 4)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "mfma.tiles must be 1, 2 or 4");
        g_over_mfma_tiles = v;
    } else if (!std::strcmp(key, "auto.mfma_min_k")) {
        g_over_auto_min_k = v;
    } else {
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "unknown tuning key '%s'", key);
    }
    return SCIR_HIP_OK;
}

extern "C" int scir_hip_autotune(const char *, int32_t)
{
    return fail(SCIR_HIP_ERR_UNSUPPORTED, "autotune not built yet");
}

extern "C" int scir_hip_autotune_load(const char *)
{
    return fail(SCIR_HIP_ERR_UNSUPPORTED, "autotune not built yet");
}
