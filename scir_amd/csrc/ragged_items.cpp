// Work items of ragged (CSR) batches, built on the host once per plan (host
// code only; fir_ragged.hip uploads them and runs the persistent kernels).
//
// Signal s is x[offs[s] .. offs[s+1]) with taps b[b_offs[s] .. b_offs[s+1]).
//   * Direct form: a signal is cut into items of consecutive 2048-output
//     tiles, enough tiles per item that an item carries at least
//     direct_min_item_kmacs * 1024 multiply-adds (so the queue overhead stays
//     small next to the work).
//   * MFMA: a signal is cut into items of at most mfma_segs_per_item
//     consecutive 8192-output segments; an item costs its segments times the
//     loop's half-steps (K-1)/16 + 4.
// Items are sorted by cost, longest first (LPT order), so the longest work
// starts first and short items fill the tail of the persistent launch.
#include <algorithm>
#include <numeric>

#include "host_common.hpp"

namespace scir {

int validate_csr(const int64_t *offs, const int64_t *b_offs, int64_t S)
{
    if (S < 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative signal count");
    if (!offs || !b_offs) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null offsets");
    if (offs[0] != 0 || b_offs[0] != 0)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "offsets must start at 0");
    for (int64_t s = 0; s < S; ++s) {
        if (offs[s + 1] < offs[s] || b_offs[s + 1] < b_offs[s])
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "offsets decrease at signal %lld",
                        (long long)s);
        if (b_offs[s + 1] - b_offs[s] > 0x7fffffff)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "signal %lld has too many taps",
                        (long long)s);
    }
    return SCIR_HIP_OK;
}

void ragged_batch_stats(const int64_t *offs, const int64_t *b_offs, int64_t S, int *kmean,
                        int64_t *nmean, int *kmax)
{
    double macs = 0.0, kmacs = 0.0, total = 0.0;
    int km = 0;
    for (int64_t s = 0; s < S; ++s) {
        const double len = (double)(offs[s + 1] - offs[s]);
        const double K = (double)(b_offs[s + 1] - b_offs[s]);
        macs += len * K;
        kmacs += len * K * K;
        total += len;
        if (len > 0) km = std::max(km, (int)(b_offs[s + 1] - b_offs[s]));
    }
    *kmean = macs > 0.0 ? (int)(kmacs / macs) : 0;
    *nmean = S > 0 ? (int64_t)(total / (double)S) : 0;
    *kmax = km;
}

int build_ragged_items(const int64_t *offs, const int64_t *b_offs, int64_t S, bool mfma,
                       const RaggedTune &tune, std::vector<RaggedItem> *items,
                       std::vector<MfmaRaggedItem> *mitems, int *kmax)
{
    host_alloc_point();
    std::vector<RaggedItem> dir;
    std::vector<MfmaRaggedItem> mf;
    std::vector<int64_t> cost;
    int km = 0;
    const int segs_per_item = std::max(1, tune.mfma_segs_per_item);
    const int64_t min_work = (int64_t)std::max(1, tune.direct_min_item_kmacs) * 1024;
    for (int64_t s = 0; s < S; ++s) {
        const int64_t len = offs[s + 1] - offs[s];
        const int K = (int)(b_offs[s + 1] - b_offs[s]);
        if (len == 0) continue;
        km = std::max(km, K);
        if (mfma) {
            const int64_t segs = (len + kMfmaRaggedSegLen - 1) / kMfmaRaggedSegLen;
            if (segs > 0x7fffffff)
                return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "signal %lld is too long",
                            (long long)s);
            const int64_t per_seg = (K + 14) / 16 + 4;  // half-steps of the loop
            for (int64_t g = 0; g < segs; g += segs_per_item) {
                MfmaRaggedItem it{};
                it.x0 = offs[s];
                it.len = len;
                it.b0 = b_offs[s];
                it.K = K;
                it.seg0 = (int32_t)g;
                it.nseg = (int32_t)std::min<int64_t>(segs_per_item, segs - g);
                mf.push_back(it);
                cost.push_back((int64_t)it.nseg * per_seg);
            }
            continue;
        }
        const int64_t tiles = (len + kRaggedDirectTile - 1) / kRaggedDirectTile;
        if (tiles > 0x7fffffff)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "signal %lld is too long", (long long)s);
        const int64_t per_tile = (int64_t)kRaggedDirectTile * std::max(K, 1);
        const int64_t tpi = std::max<int64_t>(1, std::min<int64_t>(tiles, min_work / per_tile));
        for (int64_t t = 0; t < tiles; t += tpi) {
            RaggedItem it;
            it.x0 = offs[s];
            it.t0 = t * kRaggedDirectTile;
            it.len = len;
            it.b0 = b_offs[s];
            it.K = K;
            it.ntiles = (int32_t)std::min<int64_t>(tpi, tiles - t);
            dir.push_back(it);
            cost.push_back((int64_t)it.ntiles * per_tile);
        }
    }
    const size_t n = mfma ? mf.size() : dir.size();
    if (n > 0x7fffffff) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "too many work items");
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t c) { return cost[a] > cost[c]; });
    items->clear();
    mitems->clear();
    if (mfma) {
        mitems->reserve(n);
        for (size_t i = 0; i < n; ++i) mitems->push_back(mf[order[i]]);
    } else {
        items->reserve(n);
        for (size_t i = 0; i < n; ++i) items->push_back(dir[order[i]]);
    }
    *kmax = km;
    return SCIR_HIP_OK;
}

}  // namespace scir
