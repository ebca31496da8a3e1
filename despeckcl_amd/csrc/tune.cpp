// The autotuner's records: the tile variant that ran fastest for a parameter set and an
// image size, measured by dkx_nlsar_insar_device_autotune, dkx_nlinsar_device_autotune and
// dkx_nlsar_batch_autotune.  Automatic tile choices (single calls, batches, the server) use a
// record when one exists for their key, else the built-in rule.
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"

#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace dkx {

using TuneKey = std::tuple<int, int, int, std::vector<int>, std::vector<int>, int>;

static std::mutex g_mu;
static std::map<TuneKey, int> g_records;

static TuneKey tune_key(int filter, int variant, int sw, const int* P, int nP, const int* S, int nS,
                        double pixels)
{
    const int bucket = (int)std::floor(std::log2(pixels > 1.0 ? pixels : 1.0));
    return TuneKey{filter, variant, sw, std::vector<int>(P, P + (P ? nP : 0)),
                   std::vector<int>(S, S + (S ? nS : 0)), bucket};
}

int tuned_tile(int filter, int variant, int sw, const int* P, int nP, const int* S, int nS, double pixels)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_records.find(tune_key(filter, variant, sw, P, nP, S, nS, pixels));
    return it == g_records.end() ? -1 : it->second;
}

void record_tuned_tile(int filter, int variant, int sw, const int* P, int nP, const int* S, int nS,
                       double pixels, int tile)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_records[tune_key(filter, variant, sw, P, nP, S, nS, pixels)] = tile;
}

void clear_tuned_tiles()
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_records.clear();
}

}  // namespace dkx

extern "C" {

int dkx_get_tuned_tile(int filter, int height, int width, int search_window_size, const int* patch_sizes,
                       int n_patch_sizes, const int* scale_sizes, int n_scale_sizes, int every_key)
{
    return dkx::guarded<int>([&]() -> int {
        if (filter < 0 || filter > 1 || height < 1 || width < 1) return -1;
        return dkx::tuned_tile(filter, filter == 0 ? (every_key ? 1 : 0) : 0, search_window_size, patch_sizes,
                               n_patch_sizes, filter == 0 ? scale_sizes : nullptr, filter == 0 ? n_scale_sizes : 0,
                               (double)height * width);
    });
}

void dkx_clear_tuned_tiles(void)
{
    dkx::guarded_void([] { dkx::clear_tuned_tiles(); });
}

}  // extern "C"
