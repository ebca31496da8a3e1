// h of NL-InSAR for patch sizes without a tabulated value: nlinsar::simu::quantile_insar
// (src/compute/nlinsar/insarsim_simu.cpp:62-82), called by nlinsar.cpp:65-76 with alpha 0.92.
// The reference simulates 2000 patches of patch_size^2 pairs of independent unit-reflectivity,
// zero-coherence single-look pixels (insar_gen, :33-60), takes each patch's negated mean
// pixel similarity (sim_measures.h:27-40) and returns the alpha quantile minus the mean.  It
// seeds every draw from std::random_device, so its h varies from call to call; here one
// std::mt19937 with a fixed seed drives the same draws, so h is reproducible.  The simulation
// gives the reference's tabulated values (0.58 for patch 3, 0.244 for patch 7) within its
// Monte-Carlo spread.
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_nlinsar.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <numeric>
#include <random>
#include <vector>

namespace dkx {

// pixel_similarity (sim_measures.h:27-40) as the host compiles it: float operations in the
// reference's order, with the float overloads of pow, cos, sqrt, asin and log.
static float host_pixel_similarity(float am1, float as1, float dp1, float am2, float as2, float dp2)
{
    const float A = std::pow(am1 * am1 + as1 * as1 + am2 * am2 + as2 * as2, 2.0f);
    const float B = 4.0f * (am1 * am1 * as1 * as1 + am2 * am2 * as2 * as2 +
                            2.0f * am1 * as1 * am2 * as2 * std::cos(dp1 - dp2));
    const float C = am1 * as1 * am2 * as2;
    const float f_el = std::pow(C / B, 1.5f) * ((A + B) / A * std::sqrt(B / (A - B)) - std::asin(std::sqrt(B / A)));
    if (f_el == 0 || std::isnan(f_el)) return 0.0f;
    return std::log(f_el);
}

// insar_gen (:33-60) with R = 1, D = 0: the master and slave are independent circular
// Gaussian pixels.  Four normal draws, in the reference's order (the real and imaginary
// parts of x1, then of x2).
static void insar_gen(std::mt19937& gen, float* am, float* as, float* dp)
{
    std::normal_distribution<double> randn(0.0, 1.0);
    std::complex<double> x1{randn(gen), randn(gen)};
    std::complex<double> x2{randn(gen), randn(gen)};
    x1 /= std::sqrt(2.0);
    x2 /= std::sqrt(2.0);
    const std::complex<double> slc1 = 1.0 * x1;
    const std::complex<double> slc2 = std::complex<double>(0.0, 0.0) * x1 + 1.0 * x2;
    *am = (float)std::abs(slc1);
    *as = (float)std::abs(slc2);
    *dp = (float)std::arg(slc1 * std::conj(slc2));
}

float nlinsar_simulate_h(int patch_size, float alpha, unsigned int seed)
{
    constexpr int size = 2000;
    std::mt19937 gen(seed);
    std::vector<float> similarities(size);
    const float patch_area = std::pow((float)patch_size, 2.0f);
    for (int i = 0; i < size; i++) {
        float similarity = 0;
        for (int j = 0; j < patch_area; j++) {
            float am1, as1, dp1, am2, as2, dp2;
            insar_gen(gen, &am1, &as1, &dp1);
            insar_gen(gen, &am2, &as2, &dp2);
            similarity += host_pixel_similarity(am1, as1, dp1, am2, as2, dp2);
        }
        similarities[i] = -similarity / patch_area;
    }
    // quantile (:28-32) on a sorted copy; the mean accumulated in float as std::accumulate
    // with a float initial value does
    std::vector<float> sorted = similarities;
    std::sort(sorted.begin(), sorted.end());
    const float q = sorted[(size_t)(alpha * sorted.size())];
    const float mu = std::accumulate(similarities.begin(), similarities.end(), 0.0f) / similarities.size();
    return q - mu;
}

float nlinsar_h_theo(int patch_size)
{
    // nlinsar.cpp:65-76: tabulated for 3 and 7 (alpha = 0.92), simulated otherwise
    if (patch_size == 3) return 0.58f;
    if (patch_size == 7) return 0.244f;
    static std::mutex mu;
    static std::map<int, float> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(patch_size);
    if (it != cache.end()) return it->second;
    const float h = nlinsar_simulate_h(patch_size, 0.92f, NLINSAR_H_SEED);
    cache[patch_size] = h;
    return h;
}

}  // namespace dkx

extern "C" int dkx_nlinsar_simulate_h(int patch_size, unsigned int seed, float* h_theo)
{
    return dkx::guarded<int>([&]() -> int {
        if (!h_theo) { dkx::set_error("nlinsar: null output pointer"); return DKX_EINVAL; }
        if (patch_size < 1 || patch_size % 2 == 0 || patch_size > 2 * dkx::NLI_MAX_PSH + 1) {
            dkx::set_error("nlinsar: patch sizes must be odd, up to 21");
            return DKX_EINVAL;
        }
        *h_theo = dkx::nlinsar_simulate_h(patch_size, 0.92f, seed);
        return 0;
    });
}
