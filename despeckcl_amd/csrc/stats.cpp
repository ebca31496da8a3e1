// NL-SAR dissimilarity statistics: the quantile and inverse chi-square LUTs of
// nlsar::stats (host code, as in the reference; the training that feeds them runs on the
// device, training.hip).
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_stats.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace dkx {

// ---- inverse chi-square CDF --------------------------------------------------------
// P(a, x), the regularised lower incomplete gamma function: power series below a + 1,
// modified Lentz continued fraction for the upper function above.
static double lower_gamma_reg(double a, double x)
{
    if (x <= 0.0) return 0.0;
    const double log_front = a * std::log(x) - x - std::lgamma(a);
    if (x < a + 1.0) {
        double term = 1.0 / a, sum = term;
        for (int k = 1; k < 10000; k++) {
            term *= x / (a + k);
            sum += term;
            if (std::fabs(term) <= std::fabs(sum) * 1e-17) break;
        }
        return sum * std::exp(log_front);
    }
    const double tiny = 1e-300;
    double b = x + 1.0 - a;
    double c = 1.0 / tiny;
    double d = 1.0 / b;
    double f = d;
    for (int k = 1; k < 10000; k++) {
        const double an = -k * (k - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (std::fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double delta = d * c;
        f *= delta;
        if (std::fabs(delta - 1.0) <= 1e-17) break;
    }
    return 1.0 - std::exp(log_front) * f;
}

// Bracketed Newton iteration on P(nu/2, x/2) = p, started from the Wilson-Hilferty
// approximation; converges to a few ulp of double.
double chi2_inv(double p, double nu)
{
    if (!(p > 0.0)) return 0.0;
    if (p >= 1.0) return std::numeric_limits<double>::infinity();
    const double a = 0.5 * nu;
    // normal quantile by bisection on erfc (only a starting value)
    double zlo = -40.0, zhi = 40.0;
    for (int i = 0; i < 200; i++) {
        const double zm = 0.5 * (zlo + zhi);
        if (0.5 * std::erfc(-zm / std::sqrt(2.0)) < p) zlo = zm; else zhi = zm;
    }
    const double z = 0.5 * (zlo + zhi);
    const double v = 2.0 / (9.0 * nu);
    double x = nu * std::pow(std::max(1e-3, 1.0 - v + z * std::sqrt(v)), 3.0);
    double lo = 0.0, hi = std::numeric_limits<double>::infinity();
    for (int it = 0; it < 300; it++) {
        const double f = lower_gamma_reg(a, 0.5 * x) - p;
        if (f > 0) hi = x; else lo = x;
        const double pdf = 0.5 * std::exp((a - 1.0) * std::log(0.5 * x) - 0.5 * x - std::lgamma(a));
        double xn = x - f / pdf;
        if (!(xn > lo && xn < hi) || !std::isfinite(xn))
            xn = std::isfinite(hi) ? 0.5 * (lo + hi) : 2.0 * x + 1.0;
        if (std::fabs(xn - x) <= 4e-16 * x) { x = xn; break; }
        x = xn;
    }
    return x;
}

// Quantile thresholds dmin + i*(dmax - dmin)/lut (src/compute/nlsar/stats.cpp:52-62).
void stats_thresholds(float dmin, float dmax, int lut_size, float* thr)
{
    const float step = (dmax - dmin) / (float)lut_size;
    for (int i = 0; i < lut_size; i++) thr[i] = dmin + (float)i * step;
}

// Inverse chi-square(49) LUT on lut_size equal probability steps, the last entry fixed at
// 1e5 (src/compute/nlsar/stats.cpp:64-80).
void stats_chi2_lut(int lut_size, float* chi2cdf_inv)
{
    const float pstep = 1.0f / (float)lut_size;
    for (int i = 0; i + 1 < lut_size; i++) chi2cdf_inv[i] = (float)chi2_inv((double)((float)i * pstep), 49.0);
    chi2cdf_inv[lut_size - 1] = 100000.0f;
}

// nlsar::stats::stats (src/compute/nlsar/stats.cpp:28-80): drop NaN / inf, sort, quantile
// LUT on lut_size equal steps of [min, max] (std::lower_bound counts), chi-square(49)
// inverse LUT.
long long stats_build(const float* dissims, long long n, int lut_size, float* dmin, float* dmax,
                      float* quantilles, float* chi2cdf_inv)
{
    std::vector<float> v;
    v.reserve((size_t)std::max<long long>(n, 0));
    for (long long i = 0; i < n; i++)
        if (std::isfinite(dissims[i])) v.push_back(dissims[i]);
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    *dmin = v.front();
    *dmax = v.back();
    std::vector<float> thr(lut_size);
    stats_thresholds(*dmin, *dmax, lut_size, thr.data());
    for (int i = 0; i < lut_size; i++) {
        const size_t below = (size_t)(std::lower_bound(v.begin(), v.end(), thr[i]) - v.begin());
        quantilles[i] = ((float)below) / (float)v.size();
    }
    stats_chi2_lut(lut_size, chi2cdf_inv);
    return (long long)v.size();
}

}  // namespace dkx

extern "C" {

double dkx_chi2_inv(double p, double nu) { return dkx::chi2_inv(p, nu); }

long long dkx_nlsar_stats_build(const float* dissims, long long n, int lut_size, float* dmin,
                                float* dmax, float* quantilles, float* chi2cdf_inv)
{
    return dkx::guarded<long long>([&]() -> long long {
        if (!dissims || n < 0 || lut_size < 2 || !dmin || !dmax || !quantilles || !chi2cdf_inv) {
            dkx::set_error("nlsar_stats_build: invalid arguments");
            return DKX_EINVAL;
        }
        return dkx::stats_build(dissims, n, lut_size, dmin, dmax, quantilles, chi2cdf_inv);
    });
}

}  // extern "C"
