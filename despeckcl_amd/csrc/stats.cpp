// NL-SAR dissimilarity statistics (quantile and inverse chi-square LUTs) and the training
// that produces them from a homogeneous area.  Host code: the training area is small (the
// reference's documentation suggests about 30x30 pixels, sphinx/cpp_api.rst:82) and the
// pairwise patch comparison is the whole cost.
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_nlsar.h"

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

// nlsar::stats::stats (src/compute/nlsar/stats.cpp:28-80): drop NaN / inf, sort, quantile
// LUT on lut_size equal steps of [min, max] (std::lower_bound counts), chi-square(49)
// inverse LUT with the last entry fixed at 1e5.
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
    const float step = (*dmax - *dmin) / (float)lut_size;
    for (int i = 0; i < lut_size; i++) {
        const float thr = *dmin + (float)i * step;
        const size_t below = (size_t)(std::lower_bound(v.begin(), v.end(), thr) - v.begin());
        quantilles[i] = ((float)below) / (float)v.size();
    }
    const float pstep = 1.0f / (float)lut_size;
    for (int i = 0; i + 1 < lut_size; i++) chi2cdf_inv[i] = (float)chi2_inv((double)((float)i * pstep), 49.0);
    chi2cdf_inv[lut_size - 1] = 100000.0f;
    return (long long)v.size();
}

// ---- training ----------------------------------------------------------------------

// Dissimilarity of two covariance matrices as compiled into the reference's host training
// code: sim_measures.h is included there without <math.h>, so log() is ::log(double) and
// the expression is evaluated in double, rounded to float on return
// (training/patches.cpp:96-177, sim_measures.h:37-65 and 100-158).
static float host_det2(const float* p) { return (p[0] * p[3]) - (p[1] * p[1] + p[2] * p[2]); }

static float host_det3(const float* p)
{
    // p = {a00, a11, a22, a01r, a01i, a02r, a02i, a12r, a12i}
    auto ab2 = [](float r, float i) { return r * r + i * i; };
    return p[0] * p[1] * p[2] - p[0] * ab2(p[7], p[8]) - p[1] * ab2(p[5], p[6]) -
           p[2] * ab2(p[3], p[4]) +
           2 * (p[3] * p[7] * p[5] - p[4] * p[8] * p[5] + p[3] * p[8] * p[6] + p[4] * p[7] * p[6]);
}

static float host_dissim(const float* p1, const float* p2, int dim)
{
    float s[9];
    const int k = dim == 2 ? 4 : 9;
    for (int i = 0; i < k; i++) s[i] = p1[i] + p2[i];
    const float nom1 = dim == 2 ? host_det2(p1) : host_det3(p1);
    const float nom2 = dim == 2 ? host_det2(p2) : host_det3(p2);
    const float det = dim == 2 ? host_det2(s) : host_det3(s);
    float r = (float)(-1 * (2 * dim * std::log((double)2.0f) + std::log((double)nom1) +
                            std::log((double)nom2) - 2 * std::log((double)det)));
    if (std::isnan(r)) r = 0;
    return r;
}

// Training dissimilarities of one (patch, scale): rescale and average the whole training
// covariance (get_dissims.cpp:32-117, averaging with the scale as its own maximum), cut it
// into non-overlapping clamped p x p patches (patches.cpp:32-41) and compare every
// unordered pair pixel by pixel (combinations.cpp:26-41, patches.cpp:43-94).
static std::vector<float> training_dissims(const float* cov, int H, int W, int dim, int p, int s)
{
    const int np = 2 * dim * dim;
    const size_t n = (size_t)H * W;
    const float gamma = std::pow(std::min(1.0f / (float)dim, 1.0f), 0.33333f);
    const int ho = H - s + 1, wo = W - s + 1;
    std::vector<float> out;
    if (ho <= 0 || wo <= 0) return out;
    std::vector<float> g = gen_gauss(s);
    // averaged operand planes of every pixel: {a00, a01r, a01i, a11} (dim 2) or the nine
    // Hermitian parameters (dim 3); rescaled off-diagonals
    const int nk = dim == 2 ? 4 : 9;
    std::vector<int> planes = dim == 2 ? std::vector<int>{0, 2, 3, 6}
                                       : std::vector<int>{0, 8, 16, 2, 3, 4, 5, 10, 11};
    std::vector<float> avg((size_t)nk * ho * wo);
    for (int k = 0; k < nk; k++) {
        const int pl = planes[k];
        const int r = pl / 2 / dim, c = (pl / 2) % dim;
        const bool off = r != c;
        for (int y = 0; y < ho; y++)
            for (int x = 0; x < wo; x++) {
                float acc = 0.f;
                for (int ky = 0; ky < s; ky++)
                    for (int kx = 0; kx < s; kx++) {
                        float v = cov[(size_t)pl * n + (size_t)(y + ky) * W + x + kx];
                        if (off) v = v * gamma;
                        acc = std::fma(g[ky * s + kx], v, acc);
                    }
                avg[((size_t)k * ho + y) * wo + x] = acc;
            }
    }
    (void)np;
    const int nph = (ho + p - 1) / p, npw = (wo + p - 1) / p;
    const int npatch = nph * npw;
    const int pp = p * p;
    // patch-major operand storage: patch, pixel, operand
    std::vector<float> pat((size_t)npatch * pp * nk);
    for (int ty = 0; ty < nph; ty++)
        for (int tx = 0; tx < npw; tx++)
            for (int y = 0; y < p; y++)
                for (int x = 0; x < p; x++) {
                    const int yy = std::min(ho - 1, ty * p + y), xx = std::min(wo - 1, tx * p + x);
                    for (int k = 0; k < nk; k++)
                        pat[(((size_t)(ty * npw + tx) * pp) + y * p + x) * nk + k] =
                            avg[((size_t)k * ho + yy) * wo + xx];
                }
    out.reserve((size_t)npatch * (npatch - 1) / 2);
    for (int head = npatch - 1; head >= 1; head--)
        for (int j = 0; j < head; j++) {
            float sum = 0.0f;
            for (int i = 0; i < pp; i++)
                sum += host_dissim(&pat[((size_t)head * pp + i) * nk], &pat[((size_t)j * pp + i) * nk], dim);
            out.push_back(sum);
        }
    return out;
}

int training(const float* cov, int H, int W, int dim, const int* P, int nP, const int* S, int nS,
             dkx_nlsar_stats* out, float* lut_storage)
{
    if (!cov || !P || !S || !out || !lut_storage || H <= 0 || W <= 0 || nP <= 0 || nS <= 0) {
        set_error("nlsar_training: invalid arguments");
        return DKX_EINVAL;
    }
    if (dim != 2 && dim != 3) {
        // routines.cpp:77-80
        set_error("currently only 2x2 or 3x3 covariance matrices are supported");
        return DKX_EINVAL;
    }
    std::vector<int> ps(P, P + nP), ss(S, S + nS);
    std::sort(ps.begin(), ps.end());
    std::sort(ss.begin(), ss.end());
    const int lut = 256;  // get_stats.cpp:31
    int k = 0;
    for (int p : ps)
        for (int s : ss) {
            if (p < 1 || p % 2 == 0 || s < 1 || s % 2 == 0 || s > 31) {
                set_error("nlsar_training: patch and scale sizes must be odd");
                return DKX_EINVAL;
            }
            std::vector<float> d = training_dissims(cov, H, W, dim, p, s);
            float* q = lut_storage + (size_t)k * 2 * lut;
            float* c = q + lut;
            dkx_nlsar_stats& o = out[k];
            o.patch_size = p;
            o.scale_size = s;
            o.lut_size = lut;
            o.quantilles = q;
            o.chi2cdf_inv = c;
            if (stats_build(d.data(), (long long)d.size(), lut, &o.dissims_min, &o.dissims_max, q, c) == 0) {
                set_error("nlsar_training: the training area yields no finite dissimilarity for patch " +
                          std::to_string(p) + ", scale " + std::to_string(s));
                return DKX_EINVAL;
            }
            k++;
        }
    return 0;
}

}  // namespace dkx

extern "C" {

double dkx_chi2_inv(double p, double nu) { return dkx::chi2_inv(p, nu); }

long long dkx_nlsar_stats_build(const float* dissims, long long n, int lut_size, float* dmin,
                                float* dmax, float* quantilles, float* chi2cdf_inv)
{
    if (!dissims || n < 0 || lut_size < 2 || !dmin || !dmax || !quantilles || !chi2cdf_inv) {
        dkx::set_error("nlsar_stats_build: invalid arguments");
        return DKX_EINVAL;
    }
    return dkx::stats_build(dissims, n, lut_size, dmin, dmax, quantilles, chi2cdf_inv);
}

int dkx_nlsar_training_covmat(const float* covmat, int height, int width, int dim, const int* P,
                              int nP, const int* S, int nS, dkx_nlsar_stats* out, float* lut_storage)
{
    return dkx::training(covmat, height, width, dim, P, nP, S, nS, out, lut_storage);
}

int dkx_nlsar_training_insar(const float* a1, const float* a2, const float* phi, int height,
                             int width, const int* P, int nP, const int* S, int nS,
                             dkx_nlsar_stats* out, float* lut_storage)
{
    if (!a1 || !a2 || !phi || height <= 0 || width <= 0) {
        dkx::set_error("nlsar_training: invalid arguments");
        return DKX_EINVAL;
    }
    // covmat_data(insar_data), src/data/data.cpp:60-89 (host float cos / sin)
    const size_t n = (size_t)height * width;
    std::vector<float> cov(8 * n);
    for (size_t i = 0; i < n; i++) {
        const float re = a1[i] * a2[i] * std::cos(phi[i]);
        const float im = a1[i] * a2[i] * std::sin(phi[i]);
        cov[i] = a1[i] * a1[i];
        cov[n + i] = 0.f;
        cov[2 * n + i] = re;
        cov[3 * n + i] = im;
        cov[4 * n + i] = re;
        cov[5 * n + i] = -im;
        cov[6 * n + i] = a2[i] * a2[i];
        cov[7 * n + i] = 0.f;
    }
    return dkx::training(cov.data(), height, width, 2, P, nP, S, nS, out, lut_storage);
}

}  // extern "C"
