// Host helpers of the NL-SAR statistics (nlsar::stats, src/compute/nlsar/stats.cpp:28-80).
#pragma once

namespace dkx {

// inverse CDF of the chi-square distribution with nu degrees of freedom
double chi2_inv(double p, double nu);
// quantile thresholds dmin + i*(dmax - dmin)/lut_size, i < lut_size
void stats_thresholds(float dmin, float dmax, int lut_size, float* thr);
// inverse chi-square(49) LUT
void stats_chi2_lut(int lut_size, float* chi2cdf_inv);
// the whole constructor on a host array of dissimilarities; returns the finite count
long long stats_build(const float* dissims, long long n, int lut_size, float* dmin, float* dmax,
                      float* quantilles, float* chi2cdf_inv);

}  // namespace dkx
