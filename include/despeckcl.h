/*
 * despeckcl.h -- C++ API of the MI355X NL-SAR / NL-InSAR library.
 *
 * Same names, signatures and semantics as the reference's include/despeckcl.h for the
 * nonlocal filters (lines 30 and 45-109) and the statistics types of
 * include/compute/nlsar/parameters.h:24-56 and include/compute/nlsar/stats.h:27-55, so
 * code written against despeckCL compiles and links against libdespeckcl_amd.so.
 * Differences:
 *  - no Boost: store/load_nlsar_stats_collection write and read the reference's Boost text
 *    archive format with a reader / writer of their own;
 *  - errors throw std::runtime_error / std::invalid_argument instead of terminating;
 *  - stats::operator= also copies lut_size (the reference leaves it unset);
 *  - goldstein reproduces the reference's results, including its quirks (see
 *    despeckcl_amd.h).
 */
#ifndef DESPECKCL_H
#define DESPECKCL_H

#include <map>
#include <string>
#include <vector>

namespace nlsar {

struct params {
    int patch_size;
    int scale_size;

    bool operator==(const params& other) const
    {
        return (patch_size == other.patch_size) && (scale_size == other.scale_size);
    }
    // map key order: patch size, then scale size (parameters.h:37-40)
    bool operator<(const params& other) const
    {
        return (patch_size < other.patch_size) ||
               ((patch_size == other.patch_size) && (scale_size < other.scale_size));
    }
};

class stats {
public:
    stats();
    stats(std::vector<float> dissims, unsigned int lut_size);
    stats(const stats& other) = default;
    stats& operator=(const stats& other);
    unsigned int lut_size;
    float dissims_min;
    float dissims_max;
    std::vector<float> quantilles;
    std::vector<float> chi2cdf_inv;
};

}  // namespace nlsar

typedef std::map<nlsar::params, nlsar::stats> nlsar_stats_collection;

namespace despeckcl {

int boxcar(float* ampl_master, float* ampl_slave, float* phase, float* ref_filt,
           float* phase_filt, float* coh_filt, const int height, const int width,
           const int window_width, std::vector<std::string> enabled_log_levels);

int goldstein(float* ampl_master, float* ampl_slave, float* phase, float* ref_filt,
              float* phase_filt, float* coh_filt, const unsigned int height,
              const unsigned int width, const unsigned int patch_size,
              const unsigned int overlap, const float alpha,
              std::vector<std::string> enabled_log_levels);

int nlsar(float* ampl_master, float* ampl_slave, float* phase, float* ref_filt,
          float* phase_filt, float* coh_filt, const int height, const int width,
          const int search_window_size, const std::vector<int> patch_sizes,
          const std::vector<int> scale_sizes, std::map<nlsar::params, nlsar::stats> nlsar_stats,
          const float h_param, const float c_param, std::vector<std::string> enabled_log_levels);

int nlsar(float* covmat_raw, float* covmat_filt, const int height, const int width, const int dim,
          const int search_window_size, const std::vector<int> patch_sizes,
          const std::vector<int> scale_sizes, std::map<nlsar::params, nlsar::stats> nlsar_stats,
          const float h_param, const float c_param, std::vector<std::string> enabled_log_levels);

nlsar_stats_collection nlsar_training(float* ampl_master, float* ampl_slave, float* phase,
                                      const int height, const int width,
                                      const std::vector<int> patch_sizes,
                                      const std::vector<int> scale_sizes,
                                      std::vector<std::string> enabled_log_levels);

nlsar_stats_collection nlsar_training(float* covmat_raw, const int height, const int width,
                                      const int dim, const std::vector<int> patch_sizes,
                                      const std::vector<int> scale_sizes,
                                      std::vector<std::string> enabled_log_levels);

void store_nlsar_stats_collection(nlsar_stats_collection nsc, std::string filename);
nlsar_stats_collection load_nlsar_stats_collection(std::string filename);

int nlinsar(float* ampl_master, float* ampl_slave, float* phase, float* ref_filt,
            float* phase_filt, float* coh_filt, const int height, const int width,
            const int search_window_size, const int patch_size, const int niter, const int lmin,
            std::vector<std::string> enabled_log_levels);

}  // namespace despeckcl

#endif
