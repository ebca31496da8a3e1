// C++ API of include/despeckcl.h implemented over the C ABI of include/despeckcl_amd.h.
#include "../../include/despeckcl.h"
#include "../../include/despeckcl_amd.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace {

void check(int rc)
{
    if (rc == DKX_OK) return;
    const std::string msg = dkx_last_error();
    if (rc == DKX_EINVAL || rc == DKX_ESTATS || rc == DKX_EUNSUPPORTED) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

std::vector<dkx_nlsar_stats> views(const nlsar_stats_collection& nsc)
{
    std::vector<dkx_nlsar_stats> v;
    for (const auto& kv : nsc)
        v.push_back(dkx_nlsar_stats{kv.first.patch_size, kv.first.scale_size,
                                    (int)kv.second.lut_size, kv.second.dissims_min,
                                    kv.second.dissims_max, kv.second.quantilles.data(),
                                    kv.second.chi2cdf_inv.data()});
    return v;
}

nlsar_stats_collection collect(const std::vector<dkx_nlsar_stats>& out)
{
    nlsar_stats_collection nsc;
    for (const auto& s : out) {
        nlsar::stats st;
        st.lut_size = (unsigned int)s.lut_size;
        st.dissims_min = s.dissims_min;
        st.dissims_max = s.dissims_max;
        st.quantilles.assign(s.quantilles, s.quantilles + s.lut_size);
        st.chi2cdf_inv.assign(s.chi2cdf_inv, s.chi2cdf_inv + s.lut_size);
        nsc.emplace(nlsar::params{s.patch_size, s.scale_size}, st);
    }
    return nsc;
}

}  // namespace

// nlsar::stats (src/compute/nlsar/stats.cpp:28-49)
nlsar::stats::stats() : lut_size((unsigned int)-1), dissims_min(0.f), dissims_max(0.f) {}

nlsar::stats::stats(std::vector<float> dissims, unsigned int lut) : lut_size(lut)
{
    quantilles.resize(lut);
    chi2cdf_inv.resize(lut);
    const long long m = dkx_nlsar_stats_build(dissims.data(), (long long)dissims.size(), (int)lut,
                                              &dissims_min, &dissims_max, quantilles.data(),
                                              chi2cdf_inv.data());
    if (m <= 0) throw std::invalid_argument("nlsar::stats: no finite dissimilarity");
}

nlsar::stats& nlsar::stats::operator=(const stats& other)
{
    lut_size = other.lut_size;
    dissims_min = other.dissims_min;
    dissims_max = other.dissims_max;
    quantilles = other.quantilles;
    chi2cdf_inv = other.chi2cdf_inv;
    return *this;
}

int despeckcl::boxcar(float* ampl_master, float* ampl_slave, float* phase, float* ref_filt,
                      float* phase_filt, float* coh_filt, const int height, const int width,
                      const int window_width, std::vector<std::string>)
{
    check(dkx_boxcar(ampl_master, ampl_slave, phase, ref_filt, phase_filt, coh_filt, height, width,
                     window_width));
    return 0;
}

int despeckcl::goldstein(float* ampl_master, float* ampl_slave, float* phase, float* ref_filt,
                         float* phase_filt, float* coh_filt, const unsigned int height,
                         const unsigned int width, const unsigned int patch_size,
                         const unsigned int overlap, const float alpha, std::vector<std::string>)
{
    if (height > 0x7fffffffu || width > 0x7fffffffu || patch_size > 0x7fffffffu || overlap > 0x7fffffffu)
        throw std::invalid_argument("goldstein: sizes out of range");
    check(dkx_goldstein(ampl_master, ampl_slave, phase, ref_filt, phase_filt, coh_filt, (int)height,
                        (int)width, (int)patch_size, (int)overlap, alpha));
    return 0;
}

int despeckcl::nlsar(float* ampl_master, float* ampl_slave, float* phase, float* ref_filt,
                     float* phase_filt, float* coh_filt, const int height, const int width,
                     const int search_window_size, const std::vector<int> patch_sizes,
                     const std::vector<int> scale_sizes,
                     std::map<nlsar::params, nlsar::stats> nlsar_stats, const float h_param,
                     const float c_param, std::vector<std::string>)
{
    std::vector<dkx_nlsar_stats> v = views(nlsar_stats);
    check(dkx_nlsar_insar(ampl_master, ampl_slave, phase, ref_filt, phase_filt, coh_filt, height,
                          width, search_window_size, patch_sizes.data(), (int)patch_sizes.size(),
                          scale_sizes.data(), (int)scale_sizes.size(), v.data(), (int)v.size(),
                          h_param, c_param));
    return 0;
}

// covariance overload (src/compute/nlsar/nlsar.cpp:170-193): 2*dim*dim interlaced planes in
// and out
int despeckcl::nlsar(float* covmat_raw, float* covmat_filt, const int height, const int width,
                     const int dim, const int search_window_size,
                     const std::vector<int> patch_sizes, const std::vector<int> scale_sizes,
                     std::map<nlsar::params, nlsar::stats> nlsar_stats, const float h_param,
                     const float c_param, std::vector<std::string>)
{
    std::vector<dkx_nlsar_stats> v = views(nlsar_stats);
    check(dkx_nlsar_covmat(covmat_raw, covmat_filt, height, width, dim, search_window_size,
                           patch_sizes.data(), (int)patch_sizes.size(), scale_sizes.data(),
                           (int)scale_sizes.size(), v.data(), (int)v.size(), h_param, c_param));
    return 0;
}

nlsar_stats_collection despeckcl::nlsar_training(float* ampl_master, float* ampl_slave,
                                                 float* phase, const int height, const int width,
                                                 const std::vector<int> patch_sizes,
                                                 const std::vector<int> scale_sizes,
                                                 std::vector<std::string>)
{
    const size_t n = patch_sizes.size() * scale_sizes.size();
    std::vector<dkx_nlsar_stats> out(n);
    std::vector<float> store(n * 2 * 256);
    check(dkx_nlsar_training_insar(ampl_master, ampl_slave, phase, height, width,
                                   patch_sizes.data(), (int)patch_sizes.size(), scale_sizes.data(),
                                   (int)scale_sizes.size(), out.data(), store.data()));
    return collect(out);
}

nlsar_stats_collection despeckcl::nlsar_training(float* covmat_raw, const int height,
                                                 const int width, const int dim,
                                                 const std::vector<int> patch_sizes,
                                                 const std::vector<int> scale_sizes,
                                                 std::vector<std::string>)
{
    const size_t n = patch_sizes.size() * scale_sizes.size();
    std::vector<dkx_nlsar_stats> out(n);
    std::vector<float> store(n * 2 * 256);
    check(dkx_nlsar_training_covmat(covmat_raw, height, width, dim, patch_sizes.data(),
                                    (int)patch_sizes.size(), scale_sizes.data(),
                                    (int)scale_sizes.size(), out.data(), store.data()));
    return collect(out);
}

// Text format: "despeckcl_amd-nlsar-stats 1", the entry count, then per entry
// "patch scale lut_size min max" followed by the two LUTs; floats with 9 significant digits
// so that every value reads back bit-exactly.
void despeckcl::store_nlsar_stats_collection(nlsar_stats_collection nsc, std::string filename)
{
    std::vector<dkx_nlsar_stats> v = views(nsc);
    check(dkx_nlsar_stats_store(filename.c_str(), v.data(), (int)v.size()));
}

nlsar_stats_collection despeckcl::load_nlsar_stats_collection(std::string filename)
{
    long long nf = 0;
    const long long n = dkx_nlsar_stats_load(filename.c_str(), nullptr, 0, nullptr, 0, &nf);
    if (n < 0) check((int)n);
    std::vector<dkx_nlsar_stats> e((size_t)n);
    std::vector<float> store((size_t)nf);
    check((int)std::min(0ll, dkx_nlsar_stats_load(filename.c_str(), e.data(), (int)n, store.data(), nf, &nf)));
    nlsar_stats_collection nsc;
    for (const auto& x : e) {
        nlsar::params p;
        p.patch_size = x.patch_size;
        p.scale_size = x.scale_size;
        nlsar::stats s;
        s.lut_size = (unsigned int)x.lut_size;
        s.dissims_min = x.dissims_min;
        s.dissims_max = x.dissims_max;
        s.quantilles.assign(x.quantilles, x.quantilles + x.lut_size);
        s.chi2cdf_inv.assign(x.chi2cdf_inv, x.chi2cdf_inv + x.lut_size);
        nsc.emplace(p, s);
    }
    return nsc;
}

int despeckcl::nlinsar(float* ampl_master, float* ampl_slave, float* phase, float* ref_filt,
                       float* phase_filt, float* coh_filt, const int height, const int width,
                       const int search_window_size, const int patch_size, const int niter,
                       const int lmin, std::vector<std::string>)
{
    check(dkx_nlinsar(ampl_master, ampl_slave, phase, ref_filt, phase_filt, coh_filt, height, width,
                      search_window_size, patch_size, niter, lmin));
    return 0;
}
