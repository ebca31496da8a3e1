// Host interface of the NL-SAR HIP path (planning + launch on device buffers).
#pragma once

#include "dkx_common.h"

#include <vector>

namespace dkx {

constexpr int NLSAR_MAX_NP = 3;        // patch sizes per kernel pass
constexpr int NLSAR_MAX_PATCHES = 16;  // patch sizes per call
constexpr int NLSAR_MAX_NS = 8;

// One (patch, scale) entry of the dissimilarity statistics (include/compute/nlsar/stats.h:27-55)
struct NlsarStatsView {
    int patch_size, scale_size, lut_size;
    float dissims_min, dissims_max;
    const float* quantilles;
    const float* chi2cdf_inv;
};

struct NlsarProblem {
    int H, W;
    int dim;          // 0: InSAR input (a1, a2, phase); 2: dim-2 covariance input
    int sw;
    const int* P; int nP;
    const int* S; int nS;
    const NlsarStatsView* stats; int nst;
    float h, c;
    int edge_h, edge_w;  // virtual distance of the last row/column; < 0: the overlap
    int tile;            // tile shape index (autotuned)
};

struct NlsarPlan {
    int tile, TY, TX;
    int wsh, P, smax, overlap, eh, ew;
    int nparam, lut_stride;
    Grid C, A;
    size_t off_C, off_A, off_flag, off_ub, ws_bytes;
    size_t const_floats;  // weight LUTs + Gaussian taps, see nlsar_build_constants
};

std::vector<float> gen_gauss(int scale_size);
std::vector<float> weight_lut(const NlsarStatsView& st, float h, float c);

int nlsar_plan(const NlsarProblem& pb, NlsarPlan* plan);
int nlsar_param_stats(const NlsarProblem& pb, const NlsarPlan& plan,
                      std::vector<const NlsarStatsView*>& pst);
// host image of the device constants of a call (weight LUTs and Gaussian taps)
int nlsar_build_constants(const NlsarProblem& pb, const NlsarPlan& plan, std::vector<float>& blob);
// d_const: device copy of nlsar_build_constants' blob, which must stay unchanged until the
// enqueued work has run
int nlsar_run(const NlsarProblem& pb, const NlsarPlan& plan, const float* d_a1, const float* d_a2,
              const float* d_phi, const float* d_cov, float* d_ref, float* d_phase, float* d_coh,
              void* d_ws, const float* d_const, float* d_enl_maps, float* d_alpha_maps,
              unsigned long long* ub_host, hipStream_t stream);

}  // namespace dkx
