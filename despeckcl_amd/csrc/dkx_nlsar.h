// Host interface of the NL-SAR HIP path: parameter validation, per-image workspace plans,
// image descriptors and the launch sequence, for one image or a batch of images that share
// the filter parameters.
#pragma once

#include "dkx_common.h"
#include "../../include/despeckcl_amd.h"

#include <vector>

namespace dkx {

constexpr int NLSAR_MAX_NP = 3;        // patch sizes per kernel pass
constexpr int NLSAR_MAX_PATCHES = 16;  // patch sizes per call
constexpr int NLSAR_MAX_NS = 8;
constexpr int NLSAR_TILE = 32;         // every pass variant computes 32 x 32 output tiles
constexpr int NLSAR_NTILES = 3;        // pass variants: 4, 2 or 1 rows per thread

// The autotuner's choice among the pass variants' mean times: the fastest, or, among the
// variants within 2 % of it (the spread of such timings from run to run), the one with the
// fewest rows per thread.  Its blocks hold the smallest neighbourhood each, so the blocks
// resident on an XCD keep theirs in its L2: at 2048^2 the every-key last pass fetches
// 1.9e9 B from beyond L2 with one row per thread against 1.07e10 B with two, in the same time.
inline int nlsar_pick_variant(const float (&ms)[NLSAR_NTILES])
{
    int fastest = 0;
    for (int t = 1; t < NLSAR_NTILES; t++)
        if (ms[t] < ms[fastest]) fastest = t;
    int pick = fastest;
    for (int t = fastest + 1; t < NLSAR_NTILES; t++)
        if (ms[t] <= 1.02f * ms[fastest]) pick = t;
    return pick;
}
constexpr int NLSAR_TILE_AUTO = -1;    // per pass: 1 row per thread for one patch size, else 2

// Records of the grids, per mode (0: InSAR input; 2, 3: covariance input of that dimension).
// C: the input covariance; mode 0 keeps (a1^2, a2^2, Re C01, Im C01), modes 2 / 3 every plane
// of the 2*dim*dim (padded to float4).  A: the scale average; modes 0 / 2 keep (a00, a11,
// Re a01, Im a01) and a separate plane of log det, mode 3 (a00, a11, a22, Re/Im a01, a02, a12,
// log det).
__host__ __device__ constexpr int nlsar_c_float4(int mode) { return mode == 0 ? 1 : mode == 2 ? 2 : 5; }
__host__ __device__ constexpr int nlsar_a_float4(int mode) { return mode == 3 ? 3 : 1; }
__host__ __device__ constexpr int nlsar_a_bytes(int mode) { return mode == 3 ? 48 : 20; }

// One (patch, scale) entry of the dissimilarity statistics (include/compute/nlsar/stats.h:27-55)
struct NlsarStatsView {
    int patch_size, scale_size, lut_size;
    float dissims_min, dissims_max;
    const float* quantilles;
    const float* chi2cdf_inv;
};

// The filter parameters of a call, or of every image of a batch.
struct NlsarParams {
    int dim;          // 0: InSAR input (a1, a2, phase); 2 / 3: covariance input of that dimension
    int sw;
    const int* P; int nP;
    const int* S; int nS;
    const NlsarStatsView* stats; int nst;
    float h, c;
    int tile;         // pass variant (0..NLSAR_NTILES-1), or NLSAR_TILE_AUTO
    // 1: statistics of every (patch, scale) key, for the ENL / alpha maps and the count of
    // pixels whose best-parameter selection is undefined; 0: only the last key, the one the
    // selection returns whenever it is defined (best_params.cpp:27-53), so the filtered
    // output is the same either way
    int all_keys;
    // 1: the statistics collection holds keys beyond patch x scale and its last key is one
    // of them.  The reference then selects that unused key everywhere and reads defaults for
    // it: the weights of parameter index 0 (first scale, first patch) and alpha 0
    // (nlsar_filter_sub_image.cpp:224-228, 345-353; best_alpha_copy.cpp:26-35).
    int last_unused;
};

// Validated parameters and what follows from them.
struct NlsarCommon {
    int mode;         // as NlsarParams::dim
    int tile;
    int wsh, P, smax, overlap;
    int nparam, lut_stride;
    size_t const_floats;   // weight LUTs + Gaussian taps, see nlsar_build_constants
};

// Workspace plan of one image: padded grids C (input covariance) and A (scale average).
struct NlsarImagePlan {
    int H, W, eh, ew;
    Grid C, A;
    size_t off_C, off_A, off_flag, off_ub, bytes;   // relative to the image's workspace
};

// What the kernels read about one image (device-visible; a batch keeps a table of them in
// device memory, a single call passes one by value).
struct NlsarImageDesc {
    const float* a1; const float* a2; const float* phi; const float* cov;
    float* ref; float* phase; float* coh;
    float* cov_out;                      // covariance input: the 2*dim*dim filtered planes
    float* enl_nb; float* alpha;         // [nparam][H*W] or null
    float* C;                            // float4 per element
    float* A;                            // float4 per element, then the log det plane
    unsigned char* ubflag;               // [H*W]
    unsigned long long* ub_count;
    int H, W;
    int cpitch, coy, cox, ch, cw;
    int apitch, aoy, aox, ah, aw, aplane;
    Launch L[4];                         // interior, last row, last column, corner
    int gx[4], gy[4], bstart[5];         // pass blocks of each mapping, [4] = image total
    int row0_c, row0_a, block0;          // first row / pass block of this image in a batch
};

// Validates a statistics collection for the patch x scale set: every pair needs exactly
// one entry; entries for other keys are allowed, as in the reference, which keeps them in its
// maps.  views receives the entries of the set; *last_unused (see NlsarParams) whether the
// collection's last key lies outside it.
int check_stats(const dkx_nlsar_stats* stats, int n_stats, const int* P, int nP, const int* S,
                int nS, std::vector<NlsarStatsView>& views, int* last_unused);

// Everything about one NL-SAR call that can be checked and planned without a device.
struct NlsarCall {
    std::vector<NlsarStatsView> views;
    NlsarParams pp;
    NlsarCommon cm;
    NlsarImagePlan ip;
};
// dim 0: InSAR input, 2 / 3: covariance input.  all_keys as NlsarParams::all_keys.
int nlsar_prepare(NlsarCall& call, int dim, int all_keys, int height, int width, int sw,
                  const int* P, int nP, const int* S, int nS, const dkx_nlsar_stats* stats,
                  int n_stats, float h, float c);

std::vector<float> gen_gauss(int scale_size);
std::vector<float> weight_lut(const NlsarStatsView& st, float h, float c);

int nlsar_common(const NlsarParams& pp, NlsarCommon* cm);
// edge_h / edge_w: virtual distance of the last row / column (< 0: the overlap)
int nlsar_image_plan(const NlsarCommon& cm, int H, int W, int edge_h, int edge_w,
                     NlsarImagePlan* ip);
int nlsar_param_stats(const NlsarParams& pp, const NlsarCommon& cm,
                      std::vector<const NlsarStatsView*>& pst);
// host image of the device constants (weight LUTs and Gaussian taps)
int nlsar_build_constants(const NlsarParams& pp, const NlsarCommon& cm, std::vector<float>& blob);
// descriptor of one image whose workspace starts at ws_image
void nlsar_image_desc(const NlsarImagePlan& ip, char* ws_image, const float* a1, const float* a2,
                      const float* phi, const float* cov, float* ref, float* phase, float* coh,
                      float* cov_out, float* enl_maps, float* alpha_maps, NlsarImageDesc* d);
// numbers the rows and pass blocks of a batch; returns the totals
void nlsar_batch_layout(NlsarImageDesc* d, int n, int* rows_c, int* rows_a, int* blocks,
                        int* max_cw, int* max_aw);
// Tile-queue counters a persistent batch needs (8 per pass, NLSAR_MAX_PASSES passes).
constexpr int NLSAR_MAX_PASSES = 64;
constexpr size_t NLSAR_QUEUE_BYTES = sizeof(unsigned int) * 8 * NLSAR_MAX_PASSES;
// Enqueues the filter on `stream`.  d_table: device copy of descs[0..n) (after
// nlsar_batch_layout), or null for a single image (n == 1), which is passed by value.
// d_const: device copy of nlsar_build_constants' blob, unchanged until the work has run.
// d_queues: NLSAR_QUEUE_BYTES of device memory for a batch's persistent passes (cleared
// here by a kernel), or null for one tile per block.
int nlsar_enqueue(const NlsarParams& pp, const NlsarCommon& cm, const NlsarImageDesc* descs,
                  int n, const NlsarImageDesc* d_table, const float* d_const, hipStream_t stream,
                  unsigned int* d_queues = nullptr);

}  // namespace dkx
