// Host interface of the NL-InSAR HIP path.
#pragma once

#include "dkx_common.h"

namespace dkx {

struct NlinsarProblem {
    int H, W;
    int sw, p;
    int niter, lmin;
    float h, T;       // weight parameters (nlinsar.cpp:65-78)
    int ref_tile;     // side of the reference's fixed tiles (80, nlinsar.cpp:106)
    int tile;         // tile shape index, or NLINSAR_TILE_AUTO
};

// Tile shape chosen from the image size: 54 x 54 with 1024 threads from 2^20 pixels up,
// 18 x 54 with 512 threads below (where the larger blocks would leave CUs idle).
constexpr int NLINSAR_TILE_AUTO = -1;

struct NlinsarPlan {
    int tile, TY, TX;
    int wsh, psh, overlap, step;   // step = ref_tile - 2*overlap (interior of a reference tile)
    int eh, ew;                    // virtual distance of the last row / column
    Grid R, E, F;                  // float4 grids: raw (a1, a2, phase), estimates (ref, phase,
                                   // coh), filter values (a, Re x, Im x)
    int heap_block;                // threads per block of the smoothing kernel
    int blocks;                    // blocks of the weights kernel (weight / cache chunks)
    size_t off_R, off_A, off_E, off_F, off_est, off_w, off_s, off_list, off_xlist, off_cnt, ws_bytes;
};

int nlinsar_plan(const NlinsarProblem& pb, NlinsarPlan* plan);
// d_ref/d_phase/d_coh: in = first-iteration estimates, out = result.
int nlinsar_run(const NlinsarProblem& pb, const NlinsarPlan& plan, const float* d_a1,
                const float* d_a2, const float* d_phi, float* d_ref, float* d_phase,
                float* d_coh, void* d_ws, unsigned long long* n_smoothed_host,
                hipStream_t stream);

}  // namespace dkx
