// Host interface of the Goldstein filter (goldstein.hip).
#pragma once

#include "dkx_common.h"

namespace dkx {

struct GoldsteinParams {
    int height, width, patch_size, overlap;
    float alpha;
};

// Tiling of the reference (at most two tiles per direction) and the patches each tile
// transforms; ws_bytes: the weighted patches kept between the kernels (u_bytes), then the raw
// interferogram of the image (8 bytes per pixel).
struct GoldsteinPlan {
    struct Tile {
        int h_low, w_low;
        int py0, npy, px0, npx;
        long long u_off;
        int job0, jobs;
    };
    int P, ov, m, T, n, step, nty, ntx, jobs;
    Tile tile[4];
    size_t u_bytes, ws_bytes;
};

int goldstein_plan(const GoldsteinParams& gp, GoldsteinPlan* pl);
// ref / phase: device outputs; ws: at least pl.ws_bytes of device memory
int goldstein_run(const GoldsteinParams& gp, const GoldsteinPlan& pl, const float* a1,
                  const float* a2, const float* phi, float* ref, float* phase, void* ws,
                  hipStream_t stream);

}  // namespace dkx
