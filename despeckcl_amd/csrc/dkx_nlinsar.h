// Host interface of the NL-InSAR HIP path.
#pragma once

#include "dkx_common.h"

#include <vector>

namespace dkx {

// Which weights the weights launches of a single call store for the lmin smoothing, which
// needs the weights of the pixels it queues (ENL < lmin); the queued pixels whose weights
// were not stored get them recomputed, the same bits, so the setting changes the memory
// traffic and not the result.  Automatic: see nli_store_mode (nlinsar.hip); batches follow
// the same rule (nlinsar_batch_enqueue).
constexpr int NLINSAR_STORE_AUTO = -1;
constexpr int NLINSAR_STORE_ALL = 0;    // every weight
constexpr int NLINSAR_STORE_PRED = 1;   // the pixels the previous iteration queued and their
                                        // neighbours (none in the first iteration)
constexpr int NLINSAR_STORE_NONE = 2;   // none: every queued pixel's weights recomputed

struct NlinsarProblem {
    int H, W;
    int sw, p;
    int niter, lmin;
    float h, T;       // weight parameters (nlinsar.cpp:65-78)
    int ref_tile;     // side of the reference's fixed tiles (80, nlinsar.cpp:106)
    int tile;         // tile shape index, or NLINSAR_TILE_AUTO
    int thin_edges = 0;   // 1: the last row / column in thin blocks of their own (batches)
    int store = NLINSAR_STORE_AUTO;   // which weights the weights launches store (NLINSAR_STORE_*)
};

// Tile shape of the built-in rule: 18 x 54 with 512 threads (the autotuner may record
// another per parameter set and image size).
constexpr int NLINSAR_TILE_AUTO = -1;

// Block shapes of the weights kernel per launch mapping.  Shape 0 is the tile shape.  The
// last row and the corner (shape 1) and the last column (shape 2), which the reference's
// tiling computes at virtual positions one row / column wide, either run in shape 0 with
// the interior (single calls: one launch, the edge blocks' latency hidden among the
// interior's) or, for batches (NlinsarProblem::thin_edges), in thin blocks of their own
// (1 x 54 and 54 x 3), whose launches cover the edges of all images at once.
// largest patch half width the weights kernels are instantiated for (patch sizes 1-21)
constexpr int NLI_MAX_PSH = 10;
constexpr int NLI_SHAPES = 3;
// A queued pixel's weight slot: the shape in the bits from NLI_SLOT_BITS up, below them the
// slot within that shape's weight region.
constexpr int NLI_SLOT_BITS = 29;
// Slot tag of a compact weight slot (single calls that store only some pixels' weights): the
// sw*sw weights of the pixel contiguous at slot * sw*sw of the shape-0 weight region.
constexpr int NLI_COMPACT = NLI_SHAPES;

struct NlinsarPlan {
    int tile, TY, TX;
    int wsh, psh, overlap, step;   // step = ref_tile - 2*overlap (interior of a reference tile)
    int eh, ew;                    // virtual distance of the last row / column
    Grid R, E, F;                  // float4 grids: raw (a1, a2, phase), estimates (ref, phase,
                                   // coh); 12-byte records: filter values (a, Re x, Im x)
    int heap_block;                // threads per block of the smoothing kernel
    int blocks[NLI_SHAPES];        // blocks of the weights kernel per shape (weight / cache chunks)
    size_t off_w[NLI_SHAPES], off_s[NLI_SHAPES];   // weight and similarity-cache regions per shape
    size_t off_R, off_A, off_E, off_F, off_est, off_pred, off_caux, off_list, off_xlist, off_cnt, ws_bytes;
};

int nlinsar_plan(const NlinsarProblem& pb, NlinsarPlan* plan);

// Seed of the Monte-Carlo h for patch sizes other than 3 and 7 (std::mt19937's default).
constexpr unsigned int NLINSAR_H_SEED = 5489u;
// quantile_insar (insarsim_simu.cpp:62-82) with a seeded generator (nlinsar_simu.cpp).
float nlinsar_simulate_h(int patch_size, float alpha, unsigned int seed);
// h_theo of nlinsar.cpp:65-76: 0.58 for patch 3, 0.244 for patch 7, else the simulation at
// alpha 0.92 with NLINSAR_H_SEED (computed once per patch size).
float nlinsar_h_theo(int patch_size);

// The reference's parameters for a search window, patch size, iteration count and lmin
// (h and T from the patch size, 80-pixel tiles), with the given tile shape.
int nlinsar_problem(int height, int width, int sw, int p, int niter, int lmin, int tile,
                    NlinsarProblem* pb);

// d_ref/d_phase/d_coh: in = first-iteration estimates, out = result.
int nlinsar_run(const NlinsarProblem& pb, const NlinsarPlan& plan, const float* d_a1,
                const float* d_a2, const float* d_phi, float* d_ref, float* d_phase,
                float* d_coh, void* d_ws, unsigned long long* n_smoothed_host,
                hipStream_t stream);


// ---- batches: many images with the same parameters, each stage one launch over all ----

// One image of a batch: device arrays and its workspace (NlinsarPlan::ws_bytes).
struct NlinsarBatchItem {
    const float* a1; const float* a2; const float* phi;
    float* ref; float* phase; float* coh;   // in: first-iteration estimates, out: the result
    int H, W;
    char* ws;
};

// What a batch's launches need on the host, and the layout of its device tables.
struct NlinsarBatchLaunch {
    int nimg, niter, tile, psh, sw, lmin;
    int store;              // NlinsarProblem::store of the batch
    int total_tiles[NLI_SHAPES];   // weights tiles of all images, per shape
    int persistent_blocks;  // grid of the persistent weights kernel
    long long max_n;        // pixels of the largest image
    int heap_block;
    size_t off_img, off_w, off_s, off_q, bytes;   // device tables, relative to their base:
                                                  // weights args [shape][kind][image],
                                                  // smoothing args [kind][image], 8 queue
                                                  // counters per shape
};

// Per-image plans of a batch (all with the batch's tile shape; their ws_bytes give each
// image's workspace) and the launch description.  common: the parameters (H, W unused).
int nlinsar_batch_plan(const NlinsarProblem& common, const int* H, const int* W, int n,
                       std::vector<NlinsarPlan>& plans, NlinsarBatchLaunch* L);
// Host image of the device tables for tables placed at d_tables (L.bytes of them).
int nlinsar_batch_tables(const NlinsarProblem& common, const std::vector<NlinsarPlan>& plans,
                         const NlinsarBatchItem* items, const NlinsarBatchLaunch& L,
                         char* d_tables, std::vector<unsigned char>& host);
// Enqueues every iteration of every image on `stream` (the tables must be on the device);
// zero_init: first clear the images' estimate arrays (a run from zero estimates).
int nlinsar_batch_enqueue(const NlinsarProblem& common, const NlinsarBatchLaunch& L,
                          char* d_tables, hipStream_t stream, int zero_init = 0);
// Bytes of the device tables of a batch of n images (NlinsarBatchLaunch::bytes).
size_t nlinsar_batch_table_bytes(int n);
// Upper bound of the workspace of an image under any tile shape a batch may pick.
int nlinsar_batch_image_bytes(const NlinsarProblem& common, int H, int W, size_t* bytes);

}  // namespace dkx
