// The fused NL-SAR scale pass: for one scale and up to three patch sizes, evaluates per
// search-window offset the pixel dissimilarities, the separable patch box sums, the LUT
// weights and the nonlocal statistics; the last pass also forms the weighted mean, the
// alpha blend and the decomposition.  Included by the translation units that instantiate
// it (one per (last pass, tile width) pair, so that they compile in parallel).
//
// Reference path replaced: src/compute/nlsar/nlsar_filter_sub_image.cpp:146-390 with the
// kernels of src/compute/nlsar/cl_kernels (one materialised 220/441-plane buffer per stage
// there; here every stage of one search-window offset stays in LDS).
#pragma once

#include "dkx_common.h"
#include "dkx_nlsar.h"

#include <algorithm>
#include <type_traits>

namespace dkx {

// Patch sets are compile-time bit masks of patch half widths (bit k: patch size 2k+1), so
// that every box sum is a straight chain of additions.
__host__ __device__ constexpr int pmask_count(int m) { return m ? (m & 1) + pmask_count(m >> 1) : 0; }
__host__ __device__ constexpr int pmask_top(int m) { return m > 1 ? 1 + pmask_top(m >> 1) : 0; }
// half width of slot k (slots in ascending patch size)
__host__ __device__ constexpr int pmask_bit(int m, int k, int b = 0)
{
    return (m & 1) ? (k == 0 ? b : pmask_bit(m >> 1, k - 1, b + 1)) : pmask_bit(m >> 1, k, b + 1);
}

constexpr int NLSAR_BAND = 8;   // tile rows per band of the block order

struct ScaleArgs {
    const NlsarImageDesc* table;         // batch: descriptors in device memory, else null
    NlsarImageDesc one;                  // the image of a single call (table == null)
    int nimg;
    int total_blocks;                    // blocks of all images
    int wsh;                             // search-window half width
    int pidx[NLSAR_MAX_NP];              // index of (p, s) in the per-parameter maps
    const float* lutw[NLSAR_MAX_NP];     // weight per quantile index
    float dmin[NLSAR_MAX_NP], dmax[NLSAR_MAX_NP];
    float rng[NLSAR_MAX_NP];             // dmax - dmin
    float rcp[NLSAR_MAX_NP];             // 1 / rng, correctly rounded
    float scl[NLSAR_MAX_NP];             // lut_size - 1
    int lut_size[NLSAR_MAX_NP];
    int lut_stride;                      // LDS floats per LUT
    int first_scale;                     // 1: first pass (ubflag not yet written)
    int alpha_zero;                      // last pass: blend with alpha 0 (NlsarParams::last_unused)
    unsigned int* queue;                 // batch: 8 tile counters (zero before the launch) of a
                                         // persistent grid; null: one tile per block
};

// LDS layout (floats) of a pass over TX x TX output tiles, shared by the kernel and its
// launcher, fixed at compile time for search windows of half width up to WMAX so that every
// LDS access takes an immediate offset.  XU/YU: the union region of a tile and the tile
// shifted by an offset, rounded up to runs of 4.
template <int TX, int P, int WMAX>
struct PassLds {
    static constexpr int W2 = 4 + 2 * P;          // a run of 4 plus the patch halo
    static constexpr int W2R = (W2 + 3) & ~3;     // the same in float4 loads
    static constexpr int XU = (TX + WMAX + 3) & ~3;
    static constexpr int YU = (TX + WMAX + 3) & ~3;
    static constexpr int DXM = XU + W2R;           // dissimilarity row
    static constexpr int DYM = TX + WMAX + 2 * P;
    static constexpr int RXM = XU;                 // row-sum / weight row
    static constexpr int RRows = YU + 2 * P;
    static constexpr int WRows = YU;
    static constexpr int RPL = RRows * RXM;        // row-sum plane
    static constexpr int WPL = WRows * RXM;        // weight plane
    // the dissimilarities (steps 1-2) and the weights (steps 3-4) share one region; one
    // such region and one row-sum region per offset slot of a barrier phase
    static constexpr int shared_floats(int np) { return DYM * DXM > np * WPL ? DYM * DXM : np * WPL; }
    static constexpr int off_D(int np, int slot) { return slot * shared_floats(np); }
    static constexpr int off_R(int np, int opb, int slot) { return opb * shared_floats(np) + slot * np * RPL; }
    static constexpr int off_L(int np, int opb) { return off_R(np, opb, opb); }
    static constexpr int total(int np, int opb, int lut_stride) { return off_L(np, opb) + np * lut_stride; }
};

constexpr size_t NLSAR_LDS_PER_CU = 160 * 1024;   // gfx950

// Search-window offsets per barrier phase of the scale pass: two for one patch size (more
// independent work between barriers, in LDS a single patch size leaves free), else one.
// Dimension-3 covariance input keeps one: its register-heavy pass ran 4 % slower with two.
__host__ __device__ constexpr int nlsar_offsets_per_phase(int np, int mode)
{
    return np == 1 && mode != 3 ? 2 : 1;
}

// Window classes: the largest search-window half width each compiled layout serves.
constexpr int NLSAR_WMAX_SMALL = 10;   // search windows up to 21
constexpr int NLSAR_WMAX_LARGE = 20;   // search windows up to 41

// Natural log of a positive normal float: the device library's logf without its subnormal
// branch (log2 by v_log_f32, then the product with ln 2 carried in two parts), so the result
// is the same bits for every normal input; non-finite log2 values pass through as the
// plain product.  Determinants of averaged covariances are never subnormal in practice
// (a subnormal input reads as 0 here and gives -inf).
__device__ __forceinline__ float nlsar_logf(float x)
{
    const float l2 = __builtin_amdgcn_logf(x);
    const float hi = l2 * 0x1.62e42ep-1f;            // ln 2, leading part
    float e = fmaf(l2, 0x1.62e42ep-1f, -hi);
    e = fmaf(l2, 0x1.efa39ep-25f, e);                // ln 2, trailing part
    return fabsf(l2) < __builtin_inff() ? hi + e : hi;
}

// Reciprocal of a small positive integer d by v_rcp_f32 (1 ulp).  Index quotients are
// formed as q = (unsigned)((n + 1/2) * rcp(d)) for 0 <= n < 2^16: the exact (n + 1/2) / d lies
// at least 1/(2d) from an integer, while the float product errs by at most
// (n + 1/2) / d * 1.5 * 2^-23, below 1/(2d) for every n < 2^21, so the truncation is exact.
__device__ __forceinline__ float nlsar_rcp(int d) { return __builtin_amdgcn_rcpf((float)d); }

// 2*dim*log(2) for dim 2, the constant term of pixel_similarity_2x2 (sim_measures.h:58)
__device__ __forceinline__ float nlsar_log2_term() { return 4.0f * 0.693147182464599609375f; }

// LUT weight of a patch dissimilarity (compute_weights.cl:37-52): 0 above the trained
// maximum, else the table entry of index (uint)((d - min) / (max - min) * (lut - 1)).  The
// quotient is formed as q0 = a * (1/rng) followed by one FMA residual correction, which
// gives the correctly rounded quotient for every normal result (Markstein); subnormal
// quotients can differ by an ulp but truncate to index 0 either way.
__device__ __forceinline__ float lut_weight(float s, float dmin, float dmax, float rng,
                                            float rcp, float scl, const float* lut)
{
    // the clamp as one median; a NaN s gives dmin (or NaN) and index 0 either way, as the
    // min/max form does
    const float d = __builtin_amdgcn_fmed3f(s, dmin, dmax);
    const float a = d - dmin;
    const float q0 = a * rcp;
    const float e = fmaf(-q0, rng, a);
    const float q = fmaf(e, rcp, q0);
    const float w = lut[(unsigned int)(q * scl)];
    return s > dmax ? 0.f : w;
}

__device__ __forceinline__ const NlsarImageDesc& nlsar_image(const ScaleArgs& a, int img)
{
    return a.table ? a.table[img] : a.one;
}

// What the offset loop of one block needs to know about its image.
struct PassImage {
    const DKX_GLOBAL char* A4b;   // A grid (float4 per element)
    const DKX_GLOBAL char* LDb;   // log det plane of the A grid
    const DKX_GLOBAL char* C4b;   // C grid (float4 per element)
    int apitch, aorg;     // aorg: element offset of image position (0, 0) in the A grid
    int cpitch, corg;
    int img, m;           // image index, launch mapping of the tile
};

// Tile t -> (image, tile).  The images' tiles are laid end to end, each image's four launch
// mappings end to end within it.  The tiles the blocks of one XCD (block id mod 8) take are
// a contiguous run (the x-th eighth for XCD x), and tiles are ordered column by column within
// bands of NLSAR_BAND tile rows, so the blocks resident on one XCD cover a compact 2-D patch
// whose halos overlap in that XCD's L2.  One tile per block: block b takes tile k = b / 8 of
// its XCD's run.  Persistent grid (batches): the blocks of an XCD take the tiles of its run
// from its queue counter until the run is exhausted; every XCD has blocks (grid >= 8).
__device__ __forceinline__ int nlsar_xcd_tile(const ScaleArgs& a, int k)
{
    const int nb = a.total_blocks;
    const int q = nb >> 3, r = nb & 7;
    const int xcd = blockIdx.x & 7;
    return k < q + (xcd < r ? 1 : 0) ? xcd * q + min(xcd, r) + k : -1;
}

// The next tile of a persistent block, -1 when its XCD's run is exhausted.  slot: one int of
// LDS; the barriers also order the previous tile's last LDS reads before the next tile's
// writes.
__device__ __forceinline__ int nlsar_next_tile(const ScaleArgs& a, int* slot)
{
    if (threadIdx.x == 0) *slot = (int)atomicAdd(a.queue + (blockIdx.x & 7), 1u);
    __syncthreads();
    const int k = uni(*slot);
    __syncthreads();
    return nlsar_xcd_tile(a, k);
}

__device__ __forceinline__ void nlsar_block_tile(const ScaleArgs& a, int t, int TX, PassImage& v, int& ty0,
                                                 int& tx0)
{
    int img = 0;
    if (a.table) {
        int lo = 0, hi = a.nimg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (uni(a.table[mid].block0) <= t) lo = mid;
            else hi = mid - 1;
        }
        img = lo;
    }
    const NlsarImageDesc& d = nlsar_image(a, img);
    const int tl = t - uni(d.block0);
    int m = 0;
    while (m < 3 && tl >= uni(d.bstart[m + 1])) m++;
    const int local = tl - uni(d.bstart[m]);
    const int gx = uni(d.gx[m]), gy = uni(d.gy[m]);
    const int band = local / (NLSAR_BAND * gx);
    const int within = local - band * NLSAR_BAND * gx;
    const int bh = min(NLSAR_BAND, gy - band * NLSAR_BAND);
    const int bx = within / bh;
    const int by = band * NLSAR_BAND + (within - bx * bh);
    ty0 = uni(d.L[m].vy0) + by * TX;
    tx0 = uni(d.L[m].vx0) + bx * TX;
    v.img = img;
    v.m = m;
    v.A4b = gptr(uni(reinterpret_cast<const char*>(d.A)));
    v.LDb = v.A4b + 16ull * (unsigned long long)uni(d.aplane);
    v.C4b = gptr(uni(reinterpret_cast<const char*>(d.C)));
    v.apitch = uni(d.apitch);
    v.aorg = uni(d.aoy) * v.apitch + uni(d.aox);
    v.cpitch = uni(d.cpitch);
    v.corg = uni(d.coy) * v.cpitch + uni(d.cox);
}

// Occupancy target: blocks per CU (three fit the LDS budget of the small window class; two
// for single-row threads, whose 16-wave blocks would otherwise leave 40 registers), which
// bounds the registers the compiler may use.  Covariance input holds more statistics and,
// in the last pass, every plane's weighted sum: fewer blocks leave room for them.
constexpr int nlsar_blocks_per_cu(int TX, int RPT, int WMAX, int MODE, bool LAST)
{
    return WMAX > NLSAR_WMAX_SMALL ? 1
         : MODE == 3 && LAST     ? 1
         : MODE != 0             ? 2
         : RPT == 1              ? 2
                                 : 3;
}

// Covariance of a 3x3 Hermitian matrix (det_covmat_3x3, sim_measures.h:79-97), in the
// reference's operation order.
__device__ __forceinline__ float nlsar_det3(float a00, float a11, float a22, float a01r, float a01i,
                                            float a02r, float a02i, float a12r, float a12i)
{
    return a00 * a11 * a22 - a00 * (a12r * a12r + a12i * a12i) - a11 * (a02r * a02r + a02i * a02i) -
           a22 * (a01r * a01r + a01i * a01i) +
           2 * (a01r * a12r * a02r - a01i * a12i * a02r + a01r * a12i * a02i + a01i * a12r * a02i);
}

// MODE: 0 InSAR input, 2 / 3 covariance input of that dimension (all planes filtered).
// PERSIST: a persistent grid over a batch's tile queues (ScaleArgs::queue), else one tile per
// block.
template <int TX, int RPT, int PMASK, bool LAST, int WMAX, int MODE, bool PERSIST>
__global__ __launch_bounds__(TX * TX / RPT, nlsar_blocks_per_cu(TX, RPT, WMAX, MODE, LAST) * TX * TX / RPT / 256)
void nlsar_scale_pass(const ScaleArgs a)
{
    constexpr int CF4 = nlsar_c_float4(MODE);    // float4 per C element
    constexpr int AF4 = nlsar_a_float4(MODE);    // float4 per A element
    constexpr int DIM = MODE == 0 ? 2 : MODE;
    constexpr int NSTAT = 2 + 2 * DIM;           // sum w, sum w^2, sum w I_d, sum w I_d^2
    constexpr int NPL = 2 * DIM * DIM;           // planes of a covariance input
    // planes averaged by the weighted mean besides the diagonal intensities
    constexpr int NMACC = MODE == 0 ? 2 : NPL - DIM;
    constexpr int NT = TX * TX / RPT;   // each thread owns RPT rows of one column
    constexpr int TY = TX;              // square tiles
    constexpr int NP = pmask_count(PMASK);
    constexpr int P = pmask_top(PMASK);
    using Lay = PassLds<TX, P, WMAX>;
    constexpr int W2 = Lay::W2, W2R = Lay::W2R;
    constexpr int DXM = Lay::DXM, RXM = Lay::RXM, RPL = Lay::RPL, WPL = Lay::WPL;
    constexpr int DYM_ = Lay::DYM;
    extern __shared__ float4 smem4[];
    float* smem = reinterpret_cast<float*>(smem4);
    const int wsh = a.wsh;
    constexpr int OPB = nlsar_offsets_per_phase(NP, MODE);
    float* sL = smem + Lay::off_L(NP, OPB);

    const int tid = threadIdx.x;
    const int px = tid % TX, g = tid / TX;  // column, group of RPT rows
    const float ft = (float)tid + 0.5f;     // numerator of the index quotients
    const unsigned int tid4 = 4u * tid, tid8 = 8u * tid, tid16 = 16u * tid;
#pragma unroll
    for (int k = 0; k < NP; k++)
        for (int i = tid; i < a.lut_size[k]; i += NT) sL[k * a.lut_stride + i] = a.lutw[k][i];
    // persistent grid: the tile slot follows the LUTs
    int* const tslot = reinterpret_cast<int*>(sL + NP * a.lut_stride);
    int tq = PERSIST ? -1 : nlsar_xcd_tile(a, blockIdx.x >> 3);

    for (;;) {
        if constexpr (PERSIST) {
            tq = nlsar_next_tile(a, tslot);
            if (tq < 0) break;
        }
        PassImage im;
        int ty0, tx0;
        nlsar_block_tile(a, tq, TX, im, ty0, tx0);

        float acc[RPT][NP][NSTAT];
        float macc[RPT][LAST ? NMACC : 1];
#pragma unroll
        for (int q = 0; q < RPT; q++) {
#pragma unroll
            for (int k = 0; k < NP; k++)
#pragma unroll
                for (int j = 0; j < NSTAT; j++) acc[q][k][j] = 0.f;
#pragma unroll
            for (int j = 0; j < (LAST ? NMACC : 1); j++) macc[q][j] = 0.f;
        }

        const float K = (float)(2 * DIM) * 0.693147182464599609375f;   // 2*dim*log(2), sim_measures.h:58
        // C-grid element and weight-plane position of the thread's first pixel
        const unsigned int c_own = (unsigned int)(im.corg + (ty0 + g * RPT) * im.cpitch + (tx0 + px));
        const int w_own = g * RPT * RXM + px;
        const int apitch = im.apitch, cpitch = im.cpitch;
        const DKX_GLOBAL char* A4b = im.A4b;
        const DKX_GLOBAL char* LDb = im.LDb;
        const DKX_GLOBAL char* C4b = im.C4b;

        // The four steps of one half-window offset d = (dy, dx), each on its own LDS buffers.
        auto step1 = [&](const int dy, const int dx, float* sD) {
            [[maybe_unused]] const int adx = dx < 0 ? -dx : dx;
            // union of the tile (for w(c, d)) and the tile shifted by -d (for w(c, -d))
            [[maybe_unused]] const int RY = TY - dy, RX = TX + adx;
            [[maybe_unused]] const int rxo = dx > 0 ? -dx : 0;   // region x origin relative to tx0
            [[maybe_unused]] const int DY = RY + 2 * P, DX = RX + 2 * P;
            // 1. pixel dissimilarities delta(g, g + d) of the later pixel g against the
            //    earlier one (pixel_similarity_2x2, sim_measures.h:37-65).  Element i of the
            //    DY x DX region: row r, column c.  Grid and LDS offsets are taken from
            //    wave-uniform bases of this offset: element i lies r * (pitch - DX) + i
            //    elements past the region's first element, in the grid as in LDS.
            {
                const int n = DY * DX;
                const float inv = nlsar_rcp(DX);
                const int gb0 = im.aorg + (ty0 - P) * apitch + (tx0 + rxo - P);
                const int dofs = dy * apitch + dx;
                const DKX_GLOBAL char* const Au = A4b + (long long)(16 * AF4) * gb0;
                const DKX_GLOBAL char* const Av = Au + (long long)(16 * AF4) * dofs;
                const DKX_GLOBAL char* const Lu = LDb + 4ll * gb0;
                const DKX_GLOBAL char* const Lv = Lu + 4ll * dofs;
                const unsigned int gstep = (unsigned int)(apitch - DX);       // grid pitch >= DX
                const unsigned int sstep = (unsigned int)(4 * (DXM - DX));    // bytes
                // elements tid + k*NT, at most KMAX of them, unrolled
                constexpr int KMAX = (DYM_ * (TX + WMAX + 2 * P) + NT - 1) / NT;
#pragma unroll
                for (int k = 0; k < KMAX; k++) {
                    const int i = tid + k * NT;
                    if (i >= n) break;
                    const unsigned int r = (unsigned int)((ft + (float)(k * NT)) * inv);
                    const unsigned int rel = __umul24(r, gstep) + (unsigned int)i;
                    float sim;
                    if constexpr (MODE == 3) {
                        // pixel_similarity_3x3, sim_measures.h:100-158
                        const unsigned int o = 48u * rel;
                        const float4 u0 = gload4(Au + o), u1 = gload4(Au + o + 16u);
                        const float4 u2 = gload4(Au + o + 32u);
                        const float4 v0 = gload4(Av + o), v1 = gload4(Av + o + 16u);
                        const float4 v2 = gload4(Av + o + 32u);
                        const float det = nlsar_det3(u0.x + v0.x, u0.y + v0.y, u0.z + v0.z, u0.w + v0.w,
                                                     u1.x + v1.x, u1.y + v1.y, u1.z + v1.z, u1.w + v1.w,
                                                     u2.x + v2.x);
                        sim = -(((K + u2.y) + v2.y) - 2.0f * nlsar_logf(det));
                    } else {
                        const float4 u = gload4(Au + 16u * rel);
                        const float4 v = gload4(Av + 16u * rel);
                        const float l1 = *reinterpret_cast<const DKX_GLOBAL float*>(Lu + 4u * rel);
                        const float l2 = *reinterpret_cast<const DKX_GLOBAL float*>(Lv + 4u * rel);
                        const float s00 = u.x + v.x;
                        const float s01r = u.z + v.z;
                        const float s01i = u.w + v.w;
                        const float s11 = u.y + v.y;
                        const float det = (s00 * s11) - (s01r * s01r + s01i * s01i);
                        sim = -(((K + l1) + l2) - 2.0f * nlsar_logf(det));
                    }
                    if (isnan(sim)) sim = 0.f;
                    *reinterpret_cast<float*>(reinterpret_cast<char*>(sD) + (__umul24(r, sstep) + tid4) +
                                              k * NT * 4) = sim;
                }
            }
        };
        auto step2 = [&](const int dy, const int dx, float* sD, float* sR) {
            [[maybe_unused]] const int adx = dx < 0 ? -dx : dx;
            // union of the tile (for w(c, d)) and the tile shifted by -d (for w(c, -d))
            [[maybe_unused]] const int RY = TY - dy, RX = TX + adx;
            [[maybe_unused]] const int rxo = dx > 0 ? -dx : 0;   // region x origin relative to tx0
            [[maybe_unused]] const int DY = RY + 2 * P, DX = RX + 2 * P;
            // 2. row sums (patch_similarities_row_pass.cl:52-60), left to right, of every
            //    patch size, for runs of 4 columns held in registers.  Element i: row r, run
            //    xi = i - r * nrx; its LDS byte offsets are r * (row bytes - 16 nrx) + 16 i.
            {
                const int nrx = (RX + 3) >> 2;
                const int n = DY * nrx;
                const float inv = nlsar_rcp(nrx);
                const unsigned int sstep = (unsigned int)(4 * DXM - 16 * nrx);
                const unsigned int rstep = (unsigned int)(4 * RXM - 16 * nrx);
                constexpr int KR = (DYM_ * ((TX + WMAX + 3) / 4) + NT - 1) / NT;
#pragma unroll
                for (int kk = 0; kk < KR; kk++) {
                    const int i = tid + kk * NT;
                    if (i >= n) break;
                    const unsigned int r = (unsigned int)((ft + (float)(kk * NT)) * inv);
                    float w[W2R];
                    const float4* src = reinterpret_cast<const float4*>(
                        reinterpret_cast<const char*>(sD) + (__umul24(r, sstep) + tid16) + kk * NT * 16);
#pragma unroll
                    for (int t = 0; t < W2R / 4; t++) {
                        const float4 v = src[t];
                        w[4 * t] = v.x;
                        w[4 * t + 1] = v.y;
                        w[4 * t + 2] = v.z;
                        w[4 * t + 3] = v.w;
                    }
                    float* dst = reinterpret_cast<float*>(reinterpret_cast<char*>(sR) + (__umul24(r, rstep) + tid16) +
                                                          kk * NT * 16);
#pragma unroll
                    for (int k = 0; k < NP; k++) {
                        const int psh = pmask_bit(PMASK, k);
                        float o[4];
#pragma unroll
                        for (int m = 0; m < 4; m++) {
                            float sum = w[m + P - psh];
#pragma unroll
                            for (int jj = -P + 1; jj <= P; jj++)
                                if (jj > -psh && jj <= psh) sum = sum + w[m + P + jj];
                            o[m] = sum;
                        }
                        *reinterpret_cast<float4*>(dst + k * RPL) = make_float4(o[0], o[1], o[2], o[3]);
                    }
                }
            }
        };
        auto step3_slot = [&](const int dy, const int dx, float* sR, float* sW) {
            [[maybe_unused]] const int adx = dx < 0 ? -dx : dx;
            // union of the tile (for w(c, d)) and the tile shifted by -d (for w(c, -d))
            [[maybe_unused]] const int RY = TY - dy, RX = TX + adx;
            [[maybe_unused]] const int rxo = dx > 0 ? -dx : 0;   // region x origin relative to tx0
            [[maybe_unused]] const int DY = RY + 2 * P, DX = RX + 2 * P;
            // 3. (one patch size, one offset slot at a time) column sums and LUT weights as in
            //    step3 below, element i of the slot taken by thread i.
            {
                const int nrx = (RX + 1) >> 1;
                const int nry = (RY + 3) >> 2;
                const int n = nry * nrx;
                const float inv = nlsar_rcp(nrx);
                const unsigned int ostep = (unsigned int)(16 * RXM - 8 * nrx);
                constexpr int KC = (((TX + WMAX + 3) / 4) * ((TX + WMAX + 1) / 2) + NT - 1) / NT;
#pragma unroll
                for (int kk = 0; kk < KC; kk++) {
                    const int i = tid + kk * NT;
                    if (i >= n) break;
                    const unsigned int ry = (unsigned int)((ft + (float)(kk * NT)) * inv);
                    const unsigned int offb = __umul24(ry, ostep) + tid8 + (unsigned int)(kk * NT * 8);
#pragma unroll
                    for (int k = 0; k < NP; k++) {
                        const int psh = pmask_bit(PMASK, k);
                        const float* src = reinterpret_cast<const float*>(reinterpret_cast<const char*>(sR + k * RPL) + offb);
                        float w0[W2], w1[W2];
#pragma unroll
                        for (int t = 0; t < W2; t++) {
                            const float2 v = *reinterpret_cast<const float2*>(src + t * RXM);
                            w0[t] = v.x;
                            w1[t] = v.y;
                        }
                        const float* lut = sL + k * a.lut_stride;
                        float* dst = reinterpret_cast<float*>(reinterpret_cast<char*>(sW + k * WPL) + offb);
#pragma unroll
                        for (int m = 0; m < 4; m++) {
                            float s0 = w0[m + P - psh], s1 = w1[m + P - psh];
#pragma unroll
                            for (int jj = -P + 1; jj <= P; jj++)
                                if (jj > -psh && jj <= psh) {
                                    s0 = s0 + w0[m + P + jj];
                                    s1 = s1 + w1[m + P + jj];
                                }
                            const float x0 = lut_weight(s0, a.dmin[k], a.dmax[k], a.rng[k], a.rcp[k], a.scl[k], lut);
                            const float x1 = lut_weight(s1, a.dmin[k], a.dmax[k], a.rng[k], a.rcp[k], a.scl[k], lut);
                            *reinterpret_cast<float2*>(dst + m * RXM) = make_float2(x0, x1);
                        }
                    }
                }
            }
        };
        // 3. column sums (patch_similarities_col_pass.cl:52-60), top to bottom, for runs of 4
        //    rows in pairs of columns, and the LUT weight (compute_weights.cl:19-56), of every
        //    offset slot of the phase and every patch size.  The (slot, patch) groups of
        //    elements are laid end to end, each padded to whole waves, so every wave of the
        //    block takes a share and works on one group (its patch size is a compile-time
        //    branch, uniform in the wave).  Element j of a group: run ry, column pair
        //    xi = j - ry * nrx, at LDS byte offset ry * (16 row bytes - 8 nrx) + 8 j.
        auto step3 = [&](const int* oys, const int* oxs) {
            constexpr int NG = OPB * NP;
            int n1[OPB], base[NG];
            unsigned int ostep[OPB];
            float inv[OPB];
            int total = 0;
#pragma unroll
            for (int o = 0; o < OPB; o++) {
                const int ady = oys[o] < 0 ? -oys[o] : oys[o];
                const int adx = oxs[o] < 0 ? -oxs[o] : oxs[o];
                const int RY = TY + ady, RX = TX + adx;
                const int nrx = (RX + 1) >> 1;
                n1[o] = ((RY + 3) >> 2) * nrx;
                inv[o] = nlsar_rcp(nrx);
                ostep[o] = (unsigned int)(16 * RXM - 8 * nrx);
#pragma unroll
                for (int k = 0; k < NP; k++) {
                    base[o * NP + k] = total;
                    total += (n1[o] + 63) & ~63;
                }
            }
            auto body = [&](auto kc, const int o, const unsigned int offb) {
                constexpr int k = decltype(kc)::value;
                constexpr int psh = pmask_bit(PMASK, k);
                float* sR = smem + Lay::off_R(NP, OPB, 0) + o * (Lay::off_R(NP, OPB, 1) - Lay::off_R(NP, OPB, 0));
                float* sW = smem + o * (Lay::off_D(NP, 1) - Lay::off_D(NP, 0));
                const float* src = reinterpret_cast<const float*>(reinterpret_cast<const char*>(sR + k * RPL) + offb);
                float w0[W2], w1[W2];
#pragma unroll
                for (int t = 0; t < W2; t++) {
                    const float2 v = *reinterpret_cast<const float2*>(src + t * RXM);
                    w0[t] = v.x;
                    w1[t] = v.y;
                }
                const float* lut = sL + k * a.lut_stride;
                float* dst = reinterpret_cast<float*>(reinterpret_cast<char*>(sW + k * WPL) + offb);
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    float s0 = w0[m + P - psh], s1 = w1[m + P - psh];
#pragma unroll
                    for (int jj = -P + 1; jj <= P; jj++)
                        if (jj > -psh && jj <= psh) {
                            s0 = s0 + w0[m + P + jj];
                            s1 = s1 + w1[m + P + jj];
                        }
                    const float x0 = lut_weight(s0, a.dmin[k], a.dmax[k], a.rng[k], a.rcp[k], a.scl[k], lut);
                    const float x1 = lut_weight(s1, a.dmin[k], a.dmax[k], a.rng[k], a.rcp[k], a.scl[k], lut);
                    *reinterpret_cast<float2*>(dst + m * RXM) = make_float2(x0, x1);
                }
            };
            constexpr int N1MAX = (((TX + WMAX + 3) / 4) * ((TX + WMAX + 1) / 2) + 63) & ~63;
            constexpr int KC = (NG * N1MAX + NT - 1) / NT;
#pragma unroll
            for (int kk = 0; kk < KC; kk++) {
                const int e = tid + kk * NT;
                // the wave's first element: groups start at multiples of 64, so the group is
                // the same for the whole wave
                const int ew = uni(e) & ~63;
                if (ew >= total) break;
                int g = 0;
#pragma unroll
                for (int q = 1; q < NG; q++)
                    if (ew >= base[q]) g = q;
                const int o = OPB == 1 ? 0 : g / NP;
                const int k = g - o * NP;
                const int j = e - base[g];
                if (j >= n1[o]) continue;
                const unsigned int ry = (unsigned int)(((float)j + 0.5f) * inv[o]);
                const unsigned int offb = __umul24(ry, ostep[o]) + 8u * (unsigned int)j;
                if (k == 0) body(std::integral_constant<int, 0>{}, o, offb);
                if constexpr (NP > 1) if (k == 1) body(std::integral_constant<int, (NP > 1 ? 1 : 0)>{}, o, offb);
                if constexpr (NP > 2) if (k == 2) body(std::integral_constant<int, (NP > 2 ? 2 : 0)>{}, o, offb);
            }
        };
        auto step4 = [&](const int dy, const int dx, float* sW) {
            [[maybe_unused]] const int adx = dx < 0 ? -dx : dx;
            // union of the tile (for w(c, d)) and the tile shifted by -d (for w(c, -d))
            [[maybe_unused]] const int RY = TY - dy, RX = TX + adx;
            [[maybe_unused]] const int rxo = dx > 0 ? -dx : 0;   // region x origin relative to tx0
            [[maybe_unused]] const int DY = RY + 2 * P, DX = RX + 2 * P;
            // 4. accumulate w(c, d) (upper half) and w(c, -d) = w_half(c - d, d) (lower
            //    half, copy_symm_weights.cl:53-74) into the nonlocal statistics
            //    (compute_nl_statistics.cl:69-84) of every patch size
            {
                unsigned int cu0, cl0;
                const float *wu0, *wl0;
                if constexpr (NP == 1) {
                    // the thread's own pixel plus a wave-uniform offset: no per-offset vector
                    // multiply (with several patch sizes this schedules worse, 0.7 %)
                    const int cofs = dy * cpitch + dx;
                    cu0 = c_own + (unsigned int)cofs;
                    cl0 = c_own - (unsigned int)cofs;
                    wu0 = sW + (w_own - rxo);
                    wl0 = sW + (w_own - dy * RXM - dx - rxo);
                } else {
                    const int py0 = g * RPT;
                    const int cy = ty0 + py0, cx = tx0 + px;
                    cu0 = (unsigned int)(im.corg + (cy + dy) * cpitch + (cx + dx));
                    cl0 = (unsigned int)(im.corg + (cy - dy) * cpitch + (cx - dx));
                    wu0 = sW + py0 * RXM + (px - rxo);
                    wl0 = sW + (py0 - dy) * RXM + (px - dx - rxo);
                }
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    const unsigned int cu = cu0 + (unsigned int)(q * cpitch);
                    const unsigned int cl = cl0 + (unsigned int)(q * cpitch);
                    if constexpr (MODE == 0) {
                        float i0u, i1u, i0l, i1l, cru = 0.f, ciu = 0.f, crl = 0.f, cil = 0.f;
                        if (LAST) {
                            const float4 u = gload4(C4b + 16u * cu);
                            const float4 l = gload4(C4b + 16u * cl);
                            i0u = u.x; i1u = u.y; cru = u.z; ciu = u.w;
                            i0l = l.x; i1l = l.y; crl = l.z; cil = l.w;
                        } else {
                            const float2 u = gload2(C4b + 16u * cu);
                            const float2 l = gload2(C4b + 16u * cl);
                            i0u = u.x; i1u = u.y;
                            i0l = l.x; i1l = l.y;
                        }
                        const float i0u2 = i0u * i0u, i1u2 = i1u * i1u;
                        const float i0l2 = i0l * i0l, i1l2 = i1l * i1l;
#pragma unroll
                        for (int k = 0; k < NP; k++) {
                            const float wu = wu0[k * WPL + q * RXM];
                            const float wl = wl0[k * WPL + q * RXM];
                            acc[q][k][0] += wu;
                            acc[q][k][1] = fmaf(wu, wu, acc[q][k][1]);
                            acc[q][k][2] = fmaf(wu, i0u, acc[q][k][2]);
                            acc[q][k][3] = fmaf(wu, i0u2, acc[q][k][3]);
                            acc[q][k][4] = fmaf(wu, i1u, acc[q][k][4]);
                            acc[q][k][5] = fmaf(wu, i1u2, acc[q][k][5]);
                            acc[q][k][0] += wl;
                            acc[q][k][1] = fmaf(wl, wl, acc[q][k][1]);
                            acc[q][k][2] = fmaf(wl, i0l, acc[q][k][2]);
                            acc[q][k][3] = fmaf(wl, i0l2, acc[q][k][3]);
                            acc[q][k][4] = fmaf(wl, i1l, acc[q][k][4]);
                            acc[q][k][5] = fmaf(wl, i1l2, acc[q][k][5]);
                            if (LAST && k == NP - 1) {
                                macc[q][0] = fmaf(wu, cru, macc[q][0]);
                                macc[q][1] = fmaf(wu, ciu, macc[q][1]);
                                macc[q][0] = fmaf(wl, crl, macc[q][0]);
                                macc[q][1] = fmaf(wl, cil, macc[q][1]);
                            }
                        }
                    } else {
                        // covariance input: the diagonal intensities (planes (2*dim+2)*d,
                        // compute_nl_statistics.cl:69-84) and, for the mean, every other plane
                        float cu_[4 * CF4], cl_[4 * CF4];
#pragma unroll
                        for (int j = 0; j < CF4; j++) {
                            const float4 u = gload4(C4b + (16u * CF4) * cu + 16u * j);
                            const float4 l = gload4(C4b + (16u * CF4) * cl + 16u * j);
                            cu_[4 * j] = u.x; cu_[4 * j + 1] = u.y; cu_[4 * j + 2] = u.z; cu_[4 * j + 3] = u.w;
                            cl_[4 * j] = l.x; cl_[4 * j + 1] = l.y; cl_[4 * j + 2] = l.z; cl_[4 * j + 3] = l.w;
                        }
#pragma unroll
                        for (int k = 0; k < NP; k++) {
                            const float wu = wu0[k * WPL + q * RXM];
                            const float wl = wl0[k * WPL + q * RXM];
                            acc[q][k][0] += wu;
                            acc[q][k][1] = fmaf(wu, wu, acc[q][k][1]);
#pragma unroll
                            for (int d = 0; d < DIM; d++) {
                                const float I = cu_[(2 * DIM + 2) * d];
                                acc[q][k][2 + 2 * d] = fmaf(wu, I, acc[q][k][2 + 2 * d]);
                                acc[q][k][3 + 2 * d] = fmaf(wu, I * I, acc[q][k][3 + 2 * d]);
                            }
                            acc[q][k][0] += wl;
                            acc[q][k][1] = fmaf(wl, wl, acc[q][k][1]);
#pragma unroll
                            for (int d = 0; d < DIM; d++) {
                                const float I = cl_[(2 * DIM + 2) * d];
                                acc[q][k][2 + 2 * d] = fmaf(wl, I, acc[q][k][2 + 2 * d]);
                                acc[q][k][3 + 2 * d] = fmaf(wl, I * I, acc[q][k][3 + 2 * d]);
                            }
                            if (LAST && k == NP - 1) {
                                int j = 0;
#pragma unroll
                                for (int pl = 0; pl < NPL; pl++) {
                                    if (pl % (2 * DIM + 2) == 0) continue;   // a diagonal intensity
                                    macc[q][j] = fmaf(wu, cu_[pl], macc[q][j]);
                                    macc[q][j] = fmaf(wl, cl_[pl], macc[q][j]);
                                    j++;
                                }
                            }
                        }
                    }
                }
            }
        };

        // Half-window offsets in plane order: rows above the centre, then the left half of the
        // centre row (compute_pixel_similarities_2x2.cpp:29-47), 2 wsh (wsh + 1) of them, an even
        // count.  OPB consecutive offsets share each barrier phase, in separate LDS slots; their
        // statistics are accumulated in plane order, as one offset per phase would.
        const int nofs = 2 * wsh * (wsh + 1);
        int ny = -wsh, nx = -wsh;
        for (int t = 0; t < nofs; t += OPB) {
            int oy[OPB], ox[OPB];
#pragma unroll
            for (int o = 0; o < OPB; o++) {
                oy[o] = ny;
                ox[o] = nx;
                if (++nx > wsh) {
                    nx = -wsh;
                    ny++;
                }
            }
#pragma unroll
            for (int o = 0; o < OPB; o++) step1(oy[o], ox[o], smem + Lay::off_D(NP, o));
            __syncthreads();
#pragma unroll
            for (int o = 0; o < OPB; o++) step2(oy[o], ox[o], smem + Lay::off_D(NP, o), smem + Lay::off_R(NP, OPB, o));
            __syncthreads();
            // a single patch size keeps the per-slot form (its two slots spread over the block's
            // waves measured 3 % slower); several patch sizes are spread over every wave
            if constexpr (NP == 1) {
#pragma unroll
                for (int o = 0; o < OPB; o++) step3_slot(oy[o], ox[o], smem + Lay::off_R(NP, OPB, o), smem + Lay::off_D(NP, o));
            } else {
                step3(oy, ox);
            }
            __syncthreads();
#pragma unroll
            for (int o = 0; o < OPB; o++) step4(oy[o], ox[o], smem + Lay::off_D(NP, o));
            // the next phase's step 1 overwrites these weights
            __syncthreads();
        }

        // centre weight 1 (routines.cpp:151-157), then the per-pixel statistics
        const NlsarImageDesc& dsc = nlsar_image(a, im.img);
        const int mm = im.m;
        const Launch L{uni(dsc.L[mm].vy0), uni(dsc.L[mm].vx0), uni(dsc.L[mm].ny), uni(dsc.L[mm].nx),
                       uni(dsc.L[mm].oy0), uni(dsc.L[mm].ox0)};
        const int imH = uni(dsc.H), imW = uni(dsc.W);
        DKX_GLOBAL float* const o_enl = gptr(uni(dsc.enl_nb));
        DKX_GLOBAL float* const o_alpha = gptr(uni(dsc.alpha));
        DKX_GLOBAL unsigned char* const o_flag = gptr(uni(dsc.ubflag));
        DKX_GLOBAL float* const o_ref = gptr(uni(dsc.ref));
        DKX_GLOBAL float* const o_phase = gptr(uni(dsc.phase));
        DKX_GLOBAL float* const o_coh = gptr(uni(dsc.coh));
        unsigned long long* const o_ub = uni(dsc.ub_count);
        unsigned int ubl = 0;
        DKX_GLOBAL float* const o_cov = gptr(uni(dsc.cov_out));
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int py = g * RPT + q;
            const int iy = ty0 + py - L.vy0, ix = tx0 + px - L.vx0;
            const bool valid = iy < L.ny && ix < L.nx;
            // the pixel's own covariance (C record) and its diagonal intensities
            float cc[4 * CF4];
            const unsigned int cidx = (unsigned int)(im.corg + (ty0 + py) * cpitch + (tx0 + px));
#pragma unroll
            for (int j = 0; j < CF4; j++) {
                const float4 v = gload4(C4b + (16u * CF4) * cidx + 16u * j);
                cc[4 * j] = v.x; cc[4 * j + 1] = v.y; cc[4 * j + 2] = v.z; cc[4 * j + 3] = v.w;
            }
            float ic[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) ic[d] = MODE == 0 ? cc[d] : cc[(2 * DIM + 2) * d];
            const long long opix = valid ? (long long)(L.oy0 + iy) * imW + (L.ox0 + ix) : 0;
            bool fail = false;
            float alpha_mean = 0.f, s_mean = 1.f;
            float dsum[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) dsum[d] = 0.f;
#pragma unroll
            for (int k = 0; k < NP; k++) {
                float* s = acc[q][k];
                s[0] += 1.0f;
                s[1] = fmaf(1.0f, 1.0f, s[1]);
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    s[2 + 2 * d] = fmaf(1.0f, ic[d], s[2 + 2 * d]);
                    s[3 + 2 * d] = fmaf(1.0f, ic[d] * ic[d], s[3 + 2 * d]);
                }
                // compute_nl_statistics.cl:86-102, compute_alphas.cl:19-44,
                // compute_enls_nobias.cl:19-39
                const float ws = s[0];
                const float enl = ws * ws / s[1];
                float alpha = 0.0f;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    const float inl = s[2 + 2 * d] / ws;
                    const float wv = s[3 + 2 * d] / ws;
                    const float var = wv - inl * inl;
                    alpha = fmaxf(alpha, fmaxf(0.0f, (var - inl * inl / 1.0f) / var));
                }
                const float oma = 1.0f - alpha;
                const float enb = enl / (oma * oma + (alpha * alpha + 2.0f * alpha * (1.0f - alpha) / ws) * enl);
                if (!(enb > -1.0f)) fail = true;
                if (valid) {
                    if (o_enl) o_enl[(long long)a.pidx[k] * imH * imW + opix] = enb;
                    if (o_alpha) o_alpha[(long long)a.pidx[k] * imH * imW + opix] = alpha;
                }
                if (LAST && k == NP - 1) {
                    alpha_mean = alpha;
                    s_mean = ws;
#pragma unroll
                    for (int d = 0; d < DIM; d++) dsum[d] = s[2 + 2 * d];
                    // centre weight 1 on the other planes' sums as well
                    if constexpr (MODE == 0) {
                        macc[q][0] = fmaf(1.0f, cc[2], macc[q][0]);
                        macc[q][1] = fmaf(1.0f, cc[3], macc[q][1]);
                    } else {
                        int j = 0;
#pragma unroll
                        for (int pl = 0; pl < NPL; pl++) {
                            if (pl % (2 * DIM + 2) == 0) continue;
                            macc[q][j] = fmaf(1.0f, cc[pl], macc[q][j]);
                            j++;
                        }
                    }
                }
            }
            if (!valid) continue;
            if (!LAST) {
                if (a.first_scale) o_flag[opix] = fail ? 1 : 0;
                else if (fail) o_flag[opix] = 1;
            } else {
                // best-parameter selection (best_params.cpp:27-53) keeps the last key of the
                // (patch, scale) map whenever every ENL exceeds -1; pixels where the reference's
                // selection loop is undefined are counted and also take the last key.
                if (fail || (!a.first_scale && o_flag[opix])) ubl++;
                // weighted_means.cl:62-80
                const float al = a.alpha_zero ? 0.0f : alpha_mean;
                if constexpr (MODE == 0) {
                    // and covmat_decompose.cl:19-39
                    const float m00 = (1 - al) * dsum[0] / s_mean + al * ic[0];
                    const float m11 = (1 - al) * dsum[1] / s_mean + al * ic[1];
                    const float mr = (1 - al) * macc[q][0] / s_mean + al * cc[2];
                    const float mi = (1 - al) * macc[q][1] / s_mean + al * cc[3];
                    const float am = sqrtf(m00), as = sqrtf(m11);
                    const float t = 0.5f * (am + as);
                    o_ref[opix] = t * t;
                    o_phase[opix] = atan2f(mi, mr);
                    o_coh[opix] = sqrtf(mi * mi + mr * mr) / (am * as);
                } else {
                    // every plane of the covariance, plane-major in the output
                    const long long HW = (long long)imH * imW;
                    int j = 0;
#pragma unroll
                    for (int pl = 0; pl < NPL; pl++) {
                        float sum;
                        if (pl % (2 * DIM + 2) == 0) {
                            sum = dsum[pl / (2 * DIM + 2)];
                        } else {
                            sum = macc[q][j];
                            j++;
                        }
                        o_cov[(long long)pl * HW + opix] = (1 - al) * sum / s_mean + al * cc[pl];
                    }
                }
            }
        }
        if (LAST && ubl) atomicAdd(o_ub, (unsigned long long)ubl);
        if constexpr (!PERSIST) break;
    }
}

// Launch of one pass; defined in the instantiating translation units.
template <bool LAST, int TX, int RPT, int WMAX, int MODE>
int nlsar_pass_launch(int pmask, const ScaleArgs& a, hipStream_t stream);

// Every patch set of one to three half widths in [0, 5].
#define DKX_NLSAR_PMASKS(X)                                                                  \
    X(1) X(2) X(4) X(8) X(16) X(32)                                                          \
    X(3) X(5) X(9) X(17) X(33) X(6) X(10) X(18) X(34) X(12) X(20) X(36) X(24) X(40) X(48)   \
    X(7) X(11) X(19) X(35) X(13) X(21) X(37) X(25) X(41) X(49) X(14) X(22) X(38) X(26)      \
    X(42) X(50) X(28) X(44) X(52) X(56)

template <bool LAST, int TX, int RPT, int WMAX, int MODE, int PMASK>
int nlsar_pass_launch_mask(const ScaleArgs& a, hipStream_t stream)
{
    constexpr int NP = pmask_count(PMASK);
    constexpr int P = pmask_top(PMASK);
    using Lay = PassLds<TX, P, WMAX>;
    size_t lds = sizeof(float) * (size_t)Lay::total(NP, nlsar_offsets_per_phase(NP, MODE), a.lut_stride);
    if (a.queue) lds += 16;   // the persistent grid's tile slot
    if (a.wsh > WMAX || lds > NLSAR_LDS_PER_CU) {
        set_error("nlsar: search window / patch sizes exceed the LDS budget");
        return -2;
    }
    // Residency cap of the InSAR two-row variant: where its registers and LDS would let
    // more than nlsar_blocks_per_cu blocks share a CU, the blocks resident on one XCD read
    // more A / C grid than its 4 MB L2 holds and the offset loop refetches it (a single
    // patch size at 2048 x 2048: 26 % L2 misses and 5.1 ms at 4 blocks per CU, 4.9 ms at
    // 3); an LDS request just above a (blocks + 1)-th of the CU's LDS holds it there.
    if (MODE == 0 && RPT == 2) {
        constexpr size_t lds_floor = NLSAR_LDS_PER_CU / (nlsar_blocks_per_cu(TX, RPT, WMAX, MODE, LAST) + 1) + 1024;
        if (lds < lds_floor) lds = lds_floor;
    }
    const void* fn = a.queue ? (const void*)nlsar_scale_pass<TX, RPT, PMASK, LAST, WMAX, MODE, true>
                             : (const void*)nlsar_scale_pass<TX, RPT, PMASK, LAST, WMAX, MODE, false>;
    DKX_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    static_assert(TX == NLSAR_TILE, "block counts assume 32 x 32 tiles");
    int total = a.total_blocks;
    if (total <= 0) return 0;
    if (a.queue) {
        // persistent: the blocks resident at once, at least one per XCD
        int per_cu = 0, dev = 0, ncu = 256;
        DKX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TX * TX / RPT, lds));
        DKX_HIP_CHECK(hipGetDevice(&dev));
        DKX_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        total = std::max(8, std::min(std::max(per_cu, 1) * ncu, (total + 7) / 8 * 8));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = timing_on();
    if (timed) {
        DKX_HIP_CHECK(hipEventCreate(&e0));
        DKX_HIP_CHECK(hipEventCreate(&e1));
        DKX_HIP_CHECK(hipEventRecord(e0, stream));
    }
    if (a.queue)
        hipLaunchKernelGGL((nlsar_scale_pass<TX, RPT, PMASK, LAST, WMAX, MODE, true>), dim3(total),
                           dim3(TX * TX / RPT), lds, stream, a);
    else
        hipLaunchKernelGGL((nlsar_scale_pass<TX, RPT, PMASK, LAST, WMAX, MODE, false>), dim3(total),
                           dim3(TX * TX / RPT), lds, stream, a);
    DKX_HIP_CHECK(hipGetLastError());
    if (timed) {
        DKX_HIP_CHECK(hipEventRecord(e1, stream));
        timing_push(LAST ? STAGE_NLSAR_LAST : STAGE_NLSAR_SCALE, e0, e1);
    }
    return 0;
}

#define DKX_NLSAR_PASS_CASE(M) case M: return nlsar_pass_launch_mask<LAST_, TX_, RPT_, WMAX_, MODE_, M>(a, stream);
#define DKX_NLSAR_PASS_INSTANTIATE(LAST_V, TX_V, RPT_V, WMAX_V, MODE_V)                       \
    template <>                                                                              \
    int nlsar_pass_launch<LAST_V, TX_V, RPT_V, WMAX_V, MODE_V>(int pmask, const ScaleArgs& a, \
                                                               hipStream_t stream)           \
    {                                                                                        \
        constexpr bool LAST_ = LAST_V;                                                       \
        constexpr int TX_ = TX_V;                                                            \
        constexpr int RPT_ = RPT_V;                                                          \
        constexpr int MODE_ = MODE_V;                                                        \
        constexpr int WMAX_ = WMAX_V;                                                        \
        switch (pmask) {                                                                     \
            DKX_NLSAR_PMASKS(DKX_NLSAR_PASS_CASE)                                            \
        default:                                                                             \
            set_error("nlsar: internal error, unsupported patch set");                      \
            return -1;                                                                       \
        }                                                                                    \
    }

}  // namespace dkx
