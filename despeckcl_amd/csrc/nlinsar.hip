// NL-InSAR on CDNA4.  Per iteration:
//   nlinsar_weights   fused pixel likelihood similarity + KL divergence -> 7x7 patch sums
//                     -> weights exp(S/h + KL/T), one search-window offset at a time in LDS,
//                     with the equivalent number of looks and the weighted means accumulated
//                     in the reference's plane order; pixels whose ENL is below lmin are
//                     queued for smoothing, all others get their new estimate here.
//   nlinsar_smooth    the lmin smoothing of the queued pixels (libstdc++ heap semantics)
//                     followed by their weighted means.
// Reference path replaced: src/compute/nlinsar/nlinsar.cpp:42-146 and nlinsar_sub_image.cpp
// with the kernels of src/compute/nlinsar/cl_kernels and the host smoothing of
// smoothing.cpp (a device<->host round trip per tile there).
#include "dkx_common.h"
#include "dkx_nlinsar.h"

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <cmath>

namespace dkx {

// The estimates as the KL divergence uses them (nlinsar/sim_measures.h:42-49): per element
// of the padded grid E, float4 (ref, coh * cos phase, coh * sin phase,
// u = 1 / (ref * (1 - coh^2))).
__device__ __forceinline__ void nl_estimate_record(const float* __restrict__ ref, const float* __restrict__ pha,
                                                   const float* __restrict__ coh, int H, int W,
                                                   const Grid& E, int gy, int gx)
{
    const int y = min(H - 1, max(0, gy - E.oy));
    const int x = min(W - 1, max(0, gx - E.ox));
    const long long i = (long long)y * W + x;
    const float r = ref[i], p = pha[i], c = coh[i];
    const long long o = (long long)gy * E.pitch + gx;
    reinterpret_cast<float4*>(E.base)[o] = make_float4(r, c * cosf(p), c * sinf(p), 1.0f / (r * (1.0f - c * c)));
}

__global__ void nlinsar_fill_estimates(const float* __restrict__ ref, const float* __restrict__ pha,
                                       const float* __restrict__ coh, int H, int W, Grid E)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= E.w || gy >= E.h) return;
    nl_estimate_record(ref, pha, coh, H, W, E, gy, gx);
}

// The raw data on the padded grid R: float4 (a1, a2, cos phase, sin phase), clamped like the
// reference's tile padding (src/data/sub_images.cpp:37-40); a1 alone also into A (same
// element layout as R) for the smoothing's amplitude test.
__device__ __forceinline__ void nl_raw_record(const float* __restrict__ a1, const float* __restrict__ a2,
                                              const float* __restrict__ phi, int H, int W,
                                              const Grid& R, float* __restrict__ A, int gy, int gx)
{
    const int y = min(H - 1, max(0, gy - R.oy));
    const int x = min(W - 1, max(0, gx - R.ox));
    const long long i = (long long)y * W + x;
    const float p = phi[i];
    reinterpret_cast<float4*>(R.base)[(long long)gy * R.pitch + gx] = make_float4(a1[i], a2[i], cosf(p), sinf(p));
    A[(long long)gy * R.pitch + gx] = a1[i];
}

__global__ void nlinsar_fill_raw(const float* __restrict__ a1, const float* __restrict__ a2,
                                 const float* __restrict__ phi, int H, int W, Grid R,
                                 float* __restrict__ A)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= R.w || gy >= R.h) return;
    nl_raw_record(a1, a2, phi, H, W, R, A, gy, gx);
}

// A filter-value record (a, Re x, Im x), 12 bytes: unpadded, so that the loads of the
// weights and smoothing kernels (one record per pixel and offset) move three dwords instead
// of four (4096², 10 iterations: 266.8 -> 262.8 ms per call, means of three alternating runs).
struct NlF { float x, y, z; };

// precompute_filter_values.cl:19-39: a = (a1^2 + a2^2)/2, x = a1*a2*exp(-i*phase), with
// cos(-phase) = cos(phase) and sin(-phase) = -sin(phase)
__device__ __forceinline__ void nl_filter_record(const Grid& R, const Grid& F, int gy, int gx)
{
    const float4 r = reinterpret_cast<const float4*>(R.base)[R.at(gy - F.oy, gx - F.ox)];
    const float a1 = r.x, a2 = r.y;
    reinterpret_cast<NlF*>(F.base)[(long long)gy * F.pitch + gx] =
        NlF{0.5f * (a1 * a1 + a2 * a2), a1 * a2 * r.z, a1 * a2 * -r.w};
}

__global__ void nlinsar_filter_values(Grid R, Grid F)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= F.w || gy >= F.h) return;
    nl_filter_record(R, F, gy, gx);
}

// flag[0] = 1 if any reflectivity estimate is nonzero: otherwise every KL divergence is
// -4/pi * (1 + 1 - 2) = -0 (nlinsar/sim_measures.h:42-49 with ref2 == 0) and is skipped.
__global__ void nlinsar_any_estimate(const float* __restrict__ ref, long long n, unsigned int* flag)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool nz = false;
    for (long long k = i; k < n; k += (long long)gridDim.x * blockDim.x) nz |= ref[k] != 0.0f;
    if (__any(nz) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

__global__ void nlinsar_clear_counters(unsigned int* cnt, unsigned int* flag, int clear_total)
{
    if (threadIdx.x == 0) {
        cnt[0] = 0u;
        if (clear_total) cnt[1] = 0u;
        cnt[3] = 0u;
        cnt[4] = 0u;
        cnt[6] = 0u;
        if (clear_total) cnt[5] = 0u;
        *flag = 0u;
    }
}

// Quotient and square root at the hardware's accuracy: v_rcp_f32 then one product (within
// 2 ulp), v_sqrt_f32 (1 ulp); OpenCL 1.2 allows 2.5 and 3 ulp for the reference's '/' and sqrt.
__device__ __forceinline__ float nl_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float nl_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// pixel_similarity (nlinsar/sim_measures.h:27-40) from raw records (a1, a2, cos phase,
// sin phase): cos(dp1 - dp2) = cos1*cos2 + sin1*sin2 and pow(q, 1.5) = q * sqrt(q), each a
// few ulps from the reference's expression (inside OpenCL's bounds for cos and pow); the
// quotients and square roots at the hardware's accuracy (nl_div, nl_sqrt); the other
// operations in the reference's order.
__device__ __forceinline__ float nl_pixel_similarity_pre(float4 r1, float4 r2)
{
    const float am1 = r1.x, as1 = r1.y, am2 = r2.x, as2 = r2.y;
    const float cs = fmaf(r1.z, r2.z, r1.w * r2.w);
    const float s = am1 * am1 + as1 * as1 + am2 * am2 + as2 * as2;
    const float A = s * s;
    const float B = 4.0f * (am1 * am1 * as1 * as1 + am2 * am2 * as2 * as2 + 2.0f * am1 * as1 * am2 * as2 * cs);
    const float C = am1 * as1 * am2 * as2;
    const float q = nl_div(C, B);
    const float f_el = q * nl_sqrt(q) * (nl_div(A + B, A) * nl_sqrt(nl_div(B, A - B)) - asinf(nl_sqrt(nl_div(B, A))));
    if ((f_el == 0) || isnan(f_el)) return 0.f;
    return logf(f_el);
}

// pixel_kullback_leibler (nlinsar/sim_measures.h:42-49) from the per-pixel values of
// nlinsar_fill_estimates:
//   a1 = ref1 / ref2 * (1 - coh1*coh2*cos(dp1 - dp2)) / (1 - coh2^2) = ref1 * r * u2,
//   a2 = ref2 / ref1 * r / (1 - coh1^2)                               = ref2 * r * u1,
// with coh1*coh2*cos(dp1 - dp2) = (coh1 cos1)(coh2 cos2) + (coh1 sin1)(coh2 sin2) (a few
// ulps from the reference's expression, inside OpenCL's bounds for cos and division);
// ref2 == 0 gives a1 = a2 = 1, i.e. -0.
__device__ __forceinline__ float nl_pixel_kl_pre(float4 e1, float4 e2)
{
    const float r = 1.0f - fmaf(e1.y, e2.y, e1.z * e2.z);
    const float a1 = e1.x * r * e2.w;
    const float a2 = e2.x * r * e1.w;
    return e2.x != 0.0f ? -4.0f / 3.14159274101257324219f * (a1 + a2 - 2.0f) : -0.0f;
}

__device__ __forceinline__ bool nl_is_valid(float x) { return !(isnan(x) || isinf(x) || x == 0); }

// The new estimate's reflectivity, phase and coherence (compute_insar.cl:95-101), in a
// function of its own that every kernel calls rather than inlines: inlined, the device
// library's atan2f was compiled to different instructions in different kernels (one ulp
// apart in the phase), and every kernel that writes an estimate must give the same bits
// (the single-image and batch paths compute the edge pixels in different block shapes).
__device__ __attribute__((noinline)) float4 nl_estimate_values(float N, float at, float xr, float xi)
{
    return make_float4(at / N, atan2f(xi, xr), sqrtf(xr * xr + xi * xi) / at, 0.0f);
}

// compute_insar.cl:95-117: new estimate, or the previous one where the value is invalid
__device__ __forceinline__ void nl_estimate(float N, float at, float xr, float xi, long long opix,
                                            const float* old_ref, const float* old_pha,
                                            const float* old_coh, float* nref, float* npha,
                                            float* ncoh)
{
    const float4 v = nl_estimate_values(N, at, xr, xi);
    const float reflectivity = v.x;
    const float iphase = v.y;
    const float coherence = v.z;
    const float pi_f = 3.14159274101257324219f;
    nref[opix] = nl_is_valid(reflectivity) ? reflectivity : old_ref[opix];
    if (nl_is_valid(iphase)) {
        float ph = iphase;
        if (ph > pi_f) ph -= 2.0f * pi_f;
        if (ph < -pi_f) ph += 2.0f * pi_f;
        npha[opix] = ph;
    } else {
        npha[opix] = old_pha[opix];
    }
    if (nl_is_valid(coherence)) {
        float c = coherence;
        if (c > 1.0f) c = 1.0f;
        if (c < 0.0f) c = 0.0f;
        ncoh[opix] = c;
    } else {
        ncoh[opix] = old_coh[opix];
    }
}

struct WeightArgs {
    const float4* R4;     // raw data: a1, a2, cos phase, sin phase (padded by the overlap)
    const float4* E4;     // estimates, see nlinsar_fill_estimates (same grid)
    const NlF* FV;        // filter values (padded by the window)
    int rpitch, rorg;     // rorg: element of image position (0, 0)
    int fpitch, forg;
    int wsh, step, wsim;
    float rh, rT;         // 1 / h, 1 / T
    float lmin;
    Launch L[4];          // interior, last row, last column, corner
    int gx[4], nsx[4], nsy[4];   // block columns / sub-tiles per reference tile, per mapping
    int bstart[5];        // first linear block of each mapping, [4] = total
    int H, W;
    // Per block, one contiguous chunk of [sw*sw][TY*TX] floats (block-major, so that every
    // block streams through its own memory):
    float* scache;        // patch similarities (written by the first iteration)
    float* wt;            // weights (read by the smoothing)
    const unsigned int* kl_on;
    const float* old_ref;
    const float* old_pha;
    const float* old_coh;
    float* new_ref;
    float* new_pha;
    float* new_coh;
    int* list;
    unsigned int* cnt;
    int block0;           // batch: first tile of this image in the batch's tile order
    int shape;            // block shape (NLI_SHAPES) of the launch: tags the weight slots
    // Which pixels' weights the launch stores for the lmin smoothing (NLI_STORE_*; zero, the
    // batches' value, stores all); nlinsar_fill_weights recomputes the others' when they are
    // queued.
    int store;
    int tag;                      // NLI_STORE_PRED: pixels with pred[pixel] == tag are stored
    const unsigned char* pred;    // per image pixel, row-major
    // Compact weights: the stored pixels of a wave take a run of k consecutive slots c0 ..
    // c0 + k - 1, whose weights lie offset-major at wc + c0*sw*sw (offset p of the run's
    // pixel r at p*k + r, so that a store of one offset is contiguous); caux[c] = (k << 8) | r
    // for the queued ones.
    float* wc;
    int* caux;
    const unsigned int* dense;    // NLI_STORE_PRED: nonzero when this iteration stores all instead
    int* miss;                    // the queue entries whose weights must be recomputed (cnt[6])
};

// Weight store modes of a weights launch.  The stored weights and the ones
// nlinsar_fill_weights recomputes are the same bits, so the mode changes only the traffic.
constexpr int NLI_STORE_ALL = 0;    // every pixel's
constexpr int NLI_STORE_PRED = 1;   // the pixels marked with the launch's tag
constexpr int NLI_STORE_NONE = 2;   // none
constexpr int NLI_STORE_KL = 3;     // all when the KL term is on, none when it is off (a first
                                    // iteration from zero estimates, where a weight is a
                                    // function of the cached similarity alone)

// Every pixel's weights stored, in the block-major layout of the tile slots (no compact
// slots, nothing to recompute).
__device__ __forceinline__ bool nl_store_dense(const WeightArgs& a, bool kl)
{
    return a.store == NLI_STORE_ALL || (a.store == NLI_STORE_KL && kl) ||
           (a.store == NLI_STORE_PRED && *a.dense != 0u);
}

// Otherwise: whether the weights launch stores the weights of output pixel opix, in a compact
// slot (so that sparse stores fill whole cache lines).
__device__ __forceinline__ bool nl_weight_stored(const WeightArgs& a, long long opix)
{
    return a.store == NLI_STORE_PRED && a.pred[opix] == (unsigned char)a.tag;
}

// Position of weights tile t (a block of TY x TX outputs): its first virtual output position,
// how many of its rows / columns the reference tile interior holds, the origin of that
// reference tile's interior, and the launch mapping it belongs to.
struct NlTile {
    int ty0, tx0, rows_lim, cols_lim, Y0, X0;
    Launch L;
};

constexpr int NLI_BAND = 4;   // block rows per band of the block order
constexpr int NLI_RSEG = 3;   // row sums per thread (a run of adjacent outputs)

// The four launch mappings (interior, last row, last column, corner) are laid end to end in
// one grid.  Blocks are sub-tiles of the interiors (step x step) of the reference's tiles, so
// every output of a block belongs to one reference tile: block row by = reference tile row *
// nsub_y + sub-tile row (nsub = ceil(step / TY)); the last sub-tile of a reference tile is
// cut at the tile's end.  The one-row / one-column mappings have a single block row / column
// starting at the mapping's origin.  Tile t (the launch gives each XCD a contiguous run of
// tiles) is ordered column by column within bands of NLI_BAND block rows: the blocks
// resident on one XCD then cover a compact patch of the image, whose neighbourhoods overlap
// in that XCD's L2.
__device__ __forceinline__ NlTile nl_tile(const WeightArgs& a, int t, int TY, int TX)
{
    int m = 0;
    while (m < 3 && t >= a.bstart[m + 1]) m++;
    NlTile g;
    g.L = m == 0 ? a.L[0] : m == 1 ? a.L[1] : m == 2 ? a.L[2] : a.L[3];
    const int nsy = a.nsy[m], nsx = a.nsx[m];
    const int local = t - a.bstart[m];
    const int gxm = a.gx[m];
    const int gym = (a.bstart[m + 1] - a.bstart[m]) / gxm;
    const int band = local / (NLI_BAND * gxm);
    const int within = local - band * NLI_BAND * gxm;
    const int bh = min(NLI_BAND, gym - band * NLI_BAND);
    const int blx = within / bh;
    const int bly = band * NLI_BAND + (within - blx * bh);
    const int ry = bly / nsy, sy = bly - ry * nsy;
    const int rx = blx / nsx, sx_ = blx - rx * nsx;
    g.ty0 = g.L.vy0 + ry * a.step + sy * TY;
    g.tx0 = g.L.vx0 + rx * a.step + sx_ * TX;
    g.rows_lim = min(TY, a.step - sy * TY);
    g.cols_lim = min(TX, a.step - sx_ * TX);
    // interior origin of the reference tile that computes this block's outputs
    g.Y0 = (g.ty0 / a.step) * a.step;
    g.X0 = (g.tx0 / a.step) * a.step;
    return g;
}

// Weight of a patch similarity u and KL divergence v, e^(u rh + v rT) with rh = 1/h and
// rT = 1/T (compute_weights.cl:23-27: exp(sim / h + kl / T), 0 when NaN or infinite), and 0
// for the centre offset.  e^x is the device library's expf, operation for operation: x log2 e
// carried in two parts (the product, its FMA residual plus x times the remainder of log2 e),
// its nearest integer n split off, v_exp_f32 of the rest and v_ldexp_f32 by n.  The library
// sends x < -103.27893 to 0 and x > 88.72284 to infinity, both of which end as weight 0
// here, so its two range selections and the NaN / infinity test become one range test of x
// and one class test of the result (the same bits for every float x, 2^24 of them compared
// on the device including both range ends, infinities and NaN).  Every kernel that forms
// weights calls this one function.
__device__ __forceinline__ float nl_weight(float u, float v, float rh, float rT, bool centre)
{
    const float x = u * rh + v * rT;
    const float L = 0x1.715476p+0f;        // log2 e, leading float
    const float Llo = 0x1.4ae0bep-26f;     // log2 e - L
    const float ph = x * L;
    float pl = fmaf(x, L, -ph);
    pl = fmaf(x, Llo, pl);
    const float n = __builtin_rintf(ph);
    const float r = __builtin_ldexpf(__builtin_amdgcn_exp2f((ph - n) + pl), (int)n);
    const bool in = __builtin_amdgcn_fmed3f(x, -0x1.9d1da0p+6f, 0x1.62e430p+6f) == x;
    return (in & __builtin_isfinite(r) & !centre) ? r : 0.0f;
}

// nl_pixel_kl_pre(e1, e2) with the pixel the reference passes second given by `second`:
// r, and a1 + a2, are the same bits for either argument order (products and fmaf commute,
// and a2 + a1 == a1 + a2), so only the zero test depends on the order.
__device__ __forceinline__ float nl_pixel_kl_sym(float4 e1, float4 e2, float second)
{
    const float r = 1.0f - fmaf(e1.y, e2.y, e1.z * e2.z);
    const float a1 = e1.x * r * e2.w;
    const float a2 = e2.x * r * e1.w;
    float kl = -4.0f / 3.14159274101257324219f * (a1 + a2 - 2.0f);
    // computed on every lane: otherwise the compiler moves the loads of e2 behind the zero
    // test and serialises two LDS round trips
    __asm__ volatile("" : "+v"(kl));
    return second != 0.0f ? kl : -0.0f;
}

// FIRST: evaluate the likelihood similarities and keep their patch sums (they depend on the
// raw data only); later iterations read them back and evaluate only the KL divergences.
//
// Work split (TX = 54, NTW threads): every thread keeps the same positions for all offsets,
// so per offset only the offset's displacement changes.
//   field:    positions tid + k*NTW of the (TY+2p) x (TX+2p) patch-position grid; the record
//             at the position itself stays in registers, the displaced one is read from LDS;
//   row sums: runs of NLI_RSEG adjacent outputs of one row;
//   outputs:  threads tid < NG*TX (NG row groups) own TY/NG vertically adjacent pixels of one
//             column, slot tid + m*NG*TX in the block's similarity / weight chunk.
// SPARSE: the launch may store only some pixels' weights (WeightArgs::store).
template <int TY, int TX, int PSH, bool FIRST, int NTW, bool SPARSE>
__device__ __forceinline__ void nlinsar_weights_tile(const WeightArgs& a, const int t)
{
    // output row groups: as many as the threads allow, dividing TY
    constexpr int NG = (NTW / TX) >= TY ? TY : (NTW / TX);
    static_assert(TX % NLI_RSEG == 0 && TY % NG == 0, "tile shape");
    constexpr int NPIX = TY * TX;
    constexpr int NP2 = 2 * PSH;
    constexpr int PY = TY + NP2, PX = TX + NP2;
    constexpr int NF = PY * PX;
    constexpr int K1 = (NF + NTW - 1) / NTW;
    constexpr int SEGS = TX / NLI_RSEG;
    constexpr int NRS = PY * SEGS;
    constexpr int K2 = (NRS + NTW - 1) / NTW;
    constexpr int RPT = TY / NG;
    constexpr int NCT = NG * TX;
    static_assert(NCT <= NTW, "output threads");
    extern __shared__ float4 smem4[];
    const int wsh = a.wsh;
    const int sw = 2 * wsh + 1;
    // LDS: the records of the block's whole neighbourhood (raw data in the first iteration,
    // estimates after), staged once; then the per-offset similarity / KL fields and row sums
    const int o = wsh + PSH;
    const int EW = TX + 2 * o, EH = TY + 2 * o;
    float4* stg = smem4;
    float* sS = reinterpret_cast<float*>(smem4 + EH * EW);
    float* rS = sS + (FIRST ? PY * PX : 0);
    float* sK = rS + (FIRST ? PY * TX : 0);
    float* rK = sK + PY * PX;

    const int tid = threadIdx.x;
    const bool kl = *a.kl_on != 0u;
    const bool dense = !SPARSE || nl_store_dense(a, kl);
    const NlTile tg = nl_tile(a, t, TY, TX);
    const Launch L = tg.L;
    const int ty0 = tg.ty0, tx0 = tg.tx0, rows_lim = tg.rows_lim, cols_lim = tg.cols_lim;
    const int Y0 = tg.Y0, X0 = tg.X0;

    // this block's chunk of the similarity cache and of the weights
    const long long chunk = (long long)t * (sw * sw) * NPIX;

    // output pixels of this thread: column oc, rows og*RPT .. og*RPT + RPT - 1.  Threads past
    // NCT repeat the last group's reads (nothing is stored for them), so that every load
    // issued is consumed on every path.
    const bool owner = tid < NCT;
    const int og = min(tid / TX, NG - 1), oc = tid - (tid / TX) * TX;
    const int slot0 = min(tid, NCT - 1);
    float acc[RPT][5];
    long long opix[RPT];
    bool valid[RPT];
    bool stq[RPT];            // the pixel's weights are stored
    unsigned int fofs[RPT];   // F grid element of the pixel (offset 0)
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = acc[q][4] = 0.f;
        const int py = og * RPT + q, px = oc;
        const int iy = ty0 + py - L.vy0, ix = tx0 + px - L.vx0;
        valid[q] = owner && py < rows_lim && px < cols_lim && iy < L.ny && ix < L.nx;
        opix[q] = valid[q] ? (long long)(L.oy0 + iy) * a.W + (L.ox0 + ix) : 0;
        stq[q] = valid[q] && (dense || nl_weight_stored(a, opix[q]));
        fofs[q] = (unsigned int)(a.forg + (ty0 + py) * a.fpitch + (tx0 + px));
    }
    // Compact slots of the stored pixels (not dense): a run of nst slots per wave from the
    // compact counter (cnt[4]), pixel q of a lane at rank cof[q] in the run.
    const int wlane = tid & 63;
    unsigned int cof[RPT];
    float* wcw = a.wc;
    int cslot0 = 0, nst = 0;
    if (!dense) {
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const unsigned long long bal = __ballot(stq[q]);
            cof[q] = (unsigned int)(nst + __popcll(bal & ((1ull << wlane) - 1ull)));
            nst += __popcll(bal);
        }
        unsigned int wb0 = 0;
        if (nst && wlane == 0) wb0 = atomicAdd(a.cnt + 4, (unsigned int)nst);
        cslot0 = __builtin_amdgcn_readfirstlane((int)wb0);
        nst = __builtin_amdgcn_readfirstlane(nst);
        wcw = uni(a.wc + (long long)cslot0 * (sw * sw));
    }
    // the lane's weight-store offsets from an offset's base: its tile slot (dense) or its rank
    // in the wave's run (compact)
    unsigned int soff[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) soff[q] = dense ? (unsigned int)(tid + q * NCT) : cof[q];

    // stage the neighbourhood: rows ty0 - o .. ty0 + TY + o, columns tx0 - o .. tx0 + TX + o
    {
        const float4* src = FIRST ? a.R4 : a.E4;
        const int n = EH * EW;
        for (int i = tid; i < n; i += NTW) {
            const int r = i / EW, c = i - r * EW;
            stg[i] = src[(unsigned int)(a.rorg + (ty0 - o + r) * a.rpitch + (tx0 - o + c))];
        }
    }
    // field positions of this thread: staged index, grid element, inside the reference
    // tile's similarity border (the orientation rule below), and the record at the position
    int fs[K1];
    unsigned int fg[K1];
    bool finner[K1];
    float4 fe[K1];
#pragma unroll
    for (int k = 0; k < K1; k++) {
        const int i = tid + k * NTW;
        const int r = i / PX, c = i - (i / PX) * PX;
        fs[k] = (wsh + r) * EW + (wsh + c);
        const int gy = ty0 - PSH + r, gx = tx0 - PSH + c;
        fg[k] = (unsigned int)(a.rorg + gy * a.rpitch + gx);
        const int sh = gy - Y0 + PSH, sx = gx - X0 + PSH;
        finner[k] = sh >= wsh && sx >= wsh && sx < a.wsim - wsh;
    }
    // row-sum runs of this thread: first field element and first output
    int rsrc[K2], rdst[K2];
#pragma unroll
    for (int k = 0; k < K2; k++) {
        const int j = tid + k * NTW;
        const int r = j / SEGS, g = j - (j / SEGS) * SEGS;
        rsrc[k] = r * PX + g * NLI_RSEG;
        rdst[k] = r * TX + g * NLI_RSEG;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K1; k++)
        if (tid + k * NTW < NF) fe[k] = stg[fs[k]];

    const int centre = wsh * sw + wsh;
    if constexpr (!FIRST) {
        // Later iterations: the KL field of offset p, the row sums of offset p-1 and the
        // weights of offset p-2 are computed between the same two barriers, with the fields
        // and row sums double-buffered (buffer b at sS + b*(NF + PY*TX): field, then rows).
        const int npl = sw * sw;
        float* const fbuf0 = sS;
        float* const fbuf1 = sS + NF + PY * TX;
        // Per offset only uniform values change: the staged displacement, and the base
        // pointers of the cache, weight and filter-value planes (each lane adds its own
        // 32-bit element offset, fixed for the whole kernel).
        auto field = [&](int dy, int dx, float* dst) {
            const int doff = dy * EW + dx;
            float4 vd[K1];
#pragma unroll
            for (int k = 0; k < K1; k++)
                if (!(k * NTW + NTW > NF && tid + k * NTW >= NF)) vd[k] = stg[fs[k] + doff];
#pragma unroll
            for (int k = 0; k < K1; k++) {
                const int i = tid + k * NTW;
                if (k * NTW + NTW > NF && i >= NF) continue;
                // orientation rule of the similarities, as in the first iteration's loop
                const bool swp = dy < 0 && finner[k];
                dst[i] = nl_pixel_kl_sym(fe[k], vd[k], swp ? fe[k].x : vd[k].x);
            }
        };
        auto rowsum = [&](const float* src, float* dst) {
#pragma unroll
            for (int k = 0; k < K2; k++) {
                if (k * NTW + NTW > NRS && tid + k * NTW >= NRS) continue;
                float x[NLI_RSEG + NP2];
#pragma unroll
                for (int j = 0; j < NLI_RSEG + NP2; j++) x[j] = src[rsrc[k] + j];
#pragma unroll
                for (int g = 0; g < NLI_RSEG; g++) {
                    float v = x[g];
#pragma unroll
                    for (int j = 1; j <= NP2; j++) v += x[g + j];
                    dst[rdst[k] + g] = v;
                }
            }
        };
        // loads of the cached similarity sums and filter values of one offset
        auto fetch = [&](int plane, int dy, int dx, float (&sv)[RPT], NlF (&fv)[RPT]) {
            const float* sb = a.scache + (chunk + (long long)plane * NPIX);
            const NlF* fb = a.FV + (dy * a.fpitch + dx);
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                sv[q] = sb[(unsigned int)(slot0 + q * NCT)];
                fv[q] = fb[fofs[q]];
            }
        };
        // column sums, weight (centre weight 0), ENL sums and weighted means of one offset
        auto accumulate = [&](int plane, const float (&sv)[RPT], const NlF (&fv)[RPT],
                              const float* rows) {
            float* wb = a.wt + (chunk + (long long)plane * NPIX);
            // this offset's weight stores: a wave-uniform base in the global address space and
            // the lane's offset.  Loads and stores share one counter (vmcnt), so a use of a
            // loaded value after a store waits for the store too: every load of this offset is
            // waited for before the weights are formed (s_waitcnt vmcnt(0), expcnt and lgkmcnt
            // left free), and the stores then stay in flight across the next offset's LDS phase.
            DKX_GLOBAL float* const sbase = gptr(dense ? wb : wcw + (unsigned int)(plane * nst));
            const bool self = plane == centre;
            float wq[RPT];
            float ck[RPT + NP2];
            if (rows) {
#pragma unroll
                for (int j = 0; j < RPT + NP2; j++) ck[j] = rows[(og * RPT + j) * TX + oc];
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                float v = -0.0f;
                if (rows) {
                    v = ck[q];
#pragma unroll
                    for (int j = 1; j <= NP2; j++) v += ck[q + j];
                }
                const float w = nl_weight(sv[q], v, a.rh, a.rT, self);
                wq[q] = w;
                const NlF f = fv[q];
                acc[q][0] += w;
                acc[q][1] = fmaf(w, w, acc[q][1]);
                acc[q][2] = fmaf(w, f.x, acc[q][2]);
                acc[q][3] = fmaf(w, f.y, acc[q][3]);
                acc[q][4] = fmaf(w, f.z, acc[q][4]);
            }
#pragma unroll
            for (int q = 0; q < RPT; q++)
                if (stq[q]) sbase[soff[q]] = wq[q];
        };
        if (!kl) {
            for (int plane = 0, dy = -wsh, dx = -wsh; plane < npl; plane++) {
                float sv[RPT];
                NlF fv[RPT];
                fetch(plane, dy, dx, sv, fv);
                accumulate(plane, sv, fv, nullptr);
                if (++dx > wsh) { dx = -wsh; dy++; }
            }
        } else {
            // (fy, fx): offset of the field step; (ay, ax): offset of the weights step
            int fy = -wsh, fx = -wsh, ay = -wsh, ax = -wsh;
            for (int ph = 0; ph < npl + 2; ph++) {
                const int p3 = ph - 2;
                float sv[RPT];
                NlF fv[RPT];
                if (p3 >= 0) fetch(p3, ay, ax, sv, fv);
                if (ph < npl && ph != centre) field(fy, fx, (ph & 1) ? fbuf1 : fbuf0);
                const int p2 = ph - 1;
                if (p2 >= 0 && p2 < npl && p2 != centre)
                    rowsum((p2 & 1) ? fbuf1 : fbuf0, ((p2 & 1) ? fbuf1 : fbuf0) + NF);
                if (p3 >= 0) {
                    accumulate(p3, sv, fv, p3 != centre ? ((p3 & 1) ? fbuf1 : fbuf0) + NF : nullptr);
                    if (++ax > wsh) { ax = -wsh; ay++; }
                }
                if (++fx > wsh) { fx = -wsh; fy++; }
                __syncthreads();
            }
        }
    } else {
        // First iteration: similarities (and KL only when the initial estimates are not all
        // zero) at every patch position, row sums, then per pixel the column sums, which are
        // also kept in the similarity cache for the later iterations.
        for (int hh = 0; hh < sw; hh++) {
            const int dy = hh - wsh;
            for (int ww = 0; ww < sw; ww++) {
                const int dx = ww - wsh;
                const int plane = hh * sw + ww;
                const long long pbase = chunk + (long long)plane * NPIX;
                // centre weight 0 (nlinsar_sub_image.cpp:189-197), accumulated like the others
                const bool self = plane == centre;
                // this offset's filter values, issued before the field and row-sum steps
                NlF fv[RPT];
                const int fd = dy * a.fpitch + dx;
#pragma unroll
                for (int q = 0; q < RPT; q++) fv[q] = a.FV[fofs[q] + fd];
                if (!self) {
                    __syncthreads();
                    // 1. similarity and KL at every patch position of the block.  Offsets with
                    //    dy >= 0 are evaluated centre-first everywhere; offsets with dy < 0 are
                    //    the mirrored values (the pixel at +d first) at positions inside the
                    //    reference tile's similarity border, centre-first on the border
                    //    (precompute_similarities_1st_pass.cl.in:36-42, _2nd_pass.cl:30-48).
                    const int doff = dy * EW + dx;
                    const unsigned int goff = (unsigned int)(dy * a.rpitch + dx);
                    // the displaced records first, all reads in flight together
                    float4 vd[K1];
#pragma unroll
                    for (int k = 0; k < K1; k++)
                        if (!(k * NTW + NTW > NF && tid + k * NTW >= NF)) vd[k] = stg[fs[k] + doff];
#pragma unroll
                    for (int k = 0; k < K1; k++) {
                        const int i = tid + k * NTW;
                        if (k * NTW + NTW > NF && i >= NF) continue;
                        const float4 v = vd[k];
                        const bool swp = dy < 0 && finner[k];
                        const float4 u1 = swp ? v : fe[k], u2 = swp ? fe[k] : v;
                        sS[i] = nl_pixel_similarity_pre(u1, u2);
                        if (kl) {
                            // nonzero initial estimates: read them from the grid
                            const unsigned int p1 = swp ? fg[k] + goff : fg[k];
                            const unsigned int p2 = swp ? fg[k] : fg[k] + goff;
                            sK[i] = nl_pixel_kl_pre(a.E4[p1], a.E4[p2]);
                        }
                    }
                    __syncthreads();
                    // 2. row sums along x (precompute_patch_similarities.cl:45-50)
#pragma unroll
                    for (int k = 0; k < K2; k++) {
                        if (k * NTW + NTW > NRS && tid + k * NTW >= NRS) continue;
                        {
                            float x[NLI_RSEG + NP2];
#pragma unroll
                            for (int j = 0; j < NLI_RSEG + NP2; j++) x[j] = sS[rsrc[k] + j];
#pragma unroll
                            for (int g = 0; g < NLI_RSEG; g++) {
                                float u = x[g];
#pragma unroll
                                for (int j = 1; j <= NP2; j++) u += x[g + j];
                                rS[rdst[k] + g] = u;
                            }
                        }
                        if (kl) {
                            float x[NLI_RSEG + NP2];
#pragma unroll
                            for (int j = 0; j < NLI_RSEG + NP2; j++) x[j] = sK[rsrc[k] + j];
#pragma unroll
                            for (int g = 0; g < NLI_RSEG; g++) {
                                float v = x[g];
#pragma unroll
                                for (int j = 1; j <= NP2; j++) v += x[g + j];
                                rK[rdst[k] + g] = v;
                            }
                        }
                    }
                    __syncthreads();
                }
                // 3. rows accumulated top to bottom (precompute_patch_similarities.cl:53-58),
                //    weight (compute_weights.cl:19-30), ENL sums (compute_number_of_looks.cl)
                //    and weighted means (compute_insar.cl:81-93) in plane order
                DKX_GLOBAL float* const sbase = gptr(dense ? a.wt + pbase : wcw + (unsigned int)(plane * nst));
                float cs[RPT + NP2], ck[RPT + NP2];
#pragma unroll
                for (int j = 0; j < RPT + NP2; j++) cs[j] = rS[(og * RPT + j) * TX + oc];
                if (kl) {
#pragma unroll
                    for (int j = 0; j < RPT + NP2; j++) ck[j] = rK[(og * RPT + j) * TX + oc];
                }
                // the similarity sums and weights are stored after the accumulation, this
                // offset's loads waited for first (loads and stores share one counter: see the
                // later iterations' accumulate)
                __builtin_amdgcn_s_waitcnt(0x0F70);
                float uq[RPT], wq[RPT];
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    float u = cs[q];
#pragma unroll
                    for (int j = 1; j <= NP2; j++) u += cs[q + j];
                    uq[q] = u;
                    float v = -0.0f;
                    if (kl) {
                        v = ck[q];
#pragma unroll
                        for (int j = 1; j <= NP2; j++) v += ck[q + j];
                    }
                    const float w = nl_weight(u, v, a.rh, a.rT, self);
                    wq[q] = w;
                    const NlF f = fv[q];
                    acc[q][0] += w;
                    acc[q][1] = fmaf(w, w, acc[q][1]);
                    acc[q][2] = fmaf(w, f.x, acc[q][2]);
                    acc[q][3] = fmaf(w, f.y, acc[q][3]);
                    acc[q][4] = fmaf(w, f.z, acc[q][4]);
                }
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    if (valid[q] && !self) a.scache[pbase + tid + q * NCT] = uq[q];
                    if (stq[q]) sbase[soff[q]] = wq[q];
                }
            }
        }
    }
    // pixels whose ENL is below lmin are queued for the smoothing, a wave at a time (one
    // atomic per wave keeps a wave's pixels, which are neighbours, together in the queue)
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const float ws = acc[q][0];
        const float nol = ws * ws / acc[q][1];
        const bool queued = valid[q] && nol < a.lmin;
        const unsigned long long qmask = __ballot(queued);
        if (qmask) {
            unsigned int base = 0;
            if (lane == 0) base = atomicAdd(a.cnt, (unsigned int)__popcll(qmask));
            base = __shfl(base, 0);
            if (queued) {
                const unsigned int k = base + __popcll(qmask & ((1ull << lane) - 1ull));
                a.list[2 * k] = (int)opix[q];
                // weight slot: the tile slot (dense, or recomputed by nlinsar_fill_weights), or
                // the compact slot of stored weights
                if (dense || !stq[q]) {
                    a.list[2 * k + 1] = (a.shape << NLI_SLOT_BITS) | (t * NPIX + tid + q * NCT);
                } else {
                    a.list[2 * k + 1] = (NLI_COMPACT << NLI_SLOT_BITS) | (cslot0 + (int)cof[q]);
                    a.caux[cslot0 + (int)cof[q]] = (nst << 8) | (int)cof[q];
                }
            }
        }
        if (valid[q] && !queued)
            nl_estimate(ws, acc[q][2], acc[q][3], acc[q][4], opix[q], a.old_ref, a.old_pha,
                        a.old_coh, a.new_ref, a.new_pha, a.new_coh);
    }
}

// One image per launch, one tile per block.  Linear block ids are remapped so that the
// blocks the hardware sends to one XCD (id mod 8) take a contiguous run of tiles.
template <int TY, int TX, int PSH, bool FIRST, int NTW>
__global__ __launch_bounds__(NTW) void nlinsar_weights(const WeightArgs a)
{
    const int nb = a.bstart[4];
    const int bq = nb >> 3, br = nb & 7;
    const int xcd = blockIdx.x & 7, kk = blockIdx.x >> 3;
    nlinsar_weights_tile<TY, TX, PSH, FIRST, NTW, true>(a, xcd * bq + min(xcd, br) + kk);
}

// A batch of images (table: one entry per image, the images' tiles end to end from
// entry.block0, `total` in all) on a persistent grid: each block takes tiles from the queue
// of its XCD (block id mod 8) until that queue is empty.  Queue x holds the x-th eighth of
// the tiles, a contiguous run, like the block order of a single image; the grid has at least
// eight blocks, so every queue is drained.  queue: 8 counters, zero before the launch.
template <int TY, int TX, int PSH, bool FIRST, int NTW>
__global__ __launch_bounds__(NTW) void nlinsar_weights_batch(const WeightArgs* __restrict__ table, int nimg,
                                                             int total, unsigned int* __restrict__ queue)
{
    __shared__ int s_item;
    const int bq = total >> 3, br = total & 7;
    const int xcd = blockIdx.x & 7;
    const int mine = bq + (xcd < br ? 1 : 0), first = xcd * bq + min(xcd, br);
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(queue + xcd, 1u);
        __syncthreads();
        const int k = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (k >= mine) break;
        const int T = first + k;
        // the image holding tile T: the last entry with block0 <= T
        int lo = 0, hi = nimg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (table[mid].block0 <= T) lo = mid;
            else hi = mid - 1;
        }
        nlinsar_weights_tile<TY, TX, PSH, FIRST, NTW, true>(table[lo], T - table[lo].block0);
    }
}

// The work items of a batch's smoothing or fill launch are the queued pixels of all images end
// to end (image i's items from base_i on): the first item at or after `base` that a worker `me` of
// `stride` workers takes (items me, me + stride, ...).
__device__ __forceinline__ unsigned long long nl_first_item(unsigned long long me, unsigned long long stride,
                                                            unsigned long long base)
{
    return me >= base ? me : me + (base - me + stride - 1) / stride * stride;
}

// Work split of a smoothing launch over the XCDs.  Blocks are dispatched to the 8 XCDs round
// robin (block b on XCD b mod 8); the blocks of XCD x take the x-th eighth of the work items,
// a contiguous run, and stride through it.  The queue holds the pixels in the order the
// weights kernel's waves finished them, interleaved across its XCDs, each of which works
// through a compact band of tiles: a run of items is a few compact clusters of pixels, whose
// search windows (the amplitudes and filter values the smoothing reads) then share one L2.
// A grid-wide stride would hand every XCD items from the whole resident window at once.
struct XcdItems {
    unsigned long long lo, hi;   // this XCD's items
    unsigned long long r, nw;    // the thread takes items lo + r, lo + r + nw, ...
};

// `total`: the launch's work items; launches use a multiple of 8 blocks (else one run).
__device__ __forceinline__ XcdItems nl_xcd_items(unsigned long long total)
{
    const unsigned int np = gridDim.x % 8u == 0u ? 8u : 1u;
    const unsigned int x = blockIdx.x % np;
    XcdItems s;
    s.lo = total * x / np;
    s.hi = total * (x + 1) / np;
    s.nw = (unsigned long long)(gridDim.x / np) * blockDim.x;
    s.r = (unsigned long long)(blockIdx.x / np) * blockDim.x + threadIdx.x;
    return s;
}

// The thread's first item at or after `from`.
__device__ __forceinline__ unsigned long long nl_xcd_first(const XcdItems& s, unsigned long long from)
{
    const unsigned long long r0 = s.lo + s.r;
    return r0 >= from ? r0 : r0 + (from - r0 + s.nw - 1) / s.nw * s.nw;
}

// The weights launches whose stores nlinsar_fill_scan / nlinsar_fill_weights complete: one
// image (`one`, shape 0) or a batch (`table`: shape s of image i at table[s * stride + i]).
struct FillBatch {
    const WeightArgs* table;
    WeightArgs one;
    int stride, nimg;
    int ty[NLI_SHAPES], tx[NLI_SHAPES], nt[NLI_SHAPES];   // block shape and threads per shape
};

// The queue entries whose weights the weights launch of this iteration did not store (their
// entries still carry a tile slot): appended to the image's `miss` list (count in cnt[6]), a
// wave's at once.  Grid (x, image).
__global__ __launch_bounds__(256) void nlinsar_fill_scan(const FillBatch f)
{
    const WeightArgs& a = f.table ? f.table[blockIdx.y] : f.one;
    const bool kl = *a.kl_on != 0u;
    if (nl_store_dense(a, kl)) return;
    const unsigned int total = a.cnt[0];
    const int lane = threadIdx.x & 63;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i - lane < total; i += gridDim.x * blockDim.x) {
        const bool m = i < total && (a.list[2 * i + 1] >> NLI_SLOT_BITS) != NLI_COMPACT;
        const unsigned long long b = __ballot(m);
        if (!b) continue;
        unsigned int base = 0;
        if (lane == 0) base = atomicAdd(a.cnt + 6, (unsigned int)__popcll(b));
        base = __shfl(base, 0);
        if (m) a.miss[base + __popcll(b & ((1ull << lane) - 1ull))] = (int)i;
    }
}

// The weights of those queued pixels, for the lmin smoothing: each pixel's sw*sw weights,
// recomputed from the cached similarity sums and the estimate records with the weights
// kernel's operations in its order (the KL field with its orientation rule, row sums left to
// right, then the patch rows top to bottom, the exponential, NaN / inf and the centre to 0),
// so they are the same bits, into a compact slot of its own; the queue entry then names that
// slot.  One wave per block and per pixel: the missing pixels of all images are numbered end
// to end and dealt out round robin over the blocks; each pixel's (sw + p - 1)^2 estimate
// records are staged in LDS and the lanes take the search-window offsets.  The block shape
// of a pixel's weights launch comes from its slot's shape tag.
template <int PSH>
__global__ __launch_bounds__(64) void nlinsar_fill_weights(const FillBatch f)
{
    constexpr int NP2 = 2 * PSH;
    extern __shared__ float4 stg[];
    const int lane = threadIdx.x;
    unsigned long long base = 0;
    for (int img = 0; img < f.nimg; img++) {
        const WeightArgs& a0 = f.table ? f.table[img] : f.one;
        const bool kl = *a0.kl_on != 0u;
        if (nl_store_dense(a0, kl)) continue;
        const unsigned int total = a0.cnt[6];
        const int wsh = a0.wsh, sw = 2 * wsh + 1, nfull = sw * sw, centre = wsh * sw + wsh;
        const int SE = sw + NP2;   // staged records per side
        for (unsigned long long g = nl_first_item(blockIdx.x, gridDim.x, base); g < base + total; g += gridDim.x) {
            const int e = __builtin_amdgcn_readfirstlane(a0.miss[(unsigned int)(g - base)]);
            const int tagged = __builtin_amdgcn_readfirstlane(a0.list[2 * e + 1]);
            const int sh = tagged >> NLI_SLOT_BITS;
            const int slot = tagged & ((1 << NLI_SLOT_BITS) - 1);
            const WeightArgs& a = f.table ? f.table[sh * f.stride + img] : f.one;
            // the slot of pixel (q, thread) of tile t, as the weights kernel numbers them
            const int TY = f.ty[sh], TX = f.tx[sh], NTW = f.nt[sh];
            const int NG = (NTW / TX) >= TY ? TY : (NTW / TX);
            const int RPT = TY / NG, NCT = NG * TX, NPIX = TY * TX;
            const int t = slot / NPIX, local = slot - t * NPIX;
            const int q = local / NCT, tid = local - q * NCT;
            const int og = tid / TX, oc = tid - og * TX;
            const int py = og * RPT + q, px = oc;
            const NlTile tg = nl_tile(a, t, TY, TX);
            const float* sb = a.scache + ((long long)t * nfull * NPIX + local);
            unsigned int cs = 0;
            if (lane == 0) cs = atomicAdd(a0.cnt + 4, 1u);
            cs = (unsigned int)__builtin_amdgcn_readfirstlane((int)cs);
            float* wq = a0.wc + (long long)cs * nfull;
            // patch position (0, 0) of the pixel
            const int gy0 = tg.ty0 - PSH + py, gx0 = tg.tx0 - PSH + px;
            if (kl) {
                // records of rows gy0 - wsh .. gy0 + NP2 + wsh, same columns
                const unsigned int g0 = (unsigned int)(a.rorg + (gy0 - wsh) * a.rpitch + (gx0 - wsh));
                for (int i = lane; i < SE * SE; i += 64) {
                    const int r = i / SE, c = i - r * SE;
                    stg[i] = a.E4[g0 + (unsigned int)(r * a.rpitch + c)];
                }
                __syncthreads();
            }
            for (int plane = lane; plane < nfull; plane += 64) {
                const int hh = plane / sw;
                const int dy = hh - wsh, dx = plane - hh * sw - wsh;
                float v = -0.0f;
                if (kl && plane != centre) {
                    for (int j = 0; j <= NP2; j++) {
                        const int gy = gy0 + j;
                        const int shy = gy - tg.Y0 + PSH;
                        const float4* rowp = stg + (wsh + j) * SE + wsh;
                        float x[NP2 + 1];
#pragma unroll
                        for (int jj = 0; jj <= NP2; jj++) {
                            const float4 fe = rowp[jj];
                            const float4 vd = rowp[jj + dy * SE + dx];
                            const int sx = gx0 + jj - tg.X0 + PSH;
                            // orientation rule of the similarities, as in the weights kernel
                            const bool swp = dy < 0 && shy >= wsh && sx >= wsh && sx < a.wsim - wsh;
                            x[jj] = nl_pixel_kl_sym(fe, vd, swp ? fe.x : vd.x);
                        }
                        float r = x[0];
#pragma unroll
                        for (int jj = 1; jj <= NP2; jj++) r += x[jj];
                        v = j == 0 ? r : v + r;
                    }
                }
                const float w = nl_weight(sb[(long long)plane * NPIX], v, a.rh, a.rT, plane == centre);
                wq[plane] = w;
            }
            if (lane == 0) {
                a0.list[2 * e + 1] = (NLI_COMPACT << NLI_SLOT_BITS) | (int)cs;
                a0.caux[cs] = 1 << 8;   // a run of one
            }
            // the next pixel's records overwrite these
            __syncthreads();
        }
        base += total;
    }
}

// Marks, for the next iteration's weights launch (NLI_STORE_PRED), the pixels queued in this
// iteration and their eight neighbours with `tag`: pixels the lmin smoothing queued tend to
// be queued again.  When more than `most` pixels were queued, the next launch stores all
// weights instead (cnt[5] = 1): most would be stored anyway.
__global__ void nlinsar_mark_pred(const int* __restrict__ list, unsigned int* __restrict__ cnt,
                                  unsigned char* __restrict__ pred, int H, int W, int tag, unsigned int most)
{
    const unsigned int total = cnt[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[5] = total > most ? 1u : 0u;
    if (total > most) return;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int opix = list[2 * i];
        const int y = opix / W, x = opix - y * W;
        for (int yy = max(y - 1, 0); yy <= min(y + 1, H - 1); yy++)
            for (int xx = max(x - 1, 0); xx <= min(x + 1, W - 1); xx++)
                pred[(long long)yy * W + xx] = (unsigned char)tag;
    }
}

__global__ void nlinsar_zero_bytes(unsigned char* __restrict__ p, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0;
}

struct SmoothArgs {
    Grid R, F;
    int wsh;
    int lmin;
    int H, W, eh, ew;
    int npix[NLI_SHAPES];        // weight slots per block of each shape (TY * TX)
    const float* wt[NLI_SHAPES]; // weight region of each shape
    const float* wc;             // compact weight slots (single calls), see WeightArgs
    const int* caux;
    const float* old_ref;
    const float* old_pha;
    const float* old_coh;
    float* new_ref;
    float* new_pha;
    float* new_coh;
    const int* list;    // (output pixel, weight slot) pairs
    unsigned int* cnt;  // [0] queued this iteration, [1] running total, [3] exact queue
    int* xlist;         // entries of list that need the exact queue
    const float* A;     // a1 on R's element layout
};

// The smoothing launches: one image (`one`) or a batch (`table`, one entry per image); the
// blocks work through the queue of every image in turn.
struct SmoothBatch {
    const SmoothArgs* table;
    SmoothArgs one;
    int nimg;
};

// std::priority_queue<weight_el> of smoothing.cpp:27-36 (top = smallest weight): sift
// rules of libstdc++'s push_heap / pop_heap, which decide which of several equal weights
// stays in the queue.  Entry j of a thread's heap: weight hw[j * bs], index hi[j * bs] (the
// heaps of a block are interleaved, so a wave's accesses to one entry hit distinct banks).
__device__ __forceinline__ bool hp_less(float wa, float wb) { return wa > wb; }

__device__ void hp_push(float* hw, int* hi, int bs, int& n, float w, int idx)
{
    int hole = n++;
    int parent = (hole - 1) / 2;
    while (hole > 0 && hp_less(hw[parent * bs], w)) {
        hw[hole * bs] = hw[parent * bs];
        hi[hole * bs] = hi[parent * bs];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    hw[hole * bs] = w;
    hi[hole * bs] = idx;
}

// pop_heap: the top moves to position n - 1, the heap shrinks to n - 1 entries
__device__ void hp_pop(float* hw, int* hi, int bs, int& n)
{
    const int len = n - 1;
    if (len <= 0) { n = 0; return; }
    const float vw = hw[len * bs];
    const int vi = hi[len * bs];
    hw[len * bs] = hw[0];
    hi[len * bs] = hi[0];
    int hole = 0, child = 0;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (hp_less(hw[child * bs], hw[(child - 1) * bs])) child--;
        hw[hole * bs] = hw[child * bs];
        hi[hole * bs] = hi[child * bs];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        hw[hole * bs] = hw[(child - 1) * bs];
        hi[hole * bs] = hi[(child - 1) * bs];
        hole = child - 1;
    }
    int parent = (hole - 1) / 2;
    while (hole > 0 && hp_less(hw[parent * bs], vw)) {
        hw[hole * bs] = hw[parent * bs];
        hi[hole * bs] = hi[parent * bs];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    hw[hole * bs] = vw;
    hi[hole * bs] = vi;
    n = len;
}

constexpr int NLI_SCH = 16;   // offsets whose loads the register selection issues together
// the candidate flags of one run of CH offsets, one bit per offset
template <int CH>
using NlFlags = typename std::conditional<(CH <= 8), unsigned char, unsigned int>::type;
constexpr int XSCH = 8;       // the same for the exact queue (fewer live values)
// Queues shorter than NLI_QWAVE pixels run one wave per pixel (nlinsar_smooth_wave).
constexpr unsigned int NLI_QWAVE = 1u << 15;
constexpr int NLI_TOPK = 16;  // largest lmin of the smoothing's register selection

struct SmoothPixel {
    const float* A;       // a1 on R's element layout
    const NlF* FV;
    const float* wts;     // weight of offset k at wts[k * npix]
    int npix;             // offset stride of the weights (slots per block of their shape; 1
                          // for a compact slot)
    unsigned int rb, fb;  // R / F element of window offset (0, 0) (grids hold < 2^30)
    float as2;            // 2 * amplitude of the pixel
};

// The exact queue of search_window_smoothing (smoothing.cpp:88-129) followed by the weighted
// means: the lmin largest candidate weights kept by the heap above (its tie rules decide
// which of several equal weights stay), replaced by their mean, summed in pop order.
template <int XSCH>
__device__ void nl_smooth_exact(const SmoothArgs& a, const SmoothPixel& px, float* hw, int* hi,
                                int bs, float& N, float& at_, float& xr, float& xi)
{
    const int sw = 2 * a.wsh + 1;
    const int nfull = sw * sw;
    const int lmin = a.lmin;
    int n = 0;
    float top = 0.0f;   // hw[0] once the queue holds lmin entries
    for (int k0 = 0, hh0 = 0, ww0 = 0; k0 < nfull; k0 += XSCH) {
        float at[XSCH], wv[XSCH];
        int hh = hh0, ww = ww0;
#pragma unroll
        for (int j = 0; j < XSCH; j++) {
            if (k0 + j < nfull) {
                at[j] = px.A[px.rb + (unsigned int)(hh * a.R.pitch + ww)];
                wv[j] = px.wts[(unsigned int)((k0 + j) * px.npix)];
            }
            if (++ww == sw) { ww = 0; hh++; }
        }
#pragma unroll
        for (int j = 0; j < XSCH; j++) {
            if (k0 + j < nfull && at[j] < px.as2) {
                const float w = wv[j];
                if (n < lmin) {
                    hp_push(hw, hi, bs, n, w, k0 + j);
                    top = hw[0];
                } else if (top < w) {
                    hp_pop(hw, hi, bs, n);
                    hp_push(hw, hi, bs, n, w, k0 + j);
                    top = hw[0];
                }
            }
        }
        hh0 = hh;
        ww0 = ww;
    }
    // empty the queue in its order (the order of the sum); pop_heap leaves the popped
    // entries at positions m-1 (first) .. 0 (last)
    const int m = n;
    float wsum = 0.0f;
    while (n > 0) {
        wsum += hw[0];
        hp_pop(hw, hi, bs, n);
    }
    wsum /= (float)m;
    // the selected offsets in increasing order
    for (int i = 1; i < m; i++) {
        const int v = hi[i * bs];
        int j = i - 1;
        while (j >= 0 && hi[j * bs] > v) {
            hi[(j + 1) * bs] = hi[j * bs];
            j--;
        }
        hi[(j + 1) * bs] = v;
    }
    int sel = 0;
    int nxt = m > 0 ? hi[0] : nfull;
    N = at_ = xr = xi = 0.f;
    for (int k0 = 0, hh0 = 0, ww0 = 0; k0 < nfull; k0 += XSCH) {
        float wv[XSCH];
        NlF f[XSCH];
        int hh = hh0, ww = ww0;
#pragma unroll
        for (int j = 0; j < XSCH; j++) {
            if (k0 + j < nfull) {
                wv[j] = px.wts[(unsigned int)((k0 + j) * px.npix)];
                f[j] = px.FV[px.fb + (unsigned int)(hh * a.F.pitch + ww)];
            }
            if (++ww == sw) { ww = 0; hh++; }
        }
#pragma unroll
        for (int j = 0; j < XSCH; j++) {
            if (k0 + j < nfull) {
                float w = wv[j];
                if (k0 + j == nxt) {
                    w = wsum;
                    sel++;
                    nxt = sel < m ? hi[sel * bs] : nfull;
                }
                N += w;
                at_ = fmaf(w, f[j].x, at_);
                xr = fmaf(w, f[j].y, xr);
                xi = fmaf(w, f[j].z, xi);
            }
        }
        hh0 = hh;
        ww0 = ww;
    }
}

__device__ __forceinline__ SmoothPixel nl_smooth_pixel(const SmoothArgs& a, unsigned int li, long long* opix)
{
    const int nfull = (2 * a.wsh + 1) * (2 * a.wsh + 1);
    *opix = a.list[2 * li];
    const int tagged = a.list[2 * li + 1];
    const int shape = tagged >> NLI_SLOT_BITS;
    const int slot = tagged & ((1 << NLI_SLOT_BITS) - 1);
    const int oy = (int)(*opix / a.W), ox = (int)(*opix - (long long)oy * a.W);
    const int vy = oy == a.H - 1 ? oy + a.eh : oy;
    const int vx = ox == a.W - 1 ? ox + a.ew : ox;
    SmoothPixel px;
    px.A = a.A;
    px.FV = reinterpret_cast<const NlF*>(a.F.base);
    if (shape == NLI_COMPACT) {
        // a compact slot: pixel r of a run of k, offset-major
        const int aux = a.caux[slot];
        const int k = aux >> 8, r = aux & 255;
        px.wts = a.wc + ((long long)(slot - r) * nfull + r);
        px.npix = k;
    } else {
        // a tile slot: weights of the block's pixels interleaved, offset-major
        const int npix = shape == 0 ? a.npix[0] : shape == 1 ? a.npix[1] : a.npix[2];
        const float* wbase = shape == 0 ? a.wt[0] : shape == 1 ? a.wt[1] : a.wt[2];
        px.wts = wbase + ((long long)(slot / npix) * nfull * npix + slot % npix);
        px.npix = npix;
    }
    px.rb = (unsigned int)a.R.at(vy - a.wsh, vx - a.wsh);
    px.fb = (unsigned int)a.F.at(vy - a.wsh, vx - a.wsh);
    // at < 2.0 * (double)as in the reference: 2 * as is exact in float unless it
    // overflows, and then both forms are true for every finite at
    px.as2 = 2.0f * px.A[a.R.at(vy, vx)];
    return px;
}

// search_window_smoothing (smoothing.cpp:78-129) and compute_insar for the queued pixels,
// one pixel per thread.  Offsets are visited in the reference's row-major order, CH at a time:
// their loads are issued together and then used in order.  In the heaviest iteration at
// 4096²: CH = 4 at 4 waves per SIMD, 8 at 3 waves (3 % faster), 16 at 2 waves (NLI_SCH, 3 %
// faster again: 18.2-18.4 against 18.8-18.9 ms); 24 within 0.5 % of 16 at 248 registers; 8
// with the registers capped for 4 waves per SIMD 5 % slower than 8 at 3.  Each variant
// serves the queue lengths in [qlo, qhi) and returns otherwise.  Each XCD takes a contiguous
// eighth of the queue (nl_xcd_items).
//
// The queue keeps the lmin largest candidate weights (candidates: amplitude < 2 * the
// pixel's).  Let T be the smallest kept weight.  Unless some candidate that the queue
// dropped also weighs exactly T, the kept set is exactly {candidates with weight >= T},
// whatever the heap's tie rules, and the pop-order sum is the ascending sum of those
// weights.  So for lmin < NLI_TOPK the largest weights are tracked in registers (one
// median-of-three per slot and offset); the (lmin+1)-th largest tells whether such a tie
// exists.  Pixels without one get their weighted means here; the others (and every pixel
// when lmin >= NLI_TOPK) go to a second queue that nlinsar_smooth_exact works through with
// the heap, so that neither kernel's waves wait on the other path.  The first pass keeps
// the candidate flags of each run of CH offsets in LDS (entry c of a thread at
// c * blockDim.x), which the second pass reads instead of the amplitudes.
// One image of nlinsar_smooth: its queued pixels from work item `base` on (base advances past
// them).
template <int CH>
__device__ __forceinline__ void nl_smooth_image(const SmoothArgs& a, unsigned int qlo, unsigned int qhi,
                                                const XcdItems& xs, unsigned long long& base)
{
    using Flags = NlFlags<CH>;
    static_assert(CH <= 32, "candidate flags of a run are kept in one 32-bit word");
    extern __shared__ float4 smem_flags[];
    const unsigned int total = a.cnt[0];
    if (total < qlo || total >= qhi) return;
    const unsigned long long first_item = nl_xcd_first(xs, base > xs.lo ? base : xs.lo);
    const unsigned long long end = base + total < xs.hi ? base + total : xs.hi;
    base += total;
    const int sw = 2 * a.wsh + 1;
    const int nfull = sw * sw;
    const int lmin = a.lmin;
    const int lane = threadIdx.x & 63;
    Flags* cf = reinterpret_cast<Flags*>(smem_flags) + threadIdx.x;
    const int bs = blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.cnt[1] += total;
    for (unsigned long long g = first_item; g < end; g += xs.nw) {
        const unsigned int li = (unsigned int)(g - (base - total));
        long long opix;
        const SmoothPixel px = nl_smooth_pixel(a, li, &opix);
        bool exact = lmin >= NLI_TOPK;
        float T = -INFINITY, wsum = 0.0f;
        int m = 0;
        if (!exact) {
            // the NLI_TOPK largest candidate weights, v[0] >= v[1] >= ...
            float v[NLI_TOPK];
#pragma unroll
            for (int i = 0; i < NLI_TOPK; i++) v[i] = -INFINITY;
            int cnt = 0;
            for (int k0 = 0, hh0 = 0, ww0 = 0; k0 < nfull; k0 += CH) {
                float at[CH], wv[CH];
                int hh = hh0, ww = ww0;
#pragma unroll
                for (int j = 0; j < CH; j++) {
                    if (k0 + j < nfull) {
                        at[j] = px.A[px.rb + (unsigned int)(hh * a.R.pitch + ww)];
                        wv[j] = px.wts[(unsigned int)((k0 + j) * px.npix)];
                    }
                    if (++ww == sw) { ww = 0; hh++; }
                }
                unsigned int bits = 0u;
#pragma unroll
                for (int j = 0; j < CH; j++) {
                    if (k0 + j < nfull) {
                        const bool cand = at[j] < px.as2;
                        const float x = cand ? wv[j] : -INFINITY;
                        cnt += cand;
                        bits |= (unsigned int)cand << j;
#pragma unroll
                        for (int i = NLI_TOPK - 1; i > 0; i--) v[i] = __builtin_amdgcn_fmed3f(v[i], x, v[i - 1]);
                        v[0] = fmaxf(v[0], x);
                    }
                }
                cf[(k0 / CH) * bs] = (Flags)bits;
                hh0 = hh;
                ww0 = ww;
            }
            m = min(cnt, lmin);
            float next = -INFINITY;
#pragma unroll
            for (int i = NLI_TOPK - 1; i >= 0; i--) {
                if (cnt >= lmin && i == lmin - 1) T = v[i];
                if (i == lmin) next = v[i];
                if (i < m) wsum += v[i];
            }
            wsum /= (float)m;
            exact = cnt > lmin && next == T;
        }
        const unsigned long long xm = __ballot(exact);
        if (xm) {
            unsigned int base = 0;
            const int first = __ffsll((long long)xm) - 1;
            if (lane == first) base = atomicAdd(a.cnt + 3, (unsigned int)__popcll(xm));
            base = __shfl(base, first);
            if (exact) a.xlist[base + __popcll(xm & ((1ull << lane) - 1ull))] = (int)li;
        }
        if (exact) continue;
        float N = 0.f, at_ = 0.f, xr = 0.f, xi = 0.f;
        for (int k0 = 0, hh0 = 0, ww0 = 0; k0 < nfull; k0 += CH) {
            float wv[CH];
            NlF f[CH];
            int hh = hh0, ww = ww0;
#pragma unroll
            for (int j = 0; j < CH; j++) {
                if (k0 + j < nfull) {
                    wv[j] = px.wts[(unsigned int)((k0 + j) * px.npix)];
                    f[j] = px.FV[px.fb + (unsigned int)(hh * a.F.pitch + ww)];
                }
                if (++ww == sw) { ww = 0; hh++; }
            }
            const unsigned int bits = cf[(k0 / CH) * bs];
#pragma unroll
            for (int j = 0; j < CH; j++) {
                if (k0 + j < nfull) {
                    const bool sel = ((bits >> j) & 1u) && wv[j] >= T;
                    const float w = sel ? wsum : wv[j];
                    N += w;
                    at_ = fmaf(w, f[j].x, at_);
                    xr = fmaf(w, f[j].y, xr);
                    xi = fmaf(w, f[j].z, xi);
                }
            }
            hh0 = hh;
            ww0 = ww;
        }
        nl_estimate(N, at_, xr, xi, opix, a.old_ref, a.old_pha, a.old_coh, a.new_ref, a.new_pha,
                    a.new_coh);
    }
}

template <int CH>
__global__ __launch_bounds__(256) void nlinsar_smooth(const SmoothBatch b, unsigned int qlo, unsigned int qhi)
{
    unsigned long long base = 0;
    if (!b.table) {
        const unsigned int t = b.one.cnt[0];
        const XcdItems xs = nl_xcd_items(t >= qlo && t < qhi ? t : 0u);
        nl_smooth_image<CH>(b.one, qlo, qhi, xs, base);
        return;
    }
    unsigned long long items = 0;
    for (int img = 0; img < b.nimg; img++) {
        const unsigned int t = b.table[img].cnt[0];
        if (t >= qlo && t < qhi) items += t;
    }
    const XcdItems xs = nl_xcd_items(items);
    for (int img = 0; img < b.nimg; img++) nl_smooth_image<CH>(b.table[img], qlo, qhi, xs, base);
}

// The pixels nlinsar_smooth queued for the exact heap (ties at the smallest kept weight);
// the variant with runs of CH offsets serves exact-queue lengths in [qlo, qhi).
template <int CH>
__device__ __forceinline__ void nl_smooth_exact_image(const SmoothArgs& a, unsigned int qlo, unsigned int qhi,
                                                      const XcdItems& xs, unsigned long long& base)
{
    extern __shared__ float smem[];
    const unsigned int total = a.cnt[3];
    if (total < qlo || total >= qhi) return;
    const unsigned long long first_item = nl_xcd_first(xs, base > xs.lo ? base : xs.lo);
    const unsigned long long end = base + total < xs.hi ? base + total : xs.hi;
    base += total;
    const int bs = blockDim.x;
    float* hw = smem + threadIdx.x;
    int* hi = reinterpret_cast<int*>(smem + a.lmin * bs + threadIdx.x);
    for (unsigned long long g = first_item; g < end; g += xs.nw) {
        const unsigned int xi_ = (unsigned int)(g - (base - total));
        long long opix;
        const SmoothPixel px = nl_smooth_pixel(a, (unsigned int)a.xlist[xi_], &opix);
        float N, at_, xr, xi;
        nl_smooth_exact<CH>(a, px, hw, hi, bs, N, at_, xr, xi);
        nl_estimate(N, at_, xr, xi, opix, a.old_ref, a.old_pha, a.old_coh, a.new_ref, a.new_pha,
                    a.new_coh);
    }
}

template <int CH>
__global__ __launch_bounds__(256) void nlinsar_smooth_exact(const SmoothBatch b, unsigned int qlo, unsigned int qhi)
{
    unsigned long long base = 0;
    if (!b.table) {
        const unsigned int t = b.one.cnt[3];
        const XcdItems xs = nl_xcd_items(t >= qlo && t < qhi ? t : 0u);
        nl_smooth_exact_image<CH>(b.one, qlo, qhi, xs, base);
        return;
    }
    unsigned long long items = 0;
    for (int img = 0; img < b.nimg; img++) {
        const unsigned int t = b.table[img].cnt[3];
        if (t >= qlo && t < qhi) items += t;
    }
    const XcdItems xs = nl_xcd_items(items);
    for (int img = 0; img < b.nimg; img++) nl_smooth_exact_image<CH>(b.table[img], qlo, qhi, xs, base);
}

// Short queues, one wave per pixel.  Per thread, the chain of dependent load runs of the
// kernels above sets a floor of about half a millisecond, and the heap replays of a wave's
// tied pixels add one more, however few pixels are queued.  Here the wave loads the pixel's
// window (candidate flags, weights, filter values) into LDS at once, and its first lane runs
// the reference's queue and the weighted sums in plane order from there, with the
// arithmetic of nl_smooth_exact.  One block is one wave.  TIES: the pixels of the exact
// queue that nlinsar_smooth left (when that queue is short), instead of the whole queue.
template <bool TIES>
__device__ __forceinline__ void nl_smooth_wave_image(const SmoothArgs& a, unsigned int qhi, unsigned long long& base)
{
    extern __shared__ float4 smem_wave[];
    const unsigned int total = TIES ? a.cnt[3] : a.cnt[0];
    if (total >= qhi) return;
    const unsigned long long first_item = nl_first_item(blockIdx.x, gridDim.x, base);
    base += total;
    if (!TIES && blockIdx.x == 0 && threadIdx.x == 0) a.cnt[1] += total;
    const int sw = 2 * a.wsh + 1;
    const int nfull = sw * sw;
    const int np4 = (nfull + 3) & ~3;                 // arrays padded to whole float4
    const int nwords = (nfull + 63) / 64 * 2;
    const int lane = threadIdx.x;
    float* W = reinterpret_cast<float*>(smem_wave);   // weight of offset k
    float* Fx = W + np4;                              // filter values a, Re x, Im x
    float* Fy = Fx + np4;
    float* Fz = Fy + np4;
    unsigned int* C = reinterpret_cast<unsigned int*>(Fz + np4);   // candidate flags
    float* hw = reinterpret_cast<float*>(C + nwords);
    int* hi = reinterpret_cast<int*>(hw + a.lmin);
    for (unsigned long long g = first_item; g < base; g += gridDim.x) {
        const unsigned int qi = (unsigned int)(g - (base - total));
        const unsigned int li = TIES ? (unsigned int)a.xlist[qi] : qi;
        long long opix;
        const SmoothPixel px = nl_smooth_pixel(a, li, &opix);
        for (int k0 = 0; k0 < nfull; k0 += 64) {
            const int k = k0 + lane;
            bool cand = false;
            if (k < nfull) {
                const int hh = k / sw, ww = k - (k / sw) * sw;
                cand = px.A[px.rb + (unsigned int)(hh * a.R.pitch + ww)] < px.as2;
                W[k] = px.wts[(unsigned int)(k * px.npix)];
                const NlF f = px.FV[px.fb + (unsigned int)(hh * a.F.pitch + ww)];
                Fx[k] = f.x;
                Fy[k] = f.y;
                Fz[k] = f.z;
            }
            const unsigned long long b = __ballot(cand);
            if (lane == 0) {
                C[k0 / 32] = (unsigned int)b;
                C[k0 / 32 + 1] = (unsigned int)(b >> 32);
            }
        }
        __syncthreads();
        if (lane == 0) {
            const int lmin = a.lmin;
            int n = 0;
            float top = 0.0f;
            // four offsets per LDS read; runs of four without a candidate are skipped
            for (int k0 = 0; k0 < nfull; k0 += 4) {
                const unsigned int c4 = (C[k0 >> 5] >> (k0 & 31)) & 15u;
                if (!c4) continue;
                const float4 w4 = reinterpret_cast<const float4*>(W)[k0 >> 2];
                const float wj[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (!((c4 >> j) & 1u)) continue;
                    const float w = wj[j];
                    if (n < lmin) {
                        hp_push(hw, hi, 1, n, w, k0 + j);
                        top = hw[0];
                    } else if (top < w) {
                        hp_pop(hw, hi, 1, n);
                        hp_push(hw, hi, 1, n, w, k0 + j);
                        top = hw[0];
                    }
                }
            }
            const int m = n;
            float wsum = 0.0f;
            while (n > 0) {
                wsum += hw[0];
                hp_pop(hw, hi, 1, n);
            }
            wsum /= (float)m;
            for (int i = 1; i < m; i++) {
                const int v = hi[i];
                int j = i - 1;
                while (j >= 0 && hi[j] > v) {
                    hi[j + 1] = hi[j];
                    j--;
                }
                hi[j + 1] = v;
            }
            int sel = 0;
            int nxt = m > 0 ? hi[0] : nfull;
            float N = 0.f, at_ = 0.f, xr = 0.f, xi = 0.f;
            for (int k0 = 0; k0 < nfull; k0 += 4) {
                const float4 w4 = reinterpret_cast<const float4*>(W)[k0 >> 2];
                const float4 x4 = reinterpret_cast<const float4*>(Fx)[k0 >> 2];
                const float4 y4 = reinterpret_cast<const float4*>(Fy)[k0 >> 2];
                const float4 z4 = reinterpret_cast<const float4*>(Fz)[k0 >> 2];
                const float wj[4] = {w4.x, w4.y, w4.z, w4.w};
                const float xj[4] = {x4.x, x4.y, x4.z, x4.w};
                const float yj[4] = {y4.x, y4.y, y4.z, y4.w};
                const float zj[4] = {z4.x, z4.y, z4.z, z4.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (k0 + j >= nfull) break;
                    float w = wj[j];
                    if (k0 + j == nxt) {
                        w = wsum;
                        sel++;
                        nxt = sel < m ? hi[sel] : nfull;
                    }
                    N += w;
                    at_ = fmaf(w, xj[j], at_);
                    xr = fmaf(w, yj[j], xr);
                    xi = fmaf(w, zj[j], xi);
                }
            }
            nl_estimate(N, at_, xr, xi, opix, a.old_ref, a.old_pha, a.old_coh, a.new_ref, a.new_pha,
                        a.new_coh);
        }
        __syncthreads();
    }
}

template <bool TIES>
__global__ __launch_bounds__(64) void nlinsar_smooth_wave(const SmoothBatch b, unsigned int qhi)
{
    unsigned long long base = 0;
    if (!b.table) {
        nl_smooth_wave_image<TIES>(b.one, qhi, base);
        return;
    }
    for (int img = 0; img < b.nimg; img++) nl_smooth_wave_image<TIES>(b.table[img], qhi, base);
}

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

// LDS of one weights block: the staged records of its search neighbourhood, then the
// similarity / KL fields and row sums (once each in the first iteration, KL double-buffered
// after)
static size_t nlinsar_lds_bytes(int TY, int TX, int wsh, int psh)
{
    const int o = wsh + psh, np2 = 2 * psh;
    return sizeof(float) * (4 * (size_t)(TY + 2 * o) * (TX + 2 * o) +
                            2 * ((size_t)(TY + np2) * (TX + np2) + (size_t)(TY + np2) * TX));
}

static const int kTY[3] = {27, 18, 54};
static const int kTX[3] = {54, 54, 54};
// the thin shapes of the last row / corner and of the last column (NLI_SHAPES)
// (512 threads: a thin block is a latency-bound chain of 441 offsets, so each thread takes
// at most two positions per step)
constexpr int NLI_ROW_TY = 1, NLI_ROW_TX = 54, NLI_ROW_NT = 512;
constexpr int NLI_COL_TY = 54, NLI_COL_TX = 3, NLI_COL_NT = 512;

static void shape_dims(int shape, int tile, int* TY, int* TX)
{
    *TY = shape == 0 ? kTY[tile] : shape == 1 ? NLI_ROW_TY : NLI_COL_TY;
    *TX = shape == 0 ? kTX[tile] : shape == 1 ? NLI_ROW_TX : NLI_COL_TX;
}

// The launch mappings a shape covers (the others emptied).  thin: 0 the interior, 1 the last
// row and the corner, 2 the last column; else shape 0 covers all four.
static void shape_mappings(const Launch* L, int shape, bool thin, Launch* out)
{
    for (int i = 0; i < 4; i++) {
        const bool mine = !thin ? shape == 0
                        : shape == 0 ? i == 0 : shape == 1 ? (i == 1 || i == 3) : i == 2;
        out[i] = L[i];
        if (!mine) out[i].ny = out[i].nx = 0;
    }
}

// Block grid of the weights kernel over the four launch mappings: reference tiles covered,
// times sub-tiles per reference tile; the one-row / one-column mappings use a single block
// row / column.  Returns the number of blocks.
static int nlinsar_grid(int step, int TY, int TX, const Launch* L, int* gx, int* nsx, int* nsy,
                        int* bstart)
{
    int total = 0;
    for (int i = 0; i < 4; i++) {
        bstart[i] = total;
        if (L[i].ny <= 0 || L[i].nx <= 0) {
            gx[i] = nsx[i] = nsy[i] = 1;
            continue;
        }
        nsy[i] = L[i].ny == 1 ? 1 : (step + TY - 1) / TY;
        nsx[i] = L[i].nx == 1 ? 1 : (step + TX - 1) / TX;
        const int gy = L[i].ny == 1 ? 1 : ((L[i].ny + step - 1) / step) * nsy[i];
        gx[i] = L[i].nx == 1 ? 1 : ((L[i].nx + step - 1) / step) * nsx[i];
        total += gx[i] * gy;
    }
    bstart[4] = total;
    return total;
}

static void nlinsar_mappings(int H, int W, int eh, int ew, Launch* L)
{
    L[0] = Launch{0, 0, H - 1, W - 1, 0, 0};
    L[1] = Launch{H - 1 + eh, 0, 1, W - 1, H - 1, 0};
    L[2] = Launch{0, W - 1 + ew, H - 1, 1, 0, W - 1};
    L[3] = Launch{H - 1 + eh, W - 1 + ew, 1, 1, H - 1, W - 1};
}

int nlinsar_plan(const NlinsarProblem& pb, NlinsarPlan* plan)
{
    if (pb.H <= 0 || pb.W <= 0) { set_error("nlinsar: empty image"); return -1; }
    if (pb.sw < 3 || pb.sw % 2 == 0 || pb.p < 1 || pb.p % 2 == 0) {
        set_error("nlinsar: search window and patch sizes must be odd");
        return -1;
    }
    if (pb.p > 2 * NLI_MAX_PSH + 1) { set_error("nlinsar: patch sizes must be odd, up to 21"); return -1; }
    if (pb.niter < 0 || pb.lmin < 1 || pb.lmin > pb.sw * pb.sw) {
        set_error("nlinsar: niter must be >= 0 and 1 <= lmin <= search_window_size^2");
        return -1;
    }
    plan->wsh = (pb.sw - 1) / 2;
    plan->psh = (pb.p - 1) / 2;
    plan->overlap = plan->wsh + plan->psh;
    plan->step = pb.ref_tile - 2 * plan->overlap;
    if (plan->step <= 0) { set_error("nlinsar: search window and patch too large for the 80-pixel tiles"); return -1; }
    if (pb.tile == NLINSAR_TILE_AUTO) {
        // 18 x 54 blocks of 512 threads: two per CU (LDS), and the fastest shape at every
        // measured size since the weight stores became sparse (4096^2 with scatterers, 10
        // iterations: 279 ms against 300 ms for 54 x 54 and 356 ms for 27 x 54)
        plan->tile = 1;
    } else {
        plan->tile = std::min(std::max(pb.tile, 0), 2);
    }
    plan->TY = kTY[plan->tile];
    plan->TX = kTX[plan->tile];
    // the weights kernel's LDS (16 bytes more on a batch's persistent grid)
    if (nlinsar_lds_bytes(plan->TY, plan->TX, plan->wsh, plan->psh) + 16 > 160 * 1024) {
        set_error("nlinsar: search window and patch too large for the LDS of this tile shape");
        return -2;
    }
    // virtual position written to the last row/column by the reference's 80x80 tiling
    // (tile_iterator.cpp:37-44, sub_images.cpp:47-69): the last tile's last interior row
    auto vmax = [&](int n) {
        int low = -plan->overlap, last = low;
        while (low < n) { last = low; low += plan->step; }
        return last + pb.ref_tile - plan->overlap - 1;
    };
    plan->eh = vmax(pb.H) - (pb.H - 1);
    plan->ew = vmax(pb.W) - (pb.W - 1);
    const int H = pb.H, W = pb.W;
    // rows: the interior's blocks and the last column's 54-row blocks; columns: 54 wide blocks
    const int TYg = std::max(plan->TY, NLI_COL_TY);
    const int Hx = std::max(round_up(std::max(H - 1, 1), plan->step) + TYg, H - 1 + plan->eh + TYg);
    const int Wx = std::max(round_up(std::max(W - 1, 1), plan->step) + plan->TX, W - 1 + plan->ew + plan->TX);
    const int o = plan->overlap;
    plan->R = Grid{nullptr, 0, round_up(Wx + 2 * o, 64), o, o, Hx + 2 * o, Wx + 2 * o};
    plan->R.plane = (long long)plan->R.pitch * plan->R.h;
    plan->E = plan->R;
    const int f = plan->wsh;
    plan->F = Grid{nullptr, 0, round_up(Wx + 2 * f, 64), f, f, Hx + 2 * f, Wx + 2 * f};
    plan->F.plane = (long long)plan->F.pitch * plan->F.h;
    const int nfull = pb.sw * pb.sw;
    const int per = 2 * pb.lmin * 4;
    int hb = 256;
    while (hb > 1 && (size_t)hb * per > 48 * 1024) hb /= 2;
    plan->heap_block = hb;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t q = off; off += (bytes + 255) / 256 * 256; return q; };
    const size_t n = (size_t)H * W;
    // kernels address the float4 grids with 32-bit element offsets
    if (plan->R.plane >= (1ll << 30) || plan->F.plane >= (1ll << 30)) {
        set_error("nlinsar: image too large for one call (split it into bands)");
        return -4;
    }
    plan->off_R = take(sizeof(float) * 4 * plan->R.plane);
    plan->off_A = take(sizeof(float) * plan->R.plane);
    plan->off_E = take(sizeof(float) * 4 * plan->E.plane);
    plan->off_F = take(sizeof(NlF) * plan->F.plane);
    plan->off_est = take(sizeof(float) * 6 * n);
    Launch L[4];
    nlinsar_mappings(H, W, plan->eh, plan->ew, L);
    int gx[4], nsx[4], nsy[4], bstart[5];
    for (int sh = 0; sh < NLI_SHAPES; sh++) {
        int TY, TX;
        shape_dims(sh, plan->tile, &TY, &TX);
        Launch Ls[4];
        shape_mappings(L, sh, pb.thin_edges != 0, Ls);
        plan->blocks[sh] = nlinsar_grid(plan->step, TY, TX, Ls, gx, nsx, nsy, bstart);
        const size_t slots = (size_t)plan->blocks[sh] * TY * TX;
        if (slots >= (1ull << NLI_SLOT_BITS)) {
            set_error("nlinsar: image too large for one call (split it into bands)");
            return -4;
        }
        // shape 0's weight region also holds the compact slots (at most one per pixel)
        plan->off_w[sh] = take(sizeof(float) * (size_t)nfull * (sh == 0 ? std::max(slots, n) : slots));
        plan->off_s[sh] = take(sizeof(float) * (size_t)nfull * slots);
    }
    plan->off_pred = take(n);
    plan->off_caux = take(sizeof(int) * n);
    plan->off_list = take(2 * sizeof(int) * n);
    plan->off_xlist = take(sizeof(int) * n);
    // counters: [0] queued this iteration, [1] queued in all, [2] KL on, [3] exact queue,
    // [4] compact weight slots, [5] store all (NLI_STORE_PRED), [6] missing weights
    plan->off_cnt = take(sizeof(unsigned int) * 8);
    plan->ws_bytes = off;
    return 0;
}

static int launch_smoothing(const SmoothBatch& sb, int sw, int lmin, int heap_block, long long n, int ncu,
                            hipStream_t stream);

// One weights launch of one shape over the mappings L (the others emptied); timed as `stage`.
template <int TY, int TX, int PSH, bool FIRST, int NT>
static int launch_weights(const WeightArgs& base, const Launch* L, int stage, hipStream_t stream)
{
    const size_t lds = nlinsar_lds_bytes(TY, TX, base.wsh, PSH);
    if (lds > 160 * 1024) { set_error("nlinsar: patch too large for the LDS budget"); return -2; }
    DKX_HIP_CHECK(hipFuncSetAttribute((const void*)nlinsar_weights<TY, TX, PSH, FIRST, NT>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    WeightArgs a = base;
    for (int i = 0; i < 4; i++) a.L[i] = L[i];
    const int total = nlinsar_grid(base.step, TY, TX, L, a.gx, a.nsx, a.nsy, a.bstart);
    if (total <= 0) return 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = timing_on();
    if (timed) {
        DKX_HIP_CHECK(hipEventCreate(&e0));
        DKX_HIP_CHECK(hipEventCreate(&e1));
        DKX_HIP_CHECK(hipEventRecord(e0, stream));
    }
    hipLaunchKernelGGL((nlinsar_weights<TY, TX, PSH, FIRST, NT>), dim3(total), dim3(NT), lds, stream, a);
    DKX_HIP_CHECK(hipGetLastError());
    if (timed) {
        DKX_HIP_CHECK(hipEventRecord(e1, stream));
        timing_push(stage, e0, e1);
    }
    return 0;
}

// The weights of one iteration: with thin edges the last-row / corner and last-column
// shapes, then the interior with the tile shape; else one launch of the tile shape.  a: arguments without mappings, shape or cache / weight
// regions; wt / sc: each shape's regions.
template <int PSH, bool FIRST>
static int launch_weights_all(int tile, bool thin, const WeightArgs& a, const Launch* L, float* const* wt,
                              float* const* sc, hipStream_t stream)
{
    Launch Ls[4];
    WeightArgs w = a;
    int rc;
    shape_mappings(L, 1, thin, Ls);
    w.shape = 1; w.wt = wt[1]; w.scache = sc[1];
    rc = launch_weights<NLI_ROW_TY, NLI_ROW_TX, PSH, FIRST, NLI_ROW_NT>(w, Ls, STAGE_NLINSAR_EDGE, stream);
    if (rc) return rc;
    shape_mappings(L, 2, thin, Ls);
    w.shape = 2; w.wt = wt[2]; w.scache = sc[2];
    rc = launch_weights<NLI_COL_TY, NLI_COL_TX, PSH, FIRST, NLI_COL_NT>(w, Ls, STAGE_NLINSAR_EDGE, stream);
    if (rc) return rc;
    shape_mappings(L, 0, thin, Ls);
    w.shape = 0; w.wt = wt[0]; w.scache = sc[0];
    switch (tile) {
    case 0: return launch_weights<27, 54, PSH, FIRST, 512>(w, Ls, STAGE_NLINSAR_WEIGHTS, stream);
    case 1: return launch_weights<18, 54, PSH, FIRST, 512>(w, Ls, STAGE_NLINSAR_WEIGHTS, stream);
    default: return launch_weights<54, 54, PSH, FIRST, 1024>(w, Ls, STAGE_NLINSAR_WEIGHTS, stream);
    }
}

// The shapes of a weights launch's blocks (NLI_SHAPES), for the fill kernels.
static void fill_shapes(int tile, FillBatch* f)
{
    f->ty[0] = kTY[tile]; f->tx[0] = kTX[tile]; f->nt[0] = tile == 2 ? 1024 : 512;
    f->ty[1] = NLI_ROW_TY; f->tx[1] = NLI_ROW_TX; f->nt[1] = NLI_ROW_NT;
    f->ty[2] = NLI_COL_TY; f->tx[2] = NLI_COL_TX; f->nt[2] = NLI_COL_NT;
}

// The weights the weights launches did not store, for the queued pixels: the scan over each
// image's queue, then the recomputation (timed as one stage).  grid_x: scan blocks per image.
template <int PSH>
static int launch_fill(const FillBatch& f, int wsh, int grid_x, int ncu, hipStream_t stream)
{
    const int se = 2 * wsh + 1 + 2 * PSH;
    const size_t lds = sizeof(float4) * (size_t)se * se;
    if (lds > 64 * 1024) { set_error("nlinsar: search window and patch too large"); return -2; }
    DKX_HIP_CHECK(hipFuncSetAttribute((const void*)nlinsar_fill_weights<PSH>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = timing_on();
    if (timed) {
        DKX_HIP_CHECK(hipEventCreate(&e0));
        DKX_HIP_CHECK(hipEventCreate(&e1));
        DKX_HIP_CHECK(hipEventRecord(e0, stream));
    }
    hipLaunchKernelGGL(nlinsar_fill_scan, dim3(grid_x, f.nimg), dim3(256), 0, stream, f);
    DKX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(nlinsar_fill_weights<PSH>, dim3(ncu * 16), dim3(64), lds, stream, f);
    DKX_HIP_CHECK(hipGetLastError());
    if (timed) {
        DKX_HIP_CHECK(hipEventRecord(e1, stream));
        timing_push(STAGE_NLINSAR_FILL, e0, e1);
    }
    return 0;
}

static int launch_fill_psh(int psh, const FillBatch& f, int wsh, int grid_x, int ncu, hipStream_t stream)
{
    switch (psh) {
    case 0: return launch_fill<0>(f, wsh, grid_x, ncu, stream);
    case 1: return launch_fill<1>(f, wsh, grid_x, ncu, stream);
    case 2: return launch_fill<2>(f, wsh, grid_x, ncu, stream);
    case 3: return launch_fill<3>(f, wsh, grid_x, ncu, stream);
    case 4: return launch_fill<4>(f, wsh, grid_x, ncu, stream);
    case 5: return launch_fill<5>(f, wsh, grid_x, ncu, stream);
    case 6: return launch_fill<6>(f, wsh, grid_x, ncu, stream);
    case 7: return launch_fill<7>(f, wsh, grid_x, ncu, stream);
    case 8: return launch_fill<8>(f, wsh, grid_x, ncu, stream);
    case 9: return launch_fill<9>(f, wsh, grid_x, ncu, stream);
    case 10: return launch_fill<10>(f, wsh, grid_x, ncu, stream);
    default: set_error("nlinsar: patch sizes must be odd, up to 21"); return -1;
    }
}

// Store mode of iteration it's weights launch (NlinsarProblem::store).  Automatic: the first
// iteration stores its weights only when the KL term is on (from zero estimates a weight is a
// function of the cached similarity alone); the next two store all (the lmin queue changes
// too much from one iteration to the next early on: from 0.04 % to 11 % to 88 % of the pixels
// on a 512^2 crop of the 4096^2 scatterer image); later ones store the pixels queued in the
// previous iteration and their neighbours (there 94 % / 19 % / 9 % ... of the pixels, missing
// 0.6 % / 0.3 % / 0.1 % ... that the fill recomputes).
static int nli_store_mode(int setting, int it)
{
    switch (setting) {
    case NLINSAR_STORE_ALL: return NLI_STORE_ALL;
    case NLINSAR_STORE_PRED: return NLI_STORE_PRED;
    case NLINSAR_STORE_NONE: return NLI_STORE_NONE;
    default: return it == 0 ? NLI_STORE_KL : it <= 2 ? NLI_STORE_ALL : NLI_STORE_PRED;
    }
}

int nlinsar_run(const NlinsarProblem& pb, const NlinsarPlan& plan, const float* d_a1,
                const float* d_a2, const float* d_phi, float* d_ref, float* d_phase,
                float* d_coh, void* d_ws, unsigned long long* n_smoothed_host, hipStream_t stream)
{
    char* ws = (char*)d_ws;
    const int H = pb.H, W = pb.W;
    const size_t n = (size_t)H * W;
    Grid R = plan.R, E = plan.E, F = plan.F;
    R.base = (float*)(ws + plan.off_R);
    E.base = (float*)(ws + plan.off_E);
    F.base = (float*)(ws + plan.off_F);
    float* est = (float*)(ws + plan.off_est);  // two sets of (ref, phase, coh)
    float* wt[NLI_SHAPES];
    float* scache[NLI_SHAPES];
    for (int sh = 0; sh < NLI_SHAPES; sh++) {
        wt[sh] = (float*)(ws + plan.off_w[sh]);
        scache[sh] = (float*)(ws + plan.off_s[sh]);
    }
    int* list = (int*)(ws + plan.off_list);
    unsigned int* cnt = (unsigned int*)(ws + plan.off_cnt);
    unsigned char* pred = (unsigned char*)(ws + plan.off_pred);
    int* caux = (int*)(ws + plan.off_caux);
    int hdev = 0;
    DKX_HIP_CHECK(hipGetDevice(&hdev));
    int ncu = 256;
    DKX_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, hdev));

    unsigned int* flag = cnt + 2;
    hipLaunchKernelGGL(nlinsar_clear_counters, dim3(1), dim3(64), 0, stream, cnt, flag, 1);
    DKX_HIP_CHECK(hipGetLastError());
    // no pixel marked for storing (marks are the iteration's tag, 1..255)
    hipLaunchKernelGGL(nlinsar_zero_bytes, dim3(std::max(1, std::min(ncu * 4, (int)((n + 255) / 256)))), dim3(256), 0,
                       stream, pred, (long long)n);
    DKX_HIP_CHECK(hipGetLastError());
    {
        dim3 blk(256), grd((R.w + 255) / 256, R.h);
        hipLaunchKernelGGL(nlinsar_fill_raw, grd, blk, 0, stream, d_a1, d_a2, d_phi, H, W, R,
                           (float*)(ws + plan.off_A));
        DKX_HIP_CHECK(hipGetLastError());
        dim3 grf((F.w + 255) / 256, F.h);
        hipLaunchKernelGGL(nlinsar_filter_values, grf, blk, 0, stream, R, F);
        DKX_HIP_CHECK(hipGetLastError());
    }
    // iteration 1 starts from the caller's output arrays (nlinsar.cpp:59-61)
    float* cur[3] = {d_ref, d_phase, d_coh};
    float* nxt[3] = {est, est + n, est + 2 * n};
    float* alt[3] = {est + 3 * n, est + 4 * n, est + 5 * n};
    Launch L[4];
    nlinsar_mappings(H, W, plan.eh, plan.ew, L);
    for (int it = 0; it < pb.niter; it++) {
        const int store = nli_store_mode(pb.store, it);
        const int tag = it % 255 + 1;
        if (store == NLI_STORE_PRED && it > 0) {
            // the previous iteration's queue (its length still in cnt[0]) and neighbours
            // automatic: all weights stored after an iteration that queued more than a quarter
            // of the pixels
            const unsigned int most = pb.store == NLINSAR_STORE_AUTO ? (unsigned int)(n / 4) : 0xffffffffu;
            hipLaunchKernelGGL(nlinsar_mark_pred, dim3(ncu * 4), dim3(256), 0, stream, (const int*)list, cnt,
                               pred, H, W, tag, most);
            DKX_HIP_CHECK(hipGetLastError());
        }
        {
            dim3 blk(256), grd((E.w + 255) / 256, E.h);
            hipLaunchKernelGGL(nlinsar_fill_estimates, grd, blk, 0, stream, (const float*)cur[0],
                               (const float*)cur[1], (const float*)cur[2], H, W, E);
            DKX_HIP_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(nlinsar_clear_counters, dim3(1), dim3(64), 0, stream, cnt, flag, 0);
        DKX_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(nlinsar_any_estimate, dim3(std::max(1, std::min(ncu * 4, (int)((n + 255) / 256)))),
                           dim3(256), 0, stream, (const float*)cur[0], (long long)n, flag);
        DKX_HIP_CHECK(hipGetLastError());
        WeightArgs wa;
        wa.R4 = reinterpret_cast<const float4*>(R.base);
        wa.E4 = reinterpret_cast<const float4*>(E.base);
        wa.FV = reinterpret_cast<const NlF*>(F.base);
        wa.rpitch = R.pitch;
        wa.rorg = R.oy * R.pitch + R.ox;
        wa.fpitch = F.pitch;
        wa.forg = F.oy * F.pitch + F.ox;
        wa.wsh = plan.wsh; wa.step = plan.step;
        wa.wsim = pb.ref_tile - pb.sw + 1;
        wa.rh = 1.0f / pb.h; wa.rT = 1.0f / pb.T; wa.lmin = (float)pb.lmin;
        wa.H = H; wa.W = W;
        wa.kl_on = flag;
        wa.old_ref = cur[0]; wa.old_pha = cur[1]; wa.old_coh = cur[2];
        wa.new_ref = nxt[0]; wa.new_pha = nxt[1]; wa.new_coh = nxt[2];
        wa.list = list; wa.cnt = cnt;
        wa.block0 = 0;
        wa.shape = 0;
        wa.store = store;
        wa.tag = tag;
        wa.pred = pred;
        wa.wc = wt[0];
        wa.caux = caux;
        wa.dense = cnt + 5;
        // the exact queue's list, which the smoothing fills later, holds the missing entries
        wa.miss = (int*)(ws + plan.off_xlist);
        const bool thin = pb.thin_edges != 0;
        if (thin && store != NLI_STORE_ALL) {
            set_error("nlinsar: thin edge blocks store every weight");
            return -1;
        }
        int rc;
        switch (plan.psh) {
#define DKX_NLI_PSH(P)                                                                        \
        case P:                                                                               \
            rc = it == 0 ? launch_weights_all<P, true>(plan.tile, thin, wa, L, wt, scache, stream)  \
                         : launch_weights_all<P, false>(plan.tile, thin, wa, L, wt, scache, stream); \
            break;
        DKX_NLI_PSH(0) DKX_NLI_PSH(1) DKX_NLI_PSH(2) DKX_NLI_PSH(3) DKX_NLI_PSH(4) DKX_NLI_PSH(5)
        DKX_NLI_PSH(6) DKX_NLI_PSH(7) DKX_NLI_PSH(8) DKX_NLI_PSH(9) DKX_NLI_PSH(10)
#undef DKX_NLI_PSH
        default:
            set_error("nlinsar: patch sizes must be odd, up to 21");
            rc = -1;
        }
        if (rc) return rc;
        if (store != NLI_STORE_ALL) {
            // the weights the launch did not store, for the queued pixels
            FillBatch fb;
            memset(&fb, 0, sizeof(fb));
            fb.one = wa;
            for (int i = 0; i < 4; i++) fb.one.L[i] = L[i];
            nlinsar_grid(plan.step, plan.TY, plan.TX, L, fb.one.gx, fb.one.nsx, fb.one.nsy, fb.one.bstart);
            fb.one.wt = wt[0];
            fb.one.scache = scache[0];
            fb.nimg = 1;
            fill_shapes(plan.tile, &fb);
            rc = launch_fill_psh(plan.psh, fb, plan.wsh, ncu * 2, ncu, stream);
            if (rc) return rc;
        }
        SmoothArgs sa;
        sa.R = R; sa.F = F; sa.wsh = plan.wsh; sa.lmin = pb.lmin;
        sa.H = H; sa.W = W; sa.eh = plan.eh; sa.ew = plan.ew;
        for (int sh = 0; sh < NLI_SHAPES; sh++) {
            int TY, TX;
            shape_dims(sh, plan.tile, &TY, &TX);
            sa.npix[sh] = TY * TX;
            sa.wt[sh] = wt[sh];
        }
        sa.old_ref = cur[0]; sa.old_pha = cur[1]; sa.old_coh = cur[2];
        sa.new_ref = nxt[0]; sa.new_pha = nxt[1]; sa.new_coh = nxt[2];
        sa.list = list; sa.cnt = cnt;
        sa.xlist = (int*)(ws + plan.off_xlist);
        sa.A = (const float*)(ws + plan.off_A);
        sa.wc = wt[0];
        sa.caux = caux;
        SmoothBatch sb;
        memset(&sb, 0, sizeof(sb));
        sb.one = sa;
        sb.nimg = 1;
        rc = launch_smoothing(sb, pb.sw, pb.lmin, plan.heap_block, (long long)n, ncu, stream);
        if (rc) return rc;
        // the estimates just written become the next iteration's input
        for (int k = 0; k < 3; k++) {
            float* t = nxt[k];
            if (it == 0) {
                cur[k] = t;
                nxt[k] = alt[k];
            } else {
                nxt[k] = cur[k];
                cur[k] = t;
            }
        }
    }
    // final estimates back into the caller's arrays
    if (pb.niter > 0)
        for (int k = 0; k < 3; k++) {
            float* dst = k == 0 ? d_ref : (k == 1 ? d_phase : d_coh);
            DKX_HIP_CHECK(hipMemcpyAsync(dst, cur[k], n * sizeof(float), hipMemcpyDeviceToDevice, stream));
        }
    if (n_smoothed_host) {
        unsigned int tot = 0;
        DKX_HIP_CHECK(hipMemcpyAsync(&tot, cnt + 1, sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
        DKX_HIP_CHECK(hipStreamSynchronize(stream));
        *n_smoothed_host = tot;
    }
    return 0;
}


// ---- batches ---------------------------------------------------------------------------
// Many images with the same parameters: every stage is one launch over all images (the
// elementwise stages on a (rows, image) grid, the weights on the persistent tile queue, the
// smoothing variants looping over the images' queues), so a batch of small images fills the
// GPU as one large image does.  Each image's stages, arithmetic and summation orders are the
// ones of nlinsar_run on that image alone with the batch's tile shape, so its result is the
// same bits.

// Per image: inputs, caller arrays (first-iteration estimates in, result out), its grids.
struct NliImage {
    const float* a1;
    const float* a2;
    const float* phi;
    float* io[3];
    long long n;
    int H, W;
    Grid R, E, F;
    float* A;
    unsigned int* cnt;    // the counters of nlinsar_plan (off_cnt)
    unsigned char* pred;  // store marks (NLI_STORE_PRED)
    const int* list;      // the queue
};

// The estimate arrays per iteration: iteration 0 reads the caller's arrays and writes set 0;
// odd iterations read set 0 and write set 1; even ones from 2 on read set 1 and write set 0
// (the rotation of nlinsar_run).  Weight / smoothing tables exist per kind.
static int iteration_kind(int it) { return it == 0 ? 0 : ((it & 1) ? 1 : 2); }

__global__ void nlinsar_b_clear(const NliImage* __restrict__ im, int nimg, unsigned int* __restrict__ queue,
                                int clear_total)
{
    for (int i = threadIdx.x; i < nimg; i += blockDim.x) {
        unsigned int* c = im[i].cnt;
        c[0] = 0u;
        if (clear_total) c[1] = 0u;
        c[2] = 0u;
        c[3] = 0u;
        c[4] = 0u;
        if (clear_total) c[5] = 0u;
        c[6] = 0u;
    }
    if (threadIdx.x < 8 * NLI_SHAPES) queue[threadIdx.x] = 0u;
}

// NLI_STORE_PRED iterations of a batch, before the queue counters are cleared: every
// image's weights launch stores all (dense = 1), or only the marked pixels, which the two
// kernels below mark from the image's previous queue (the marks of earlier iterations
// cleared first; tag 1).  Grid (x, image) for the marks.
__global__ void nlinsar_b_set_dense(const NliImage* __restrict__ im, int nimg)
{
    for (int i = threadIdx.x; i < nimg; i += blockDim.x) im[i].cnt[5] = 1u;
}

__global__ __launch_bounds__(256) void nlinsar_b_clear_marks(const NliImage* __restrict__ im)
{
    const NliImage& g = im[blockIdx.y];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += (long long)gridDim.x * blockDim.x)
        g.pred[i] = 0;
}

__global__ __launch_bounds__(256) void nlinsar_b_mark(const NliImage* __restrict__ im, int quarter)
{
    const NliImage& g = im[blockIdx.y];
    const unsigned int total = g.cnt[0];
    // automatic (quarter): all weights stored after an iteration that queued more than a
    // quarter of the pixels
    const bool all = quarter && (long long)total > g.n / 4;
    if (blockIdx.x == 0 && threadIdx.x == 0) g.cnt[5] = all ? 1u : 0u;
    if (all) return;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int opix = g.list[2 * i];
        const int y = opix / g.W, x = opix - y * g.W;
        for (int yy = max(y - 1, 0); yy <= min(y + 1, g.H - 1); yy++)
            for (int xx = max(x - 1, 0); xx <= min(x + 1, g.W - 1); xx++)
                g.pred[(long long)yy * g.W + xx] = 1;
    }
}

// Grid (x, image): blocks x, x + gridDim.x, ... take the rows of their image.
__global__ __launch_bounds__(256) void nlinsar_b_fill_raw(const NliImage* __restrict__ im)
{
    const NliImage& g = im[blockIdx.y];
    for (int gy = blockIdx.x; gy < g.R.h; gy += gridDim.x)
        for (int gx = threadIdx.x; gx < g.R.w; gx += blockDim.x)
            nl_raw_record(g.a1, g.a2, g.phi, g.H, g.W, g.R, g.A, gy, gx);
}

__global__ __launch_bounds__(256) void nlinsar_b_filter_values(const NliImage* __restrict__ im)
{
    const NliImage& g = im[blockIdx.y];
    for (int gy = blockIdx.x; gy < g.F.h; gy += gridDim.x)
        for (int gx = threadIdx.x; gx < g.F.w; gx += blockDim.x)
            nl_filter_record(g.R, g.F, gy, gx);
}

// The estimates that the iteration of `wt` reads (wt[i].old_*) on image i's grid E.
__global__ __launch_bounds__(256) void nlinsar_b_fill_estimates(const NliImage* __restrict__ im,
                                                                const WeightArgs* __restrict__ wt)
{
    const NliImage& g = im[blockIdx.y];
    const WeightArgs& a = wt[blockIdx.y];
    for (int gy = blockIdx.x; gy < g.E.h; gy += gridDim.x)
        for (int gx = threadIdx.x; gx < g.E.w; gx += blockDim.x)
            nl_estimate_record(a.old_ref, a.old_pha, a.old_coh, g.H, g.W, g.E, gy, gx);
}

// nlinsar_any_estimate per image
__global__ __launch_bounds__(256) void nlinsar_b_any_estimate(const NliImage* __restrict__ im,
                                                              const WeightArgs* __restrict__ wt)
{
    const NliImage& g = im[blockIdx.y];
    const float* ref = wt[blockIdx.y].old_ref;
    bool nz = false;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < g.n; k += (long long)gridDim.x * blockDim.x)
        nz |= ref[k] != 0.0f;
    if (__any(nz) && (threadIdx.x & 63) == 0) atomicOr(g.cnt + 2, 1u);
}

// Zero first-iteration estimates in the caller's arrays (a filter run from scratch).
__global__ __launch_bounds__(256) void nlinsar_b_zero(const NliImage* __restrict__ im)
{
    const NliImage& g = im[blockIdx.y];
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < g.n; k += (long long)gridDim.x * blockDim.x) {
        g.io[0][k] = 0.0f;
        g.io[1][k] = 0.0f;
        g.io[2][k] = 0.0f;
    }
}

// The last iteration's estimates (wt[i].new_*) into the caller's arrays.
__global__ __launch_bounds__(256) void nlinsar_b_copy_out(const NliImage* __restrict__ im,
                                                          const WeightArgs* __restrict__ wt)
{
    const NliImage& g = im[blockIdx.y];
    const WeightArgs& a = wt[blockIdx.y];
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < g.n; k += (long long)gridDim.x * blockDim.x) {
        g.io[0][k] = a.new_ref[k];
        g.io[1][k] = a.new_pha[k];
        g.io[2][k] = a.new_coh[k];
    }
}

// Device tables of a batch of n images: descriptors, weight and smoothing arguments of the
// three iteration kinds, the XCD queues.
static void batch_table_layout(int n, NlinsarBatchLaunch* L)
{
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t q = off; off += (bytes + 255) / 256 * 256; return q; };
    L->off_img = take(sizeof(NliImage) * (size_t)n);
    L->off_w = take(sizeof(WeightArgs) * NLI_SHAPES * 3 * (size_t)n);
    L->off_s = take(sizeof(SmoothArgs) * 3 * (size_t)n);
    L->off_q = take(sizeof(unsigned int) * 8 * NLI_SHAPES);
    L->bytes = off;
}

int nlinsar_batch_plan(const NlinsarProblem& common, const int* H, const int* W, int n,
                       std::vector<NlinsarPlan>& plans, NlinsarBatchLaunch* L)
{
    if (n < 1) { set_error("nlinsar batch: no images"); return -1; }
    if (n > 65535) { set_error("nlinsar batch: at most 65535 images"); return -1; }
    long long px = 0, max_n = 0;
    for (int i = 0; i < n; i++) {
        if (H[i] <= 0 || W[i] <= 0) { set_error("nlinsar batch: empty image " + std::to_string(i)); return -1; }
        px += (long long)H[i] * W[i];
        max_n = std::max(max_n, (long long)H[i] * W[i]);
    }
    // the batch's tile shape: the single-image rule applied to the batch's pixel count
    NlinsarProblem pb = common;
    pb.thin_edges = 1;
    pb.H = (int)std::min<long long>(px, 1ll << 20);
    pb.W = 1;
    NlinsarPlan probe;
    int rc = nlinsar_plan(pb, &probe);
    if (rc) return rc;
    const int tile = probe.tile;
    plans.assign(n, NlinsarPlan{});
    long long total[NLI_SHAPES] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        pb = common;
        pb.thin_edges = 1;
        pb.H = H[i];
        pb.W = W[i];
        pb.tile = tile;
        rc = nlinsar_plan(pb, &plans[i]);
        if (rc) return rc;
        for (int sh = 0; sh < NLI_SHAPES; sh++) {
            total[sh] += plans[i].blocks[sh];
            if (total[sh] >= (1ll << 30)) { set_error("nlinsar batch: too many tiles"); return -4; }
        }
    }
    L->nimg = n;
    L->niter = common.niter;
    L->tile = tile;
    L->psh = probe.psh;
    L->sw = common.sw;
    L->lmin = common.lmin;
    L->store = common.store;
    for (int sh = 0; sh < NLI_SHAPES; sh++) L->total_tiles[sh] = (int)total[sh];
    L->persistent_blocks = 0;   // set by the first enqueue (it needs the device)
    L->max_n = max_n;
    L->heap_block = probe.heap_block;
    batch_table_layout(n, L);
    return 0;
}

size_t nlinsar_batch_table_bytes(int n)
{
    NlinsarBatchLaunch L;
    batch_table_layout(n, &L);
    return L.bytes;
}

int nlinsar_batch_image_bytes(const NlinsarProblem& common, int H, int W, size_t* bytes)
{
    size_t best = 0;
    for (int t = 0; t < 3; t++) {
        if (common.tile != NLINSAR_TILE_AUTO && t != common.tile) continue;
        if (common.tile == NLINSAR_TILE_AUTO && t == 0) continue;   // the automatic rule picks 1 or 2
        NlinsarProblem pb = common;
        pb.thin_edges = 1;
        pb.H = H;
        pb.W = W;
        pb.tile = t;
        NlinsarPlan plan;
        const int rc = nlinsar_plan(pb, &plan);
        if (rc) return rc;
        best = std::max(best, plan.ws_bytes);
    }
    *bytes = best;
    return 0;
}

int nlinsar_batch_tables(const NlinsarProblem& common, const std::vector<NlinsarPlan>& plans,
                         const NlinsarBatchItem* items, const NlinsarBatchLaunch& L,
                         char* d_tables, std::vector<unsigned char>& host)
{
    const int n = L.nimg;
    host.assign(L.bytes, 0);
    NliImage* img = reinterpret_cast<NliImage*>(host.data() + L.off_img);
    WeightArgs* wt = reinterpret_cast<WeightArgs*>(host.data() + L.off_w);
    SmoothArgs* st = reinterpret_cast<SmoothArgs*>(host.data() + L.off_s);
    int block0[NLI_SHAPES] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        const NlinsarPlan& plan = plans[i];
        const NlinsarBatchItem& it = items[i];
        char* ws = it.ws;
        const long long npx = (long long)it.H * it.W;
        NliImage& g = img[i];
        g.a1 = it.a1; g.a2 = it.a2; g.phi = it.phi;
        g.io[0] = it.ref; g.io[1] = it.phase; g.io[2] = it.coh;
        g.n = npx;
        g.H = it.H; g.W = it.W;
        g.R = plan.R; g.E = plan.E; g.F = plan.F;
        g.R.base = (float*)(ws + plan.off_R);
        g.E.base = (float*)(ws + plan.off_E);
        g.F.base = (float*)(ws + plan.off_F);
        g.A = (float*)(ws + plan.off_A);
        g.cnt = (unsigned int*)(ws + plan.off_cnt);
        g.pred = (unsigned char*)(ws + plan.off_pred);
        g.list = (const int*)(ws + plan.off_list);
        float* est = (float*)(ws + plan.off_est);
        float* set0[3] = {est, est + npx, est + 2 * npx};
        float* set1[3] = {est + 3 * npx, est + 4 * npx, est + 5 * npx};
        Launch Lall[4];
        nlinsar_mappings(it.H, it.W, plan.eh, plan.ew, Lall);
        for (int kind = 0; kind < 3; kind++) {
            float* const* cur = kind == 0 ? g.io : (kind == 1 ? set0 : set1);
            float* const* nxt = kind == 1 ? set1 : set0;
            for (int sh = 0; sh < NLI_SHAPES; sh++) {
                int TY, TX;
                shape_dims(sh, L.tile, &TY, &TX);
                WeightArgs& a = wt[((size_t)sh * 3 + kind) * n + i];
                a.R4 = reinterpret_cast<const float4*>(g.R.base);
                a.E4 = reinterpret_cast<const float4*>(g.E.base);
                a.FV = reinterpret_cast<const NlF*>(g.F.base);
                a.rpitch = g.R.pitch;
                a.rorg = g.R.oy * g.R.pitch + g.R.ox;
                a.fpitch = g.F.pitch;
                a.forg = g.F.oy * g.F.pitch + g.F.ox;
                a.wsh = plan.wsh; a.step = plan.step;
                a.wsim = common.ref_tile - common.sw + 1;
                a.rh = 1.0f / common.h; a.rT = 1.0f / common.T; a.lmin = (float)common.lmin;
                a.H = it.H; a.W = it.W;
                shape_mappings(Lall, sh, true, a.L);
                const int nb = nlinsar_grid(plan.step, TY, TX, a.L, a.gx, a.nsx, a.nsy, a.bstart);
                if (nb != plan.blocks[sh]) { set_error("nlinsar batch: inconsistent tile grid"); return -1; }
                a.scache = (float*)(ws + plan.off_s[sh]);
                a.wt = (float*)(ws + plan.off_w[sh]);
                a.kl_on = g.cnt + 2;
                a.old_ref = cur[0]; a.old_pha = cur[1]; a.old_coh = cur[2];
                a.new_ref = nxt[0]; a.new_pha = nxt[1]; a.new_coh = nxt[2];
                a.list = (int*)(ws + plan.off_list);
                a.cnt = g.cnt;
                a.block0 = block0[sh];
                a.shape = sh;
                // store modes (NlinsarProblem::store): the first iteration's kind stores per the
                // KL flag, the others predict (all stored in iterations 2-3 of the automatic
                // mode, by the dense flag)
                a.store = L.store == NLINSAR_STORE_ALL    ? NLI_STORE_ALL
                        : L.store == NLINSAR_STORE_NONE   ? NLI_STORE_NONE
                        : L.store == NLINSAR_STORE_PRED   ? NLI_STORE_PRED
                        : kind == 0                       ? NLI_STORE_KL
                                                          : NLI_STORE_PRED;
                a.tag = 1;
                a.pred = g.pred;
                a.wc = (float*)(ws + plan.off_w[0]);
                a.caux = (int*)(ws + plan.off_caux);
                a.dense = g.cnt + 5;
                a.miss = (int*)(ws + plan.off_xlist);
            }
            SmoothArgs& sa = st[kind * n + i];
            sa.R = g.R; sa.F = g.F; sa.wsh = plan.wsh; sa.lmin = common.lmin;
            sa.H = it.H; sa.W = it.W; sa.eh = plan.eh; sa.ew = plan.ew;
            for (int sh = 0; sh < NLI_SHAPES; sh++) {
                int TY, TX;
                shape_dims(sh, L.tile, &TY, &TX);
                sa.npix[sh] = TY * TX;
                sa.wt[sh] = (const float*)(ws + plan.off_w[sh]);
            }
            sa.old_ref = cur[0]; sa.old_pha = cur[1]; sa.old_coh = cur[2];
            sa.new_ref = nxt[0]; sa.new_pha = nxt[1]; sa.new_coh = nxt[2];
            sa.list = (const int*)(ws + plan.off_list); sa.cnt = g.cnt;
            sa.xlist = (int*)(ws + plan.off_xlist);
            sa.A = g.A;
            sa.wc = (const float*)(ws + plan.off_w[0]);
            sa.caux = (const int*)(ws + plan.off_caux);
        }
        for (int sh = 0; sh < NLI_SHAPES; sh++) block0[sh] += plan.blocks[sh];
    }
    (void)d_tables;
    return 0;
}

// One persistent weights launch of one shape over `total` tiles of the batch (table: that
// shape's arguments of every image, queue: its 8 counters); timed as `stage`.
template <int TY, int TX, int PSH, bool FIRST, int NT>
static int launch_weights_batch(const NlinsarBatchLaunch& L, const WeightArgs* table, int total,
                                unsigned int* queue, int stage, int ncu, hipStream_t stream)
{
    if (total <= 0) return 0;
    const size_t lds = nlinsar_lds_bytes(TY, TX, (L.sw - 1) / 2, PSH);
    if (lds + 16 > 160 * 1024) { set_error("nlinsar: patch too large for the LDS budget"); return -2; }
    const void* fn = (const void*)nlinsar_weights_batch<TY, TX, PSH, FIRST, NT>;
    DKX_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    DKX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, lds));
    per_cu = std::max(per_cu, 1);
    // every XCD queue needs a block: at least 8, at most what is resident at once
    const int grid = std::max(8, std::min(per_cu * ncu, (total + 7) / 8 * 8));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = timing_on();
    if (timed) {
        DKX_HIP_CHECK(hipEventCreate(&e0));
        DKX_HIP_CHECK(hipEventCreate(&e1));
        DKX_HIP_CHECK(hipEventRecord(e0, stream));
    }
    hipLaunchKernelGGL((nlinsar_weights_batch<TY, TX, PSH, FIRST, NT>), dim3(grid), dim3(NT), lds, stream, table,
                       L.nimg, total, queue);
    DKX_HIP_CHECK(hipGetLastError());
    if (timed) {
        DKX_HIP_CHECK(hipEventRecord(e1, stream));
        timing_push(stage, e0, e1);
    }
    return 0;
}

// The weights of one iteration of a batch: the thin shapes, then the interior.  wk: the
// iteration kind's arguments of shape 0; shape s follows at wk + s * 3 * nimg.
template <int PSH, bool FIRST>
static int launch_weights_batch_all(const NlinsarBatchLaunch& L, const WeightArgs* wk, unsigned int* queue,
                                    int ncu, hipStream_t stream)
{
    const size_t per_shape = (size_t)3 * L.nimg;
    int rc = launch_weights_batch<NLI_ROW_TY, NLI_ROW_TX, PSH, FIRST, NLI_ROW_NT>(
        L, wk + per_shape, L.total_tiles[1], queue + 8, STAGE_NLINSAR_EDGE, ncu, stream);
    if (rc) return rc;
    rc = launch_weights_batch<NLI_COL_TY, NLI_COL_TX, PSH, FIRST, NLI_COL_NT>(
        L, wk + 2 * per_shape, L.total_tiles[2], queue + 16, STAGE_NLINSAR_EDGE, ncu, stream);
    if (rc) return rc;
    switch (L.tile) {
    case 0: return launch_weights_batch<27, 54, PSH, FIRST, 512>(L, wk, L.total_tiles[0], queue, STAGE_NLINSAR_WEIGHTS, ncu, stream);
    case 1: return launch_weights_batch<18, 54, PSH, FIRST, 512>(L, wk, L.total_tiles[0], queue, STAGE_NLINSAR_WEIGHTS, ncu, stream);
    default: return launch_weights_batch<54, 54, PSH, FIRST, 1024>(L, wk, L.total_tiles[0], queue, STAGE_NLINSAR_WEIGHTS, ncu, stream);
    }
}

// The four smoothing launches of one iteration (nlinsar_run's sequence); n: pixels of the
// largest image.
static int launch_smoothing(const SmoothBatch& sb, int sw, int lmin, int heap_block, long long n, int ncu,
                            hipStream_t stream)
{
    const int nfull = sw * sw;
    const size_t lds = (size_t)heap_block * 2 * lmin * 4;
    // candidate flags: per thread one NlFlags word per run of NLI_SCH offsets; blocks are
    // narrowed until they fit 48 KB
    const size_t fl = (size_t)(nfull + NLI_SCH - 1) / NLI_SCH * sizeof(NlFlags<NLI_SCH>);
    int fbs = 256;
    while (fbs > 64 && fbs * fl > 48 * 1024) fbs /= 2;
    // multiples of 8 blocks: each XCD takes a contiguous eighth of the queue (nl_xcd_items)
    const int blocks = (int)((std::max<long long>(1, std::min<long long>((n + fbs - 1) / fbs, ncu * 8)) + 7) / 8 * 8);
    const int xblocks =
        (int)((std::max<long long>(1, std::min<long long>((n + heap_block - 1) / heap_block, ncu * 8)) + 7) / 8 * 8);
    // one wave per pixel for short queues: window in LDS, then the heap
    const size_t wlds = 4 * (4 * (size_t)((nfull + 3) & ~3) + (size_t)(nfull + 63) / 64 * 2 + 2 * (size_t)lmin);
    const int wblocks = (int)std::min<long long>(n, NLI_QWAVE);
    if (wlds > 64 * 1024) { set_error("nlinsar: search window too large for the smoothing"); return -2; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timing_on()) {
        DKX_HIP_CHECK(hipEventCreate(&e0));
        DKX_HIP_CHECK(hipEventCreate(&e1));
        DKX_HIP_CHECK(hipEventRecord(e0, stream));
    }
    // the variants split the queue lengths between them: each returns at once outside its
    // range (the lengths are known on the device only).  Short queue: one wave per pixel;
    // else the register selection, whose tied pixels go to the waves again when they are few
    // and to the per-thread heap otherwise
    hipLaunchKernelGGL(nlinsar_smooth_wave<false>, dim3(wblocks), dim3(64), wlds, stream, sb, NLI_QWAVE);
    DKX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(nlinsar_smooth<NLI_SCH>, dim3(blocks), dim3(fbs), fbs * fl, stream, sb, NLI_QWAVE,
                       0xffffffffu);
    DKX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(nlinsar_smooth_wave<true>, dim3(wblocks), dim3(64), wlds, stream, sb, NLI_QWAVE);
    DKX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(nlinsar_smooth_exact<XSCH>, dim3(xblocks), dim3(heap_block), lds, stream, sb, NLI_QWAVE,
                       0xffffffffu);
    DKX_HIP_CHECK(hipGetLastError());
    if (e0) {
        DKX_HIP_CHECK(hipEventRecord(e1, stream));
        timing_push(STAGE_NLINSAR_SMOOTH, e0, e1);
    }
    return 0;
}

int nlinsar_batch_enqueue(const NlinsarProblem& common, const NlinsarBatchLaunch& L,
                          char* d_tables, hipStream_t stream, int zero_init)
{
    (void)common;
    const int n = L.nimg;
    const NliImage* img = reinterpret_cast<const NliImage*>(d_tables + L.off_img);
    const WeightArgs* wt = reinterpret_cast<const WeightArgs*>(d_tables + L.off_w);
    const SmoothArgs* st = reinterpret_cast<const SmoothArgs*>(d_tables + L.off_s);
    unsigned int* queue = reinterpret_cast<unsigned int*>(d_tables + L.off_q);
    int hdev = 0;
    DKX_HIP_CHECK(hipGetDevice(&hdev));
    int ncu = 256;
    DKX_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, hdev));
    // (rows, image) grids of the elementwise stages: about 8 blocks per CU in all
    const int bx = std::max(1, ncu * 8 / n);
    const dim3 egrid(bx, n), eblk(256);
    hipLaunchKernelGGL(nlinsar_b_clear, dim3(1), dim3(256), 0, stream, img, n, queue, 1);
    DKX_HIP_CHECK(hipGetLastError());
    if (zero_init) {
        hipLaunchKernelGGL(nlinsar_b_zero, egrid, eblk, 0, stream, img);
        DKX_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(nlinsar_b_fill_raw, egrid, eblk, 0, stream, img);
    DKX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(nlinsar_b_filter_values, egrid, eblk, 0, stream, img);
    DKX_HIP_CHECK(hipGetLastError());
    const bool sparse = L.store != NLINSAR_STORE_ALL;
    for (int it = 0; it < L.niter; it++) {
        const int kind = iteration_kind(it);
        const WeightArgs* wk = wt + (size_t)kind * n;   // shape 0 of this kind (the per-image
                                                         // estimate pointers are alike in all shapes)
        if (sparse && it > 0 && L.store != NLINSAR_STORE_NONE) {
            // which weights this iteration's launches store (before the counters are cleared:
            // the marks come from the previous queue)
            if (L.store == NLINSAR_STORE_AUTO && it <= 2) {
                hipLaunchKernelGGL(nlinsar_b_set_dense, dim3(1), dim3(256), 0, stream, img, n);
                DKX_HIP_CHECK(hipGetLastError());
            } else {
                hipLaunchKernelGGL(nlinsar_b_clear_marks, egrid, eblk, 0, stream, img);
                DKX_HIP_CHECK(hipGetLastError());
                hipLaunchKernelGGL(nlinsar_b_mark, egrid, eblk, 0, stream, img,
                                   L.store == NLINSAR_STORE_AUTO ? 1 : 0);
                DKX_HIP_CHECK(hipGetLastError());
            }
        }
        hipLaunchKernelGGL(nlinsar_b_fill_estimates, egrid, eblk, 0, stream, img, wk);
        DKX_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(nlinsar_b_clear, dim3(1), dim3(256), 0, stream, img, n, queue, 0);
        DKX_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(nlinsar_b_any_estimate, egrid, eblk, 0, stream, img, wk);
        DKX_HIP_CHECK(hipGetLastError());
        int rc;
        switch (L.psh) {
#define DKX_NLI_PSH(P)                                                                        \
        case P:                                                                               \
            rc = it == 0 ? launch_weights_batch_all<P, true>(L, wk, queue, ncu, stream)       \
                         : launch_weights_batch_all<P, false>(L, wk, queue, ncu, stream);     \
            break;
        DKX_NLI_PSH(0) DKX_NLI_PSH(1) DKX_NLI_PSH(2) DKX_NLI_PSH(3) DKX_NLI_PSH(4) DKX_NLI_PSH(5)
        DKX_NLI_PSH(6) DKX_NLI_PSH(7) DKX_NLI_PSH(8) DKX_NLI_PSH(9) DKX_NLI_PSH(10)
#undef DKX_NLI_PSH
        default:
            set_error("nlinsar: patch sizes must be odd, up to 21");
            rc = -1;
        }
        if (rc) return rc;
        if (sparse) {
            // the weights the launches did not store, for the queued pixels of every image
            FillBatch fb;
            memset(&fb, 0, sizeof(fb));
            fb.table = wk;
            fb.stride = 3 * n;
            fb.nimg = n;
            fill_shapes(L.tile, &fb);
            rc = launch_fill_psh(L.psh, fb, (L.sw - 1) / 2, bx, ncu, stream);
            if (rc) return rc;
        }
        SmoothBatch sb;
        memset(&sb, 0, sizeof(sb));
        sb.table = st + (size_t)kind * n;
        sb.nimg = n;
        rc = launch_smoothing(sb, L.sw, L.lmin, L.heap_block, L.max_n, ncu, stream);
        if (rc) return rc;
    }
    if (L.niter > 0) {
        hipLaunchKernelGGL(nlinsar_b_copy_out, egrid, eblk, 0, stream, img,
                           wt + (size_t)iteration_kind(L.niter - 1) * n);
        DKX_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

}  // namespace dkx
