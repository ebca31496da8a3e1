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
#include <cmath>

namespace dkx {

// replicate an H x W image set (nplanes planes) into a padded grid, clamping coordinates
__global__ void grid_fill(const float* __restrict__ p0, const float* __restrict__ p1,
                          const float* __restrict__ p2, int H, int W, Grid G)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= G.w || gy >= G.h) return;
    const int y = min(H - 1, max(0, gy - G.oy));
    const int x = min(W - 1, max(0, gx - G.ox));
    const long long i = (long long)y * W + x;
    const long long o = (long long)gy * G.pitch + gx;
    G.base[o] = p0[i];
    G.base[G.plane + o] = p1[i];
    G.base[2 * G.plane + o] = p2[i];
}

// precompute_filter_values.cl:19-39: a = (a1^2 + a2^2)/2, x = a1*a2*exp(-i*phase)
__global__ void nlinsar_filter_values(Grid R, Grid F)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= F.w || gy >= F.h) return;
    const long long r = R.at(gy - F.oy, gx - F.ox);
    const float a1 = R.base[r], a2 = R.base[R.plane + r], dp = R.base[2 * R.plane + r];
    const long long o = (long long)gy * F.pitch + gx;
    F.base[o] = 0.5f * (a1 * a1 + a2 * a2);
    F.base[F.plane + o] = a1 * a2 * cosf(-dp);
    F.base[2 * F.plane + o] = a1 * a2 * sinf(-dp);
}

// pixel_similarity, nlinsar/sim_measures.h:27-40
__device__ __forceinline__ float nl_pixel_similarity(float am1, float as1, float dp1, float am2,
                                                     float as2, float dp2)
{
    const float s = am1 * am1 + as1 * as1 + am2 * am2 + as2 * as2;
    const float A = s * s;
    const float B = 4.0f * (am1 * am1 * as1 * as1 + am2 * am2 * as2 * as2 +
                            2.0f * am1 * as1 * am2 * as2 * cosf(dp1 - dp2));
    const float C = am1 * as1 * am2 * as2;
    const float f_el = powf(C / B, 1.5f) * ((A + B) / A * sqrtf(B / (A - B)) - asinf(sqrtf(B / A)));
    if ((f_el == 0) || isnan(f_el)) return 0.f;
    return logf(f_el);
}

// pixel_kullback_leibler, nlinsar/sim_measures.h:42-49
__device__ __forceinline__ float nl_pixel_kl(float ref1, float dp1, float coh1, float ref2,
                                             float dp2, float coh2)
{
    const float cs = cosf(dp1 - dp2);
    const float a1 = (ref2 != 0.0f) ? ref1 / ref2 * (1.0f - coh1 * coh2 * cs) / (1.0f - coh2 * coh2) : 1.0f;
    const float a2 = (ref2 != 0.0f) ? ref2 / ref1 * (1.0f - coh1 * coh2 * cs) / (1.0f - coh1 * coh1) : 1.0f;
    return -4.0f / 3.14159274101257324219f * (a1 + a2 - 2.0f);
}

__device__ __forceinline__ bool nl_is_valid(float x) { return !(isnan(x) || isinf(x) || x == 0); }

// compute_insar.cl:95-117: new estimate, or the previous one where the value is invalid
__device__ __forceinline__ void nl_estimate(float N, float at, float xr, float xi, long long opix,
                                            const float* old_ref, const float* old_pha,
                                            const float* old_coh, float* nref, float* npha,
                                            float* ncoh)
{
    const float reflectivity = at / N;
    const float iphase = atan2f(xi, xr);
    const float coherence = sqrtf(xr * xr + xi * xi) / at;
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
    Grid R, E, F;
    int wsh, psh, step, wsim;
    float h, T;
    float lmin;
    Launch L;
    int H, W;
    float* wt;  // [sw*sw][H*W]
    const float* old_ref;
    const float* old_pha;
    const float* old_coh;
    float* new_ref;
    float* new_pha;
    float* new_coh;
    int* list;
    unsigned int* cnt;
};

constexpr int NTW = 256;

// EDGE: separate kernel symbol for the thin last-row / last-column launches
template <int TY, int TX, bool EDGE>
__global__ __launch_bounds__(NTW) void nlinsar_weights(const WeightArgs a)
{
    constexpr int NPIX = TY * TX;
    constexpr int PPT = (NPIX + NTW - 1) / NTW;
    extern __shared__ float smem[];
    const int wsh = a.wsh, psh = a.psh;
    const int sw = 2 * wsh + 1;
    const int np2 = 2 * psh;
    const int PY = TY + np2, PX = TX + np2;
    float* sS = smem;
    float* sK = sS + PY * PX;
    float* rS = sK + PY * PX;
    float* rK = rS + PY * TX;

    const int tid = threadIdx.x;
    // Blocks are sub-tiles of the interiors (step x step) of the reference's tiles, so every
    // output of a block belongs to one reference tile: block row by = reference tile row
    // * nsub_y + sub-tile row (nsub = ceil(step / TY)); the last sub-tile of a reference
    // tile is cut at the tile's end.  One-row / one-column launches have a single block
    // row / column starting at the launch origin.
    const int nsy = (a.step + TY - 1) / TY, nsx = (a.step + TX - 1) / TX;
    const int ry = blockIdx.y / nsy, sy = blockIdx.y - ry * nsy;
    const int rx = blockIdx.x / nsx, sx_ = blockIdx.x - rx * nsx;
    const int ty0 = a.L.vy0 + ry * a.step + sy * TY;
    const int tx0 = a.L.vx0 + rx * a.step + sx_ * TX;
    const int rows_lim = min(TY, a.step - sy * TY);
    const int cols_lim = min(TX, a.step - sx_ * TX);
    // interior origin of the reference tile that computes this block's outputs
    const int Y0 = (ty0 / a.step) * a.step;
    const int X0 = (tx0 / a.step) * a.step;

    const long long HW = (long long)a.H * a.W;
    float acc[PPT][5];
    long long opix[PPT];
    bool valid[PPT];
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = acc[q][4] = 0.f;
        const int pq = tid + q * NTW;
        const int py = pq / TX, px = pq - (pq / TX) * TX;
        const int iy = ty0 + py - a.L.vy0, ix = tx0 + px - a.L.vx0;
        valid[q] = pq < NPIX && py < rows_lim && px < cols_lim && iy < a.L.ny && ix < a.L.nx;
        opix[q] = valid[q] ? (long long)(a.L.oy0 + iy) * a.W + (a.L.ox0 + ix) : 0;
    }

    const float* Ra1 = a.R.base;
    const float* Ra2 = Ra1 + a.R.plane;
    const float* Rdp = Ra1 + 2 * a.R.plane;
    const float* Eref = a.E.base;
    const float* Epha = Eref + a.E.plane;
    const float* Ecoh = Eref + 2 * a.E.plane;
    const float* Fa = a.F.base;
    const float* Fxr = Fa + a.F.plane;
    const float* Fxi = Fa + 2 * a.F.plane;

    const int centre = wsh * sw + wsh;
    for (int hh = 0; hh < sw; hh++) {
        const int dy = hh - wsh;
        for (int ww = 0; ww < sw; ww++) {
            const int dx = ww - wsh;
            const int plane = hh * sw + ww;
            if (plane == centre) {
                // centre weight 0 (nlinsar_sub_image.cpp:189-197); adding it changes no sum
#pragma unroll
                for (int q = 0; q < PPT; q++)
                    if (valid[q]) a.wt[(long long)plane * HW + opix[q]] = 0.f;
                continue;
            }
            __syncthreads();
            // 1. similarity and KL at every patch position of the block.  Offsets with
            //    dy >= 0 are evaluated centre-first everywhere; offsets with dy < 0 are the
            //    mirrored values (the pixel at +d first) at positions inside the reference
            //    tile's similarity border, centre-first on the border
            //    (precompute_similarities_1st_pass.cl.in:36-42, _2nd_pass.cl:30-48).
            {
                const int n = PY * PX;
                int r = tid / PX, c = tid - r * PX;
                const int sr_ = NTW / PX, sc_ = NTW - sr_ * PX;
                for (int i = tid; i < n; i += NTW) {
                    const int gy = ty0 - psh + r, gx = tx0 - psh + c;
                    const int sh = gy - Y0 + psh, sx = gx - X0 + psh;
                    const bool interior = sh >= wsh && sx >= wsh && sx < a.wsim - wsh;
                    const bool swap = dy < 0 && interior;
                    const int y1 = swap ? gy + dy : gy, x1 = swap ? gx + dx : gx;
                    const int y2 = swap ? gy : gy + dy, x2 = swap ? gx : gx + dx;
                    const long long p1 = a.R.at(y1, x1), p2 = a.R.at(y2, x2);
                    sS[r * PX + c] = nl_pixel_similarity(Ra1[p1], Ra2[p1], Rdp[p1], Ra1[p2], Ra2[p2], Rdp[p2]);
                    const long long e1 = a.E.at(y1, x1), e2 = a.E.at(y2, x2);
                    sK[r * PX + c] = nl_pixel_kl(Eref[e1], Epha[e1], Ecoh[e1], Eref[e2], Epha[e2], Ecoh[e2]);
                    r += sr_;
                    c += sc_;
                    if (c >= PX) { c -= PX; r++; }
                }
            }
            __syncthreads();
            // 2. row sums along x (precompute_patch_similarities.cl:45-50)
            {
                const int n = PY * TX;
                int r = tid / TX, c = tid - r * TX;
                const int sr_ = NTW / TX, sc_ = NTW - sr_ * TX;
                for (int i = tid; i < n; i += NTW) {
                    const float* s1 = sS + r * PX + c;
                    const float* s2 = sK + r * PX + c;
                    float u = s1[0], v = s2[0];
                    for (int j = 1; j <= np2; j++) {
                        u += s1[j];
                        v += s2[j];
                    }
                    rS[r * TX + c] = u;
                    rK[r * TX + c] = v;
                    r += sr_;
                    c += sc_;
                    if (c >= TX) { c -= TX; r++; }
                }
            }
            __syncthreads();
            // 3. rows accumulated top to bottom (precompute_patch_similarities.cl:53-58),
            //    weight (compute_weights.cl:19-30), ENL sums (compute_number_of_looks.cl)
            //    and weighted means (compute_insar.cl:81-93) in plane order
#pragma unroll
            for (int q = 0; q < PPT; q++) {
                const int pq = tid + q * NTW;
                if (pq >= NPIX) continue;
                const int py = pq / TX, px = pq - (pq / TX) * TX;
                const float* s1 = rS + py * TX + px;
                const float* s2 = rK + py * TX + px;
                float u = s1[0], v = s2[0];
                for (int j = 1; j <= np2; j++) {
                    u += s1[j * TX];
                    v += s2[j * TX];
                }
                float w = expf(u / a.h + v / a.T);
                if (isnan(w) || isinf(w)) w = 0.0f;
                if (valid[q]) a.wt[(long long)plane * HW + opix[q]] = w;
                const long long f = a.F.at(ty0 + py + dy, tx0 + px + dx);
                acc[q][0] += w;
                acc[q][1] = fmaf(w, w, acc[q][1]);
                acc[q][2] = fmaf(w, Fa[f], acc[q][2]);
                acc[q][3] = fmaf(w, Fxr[f], acc[q][3]);
                acc[q][4] = fmaf(w, Fxi[f], acc[q][4]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        if (!valid[q]) continue;
        const float ws = acc[q][0];
        const float nol = ws * ws / acc[q][1];
        if (nol < a.lmin) {
            const unsigned int slot = atomicAdd(a.cnt, 1u);
            a.list[slot] = (int)opix[q];
        } else {
            nl_estimate(ws, acc[q][2], acc[q][3], acc[q][4], opix[q], a.old_ref, a.old_pha,
                        a.old_coh, a.new_ref, a.new_pha, a.new_coh);
        }
    }
}

struct SmoothArgs {
    Grid R, F;
    int wsh;
    int lmin;
    int H, W, eh, ew;
    const float* wt;
    const float* old_ref;
    const float* old_pha;
    const float* old_coh;
    float* new_ref;
    float* new_pha;
    float* new_coh;
    const int* list;
    unsigned int* cnt;  // [0] queued this iteration, [1] running total
};

// std::priority_queue<weight_el> of smoothing.cpp:27-36 (top = smallest weight): sift
// rules of libstdc++'s push_heap / pop_heap, which decide which of several equal weights
// stays in the queue.
__device__ __forceinline__ bool hp_less(float wa, float wb) { return wa > wb; }

__device__ void hp_push(float* hw, int* hi, int& n, float w, int idx)
{
    int hole = n++;
    int parent = (hole - 1) / 2;
    while (hole > 0 && hp_less(hw[parent], w)) {
        hw[hole] = hw[parent];
        hi[hole] = hi[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    hw[hole] = w;
    hi[hole] = idx;
}

__device__ void hp_pop(float* hw, int* hi, int& n)
{
    const int len = n - 1;
    if (len <= 0) { n = 0; return; }
    const float vw = hw[len];
    const int vi = hi[len];
    hw[len] = hw[0];
    hi[len] = hi[0];
    int hole = 0, child = 0;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (hp_less(hw[child], hw[child - 1])) child--;
        hw[hole] = hw[child];
        hi[hole] = hi[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        hw[hole] = hw[child - 1];
        hi[hole] = hi[child - 1];
        hole = child - 1;
    }
    int parent = (hole - 1) / 2;
    while (hole > 0 && hp_less(hw[parent], vw)) {
        hw[hole] = hw[parent];
        hi[hole] = hi[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    hw[hole] = vw;
    hi[hole] = vi;
    n = len;
}

// search_window_smoothing (smoothing.cpp:78-129) and compute_insar for the queued pixels
__global__ void nlinsar_smooth(const SmoothArgs a)
{
    extern __shared__ float smem[];
    const int sw = 2 * a.wsh + 1;
    const int nfull = sw * sw;
    const int mwords = (nfull + 31) / 32;
    const int per = 2 * a.lmin + mwords;
    float* hw = smem + threadIdx.x * per;
    int* hi = (int*)(hw + a.lmin);
    unsigned int* mask = (unsigned int*)(hi + a.lmin);
    const unsigned int total = a.cnt[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) a.cnt[1] += total;
    const long long HW = (long long)a.H * a.W;
    for (unsigned int li = blockIdx.x * blockDim.x + threadIdx.x; li < total;
         li += gridDim.x * blockDim.x) {
        const long long opix = a.list[li];
        const int oy = (int)(opix / a.W), ox = (int)(opix - (long long)oy * a.W);
        const int vy = oy == a.H - 1 ? oy + a.eh : oy;
        const int vx = ox == a.W - 1 ? ox + a.ew : ox;
        const double as = a.R.base[a.R.at(vy, vx)];
        int n = 0;
        for (int hh = 0; hh < sw; hh++)
            for (int ww = 0; ww < sw; ww++) {
                const float at = a.R.base[a.R.at(vy + hh - a.wsh, vx + ww - a.wsh)];
                if ((double)at < 2.0 * as) {
                    const int idx = hh * sw + ww;
                    const float w = a.wt[(long long)idx * HW + opix];
                    if (n < a.lmin) {
                        hp_push(hw, hi, n, w, idx);
                    } else if (hw[0] < w) {
                        hp_pop(hw, hi, n);
                        hp_push(hw, hi, n, w, idx);
                    }
                }
            }
        for (int k = 0; k < mwords; k++) mask[k] = 0u;
        int m = 0;
        float wsum = 0.0f;
        while (n > 0) {
            wsum += hw[0];
            mask[hi[0] >> 5] |= 1u << (hi[0] & 31);
            m++;
            hp_pop(hw, hi, n);
        }
        wsum /= (float)m;
        float N = 0.f, at_ = 0.f, xr = 0.f, xi = 0.f;
        const float* Fa = a.F.base;
        const float* Fxr = Fa + a.F.plane;
        const float* Fxi = Fa + 2 * a.F.plane;
        for (int k = 0; k < nfull; k++) {
            const float w = ((mask[k >> 5] >> (k & 31)) & 1u) ? wsum : a.wt[(long long)k * HW + opix];
            const int hh = k / sw, ww = k - (k / sw) * sw;
            const long long f = a.F.at(vy + hh - a.wsh, vx + ww - a.wsh);
            N += w;
            at_ = fmaf(w, Fa[f], at_);
            xr = fmaf(w, Fxr[f], xr);
            xi = fmaf(w, Fxi[f], xi);
        }
        nl_estimate(N, at_, xr, xi, opix, a.old_ref, a.old_pha, a.old_coh, a.new_ref, a.new_pha,
                    a.new_coh);
    }
}

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

static const int kTY[3] = {27, 18, 54};
static const int kTX[3] = {54, 54, 54};

int nlinsar_plan(const NlinsarProblem& pb, NlinsarPlan* plan)
{
    if (pb.H <= 0 || pb.W <= 0) { set_error("nlinsar: empty image"); return -1; }
    if (pb.sw < 3 || pb.sw % 2 == 0 || pb.p < 1 || pb.p % 2 == 0) {
        set_error("nlinsar: search window and patch sizes must be odd");
        return -1;
    }
    if (pb.niter < 0 || pb.lmin < 1 || pb.lmin > pb.sw * pb.sw) {
        set_error("nlinsar: niter must be >= 0 and 1 <= lmin <= search_window_size^2");
        return -1;
    }
    plan->wsh = (pb.sw - 1) / 2;
    plan->psh = (pb.p - 1) / 2;
    plan->overlap = plan->wsh + plan->psh;
    plan->step = pb.ref_tile - 2 * plan->overlap;
    if (plan->step <= 0) { set_error("nlinsar: search window and patch too large for the 80-pixel tiles"); return -1; }
    plan->tile = std::min(std::max(pb.tile, 0), 2);
    plan->TY = kTY[plan->tile];
    plan->TX = kTX[plan->tile];
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
    const int Hx = std::max(round_up(std::max(H - 1, 1), plan->step) + plan->TY, H - 1 + plan->eh + plan->TY);
    const int Wx = std::max(round_up(std::max(W - 1, 1), plan->step) + plan->TX, W - 1 + plan->ew + plan->TX);
    const int o = plan->overlap;
    plan->R = Grid{nullptr, 0, round_up(Wx + 2 * o, 64), o, o, Hx + 2 * o, Wx + 2 * o};
    plan->R.plane = (long long)plan->R.pitch * plan->R.h;
    plan->E = plan->R;
    const int f = plan->wsh;
    plan->F = Grid{nullptr, 0, round_up(Wx + 2 * f, 64), f, f, Hx + 2 * f, Wx + 2 * f};
    plan->F.plane = (long long)plan->F.pitch * plan->F.h;
    const int nfull = pb.sw * pb.sw;
    const int per = (2 * pb.lmin + (nfull + 31) / 32) * 4;
    int hb = 256;
    while (hb > 1 && (size_t)hb * per > 48 * 1024) hb /= 2;
    plan->heap_block = hb;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t q = off; off += (bytes + 255) / 256 * 256; return q; };
    const size_t n = (size_t)H * W;
    plan->off_R = take(sizeof(float) * 3 * plan->R.plane);
    plan->off_E = take(sizeof(float) * 3 * plan->E.plane);
    plan->off_F = take(sizeof(float) * 3 * plan->F.plane);
    plan->off_est = take(sizeof(float) * 6 * n);
    plan->off_w = take(sizeof(float) * (size_t)nfull * n);
    plan->off_list = take(sizeof(int) * n);
    plan->off_cnt = take(sizeof(unsigned int) * 2);
    plan->ws_bytes = off;
    return 0;
}

template <int TY, int TX>
static int launch_weights(const WeightArgs& base, const Launch* L, int nl, hipStream_t stream)
{
    const int np2 = 2 * base.psh;
    const size_t lds = sizeof(float) * (2 * (size_t)(TY + np2) * (TX + np2) + 2 * (size_t)(TY + np2) * TX);
    if (lds > 160 * 1024) { set_error("nlinsar: patch too large for the LDS budget"); return -2; }
    DKX_HIP_CHECK(hipFuncSetAttribute((const void*)nlinsar_weights<TY, TX, false>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    DKX_HIP_CHECK(hipFuncSetAttribute((const void*)nlinsar_weights<TY, TX, true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    for (int i = 0; i < nl; i++) {
        if (L[i].ny <= 0 || L[i].nx <= 0) continue;
        WeightArgs a = base;
        a.L = L[i];
        // reference tiles covered, times sub-tiles per reference tile; thin launches
        // (one row / column) use a single block row / column
        const int nsy = (base.step + TY - 1) / TY, nsx = (base.step + TX - 1) / TX;
        const int gy = L[i].ny == 1 ? 1 : ((L[i].ny + base.step - 1) / base.step) * nsy;
        const int gx = L[i].nx == 1 ? 1 : ((L[i].nx + base.step - 1) / base.step) * nsx;
        dim3 grid(gx, gy);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool timed = i == 0 && timing_on();
        if (timed) {
            DKX_HIP_CHECK(hipEventCreate(&e0));
            DKX_HIP_CHECK(hipEventCreate(&e1));
            DKX_HIP_CHECK(hipEventRecord(e0, stream));
        }
        if (i == 0)
            hipLaunchKernelGGL((nlinsar_weights<TY, TX, false>), grid, dim3(NTW), lds, stream, a);
        else
            hipLaunchKernelGGL((nlinsar_weights<TY, TX, true>), grid, dim3(NTW), lds, stream, a);
        DKX_HIP_CHECK(hipGetLastError());
        if (timed) {
            DKX_HIP_CHECK(hipEventRecord(e1, stream));
            timing_push(STAGE_NLINSAR_WEIGHTS, e0, e1);
        }
    }
    return 0;
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
    float* wt = (float*)(ws + plan.off_w);
    int* list = (int*)(ws + plan.off_list);
    unsigned int* cnt = (unsigned int*)(ws + plan.off_cnt);

    DKX_HIP_CHECK(hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned int), stream));
    {
        dim3 blk(256), grd((R.w + 255) / 256, R.h);
        hipLaunchKernelGGL(grid_fill, grd, blk, 0, stream, d_a1, d_a2, d_phi, H, W, R);
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
    L[0] = Launch{0, 0, H - 1, W - 1, 0, 0};
    L[1] = Launch{H - 1 + plan.eh, 0, 1, W - 1, H - 1, 0};
    L[2] = Launch{0, W - 1 + plan.ew, H - 1, 1, 0, W - 1};
    L[3] = Launch{H - 1 + plan.eh, W - 1 + plan.ew, 1, 1, H - 1, W - 1};
    int hdev = 0;
    DKX_HIP_CHECK(hipGetDevice(&hdev));
    int ncu = 256;
    DKX_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, hdev));
    for (int it = 0; it < pb.niter; it++) {
        {
            dim3 blk(256), grd((E.w + 255) / 256, E.h);
            hipLaunchKernelGGL(grid_fill, grd, blk, 0, stream, (const float*)cur[0],
                               (const float*)cur[1], (const float*)cur[2], H, W, E);
            DKX_HIP_CHECK(hipGetLastError());
        }
        DKX_HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned int), stream));
        WeightArgs wa;
        wa.R = R; wa.E = E; wa.F = F;
        wa.wsh = plan.wsh; wa.psh = plan.psh; wa.step = plan.step;
        wa.wsim = pb.ref_tile - pb.sw + 1;
        wa.h = pb.h; wa.T = pb.T; wa.lmin = (float)pb.lmin;
        wa.H = H; wa.W = W;
        wa.wt = wt;
        wa.old_ref = cur[0]; wa.old_pha = cur[1]; wa.old_coh = cur[2];
        wa.new_ref = nxt[0]; wa.new_pha = nxt[1]; wa.new_coh = nxt[2];
        wa.list = list; wa.cnt = cnt;
        int rc;
        switch (plan.tile) {
        case 0: rc = launch_weights<27, 54>(wa, L, 4, stream); break;
        case 1: rc = launch_weights<18, 54>(wa, L, 4, stream); break;
        default: rc = launch_weights<54, 54>(wa, L, 4, stream); break;
        }
        if (rc) return rc;
        SmoothArgs sa;
        sa.R = R; sa.F = F; sa.wsh = plan.wsh; sa.lmin = pb.lmin;
        sa.H = H; sa.W = W; sa.eh = plan.eh; sa.ew = plan.ew;
        sa.wt = wt;
        sa.old_ref = cur[0]; sa.old_pha = cur[1]; sa.old_coh = cur[2];
        sa.new_ref = nxt[0]; sa.new_pha = nxt[1]; sa.new_coh = nxt[2];
        sa.list = list; sa.cnt = cnt;
        const int nfull = pb.sw * pb.sw;
        const size_t lds = (size_t)plan.heap_block * (2 * pb.lmin + (nfull + 31) / 32) * 4;
        const int blocks = std::max(1, std::min((int)((n + plan.heap_block - 1) / plan.heap_block), ncu * 8));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing_on()) {
            DKX_HIP_CHECK(hipEventCreate(&e0));
            DKX_HIP_CHECK(hipEventCreate(&e1));
            DKX_HIP_CHECK(hipEventRecord(e0, stream));
        }
        hipLaunchKernelGGL(nlinsar_smooth, dim3(blocks), dim3(plan.heap_block), lds, stream, sa);
        DKX_HIP_CHECK(hipGetLastError());
        if (e0) {
            DKX_HIP_CHECK(hipEventRecord(e1, stream));
            timing_push(STAGE_NLINSAR_SMOOTH, e0, e1);
        }
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

}  // namespace dkx
