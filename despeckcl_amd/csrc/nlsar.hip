// NL-SAR on CDNA4: covariance preparation, Gaussian scale average, planning and the pass
// sequence.  The fused scale pass itself lives in dkx_nlsar_pass.h.
//
// Reference path replaced: src/compute/nlsar/nlsar_filter_sub_image.cpp:146-390 with the
// kernels of src/compute/nlsar/cl_kernels.
#include "dkx_common.h"
#include "dkx_nlsar.h"
#include "dkx_nlsar_pass.h"

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstring>
#include <vector>

namespace dkx {

// ------------------------------------------------------------------------------------
// preparation kernels (one thread per grid element; blocks cover the rows of every image)
// ------------------------------------------------------------------------------------

struct PrepArgs {
    const NlsarImageDesc* table;   // batch: descriptors in device memory, else null
    NlsarImageDesc one;            // the image of a single call
    int nimg;
    unsigned int* queues;          // persistent passes' tile counters (nqueues), or null
    int nqueues;
    int cblocks;                   // blocks per grid row
    const float* gauss;            // scale average: s x s taps
    int s;                         // scale size
    float gamma;                   // off-diagonal rescaling factor
};

// image owning batch row `row` of the C grids (A: of the A grids)
__device__ __forceinline__ const NlsarImageDesc& prep_image(const PrepArgs& a, int row, bool A)
{
    if (!a.table) return a.one;
    int lo = 0, hi = a.nimg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((A ? a.table[mid].row0_a : a.table[mid].row0_c) <= row) lo = mid;
        else hi = mid - 1;
    }
    return a.table[lo];
}

__global__ void nlsar_zero_counts(const PrepArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (!a.table) {
        if (i == 0) *a.one.ub_count = 0ull;
    } else if (i < a.nimg) {
        *a.table[i].ub_count = 0ull;
    }
    if (a.queues && i < a.nqueues) a.queues[i] = 0u;
}

// Covariance of the input pixel on the padded grid C (records of nlsar_c_float4(MODE)
// float4): InSAR input (covmat_create.cl:19-56): a1^2, a2^2, a1*a2*cos(phi), a1*a2*sin(phi);
// covariance input (2*dim*dim interlaced planes, nlsar.i:146-153): every plane, in order.
template <int MODE>
__global__ void nlsar_prep(const PrepArgs a)
{
    const int row = blockIdx.x / a.cblocks;
    const int gx = (blockIdx.x - row * a.cblocks) * blockDim.x + threadIdx.x;
    const NlsarImageDesc& d = prep_image(a, row, false);
    const int gy = row - d.row0_c;
    if (gx >= d.cw || gy >= d.ch) return;
    const int H = d.H, W = d.W;
    const int y = min(H - 1, max(0, gy - d.coy));
    const int x = min(W - 1, max(0, gx - d.cox));
    const long long i = (long long)y * W + x;
    constexpr int CF4 = nlsar_c_float4(MODE);
    DKX_GLOBAL float* C = gptr(d.C) + 4 * CF4 * ((long long)gy * d.cpitch + gx);
    if constexpr (MODE == 0) {
        const float m = gptr(d.a1)[i], sl = gptr(d.a2)[i], p = gptr(d.phi)[i];
        gstore4(C, make_float4(m * m, sl * sl, m * sl * cosf(p), m * sl * sinf(p)));
    } else {
        constexpr int NPL = 2 * MODE * MODE;
        const DKX_GLOBAL float* cov = gptr(d.cov);
        const long long n = (long long)H * W;
        float v[4 * CF4];
#pragma unroll
        for (int pl = 0; pl < 4 * CF4; pl++) v[pl] = pl < NPL ? cov[pl * n + i] : 0.f;
#pragma unroll
        for (int j = 0; j < CF4; j++) gstore4(C + 4 * j, make_float4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]));
    }
}

// Gaussian s x s average of the rescaled covariance (covmat_rescale.cl:19-41 then
// covmat_spatial_avg.cl:19-63; planes combined per pixel in the same operation order),
// plus the log-determinant of each averaged matrix, which every pixel dissimilarity
// involving that pixel uses (sim_measures.h:51-52, 100-158).  Only the planes the
// dissimilarity reads are averaged: (0,0), (1,1), (0,1) for dim 2; all 9 real values of the
// Hermitian matrix for dim 3.  Records as described at nlsar_a_float4.
template <int MODE>
__global__ void nlsar_scale_average(const PrepArgs a)
{
    const int row = blockIdx.x / a.cblocks;
    const int gx = (blockIdx.x - row * a.cblocks) * blockDim.x + threadIdx.x;
    const NlsarImageDesc& d = prep_image(a, row, true);
    const int gy = row - d.row0_a;
    if (gx >= d.aw || gy >= d.ah) return;
    const int y = gy - d.aoy, x = gx - d.aox;  // image coordinates
    const int s = a.s;
    const int ssh = (s - 1) / 2;
    constexpr int CF4 = nlsar_c_float4(MODE);
    const DKX_GLOBAL float* C = gptr(d.C);
    const DKX_GLOBAL float* gauss = gptr(a.gauss);
    const float gamma = a.gamma;
    // (plane of the C record, off-diagonal) of each averaged value
    constexpr int NV = MODE == 3 ? 9 : 4;
    constexpr int src[9] = {0, MODE == 0 ? 1 : MODE == 2 ? 6 : 8, MODE == 0 ? 2 : MODE == 2 ? 2 : 16,
                            MODE == 0 ? 3 : MODE == 2 ? 3 : 2, 3, 4, 5, 10, 11};
    constexpr bool offd[9] = {false, false, MODE != 3, true, true, true, true, true, true};
    float acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0.f;
    for (int kx = 0; kx < s; kx++) {
        const long long rowc = (long long)(y - ssh + kx + d.coy) * d.cpitch + (x - ssh + d.cox);
        for (int ky = 0; ky < s; ky++) {
            const float g = gauss[kx * s + ky];
            const DKX_GLOBAL float* c = C + 4 * CF4 * (rowc + ky);
#pragma unroll
            for (int v = 0; v < NV; v++) {
                const float val = offd[v] ? c[src[v]] * gamma : c[src[v]];
                acc[v] = fmaf(g, val, acc[v]);
            }
        }
    }
    const long long o = (long long)gy * d.apitch + gx;
    DKX_GLOBAL float* A = gptr(d.A);
    if constexpr (MODE == 3) {
        const float det = nlsar_det3(acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], acc[6], acc[7], acc[8]);
        DKX_GLOBAL float* r = A + 12 * o;
        gstore4(r, make_float4(acc[0], acc[1], acc[2], acc[3]));
        gstore4(r + 4, make_float4(acc[4], acc[5], acc[6], acc[7]));
        gstore4(r + 8, make_float4(acc[8], logf(det), 0.f, 0.f));
    } else {
        const float s00 = acc[0], s11 = acc[1], sr = acc[2], si = acc[3];
        gstore4(A + 4 * o, make_float4(s00, s11, sr, si));
        const float det = (s00 * s11) - (sr * sr + si * si);
        A[4ll * d.aplane + o] = logf(det);
    }
}

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------

// covmat_spatial_avg::gen_gauss, cl_wrappers/covmat_spatial_avg.cpp:35-50
std::vector<float> gen_gauss(int scale_size)
{
    std::vector<float> g;
    const int ssh = (scale_size - 1) / 2;
    for (int x = -ssh; x <= ssh; x++)
        for (int y = -ssh; y <= ssh; y++)
            g.push_back((float)std::exp(-M_PI * (x * x + y * y) / (double)std::pow(ssh + 0.5f, 2.0f)));
    float K = 0.0f;
    for (float v : g) K += v;
    for (float& v : g) v = v / K;
    return g;
}

// Weight of every quantile-LUT index: exp(-|chi2inv(q) - c| / h), the value compute_weights.cl
// (lines 47-51) derives from the index; built once per call on the host.
std::vector<float> weight_lut(const NlsarStatsView& st, float h, float c)
{
    std::vector<float> w(st.lut_size);
    const float scl = (float)(st.lut_size - 1);
    for (int i = 0; i < st.lut_size; i++) {
        const float q = st.quantilles[i];
        const float x = st.chi2cdf_inv[(unsigned int)(q * scl)];
        w[i] = std::exp(-std::fabs(x - c) / h);
    }
    return w;
}

static int round_up(int v, int m) { return (v + m - 1) / m * m; }


int nlsar_common(const NlsarParams& pp, NlsarCommon* cm)
{
    if (pp.sw < 3 || pp.sw % 2 == 0) { set_error("nlsar: search window size must be odd and >= 3"); return -1; }
    if (pp.sw > 2 * NLSAR_WMAX_LARGE + 1) { set_error("nlsar: search windows above 41 are not supported"); return -4; }
    if (pp.nP < 1 || pp.nP > NLSAR_MAX_PATCHES || pp.nS < 1 || pp.nS > NLSAR_MAX_NS) {
        set_error("nlsar: 1-16 patch sizes and 1-8 scale sizes are supported");
        return -1;
    }
    int pmax = 0, smax = 0;
    for (int i = 0; i < pp.nP; i++) {
        if (pp.P[i] < 1 || pp.P[i] % 2 == 0) { set_error("nlsar: patch sizes must be odd"); return -1; }
        if (pp.P[i] > 11) { set_error("nlsar: patch sizes above 11 are not supported"); return -4; }
        pmax = std::max(pmax, pp.P[i]);
    }
    for (int i = 0; i < pp.nS; i++) {
        if (pp.S[i] < 1 || pp.S[i] % 2 == 0 || pp.S[i] > 31) { set_error("nlsar: scale sizes must be odd and <= 31"); return -1; }
        smax = std::max(smax, pp.S[i]);
    }
    for (int i = 0; i < pp.nP; i++)
        for (int j = i + 1; j < pp.nP; j++)
            if (pp.P[i] == pp.P[j]) { set_error("nlsar: duplicate patch size"); return -1; }
    for (int i = 0; i < pp.nS; i++)
        for (int j = i + 1; j < pp.nS; j++)
            if (pp.S[i] == pp.S[j]) { set_error("nlsar: duplicate scale size"); return -1; }
    if (pp.dim != 0 && pp.dim != 2 && pp.dim != 3) {
        set_error("nlsar: covariance dimensions 2 and 3 are supported");
        return -1;
    }
    cm->mode = pp.dim;
    cm->tile = pp.tile < 0 ? NLSAR_TILE_AUTO : std::min(pp.tile, NLSAR_NTILES - 1);
    cm->wsh = (pp.sw - 1) / 2;
    cm->P = (pmax - 1) / 2;
    cm->smax = smax;
    cm->overlap = cm->P + cm->wsh + (smax - 1) / 2;
    int maxlut = 0;
    for (int i = 0; i < pp.nst; i++) maxlut = std::max(maxlut, pp.stats[i].lut_size);
    cm->lut_stride = maxlut;
    cm->nparam = pp.nP * pp.nS;
    cm->const_floats = (size_t)cm->nparam * maxlut + (size_t)NLSAR_MAX_NS * 31 * 31;
    return 0;
}

int nlsar_image_plan(const NlsarCommon& cm, int H, int W, int edge_h, int edge_w, NlsarImagePlan* ip)
{
    if (H <= 0 || W <= 0) { set_error("nlsar: empty image"); return -1; }
    ip->H = H;
    ip->W = W;
    ip->eh = edge_h < 0 ? cm.overlap : edge_h;
    ip->ew = edge_w < 0 ? cm.overlap : edge_w;
    const int T = NLSAR_TILE;
    // virtual extents covered by the four launch mappings
    const int Hx = std::max(round_up(std::max(H - 1, 1), T), H - 1 + ip->eh + T);
    const int Wx = std::max(round_up(std::max(W - 1, 1), T), W - 1 + ip->ew + T);
    const int ra = cm.wsh + cm.P;               // A grid padding
    const int rc = ra + (cm.smax - 1) / 2;      // C grid padding
    ip->C = Grid{nullptr, 0, round_up(Wx + 2 * rc, 64), rc, rc, Hx + 2 * rc, Wx + 2 * rc};
    ip->C.plane = (long long)ip->C.pitch * ip->C.h;
    ip->A = Grid{nullptr, 0, round_up(Wx + 2 * ra, 64), ra, ra, Hx + 2 * ra, Wx + 2 * ra};
    ip->A.plane = (long long)ip->A.pitch * ip->A.h;
    // kernels address grids with 32-bit byte offsets
    const double cbytes = 16.0 * nlsar_c_float4(cm.mode) * ip->C.plane;
    const double abytes = 16.0 * nlsar_a_float4(cm.mode) * ip->A.plane;
    if (cbytes >= 4294967296.0 || abytes >= 4294967296.0) {
        set_error("nlsar: image too large for one call (split it into bands)");
        return -4;
    }
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    ip->off_C = take((size_t)16 * nlsar_c_float4(cm.mode) * ip->C.plane);
    ip->off_A = take((size_t)nlsar_a_bytes(cm.mode) * ip->A.plane);
    ip->off_flag = take((size_t)H * W);
    ip->off_ub = take(sizeof(unsigned long long));
    ip->bytes = off;
    return 0;
}

// statistics entry of every parameter index (scale-major loop order)
int nlsar_param_stats(const NlsarParams& pp, const NlsarCommon& cm,
                      std::vector<const NlsarStatsView*>& pst)
{
    pst.assign(cm.nparam, nullptr);
    for (int si = 0; si < pp.nS; si++)
        for (int pi = 0; pi < pp.nP; pi++) {
            const int idx = si * pp.nP + pi;
            for (int k = 0; k < pp.nst; k++)
                if (pp.stats[k].patch_size == pp.P[pi] && pp.stats[k].scale_size == pp.S[si]) pst[idx] = &pp.stats[k];
            if (!pst[idx]) { set_error("nlsar: statistics missing for a (patch, scale) pair"); return -3; }
        }
    return 0;
}

// Host image of the call's device constants: the weight LUT of every parameter index,
// then the Gaussian taps of every scale (31*31 floats each).
int nlsar_build_constants(const NlsarParams& pp, const NlsarCommon& cm, std::vector<float>& blob)
{
    std::vector<const NlsarStatsView*> pst;
    if (nlsar_param_stats(pp, cm, pst)) return -3;
    blob.assign(cm.const_floats, 0.f);
    for (int idx = 0; idx < cm.nparam; idx++) {
        std::vector<float> w = weight_lut(*pst[idx], pp.h, pp.c);
        std::copy(w.begin(), w.end(), blob.begin() + (size_t)idx * cm.lut_stride);
    }
    const size_t goff = (size_t)cm.nparam * cm.lut_stride;
    for (int si = 0; si < pp.nS; si++) {
        std::vector<float> g = gen_gauss(pp.S[si]);
        std::copy(g.begin(), g.end(), blob.begin() + goff + (size_t)si * 31 * 31);
    }
    return 0;
}

void nlsar_image_desc(const NlsarImagePlan& ip, char* ws, const float* a1, const float* a2,
                      const float* phi, const float* cov, float* ref, float* phase, float* coh,
                      float* cov_out, float* enl_maps, float* alpha_maps, NlsarImageDesc* d)
{
    memset(d, 0, sizeof(*d));
    d->a1 = a1;
    d->a2 = a2;
    d->phi = phi;
    d->cov = cov;
    d->ref = ref;
    d->phase = phase;
    d->coh = coh;
    d->cov_out = cov_out;
    d->enl_nb = enl_maps;
    d->alpha = alpha_maps;
    d->C = (float*)(ws + ip.off_C);
    d->A = (float*)(ws + ip.off_A);
    d->ubflag = (unsigned char*)(ws + ip.off_flag);
    d->ub_count = (unsigned long long*)(ws + ip.off_ub);
    d->H = ip.H;
    d->W = ip.W;
    d->cpitch = ip.C.pitch; d->coy = ip.C.oy; d->cox = ip.C.ox; d->ch = ip.C.h; d->cw = ip.C.w;
    d->apitch = ip.A.pitch; d->aoy = ip.A.oy; d->aox = ip.A.ox; d->ah = ip.A.h; d->aw = ip.A.w;
    d->aplane = (int)ip.A.plane;
    const int H = ip.H, W = ip.W;
    d->L[0] = Launch{0, 0, H - 1, W - 1, 0, 0};
    d->L[1] = Launch{H - 1 + ip.eh, 0, 1, W - 1, H - 1, 0};
    d->L[2] = Launch{0, W - 1 + ip.ew, H - 1, 1, 0, W - 1};
    d->L[3] = Launch{H - 1 + ip.eh, W - 1 + ip.ew, 1, 1, H - 1, W - 1};
    int total = 0;
    for (int i = 0; i < 4; i++) {
        d->bstart[i] = total;
        if (d->L[i].ny <= 0 || d->L[i].nx <= 0) {
            d->gx[i] = d->gy[i] = 1;
            continue;
        }
        d->gx[i] = (d->L[i].nx + NLSAR_TILE - 1) / NLSAR_TILE;
        d->gy[i] = (d->L[i].ny + NLSAR_TILE - 1) / NLSAR_TILE;
        total += d->gx[i] * d->gy[i];
    }
    d->bstart[4] = total;
}

void nlsar_batch_layout(NlsarImageDesc* d, int n, int* rows_c, int* rows_a, int* blocks, int* max_cw,
                        int* max_aw)
{
    int rc = 0, ra = 0, nb = 0, cw = 0, aw = 0;
    for (int i = 0; i < n; i++) {
        d[i].row0_c = rc;
        d[i].row0_a = ra;
        d[i].block0 = nb;
        rc += d[i].ch;
        ra += d[i].ah;
        nb += d[i].bstart[4];
        cw = std::max(cw, d[i].cw);
        aw = std::max(aw, d[i].aw);
    }
    *rows_c = rc;
    *rows_a = ra;
    *blocks = nb;
    *max_cw = cw;
    *max_aw = aw;
}

int nlsar_enqueue(const NlsarParams& pp, const NlsarCommon& cm, const NlsarImageDesc* descs, int n,
                  const NlsarImageDesc* d_table, const float* d_const, hipStream_t stream,
                  unsigned int* d_queues)
{
    if (n < 1 || (!d_table && n != 1)) { set_error("nlsar: internal error, bad image set"); return -1; }
    const float* d_lut = d_const;
    const float* d_gauss = d_const + (size_t)cm.nparam * cm.lut_stride;
    std::vector<const NlsarStatsView*> pst;
    if (nlsar_param_stats(pp, cm, pst)) return -3;
    // totals (the batch table was numbered by nlsar_batch_layout)
    const NlsarImageDesc& last = descs[n - 1];
    const int rows_c = last.row0_c + last.ch, rows_a = last.row0_a + last.ah;
    const int blocks = last.block0 + last.bstart[4];
    int max_cw = 0, max_aw = 0;
    for (int i = 0; i < n; i++) {
        max_cw = std::max(max_cw, descs[i].cw);
        max_aw = std::max(max_aw, descs[i].aw);
    }

    PrepArgs pa;
    memset(&pa, 0, sizeof(pa));
    pa.table = d_table;
    if (!d_table) pa.one = descs[0];
    pa.nimg = n;
    if (d_table && d_queues) {
        pa.queues = d_queues;
        pa.nqueues = 8 * NLSAR_MAX_PASSES;
    }
    hipLaunchKernelGGL(nlsar_zero_counts, dim3((std::max(n, pa.nqueues) + 255) / 256), dim3(256), 0, stream, pa);
    DKX_HIP_CHECK(hipGetLastError());
    pa.cblocks = (max_cw + 255) / 256;
    const dim3 gprep((unsigned)rows_c * pa.cblocks);
    switch (cm.mode) {
    case 0: hipLaunchKernelGGL(nlsar_prep<0>, gprep, dim3(256), 0, stream, pa); break;
    case 2: hipLaunchKernelGGL(nlsar_prep<2>, gprep, dim3(256), 0, stream, pa); break;
    default: hipLaunchKernelGGL(nlsar_prep<3>, gprep, dim3(256), 0, stream, pa); break;
    }
    DKX_HIP_CHECK(hipGetLastError());
    // covmat_rescale.cl:24-25 with nlooks = 1 (the InSAR input has dimension 2)
    const int dim = cm.mode == 0 ? 2 : cm.mode;
    pa.gamma = std::pow(std::min(1.0f / dim, 1.0f), 0.33333f);
    pa.cblocks = (max_aw + 255) / 256;

    // Passes: each handles one scale and up to three patch sizes (one dissimilarity field
    // feeds all of them).  The pass holding the largest patch at the largest scale runs last:
    // it also forms the weighted mean with that (patch, scale), the key the selection keeps.
    int pmax_i = 0;
    for (int pi = 0; pi < pp.nP; pi++) if (pp.P[pi] > pp.P[pmax_i]) pmax_i = pi;
    int smax_i = 0;
    for (int si = 0; si < pp.nS; si++) if (pp.S[si] > pp.S[smax_i]) smax_i = si;
    struct Pass { int si; std::vector<int> pis; bool last; };
    std::vector<Pass> passes;
    for (int si = 0; si < pp.nS; si++) {
        std::vector<int> rest;
        for (int pi = 0; pi < pp.nP; pi++) if (!(si == smax_i && pi == pmax_i)) rest.push_back(pi);
        if (si == smax_i) {
            // the final chunk of this scale: the largest patch plus up to two others
            std::vector<int> fin{pmax_i};
            while (fin.size() < 3 && !rest.empty()) { fin.insert(fin.begin(), rest.back()); rest.pop_back(); }
            for (size_t k = 0; k < rest.size(); k += 3)
                passes.push_back(Pass{si, std::vector<int>(rest.begin() + k, rest.begin() + std::min(rest.size(), k + 3)), false});
            passes.push_back(Pass{si, fin, true});
        } else {
            for (size_t k = 0; k < rest.size(); k += 3)
                passes.push_back(Pass{si, std::vector<int>(rest.begin() + k, rest.begin() + std::min(rest.size(), k + 3)), false});
        }
    }
    // the final pass must come last
    std::stable_partition(passes.begin(), passes.end(), [](const Pass& q) { return !q.last; });
    // without diagnostics only the selected key's statistics are needed: the last key, on its
    // own (the other keys of its pass only feed the maps and the undefined-selection count)
    if (pp.last_unused) {
        // the selection lands on a key outside patch x scale: the mean uses parameter index 0
        // (first scale, first patch) with alpha 0, after every key's statistics if wanted
        if (!pp.all_keys) passes.clear();
        for (Pass& q : passes) q.last = false;
        passes.push_back(Pass{0, std::vector<int>{0}, true});
    } else if (!pp.all_keys) {
        passes.assign(1, Pass{smax_i, std::vector<int>{pmax_i}, true});
    }

    if (passes.size() > (size_t)NLSAR_MAX_PASSES) {
        set_error("nlsar: too many patch and scale sizes for one call");
        return -1;
    }
    int cur_scale = -1;
    for (size_t oi = 0; oi < passes.size(); oi++) {
        const Pass& ps = passes[oi];
        const int si = ps.si;
        if (si != cur_scale) {
            pa.gauss = d_gauss + (size_t)si * 31 * 31;
            pa.s = pp.S[si];
            const dim3 gavg((unsigned)rows_a * pa.cblocks);
            switch (cm.mode) {
            case 0: hipLaunchKernelGGL(nlsar_scale_average<0>, gavg, dim3(256), 0, stream, pa); break;
            case 2: hipLaunchKernelGGL(nlsar_scale_average<2>, gavg, dim3(256), 0, stream, pa); break;
            default: hipLaunchKernelGGL(nlsar_scale_average<3>, gavg, dim3(256), 0, stream, pa); break;
            }
            DKX_HIP_CHECK(hipGetLastError());
            cur_scale = si;
        }
        // slots in ascending patch size: the pass's patch set is a compile-time mask of
        // half widths, and the last pass's largest patch (the mean's) is its last slot
        std::vector<int> pis = ps.pis;
        std::sort(pis.begin(), pis.end(), [&](int x, int y) { return pp.P[x] < pp.P[y]; });
        ScaleArgs sa;
        memset(&sa, 0, sizeof(sa));
        sa.table = d_table;
        if (!d_table) sa.one = descs[0];
        sa.nimg = n;
        sa.total_blocks = blocks;
        sa.wsh = cm.wsh;
        int pmask = 0;
        for (size_t k = 0; k < pis.size(); k++) {
            const int pi = pis[k];
            const int idx = si * pp.nP + pi;
            pmask |= 1 << ((pp.P[pi] - 1) / 2);
            sa.pidx[k] = idx;
            sa.lutw[k] = d_lut + (size_t)idx * cm.lut_stride;
            const float dmin = pst[idx]->dissims_min, dmax = pst[idx]->dissims_max;
            sa.dmin[k] = dmin;
            sa.dmax[k] = dmax;
            sa.rng[k] = dmax - dmin;
            sa.rcp[k] = 1.0f / sa.rng[k];
            sa.scl[k] = (float)(pst[idx]->lut_size - 1);
            sa.lut_size[k] = pst[idx]->lut_size;
        }
        sa.alpha_zero = ps.last && pp.last_unused ? 1 : 0;
        sa.queue = pa.queues ? pa.queues + 8 * oi : nullptr;
        if (ps.last && !pp.last_unused && pis.back() != pmax_i) {
            set_error("nlsar: internal error, the last pass must end with the largest patch");
            return -1;
        }
        sa.lut_stride = cm.lut_stride;
        sa.first_scale = oi == 0 ? 1 : 0;
        int rc;
        if (cm.mode == 0) {
            if (cm.wsh <= NLSAR_WMAX_SMALL) {
                // automatic: one row per thread for a single patch size (its two offsets per
                // barrier phase fill 1024-thread blocks best), two rows otherwise
                const int tile = cm.tile != NLSAR_TILE_AUTO ? cm.tile : pmask_count(pmask) == 1 ? 2 : 1;
                switch (tile) {
                case 0:
                    rc = ps.last ? nlsar_pass_launch<true, 32, 4, NLSAR_WMAX_SMALL, 0>(pmask, sa, stream)
                                 : nlsar_pass_launch<false, 32, 4, NLSAR_WMAX_SMALL, 0>(pmask, sa, stream);
                    break;
                case 2:
                    rc = ps.last ? nlsar_pass_launch<true, 32, 1, NLSAR_WMAX_SMALL, 0>(pmask, sa, stream)
                                 : nlsar_pass_launch<false, 32, 1, NLSAR_WMAX_SMALL, 0>(pmask, sa, stream);
                    break;
                default:
                    rc = ps.last ? nlsar_pass_launch<true, 32, 2, NLSAR_WMAX_SMALL, 0>(pmask, sa, stream)
                                 : nlsar_pass_launch<false, 32, 2, NLSAR_WMAX_SMALL, 0>(pmask, sa, stream);
                    break;
                }
            } else {
                rc = ps.last ? nlsar_pass_launch<true, 32, 4, NLSAR_WMAX_LARGE, 0>(pmask, sa, stream)
                             : nlsar_pass_launch<false, 32, 4, NLSAR_WMAX_LARGE, 0>(pmask, sa, stream);
            }
        } else if (cm.mode == 2) {
            // covariance input: one pass variant per window class
            if (cm.wsh <= NLSAR_WMAX_SMALL)
                rc = ps.last ? nlsar_pass_launch<true, 32, 2, NLSAR_WMAX_SMALL, 2>(pmask, sa, stream)
                             : nlsar_pass_launch<false, 32, 2, NLSAR_WMAX_SMALL, 2>(pmask, sa, stream);
            else
                rc = ps.last ? nlsar_pass_launch<true, 32, 4, NLSAR_WMAX_LARGE, 2>(pmask, sa, stream)
                             : nlsar_pass_launch<false, 32, 4, NLSAR_WMAX_LARGE, 2>(pmask, sa, stream);
        } else {
            if (cm.wsh <= NLSAR_WMAX_SMALL)
                rc = ps.last ? nlsar_pass_launch<true, 32, 2, NLSAR_WMAX_SMALL, 3>(pmask, sa, stream)
                             : nlsar_pass_launch<false, 32, 2, NLSAR_WMAX_SMALL, 3>(pmask, sa, stream);
            else
                rc = ps.last ? nlsar_pass_launch<true, 32, 4, NLSAR_WMAX_LARGE, 3>(pmask, sa, stream)
                             : nlsar_pass_launch<false, 32, 4, NLSAR_WMAX_LARGE, 3>(pmask, sa, stream);
        }
        if (rc) return rc;
    }
    return 0;
}

}  // namespace dkx
