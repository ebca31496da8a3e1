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
// preparation kernels
// ------------------------------------------------------------------------------------

// Covariance of the InSAR input (covmat_create.cl:19-56) on the padded grid C, one float4
// per element: a1^2, a2^2, Re C01 = a1*a2*cos(phi), Im C01 = a1*a2*sin(phi).
__global__ void nlsar_prep_insar(const float* __restrict__ a1, const float* __restrict__ a2,
                                 const float* __restrict__ phi, int H, int W, Grid C)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= C.w || gy >= C.h) return;
    const int y = min(H - 1, max(0, gy - C.oy));
    const int x = min(W - 1, max(0, gx - C.ox));
    const long long i = (long long)y * W + x;
    const float m = a1[i], s = a2[i], p = phi[i];
    const long long o = (long long)gy * C.pitch + gx;
    reinterpret_cast<float4*>(C.base)[o] = make_float4(m * m, s * s, m * s * cosf(p), m * s * sinf(p));
}

// Covariance of a dim-2 covariance input (8 interlaced planes, nlsar.i:146-153): the
// diagonal real parts and the (0,1) element, laid out as for the InSAR input.
__global__ void nlsar_prep_covmat2(const float* __restrict__ cov, int H, int W, Grid C)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= C.w || gy >= C.h) return;
    const int y = min(H - 1, max(0, gy - C.oy));
    const int x = min(W - 1, max(0, gx - C.ox));
    const long long n = (long long)H * W;
    const long long i = (long long)y * W + x;
    const long long o = (long long)gy * C.pitch + gx;
    reinterpret_cast<float4*>(C.base)[o] = make_float4(cov[i], cov[6 * n + i], cov[2 * n + i], cov[3 * n + i]);
}

// Gaussian s x s average of the rescaled covariance (covmat_rescale.cl:19-41 then
// covmat_spatial_avg.cl:19-63; planes combined per pixel in the same operation order),
// plus the log-determinant of each averaged matrix, which every pixel dissimilarity
// involving that pixel uses (sim_measures.h:51-52).
// A holds one float4 per element (a00, a11, Re a01, Im a01) followed by the log det plane.
__global__ void nlsar_scale_average(Grid C, Grid A, const float* __restrict__ gauss, int s,
                                    float gamma)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= A.w || gy >= A.h) return;
    const int y = gy - A.oy, x = gx - A.ox;  // image coordinates
    const int ssh = (s - 1) / 2;
    const float4* C4 = reinterpret_cast<const float4*>(C.base);
    float s00 = 0.f, s11 = 0.f, sr = 0.f, si = 0.f;
    for (int kx = 0; kx < s; kx++) {
        const long long row = C.at(y - ssh + kx, x - ssh);
        for (int ky = 0; ky < s; ky++) {
            const float g = gauss[kx * s + ky];
            const float4 c = C4[row + ky];
            s00 = fmaf(g, c.x, s00);
            s11 = fmaf(g, c.y, s11);
            const float r = c.z * gamma;
            const float im = c.w * gamma;
            sr = fmaf(g, r, sr);
            si = fmaf(g, im, si);
        }
    }
    const long long o = (long long)gy * A.pitch + gx;
    reinterpret_cast<float4*>(A.base)[o] = make_float4(s00, s11, sr, si);
    const float det = (s00 * s11) - (sr * sr + si * si);
    A.base[4 * A.plane + o] = logf(det);
}

// Clears the undefined-selection counter (a kernel, not a memset node, so that the call
// can be captured in a HIP graph).
__global__ void nlsar_zero_count(unsigned long long* c)
{
    if (threadIdx.x == 0) *c = 0ull;
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

// tile shapes (rows x columns); tiles 0 / 1 / 2 give each thread 4 / 2 / 1 rows
static const int kTileTY[3] = {32, 32, 32};
static const int kTileTX[3] = {32, 32, 32};

int nlsar_plan(const NlsarProblem& pb, NlsarPlan* plan)
{
    if (pb.H <= 0 || pb.W <= 0) { set_error("nlsar: empty image"); return -1; }
    if (pb.sw < 3 || pb.sw % 2 == 0) { set_error("nlsar: search window size must be odd and >= 3"); return -1; }
    if (pb.sw > 2 * NLSAR_WMAX_LARGE + 1) { set_error("nlsar: search windows above 41 are not supported"); return -4; }
    if (pb.nP < 1 || pb.nP > NLSAR_MAX_PATCHES || pb.nS < 1 || pb.nS > NLSAR_MAX_NS) {
        set_error("nlsar: 1-16 patch sizes and 1-8 scale sizes are supported");
        return -1;
    }
    int pmax = 0, smax = 0;
    for (int i = 0; i < pb.nP; i++) {
        if (pb.P[i] < 1 || pb.P[i] % 2 == 0) { set_error("nlsar: patch sizes must be odd"); return -1; }
        if (pb.P[i] > 11) { set_error("nlsar: patch sizes above 11 are not supported"); return -4; }
        pmax = std::max(pmax, pb.P[i]);
    }
    for (int i = 0; i < pb.nS; i++) {
        if (pb.S[i] < 1 || pb.S[i] % 2 == 0 || pb.S[i] > 31) { set_error("nlsar: scale sizes must be odd and <= 31"); return -1; }
        smax = std::max(smax, pb.S[i]);
    }
    for (int i = 0; i < pb.nP; i++)
        for (int j = i + 1; j < pb.nP; j++)
            if (pb.P[i] == pb.P[j]) { set_error("nlsar: duplicate patch size"); return -1; }
    for (int i = 0; i < pb.nS; i++)
        for (int j = i + 1; j < pb.nS; j++)
            if (pb.S[i] == pb.S[j]) { set_error("nlsar: duplicate scale size"); return -1; }
    plan->tile = std::min(std::max(pb.tile, 0), 2);
    plan->TY = kTileTY[plan->tile];
    plan->TX = kTileTX[plan->tile];
    plan->wsh = (pb.sw - 1) / 2;
    plan->P = (pmax - 1) / 2;
    plan->smax = smax;
    plan->overlap = plan->P + plan->wsh + (smax - 1) / 2;
    plan->eh = pb.edge_h < 0 ? plan->overlap : pb.edge_h;
    plan->ew = pb.edge_w < 0 ? plan->overlap : pb.edge_w;
    const int H = pb.H, W = pb.W;
    // virtual extents covered by the four launch mappings
    const int Hx = std::max(round_up(std::max(H - 1, 1), plan->TY), H - 1 + plan->eh + plan->TY);
    const int Wx = std::max(round_up(std::max(W - 1, 1), plan->TX), W - 1 + plan->ew + plan->TX);
    const int ra = plan->wsh + plan->P;          // A grid padding
    const int rc = ra + (smax - 1) / 2;          // C grid padding
    plan->C = Grid{nullptr, 0, round_up(Wx + 2 * rc, 64), rc, rc, Hx + 2 * rc, Wx + 2 * rc};
    plan->C.plane = (long long)plan->C.pitch * plan->C.h;
    plan->A = Grid{nullptr, 0, round_up(Wx + 2 * ra, 64), ra, ra, Hx + 2 * ra, Wx + 2 * ra};
    plan->A.plane = (long long)plan->A.pitch * plan->A.h;
    // kernels index grid elements with 32-bit offsets
    if (plan->C.plane >= INT_MAX / 2 || plan->A.plane >= INT_MAX / 2) {
        set_error("nlsar: image too large for one call (split it into bands)");
        return -4;
    }
    int maxlut = 0;
    for (int i = 0; i < pb.nst; i++) maxlut = std::max(maxlut, pb.stats[i].lut_size);
    plan->lut_stride = maxlut;
    plan->nparam = pb.nP * pb.nS;
    // workspace layout (bytes, 256-aligned)
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    plan->off_C = take(sizeof(float) * 4 * plan->C.plane);
    plan->off_A = take(sizeof(float) * 5 * plan->A.plane);
    plan->const_floats = (size_t)plan->nparam * maxlut + (size_t)NLSAR_MAX_NS * 31 * 31;
    plan->off_flag = take((size_t)H * W);
    plan->off_ub = take(sizeof(unsigned long long));
    plan->ws_bytes = off;
    return 0;
}

// statistics entry of every parameter index (scale-major loop order)
int nlsar_param_stats(const NlsarProblem& pb, const NlsarPlan& plan,
                      std::vector<const NlsarStatsView*>& pst)
{
    pst.assign(plan.nparam, nullptr);
    for (int si = 0; si < pb.nS; si++)
        for (int pi = 0; pi < pb.nP; pi++) {
            const int idx = si * pb.nP + pi;
            for (int k = 0; k < pb.nst; k++)
                if (pb.stats[k].patch_size == pb.P[pi] && pb.stats[k].scale_size == pb.S[si]) pst[idx] = &pb.stats[k];
            if (!pst[idx]) { set_error("nlsar: statistics missing for a (patch, scale) pair"); return -3; }
        }
    return 0;
}

// Host image of the call's device constants: the weight LUT of every parameter index,
// then the Gaussian taps of every scale (31*31 floats each).
int nlsar_build_constants(const NlsarProblem& pb, const NlsarPlan& plan, std::vector<float>& blob)
{
    std::vector<const NlsarStatsView*> pst;
    if (nlsar_param_stats(pb, plan, pst)) return -3;
    blob.assign(plan.const_floats, 0.f);
    for (int idx = 0; idx < plan.nparam; idx++) {
        std::vector<float> w = weight_lut(*pst[idx], pb.h, pb.c);
        std::copy(w.begin(), w.end(), blob.begin() + (size_t)idx * plan.lut_stride);
    }
    const size_t goff = (size_t)plan.nparam * plan.lut_stride;
    for (int si = 0; si < pb.nS; si++) {
        std::vector<float> g = gen_gauss(pb.S[si]);
        std::copy(g.begin(), g.end(), blob.begin() + goff + (size_t)si * 31 * 31);
    }
    return 0;
}

int nlsar_run(const NlsarProblem& pb, const NlsarPlan& plan, const float* d_a1, const float* d_a2,
              const float* d_phi, const float* d_cov, float* d_ref, float* d_phase, float* d_coh,
              void* d_ws, const float* d_const, float* d_enl_maps, float* d_alpha_maps,
              unsigned long long* ub_host, hipStream_t stream)
{
    char* ws = (char*)d_ws;
    Grid C = plan.C, A = plan.A;
    C.base = (float*)(ws + plan.off_C);
    A.base = (float*)(ws + plan.off_A);
    const float* d_lut = d_const;
    const float* d_gauss = d_const + (size_t)plan.nparam * plan.lut_stride;
    unsigned char* d_flag = (unsigned char*)(ws + plan.off_flag);
    unsigned long long* d_ub = (unsigned long long*)(ws + plan.off_ub);
    const int H = pb.H, W = pb.W;
    std::vector<const NlsarStatsView*> pst;
    if (nlsar_param_stats(pb, plan, pst)) return -3;
    hipLaunchKernelGGL(nlsar_zero_count, dim3(1), dim3(64), 0, stream, d_ub);
    DKX_HIP_CHECK(hipGetLastError());

    {
        dim3 blk(256), grd((C.w + 255) / 256, C.h);
        if (pb.dim == 0)
            hipLaunchKernelGGL(nlsar_prep_insar, grd, blk, 0, stream, d_a1, d_a2, d_phi, H, W, C);
        else
            hipLaunchKernelGGL(nlsar_prep_covmat2, grd, blk, 0, stream, d_cov, H, W, C);
        DKX_HIP_CHECK(hipGetLastError());
    }
    // covmat_rescale.cl:24-25 with nlooks = 1 and dimension 2
    const float gamma = std::pow(std::min(1.0f / 2, 1.0f), 0.33333f);

    Launch L[4];
    L[0] = Launch{0, 0, H - 1, W - 1, 0, 0};
    L[1] = Launch{H - 1 + plan.eh, 0, 1, W - 1, H - 1, 0};
    L[2] = Launch{0, W - 1 + plan.ew, H - 1, 1, 0, W - 1};
    L[3] = Launch{H - 1 + plan.eh, W - 1 + plan.ew, 1, 1, H - 1, W - 1};

    // Passes: each handles one scale and up to three patch sizes (one dissimilarity field
    // feeds all of them).  The pass holding the largest patch at the largest scale runs last:
    // it also forms the weighted mean with that (patch, scale), the key the selection keeps.
    int pmax_i = 0;
    for (int pi = 0; pi < pb.nP; pi++) if (pb.P[pi] > pb.P[pmax_i]) pmax_i = pi;
    int smax_i = 0;
    for (int si = 0; si < pb.nS; si++) if (pb.S[si] > pb.S[smax_i]) smax_i = si;
    struct Pass { int si; std::vector<int> pis; bool last; };
    std::vector<Pass> passes;
    for (int si = 0; si < pb.nS; si++) {
        std::vector<int> rest;
        for (int pi = 0; pi < pb.nP; pi++) if (!(si == smax_i && pi == pmax_i)) rest.push_back(pi);
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

    int cur_scale = -1;
    for (size_t oi = 0; oi < passes.size(); oi++) {
        const Pass& ps = passes[oi];
        const int si = ps.si;
        const int s = pb.S[si];
        if (si != cur_scale) {
            dim3 blk(256), grd((A.w + 255) / 256, A.h);
            hipLaunchKernelGGL(nlsar_scale_average, grd, blk, 0, stream, C, A,
                               (const float*)(d_gauss + (size_t)si * 31 * 31), s, gamma);
            DKX_HIP_CHECK(hipGetLastError());
            cur_scale = si;
        }
        // slots in ascending patch size: the pass's patch set is a compile-time mask of
        // half widths, and the last pass's largest patch (the mean's) is its last slot
        std::vector<int> pis = ps.pis;
        std::sort(pis.begin(), pis.end(), [&](int x, int y) { return pb.P[x] < pb.P[y]; });
        ScaleArgs sa;
        memset(&sa, 0, sizeof(sa));
        sa.A4 = reinterpret_cast<const float4*>(A.base);
        sa.LD = A.base + 4 * A.plane;
        sa.C4 = reinterpret_cast<const float4*>(C.base);
        sa.apitch = A.pitch;
        sa.aoy = A.oy;
        sa.aox = A.ox;
        sa.cpitch = C.pitch;
        sa.coy = C.oy;
        sa.cox = C.ox;
        sa.wsh = plan.wsh;
        int pmask = 0;
        for (size_t k = 0; k < pis.size(); k++) {
            const int pi = pis[k];
            const int idx = si * pb.nP + pi;
            pmask |= 1 << ((pb.P[pi] - 1) / 2);
            sa.pidx[k] = idx;
            sa.lutw[k] = d_lut + (size_t)idx * plan.lut_stride;
            const float dmin = pst[idx]->dissims_min, dmax = pst[idx]->dissims_max;
            sa.dmin[k] = dmin;
            sa.dmax[k] = dmax;
            sa.rng[k] = dmax - dmin;
            sa.rcp[k] = 1.0f / sa.rng[k];
            sa.scl[k] = (float)(pst[idx]->lut_size - 1);
            sa.lut_size[k] = pst[idx]->lut_size;
        }
        if (ps.last && pis.back() != pmax_i) {
            set_error("nlsar: internal error, the last pass must end with the largest patch");
            return -1;
        }
        for (int i = 0; i < 4; i++) sa.L[i] = L[i];
        sa.lut_stride = plan.lut_stride;
        sa.H = H;
        sa.W = W;
        sa.enl_nb = d_enl_maps;
        sa.alpha = d_alpha_maps;
        sa.ubflag = d_flag;
        sa.first_scale = oi == 0 ? 1 : 0;
        sa.ref = d_ref;
        sa.phase = d_phase;
        sa.coh = d_coh;
        sa.ub_count = d_ub;
        int rc;
        if (plan.wsh <= NLSAR_WMAX_SMALL) {
            switch (plan.tile) {
            case 0:
                rc = ps.last ? nlsar_pass_launch<true, 32, 4, NLSAR_WMAX_SMALL>(pmask, sa, stream)
                             : nlsar_pass_launch<false, 32, 4, NLSAR_WMAX_SMALL>(pmask, sa, stream);
                break;
            case 2:
                rc = ps.last ? nlsar_pass_launch<true, 32, 1, NLSAR_WMAX_SMALL>(pmask, sa, stream)
                             : nlsar_pass_launch<false, 32, 1, NLSAR_WMAX_SMALL>(pmask, sa, stream);
                break;
            default:
                rc = ps.last ? nlsar_pass_launch<true, 32, 2, NLSAR_WMAX_SMALL>(pmask, sa, stream)
                             : nlsar_pass_launch<false, 32, 2, NLSAR_WMAX_SMALL>(pmask, sa, stream);
                break;
            }
        } else {
            rc = ps.last ? nlsar_pass_launch<true, 32, 4, NLSAR_WMAX_LARGE>(pmask, sa, stream)
                         : nlsar_pass_launch<false, 32, 4, NLSAR_WMAX_LARGE>(pmask, sa, stream);
        }
        if (rc) return rc;
    }
    if (ub_host) {
        DKX_HIP_CHECK(hipMemcpyAsync(ub_host, d_ub, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    }
    return 0;
}

}  // namespace dkx
