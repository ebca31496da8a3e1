// NL-SAR on CDNA4: covariance preparation, Gaussian scale average, and one fused
// "scale pass" kernel per scale that evaluates pixel dissimilarities, separable patch box
// sums, LUT weights and the nonlocal statistics of every patch size, and (for the last
// scale) the weighted mean, the alpha blend and the decomposition into reflectivity, phase
// and coherence.
//
// Reference path replaced: src/compute/nlsar/nlsar_filter_sub_image.cpp:146-390 with the
// kernels of src/compute/nlsar/cl_kernels (one materialised 220/441-plane buffer per stage
// there; here every stage of one search-window offset stays in LDS).
#include "dkx_common.h"
#include "dkx_nlsar.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace dkx {

// ------------------------------------------------------------------------------------
// preparation kernels
// ------------------------------------------------------------------------------------

// Covariance planes of the InSAR input (covmat_create.cl:19-56) on the padded grid C:
// plane 0 = a1^2, 1 = a2^2, 2 = Re C01 = a1*a2*cos(phi), 3 = Im C01 = a1*a2*sin(phi).
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
    C.base[o] = m * m;
    C.base[C.plane + o] = s * s;
    C.base[2 * C.plane + o] = m * s * cosf(p);
    C.base[3 * C.plane + o] = m * s * sinf(p);
}

// Covariance planes of a dim-2 covariance input (8 interlaced planes, nlsar.i:146-153):
// the diagonal real parts and the (0,1) element, as for the InSAR input.
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
    C.base[o] = cov[i];
    C.base[C.plane + o] = cov[6 * n + i];
    C.base[2 * C.plane + o] = cov[2 * n + i];
    C.base[3 * C.plane + o] = cov[3 * n + i];
}

// Gaussian s x s average of the rescaled covariance (covmat_rescale.cl:19-41 then
// covmat_spatial_avg.cl:19-63; planes combined per pixel in the same operation order),
// plus the log-determinant of each averaged matrix, which every pixel dissimilarity
// involving that pixel uses (sim_measures.h:51-52).
// Output planes of A: 0 = a00, 1 = a11, 2 = Re a01, 3 = Im a01, 4 = log det.
__global__ void nlsar_scale_average(Grid C, Grid A, const float* __restrict__ gauss, int s,
                                    float gamma)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    if (gx >= A.w || gy >= A.h) return;
    const int y = gy - A.oy, x = gx - A.ox;  // image coordinates
    const int ssh = (s - 1) / 2;
    float s00 = 0.f, s11 = 0.f, sr = 0.f, si = 0.f;
    for (int kx = 0; kx < s; kx++) {
        const long long row = C.at(y - ssh + kx, x - ssh);
        for (int ky = 0; ky < s; ky++) {
            const float g = gauss[kx * s + ky];
            const long long j = row + ky;
            s00 = fmaf(g, C.base[j], s00);
            s11 = fmaf(g, C.base[C.plane + j], s11);
            const float r = C.base[2 * C.plane + j] * gamma;
            const float im = C.base[3 * C.plane + j] * gamma;
            sr = fmaf(g, r, sr);
            si = fmaf(g, im, si);
        }
    }
    const long long o = (long long)gy * A.pitch + gx;
    A.base[o] = s00;
    A.base[A.plane + o] = s11;
    A.base[2 * A.plane + o] = sr;
    A.base[3 * A.plane + o] = si;
    const float det = (s00 * s11) - (sr * sr + si * si);
    A.base[4 * A.plane + o] = logf(det);
}

// ------------------------------------------------------------------------------------
// fused scale pass
// ------------------------------------------------------------------------------------

struct ScaleArgs {
    Grid A;        // 5 planes, see nlsar_scale_average
    Grid C;        // 4 planes, see nlsar_prep_insar
    int wsh;       // search-window half width
    int P;         // largest patch half width (also a template parameter of the kernel)
    int psh[NLSAR_MAX_NP];
    int pidx[NLSAR_MAX_NP];           // index of (p, s) in the per-parameter maps
    const float* lutw[NLSAR_MAX_NP];  // weight per quantile index
    float dmin[NLSAR_MAX_NP], dmax[NLSAR_MAX_NP];
    int lut_size[NLSAR_MAX_NP];
    int lut_stride;                // LDS floats per LUT
    int mean_slot;                 // slot of the largest patch (weighted mean), last pass
    Launch L[4];                   // interior, last row, last column, corner
    int gx[4];                     // block columns of each mapping
    int bstart[5];                 // first linear block of each mapping, bstart[4] = total
    int H, W;
    float* enl_nb;                 // [nparam][H*W] or null
    float* alpha;                  // [nparam][H*W] or null
    unsigned char* ubflag;         // [H*W]
    int first_scale;               // 1: overwrite ubflag, 0: or into it
    float* ref;                    // last pass: outputs
    float* phase;
    float* coh;
    unsigned long long* ub_count;
};

constexpr int NLSAR_TY = 32;   // tile rows; each thread owns a column run of 4 of them

// 2*dim*log(2) for dim 2, the constant term of pixel_similarity_2x2 (sim_measures.h:58)
__device__ __forceinline__ float nlsar_log2_term() { return 4.0f * 0.693147182464599609375f; }

// LUT weight of a patch dissimilarity (compute_weights.cl:37-52): 0 above the trained
// maximum, else the table entry of index (uint)((d - min) / (max - min) * (lut - 1)).  The
// index is first formed with a multiplication by s = (lut - 1) / (max - min); it differs from
// the reference's division by at most 3 roundings (6.1e-5 in index units over 256 entries),
// so it can only change the truncated index where its fraction lies within 1e-4 of an
// integer, and there the reference expression itself is evaluated.
__device__ __forceinline__ float lut_weight(float s, float dmin, float dmax, float rng,
                                            float scl, float mul, const float* lut)
{
    if (s > dmax) return 0.f;
    float d = fmaxf(s, dmin);
    d = fminf(d, dmax);
    const float a = d - dmin;
    const float t = a * mul;
    const float fr = t - floorf(t);
    unsigned int idx;
    if (fr < 1e-4f || fr > 0.9999f) idx = (unsigned int)(a / rng * scl);
    else idx = (unsigned int)t;
    return lut[idx];
}

// Block -> tile: the four launch mappings are laid end to end in one grid; linear block
// ids are remapped so that the blocks the hardware sends to one XCD (id mod 8) cover a
// contiguous run of tiles, which then share their halos in that XCD's L2.
__device__ __forceinline__ void nlsar_block_tile(const ScaleArgs& a, int TX, int& m, int& ty0,
                                                 int& tx0)
{
    const int nb = a.bstart[4];
    const int b = blockIdx.x;
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, k = b >> 3;
    const int t = xcd * q + min(xcd, r) + k;
    m = 0;
    while (m < 3 && t >= a.bstart[m + 1]) m++;
    const int local = t - a.bstart[m];
    const int by = local / a.gx[m], bx = local - by * a.gx[m];
    ty0 = a.L[m].vy0 + by * NLSAR_TY;
    tx0 = a.L[m].vx0 + bx * TX;
}

// LDS layout (floats) shared by the kernel and its launcher
template <int TX, int NP, int P>
struct ScaleLds {
    int wsh, lut_stride;
    __host__ __device__ int DXM() const { return TX + wsh + 2 * P + 4; }
    __host__ __device__ int DYM() const { return NLSAR_TY + wsh + 2 * P + 4; }
    __host__ __device__ int RXM() const { return TX + wsh + 4; }
    __host__ __device__ int WRows() const { return NLSAR_TY + wsh + 4; }
    __host__ __device__ int off_R() const { return DYM() * DXM(); }
    __host__ __device__ int off_W() const { return off_R() + NP * DYM() * RXM(); }
    __host__ __device__ int off_L() const { return off_W() + NP * WRows() * RXM(); }
    __host__ __device__ int total() const { return off_L() + NP * lut_stride; }
};

template <int TX, int NP, bool LAST, int P>
__global__ __launch_bounds__(8 * TX) void nlsar_scale_pass(const ScaleArgs a)
{
    constexpr int NT = 8 * TX;
    constexpr int TY = NLSAR_TY;
    constexpr int W2 = 4 + 2 * P;         // register window of a 4-run plus the patch halo
    extern __shared__ float smem[];
    const int wsh = a.wsh;
    const int sw = 2 * wsh + 1;
    const ScaleLds<TX, NP, P> lay{wsh, a.lut_stride};
    const int DXM = lay.DXM(), RXM = lay.RXM(), DYM = lay.DYM();
    float* sD = smem;
    float* sR = smem + lay.off_R();
    float* sW = smem + lay.off_W();
    float* sL = smem + lay.off_L();
    const int RPL = DYM * RXM;            // row-sum plane
    const int WPL = lay.WRows() * RXM;    // weight plane

    const int tid = threadIdx.x;
    const int px = tid % TX, g = tid / TX;  // column, group of 4 rows
    int m_, ty0, tx0;
    nlsar_block_tile(a, TX, m_, ty0, tx0);
    const Launch L = a.L[m_];

    for (int k = 0; k < NP; k++)
        for (int i = tid; i < a.lut_size[k]; i += NT) sL[k * a.lut_stride + i] = a.lutw[k][i];
    float mul[NP], rng[NP], scl[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
        rng[k] = a.dmax[k] - a.dmin[k];
        scl[k] = (float)(a.lut_size[k] - 1);
        mul[k] = scl[k] / rng[k];
    }

    float acc[4][NP][6];
    float macc[4][2];
#pragma unroll
    for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int k = 0; k < NP; k++)
#pragma unroll
            for (int j = 0; j < 6; j++) acc[q][k][j] = 0.f;
        macc[q][0] = macc[q][1] = 0.f;
    }

    const float* A0 = a.A.base;
    const float* A1 = A0 + a.A.plane;
    const float* A2 = A0 + 2 * a.A.plane;
    const float* A3 = A0 + 3 * a.A.plane;
    const float* A4 = A0 + 4 * a.A.plane;
    const float* I0 = a.C.base;
    const float* I1 = I0 + a.C.plane;
    const float* Cr = I0 + 2 * a.C.plane;
    const float* Ci = I0 + 3 * a.C.plane;
    const float K = nlsar_log2_term();
    const int cpitch = a.C.pitch;

    // half-window offsets d = (dy, dx) in plane order: rows above the centre, then the
    // left half of the centre row (compute_pixel_similarities_2x2.cpp:29-47)
    for (int hh = 0; hh <= wsh; hh++) {
        const int dy = hh - wsh;                  // <= 0
        const int wwe = (hh == wsh) ? wsh : sw;
        for (int ww = 0; ww < wwe; ww++) {
            const int dx = ww - wsh;
            const int adx = dx < 0 ? -dx : dx;
            // union of the tile (for w(c, d)) and the tile shifted by -d (for w(c, -d))
            const int RY = TY - dy, RX = TX + adx;
            const int rxo = dx > 0 ? -dx : 0;      // region x origin relative to tx0
            const int DY = RY + 2 * P, DX = RX + 2 * P;
            // 1. pixel dissimilarities delta(g, g + d) of the later pixel g against the
            //    earlier one (pixel_similarity_2x2, sim_measures.h:37-65)
            {
                const int gy0 = ty0 - P, gx0 = tx0 + rxo - P;
                const int n = DY * DX;
                int r = tid / DX, c = tid - r * DX;
                const int sr_ = NT / DX, sc_ = NT - sr_ * DX;
                const long long dofs = (long long)dy * a.A.pitch + dx;
                for (int i = tid; i < n; i += NT) {
                    const long long g1 = a.A.at(gy0 + r, gx0 + c);
                    const long long g2 = g1 + dofs;
                    const float s00 = A0[g1] + A0[g2];
                    const float s01r = A2[g1] + A2[g2];
                    const float s01i = A3[g1] + A3[g2];
                    const float s11 = A1[g1] + A1[g2];
                    const float det = (s00 * s11) - (s01r * s01r + s01i * s01i);
                    float sim = -(((K + A4[g1]) + A4[g2]) - 2.0f * logf(det));
                    if (isnan(sim)) sim = 0.f;
                    sD[r * DXM + c] = sim;
                    r += sr_;
                    c += sc_;
                    if (c >= DX) { c -= DX; r++; }
                }
            }
            __syncthreads();
            // 2. row sums (patch_similarities_row_pass.cl:52-60), left to right, of every
            //    patch size, for runs of 4 columns held in registers
            {
                const int nrx = (RX + 3) >> 2;
                const int n = DY * nrx;
                for (int i = tid; i < n; i += NT) {
                    const int r = i / nrx, x0 = (i - r * nrx) << 2;
                    float w[W2];
                    const float* src = sD + r * DXM + x0;
#pragma unroll
                    for (int t = 0; t < W2; t++) w[t] = src[t];
#pragma unroll
                    for (int k = 0; k < NP; k++) {
                        const int psh = a.psh[k];
                        float* dst = sR + k * RPL + r * RXM + x0;
#pragma unroll
                        for (int m = 0; m < 4; m++) {
                            float acc_s = 0.f;
                            bool first = true;
#pragma unroll
                            for (int jj = -P; jj <= P; jj++) {
                                if (jj >= -psh && jj <= psh) {
                                    const float v = w[m + P + jj];
                                    acc_s = first ? v : acc_s + v;
                                    first = false;
                                }
                            }
                            dst[m] = acc_s;
                        }
                    }
                }
            }
            __syncthreads();
            // 3. column sums (patch_similarities_col_pass.cl:52-60), top to bottom, for
            //    runs of 4 rows, and the LUT weight (compute_weights.cl:19-56)
            {
                const int nry = (RY + 3) >> 2;
                const int n = nry * RX;
                for (int i = tid; i < n; i += NT) {
                    const int ry = i / RX, x = i - ry * RX, y0 = ry << 2;
#pragma unroll
                    for (int k = 0; k < NP; k++) {
                        const int psh = a.psh[k];
                        const float* src = sR + k * RPL + y0 * RXM + x;
                        float w[W2];
#pragma unroll
                        for (int t = 0; t < W2; t++) w[t] = src[t * RXM];
                        const float* lut = sL + k * a.lut_stride;
                        float* dst = sW + k * WPL + y0 * RXM + x;
#pragma unroll
                        for (int m = 0; m < 4; m++) {
                            float acc_s = 0.f;
                            bool first = true;
#pragma unroll
                            for (int jj = -P; jj <= P; jj++) {
                                if (jj >= -psh && jj <= psh) {
                                    const float v = w[m + P + jj];
                                    acc_s = first ? v : acc_s + v;
                                    first = false;
                                }
                            }
                            dst[m * RXM] = lut_weight(acc_s, a.dmin[k], a.dmax[k], rng[k], scl[k], mul[k], lut);
                        }
                    }
                }
            }
            __syncthreads();
            // 4. accumulate w(c, d) (upper half) and w(c, -d) = w_half(c - d, d) (lower
            //    half, copy_symm_weights.cl:53-74) into the nonlocal statistics
            //    (compute_nl_statistics.cl:69-84) of every patch size
            {
                const int py0 = g * 4;
                const int cy = ty0 + py0, cx = tx0 + px;
                const long long cu0 = a.C.at(cy + dy, cx + dx);
                const long long cl0 = a.C.at(cy - dy, cx - dx);
                const int iu0 = py0 * RXM + (px - rxo);
                const int il0 = (py0 - dy) * RXM + (px - dx - rxo);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const long long cu = cu0 + (long long)q * cpitch;
                    const long long cl = cl0 + (long long)q * cpitch;
                    const float i0u = I0[cu], i1u = I1[cu];
                    const float i0l = I0[cl], i1l = I1[cl];
                    const float i0u2 = i0u * i0u, i1u2 = i1u * i1u;
                    const float i0l2 = i0l * i0l, i1l2 = i1l * i1l;
                    float cru = 0.f, ciu = 0.f, crl = 0.f, cil = 0.f;
                    if (LAST) {
                        cru = Cr[cu];
                        ciu = Ci[cu];
                        crl = Cr[cl];
                        cil = Ci[cl];
                    }
#pragma unroll
                    for (int k = 0; k < NP; k++) {
                        const float wu = sW[k * WPL + iu0 + q * RXM];
                        const float wl = sW[k * WPL + il0 + q * RXM];
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
                        if (LAST && k == a.mean_slot) {
                            macc[q][0] = fmaf(wu, cru, macc[q][0]);
                            macc[q][1] = fmaf(wu, ciu, macc[q][1]);
                            macc[q][0] = fmaf(wl, crl, macc[q][0]);
                            macc[q][1] = fmaf(wl, cil, macc[q][1]);
                        }
                    }
                }
            }
            // the next offset's step 1 writes sD only; its barrier orders step 2/3 writes
            // after this step's reads of sW
        }
    }

    // centre weight 1 (routines.cpp:151-157), then the per-pixel statistics
    unsigned int ubl = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int py = g * 4 + q;
        const int iy = ty0 + py - L.vy0, ix = tx0 + px - L.vx0;
        const bool valid = iy < L.ny && ix < L.nx;
        const long long cc = a.C.at(ty0 + py, tx0 + px);
        const float i0c = I0[cc], i1c = I1[cc];
        const long long opix = valid ? (long long)(L.oy0 + iy) * a.W + (L.ox0 + ix) : 0;
        bool fail = false;
        float alpha_mean = 0.f, s_mean = 1.f, m0_mean = 0.f, m1_mean = 0.f;
#pragma unroll
        for (int k = 0; k < NP; k++) {
            float* s = acc[q][k];
            s[0] += 1.0f;
            s[1] = fmaf(1.0f, 1.0f, s[1]);
            s[2] = fmaf(1.0f, i0c, s[2]);
            s[3] = fmaf(1.0f, i0c * i0c, s[3]);
            s[4] = fmaf(1.0f, i1c, s[4]);
            s[5] = fmaf(1.0f, i1c * i1c, s[5]);
            // compute_nl_statistics.cl:86-102, compute_alphas.cl:19-44,
            // compute_enls_nobias.cl:19-39
            const float ws = s[0];
            const float enl = ws * ws / s[1];
            float alpha = 0.0f;
            {
                const float inl = s[2] / ws;
                const float wv = s[3] / ws;
                const float var = wv - inl * inl;
                alpha = fmaxf(alpha, fmaxf(0.0f, (var - inl * inl / 1.0f) / var));
            }
            {
                const float inl = s[4] / ws;
                const float wv = s[5] / ws;
                const float var = wv - inl * inl;
                alpha = fmaxf(alpha, fmaxf(0.0f, (var - inl * inl / 1.0f) / var));
            }
            const float oma = 1.0f - alpha;
            const float enb = enl / (oma * oma + (alpha * alpha + 2.0f * alpha * (1.0f - alpha) / ws) * enl);
            if (!(enb > -1.0f)) fail = true;
            if (valid) {
                if (a.enl_nb) a.enl_nb[(long long)a.pidx[k] * a.H * a.W + opix] = enb;
                if (a.alpha) a.alpha[(long long)a.pidx[k] * a.H * a.W + opix] = alpha;
            }
            if (LAST && k == a.mean_slot) {
                alpha_mean = alpha;
                s_mean = ws;
                m0_mean = s[2];
                m1_mean = s[4];
                // centre weight 1 on the off-diagonal sums as well
                macc[q][0] = fmaf(1.0f, Cr[cc], macc[q][0]);
                macc[q][1] = fmaf(1.0f, Ci[cc], macc[q][1]);
            }
        }
        if (!valid) continue;
        if (!LAST) {
            if (a.first_scale) a.ubflag[opix] = fail ? 1 : 0;
            else if (fail) a.ubflag[opix] = 1;
        } else {
            // best-parameter selection (best_params.cpp:27-53) keeps the last key of the
            // (patch, scale) map whenever every ENL exceeds -1; pixels where the reference's
            // selection loop is undefined are counted and also take the last key.
            if (fail || a.ubflag[opix]) ubl++;
            // weighted_means.cl:62-80 and covmat_decompose.cl:19-39
            const float al = alpha_mean;
            const float crc = Cr[cc], cic = Ci[cc];
            const float m00 = (1 - al) * m0_mean / s_mean + al * i0c;
            const float m11 = (1 - al) * m1_mean / s_mean + al * i1c;
            const float mr = (1 - al) * macc[q][0] / s_mean + al * crc;
            const float mi = (1 - al) * macc[q][1] / s_mean + al * cic;
            const float am = sqrtf(m00), as = sqrtf(m11);
            const float t = 0.5f * (am + as);
            a.ref[opix] = t * t;
            a.phase[opix] = atan2f(mi, mr);
            a.coh[opix] = sqrtf(mi * mi + mr * mr) / (am * as);
        }
    }
    if (LAST && ubl) atomicAdd(a.ub_count, (unsigned long long)ubl);
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

template <int TX, int NP, bool LAST, int P>
static int launch_scale_pass(ScaleArgs a, const Launch* L, int lut_stride, hipStream_t stream)
{
    const ScaleLds<TX, NP, P> lay{a.wsh, lut_stride};
    const size_t lds = sizeof(float) * (size_t)lay.total();
    if (lds > 160 * 1024) {
        set_error("nlsar: search window / patch sizes exceed the LDS budget");
        return -2;
    }
    DKX_HIP_CHECK(hipFuncSetAttribute((const void*)nlsar_scale_pass<TX, NP, LAST, P>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int total = 0;
    for (int i = 0; i < 4; i++) {
        a.L[i] = L[i];
        a.bstart[i] = total;
        if (L[i].ny <= 0 || L[i].nx <= 0) {
            a.gx[i] = 1;
            continue;
        }
        a.gx[i] = (L[i].nx + TX - 1) / TX;
        total += a.gx[i] * ((L[i].ny + NLSAR_TY - 1) / NLSAR_TY);
    }
    a.bstart[4] = total;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = timing_on();
    if (timed) {
        DKX_HIP_CHECK(hipEventCreate(&e0));
        DKX_HIP_CHECK(hipEventCreate(&e1));
        DKX_HIP_CHECK(hipEventRecord(e0, stream));
    }
    hipLaunchKernelGGL((nlsar_scale_pass<TX, NP, LAST, P>), dim3(total), dim3(8 * TX), lds, stream, a);
    DKX_HIP_CHECK(hipGetLastError());
    if (timed) {
        DKX_HIP_CHECK(hipEventRecord(e1, stream));
        timing_push(LAST ? STAGE_NLSAR_LAST : STAGE_NLSAR_SCALE, e0, e1);
    }
    return 0;
}

template <int NP, bool LAST, int P>
static int dispatch_tile(const ScaleArgs& a, const Launch* L, int lut_stride, int tile,
                         hipStream_t stream)
{
    if (tile == 1) return launch_scale_pass<64, NP, LAST, P>(a, L, lut_stride, stream);
    return launch_scale_pass<32, NP, LAST, P>(a, L, lut_stride, stream);
}

template <int NP, bool LAST>
static int dispatch_p(const ScaleArgs& a, const Launch* L, int lut_stride, int tile,
                      hipStream_t stream)
{
    switch (a.P) {
    case 0: return dispatch_tile<NP, LAST, 0>(a, L, lut_stride, tile, stream);
    case 1: return dispatch_tile<NP, LAST, 1>(a, L, lut_stride, tile, stream);
    case 2: return dispatch_tile<NP, LAST, 2>(a, L, lut_stride, tile, stream);
    case 3: return dispatch_tile<NP, LAST, 3>(a, L, lut_stride, tile, stream);
    case 4: return dispatch_tile<NP, LAST, 4>(a, L, lut_stride, tile, stream);
    case 5: return dispatch_tile<NP, LAST, 5>(a, L, lut_stride, tile, stream);
    default: set_error("nlsar: patch sizes above 11 are not supported"); return -4;
    }
}

static int dispatch(const ScaleArgs& a, int np, bool last, const Launch* L, int nl,
                    int lut_stride, int tile, hipStream_t stream)
{
    (void)nl;
    switch (np) {
    case 1: return last ? dispatch_p<1, true>(a, L, lut_stride, tile, stream)
                        : dispatch_p<1, false>(a, L, lut_stride, tile, stream);
    case 2: return last ? dispatch_p<2, true>(a, L, lut_stride, tile, stream)
                        : dispatch_p<2, false>(a, L, lut_stride, tile, stream);
    case 3: return last ? dispatch_p<3, true>(a, L, lut_stride, tile, stream)
                        : dispatch_p<3, false>(a, L, lut_stride, tile, stream);
    default: set_error("nlsar: internal error, more than 3 patch sizes in one pass"); return -1;
    }
}

static const int kTileTY[2] = {32, 32};
static const int kTileTX[2] = {32, 64};

int nlsar_plan(const NlsarProblem& pb, NlsarPlan* plan)
{
    if (pb.H <= 0 || pb.W <= 0) { set_error("nlsar: empty image"); return -1; }
    if (pb.sw < 3 || pb.sw % 2 == 0) { set_error("nlsar: search window size must be odd and >= 3"); return -1; }
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
    plan->tile = std::min(std::max(pb.tile, 0), 1);
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
    DKX_HIP_CHECK(hipMemsetAsync(d_ub, 0, sizeof(unsigned long long), stream));

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
        ScaleArgs sa;
        memset(&sa, 0, sizeof(sa));
        sa.A = A;
        sa.C = C;
        sa.wsh = plan.wsh;
        sa.P = plan.P;
        sa.mean_slot = -1;
        for (size_t k = 0; k < ps.pis.size(); k++) {
            const int pi = ps.pis[k];
            const int idx = si * pb.nP + pi;
            sa.psh[k] = (pb.P[pi] - 1) / 2;
            sa.pidx[k] = idx;
            sa.lutw[k] = d_lut + (size_t)idx * plan.lut_stride;
            sa.dmin[k] = pst[idx]->dissims_min;
            sa.dmax[k] = pst[idx]->dissims_max;
            sa.lut_size[k] = pst[idx]->lut_size;
            if (ps.last && pi == pmax_i) sa.mean_slot = (int)k;
        }
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
        if (ps.last && oi == 0) {
            // a single pass: nothing else wrote the flags
            DKX_HIP_CHECK(hipMemsetAsync(d_flag, 0, (size_t)H * W, stream));
        }
        int rc = dispatch(sa, (int)ps.pis.size(), ps.last, L, 4, plan.lut_stride, plan.tile, stream);
        if (rc) return rc;
    }
    if (ub_host) {
        DKX_HIP_CHECK(hipMemcpyAsync(ub_host, d_ub, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    }
    return 0;
}

}  // namespace dkx
