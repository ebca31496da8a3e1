/*
 * oracle.c -- CPU restatement of despeckCL's NL-SAR / NL-InSAR hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP product path and the "port"
 * CPU baseline of bench.py.  The product never links or calls this file.
 *
 * The restatement is literal: it keeps the reference's buffers (tile-sized planes,
 * half-window similarity planes, full 441-plane weights), its index formulas, its launch
 * quadrants, its float operation order and its host-side quirks.  Citations are
 * file:line in the despeckCL source tree.
 *
 * Floating point: built with -ffp-contract=off; products accumulated into a running sum
 * ("sum += a*b" in the OpenCL kernels, which OpenCL's default FP_CONTRACT may fuse) are
 * written as fmaf() here and in the HIP kernels alike, so that the two implementations
 * differ only in transcendental rounding and summation order where stated.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ACC(acc, a, b) ((acc) = fmaf((a), (b), (acc)))

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* ====================================================================================
 * similarity measures -- src/compute/nlsar/sim_measures.h
 * ==================================================================================== */

/* det_covmat_2x2, sim_measures.h:31-34 */
float orc_det_covmat_2x2(float el_00, float el_01real, float el_01imag, float el_11)
{
    return (el_00 * el_11) - (el_01real * el_01real + el_01imag * el_01imag);
}

static float abs2f(float re, float im) { return re * re + im * im; }

/* det_covmat_3x3, sim_measures.h:79-97 (Leibniz formula for a Hermitian matrix) */
float orc_det_covmat_3x3(float a00, float a11, float a22, float a01r, float a01i, float a02r,
                         float a02i, float a12r, float a12i)
{
    return a00 * a11 * a22 - a00 * abs2f(a12r, a12i) - a11 * abs2f(a02r, a02i) -
           a22 * abs2f(a01r, a01i) +
           2 * (a01r * a12r * a02r - a01i * a12i * a02r + a01r * a12i * a02i +
                a01i * a12r * a02i);
}

/* pixel_similarity_2x2, sim_measures.h:37-65 (device version: float log).
 * p = {el_00, el_01real, el_01imag, el_11} */
float orc_pixel_similarity_2x2(const float p1[4], const float p2[4], int nlooks)
{
    const int dimensions = 2;
    float nom1 = orc_det_covmat_2x2(p1[0], p1[1], p1[2], p1[3]);
    float nom2 = orc_det_covmat_2x2(p2[0], p2[1], p2[2], p2[3]);
    float det = orc_det_covmat_2x2(p1[0] + p2[0], p1[1] + p2[1], p1[2] + p2[2], p1[3] + p2[3]);
    float similarity = (float)(-nlooks) * ((float)(2 * dimensions) * logf(2.0f) + logf(nom1) +
                                           logf(nom2) - 2 * logf(det));
    if (isnan(similarity)) similarity = 0;
    return similarity;
}

/* pixel_similarity_3x3, sim_measures.h:100-158 (device version).
 * p = {a00, a11, a22, a01r, a01i, a02r, a02i, a12r, a12i} */
float orc_pixel_similarity_3x3(const float p1[9], const float p2[9], int nlooks)
{
    const int dimensions = 3;
    float nom1 = orc_det_covmat_3x3(p1[0], p1[1], p1[2], p1[3], p1[4], p1[5], p1[6], p1[7], p1[8]);
    float nom2 = orc_det_covmat_3x3(p2[0], p2[1], p2[2], p2[3], p2[4], p2[5], p2[6], p2[7], p2[8]);
    float det = orc_det_covmat_3x3(p1[0] + p2[0], p1[1] + p2[1], p1[2] + p2[2], p1[3] + p2[3],
                                   p1[4] + p2[4], p1[5] + p2[5], p1[6] + p2[6], p1[7] + p2[7],
                                   p1[8] + p2[8]);
    float similarity = (float)(-nlooks) * ((float)(2 * dimensions) * logf(2.0f) + logf(nom1) +
                                           logf(nom2) - 2 * logf(det));
    if (isnan(similarity)) similarity = 0;
    return similarity;
}

/* The same measures as compiled into the host-side training code
 * (training/patches.cpp:96-177 includes sim_measures.h without <math.h>, so the unqualified
 * log() resolves to ::log(double): the logs and the sum are evaluated in double and the
 * result is rounded to float on return). */
static float host_similarity_2x2(const float p1[4], const float p2[4], int nlooks)
{
    const int dimensions = 2;
    float nom1 = orc_det_covmat_2x2(p1[0], p1[1], p1[2], p1[3]);
    float nom2 = orc_det_covmat_2x2(p2[0], p2[1], p2[2], p2[3]);
    float det = orc_det_covmat_2x2(p1[0] + p2[0], p1[1] + p2[1], p1[2] + p2[2], p1[3] + p2[3]);
    float similarity = (float)(-nlooks * (2 * dimensions * log(2.0f) + log(nom1) + log(nom2) -
                                          2 * log(det)));
    if (isnan(similarity)) similarity = 0;
    return similarity;
}

static float host_similarity_3x3(const float p1[9], const float p2[9], int nlooks)
{
    const int dimensions = 3;
    float nom1 = orc_det_covmat_3x3(p1[0], p1[1], p1[2], p1[3], p1[4], p1[5], p1[6], p1[7], p1[8]);
    float nom2 = orc_det_covmat_3x3(p2[0], p2[1], p2[2], p2[3], p2[4], p2[5], p2[6], p2[7], p2[8]);
    float det = orc_det_covmat_3x3(p1[0] + p2[0], p1[1] + p2[1], p1[2] + p2[2], p1[3] + p2[3],
                                   p1[4] + p2[4], p1[5] + p2[5], p1[6] + p2[6], p1[7] + p2[7],
                                   p1[8] + p2[8]);
    float similarity = (float)(-nlooks * (2 * dimensions * log(2.0f) + log(nom1) + log(nom2) -
                                          2 * log(det)));
    if (isnan(similarity)) similarity = 0;
    return similarity;
}

/* gather the similarity operands of one pixel from planar covmat storage */
static void gather2(const float* c, size_t idx, size_t plane, float out[4])
{
    out[0] = c[idx];
    out[1] = c[idx + 2 * plane];
    out[2] = c[idx + 3 * plane];
    out[3] = c[idx + 6 * plane];
}

static void gather3(const float* c, size_t idx, size_t plane, float out[9])
{
    out[0] = c[idx];
    out[1] = c[idx + 8 * plane];
    out[2] = c[idx + 16 * plane];
    out[3] = c[idx + 2 * plane];
    out[4] = c[idx + 3 * plane];
    out[5] = c[idx + 4 * plane];
    out[6] = c[idx + 5 * plane];
    out[7] = c[idx + 10 * plane];
    out[8] = c[idx + 11 * plane];
}

/* ====================================================================================
 * NL-SAR kernels -- the .cl files of src/compute/nlsar/cl_kernels, launch geometry from cl_wrappers
 * ==================================================================================== */

/* covmat_create.cl:19-56 */
void orc_covmat_create(const float* a1, const float* a2, const float* phi, float* covmat, int h,
                       int w)
{
    const size_t n = (size_t)h * w;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; i++) {
        const float el_00real = a1[i] * a1[i];
        const float el_01real = a1[i] * a2[i] * cosf(phi[i]);
        const float el_01imag = a1[i] * a2[i] * sinf(phi[i]);
        const float el_11real = a2[i] * a2[i];
        covmat[i] = el_00real;
        covmat[i + 1 * n] = 0.0f;
        covmat[i + 2 * n] = el_01real;
        covmat[i + 3 * n] = el_01imag;
        covmat[i + 4 * n] = el_01real;
        covmat[i + 5 * n] = -el_01imag;
        covmat[i + 6 * n] = el_11real;
        covmat[i + 7 * n] = 0.0f;
    }
}

/* covmat_rescale.cl:19-41: off-diagonal elements times pow(min(L/D, 1), 0.33333f) */
void orc_covmat_rescale(float* covmat, int dim, int nlooks, int h, int w)
{
    const float LD = ((float)nlooks) / dim;
    const float gamma = powf(fminf(LD, 1.0f), 0.33333f);
    const size_t n = (size_t)h * w;
    for (int r = 0; r < dim; r++)
        for (int c = 0; c < dim; c++)
            if (r != c) {
                float* re = covmat + (size_t)(2 * (r * dim + c)) * n;
                float* im = covmat + (size_t)(2 * (r * dim + c) + 1) * n;
                for (size_t i = 0; i < n; i++) {
                    re[i] *= gamma;
                    im[i] *= gamma;
                }
            }
}

/* covmat_spatial_avg::gen_gauss, cl_wrappers/covmat_spatial_avg.cpp:35-50.
 * exp of a double argument, stored as float, normalised by a float sum. */
void orc_gen_gauss(int scale_size, float* out)
{
    const int ssh = (scale_size - 1) / 2;
    int k = 0;
    for (int x = -ssh; x <= ssh; x++)
        for (int y = -ssh; y <= ssh; y++)
            out[k++] = (float)exp(-M_PI * (x * x + y * y) / (double)powf(ssh + 0.5f, 2.0f));
    float K = 0.0f;
    for (int i = 0; i < k; i++) K += out[i];
    for (int i = 0; i < k; i++) out[i] = out[i] / K;
}

/* covmat_spatial_avg.cl:19-63.  Output (h_overlap x w_overlap) from input
 * (h_overlap+smax-1 x w_overlap+smax-1), offset delta=(smax-s)/2 (quirk: scales aligned on
 * the largest one). */
void orc_covmat_spatial_avg(const float* in, float* out, int dim, int h_overlap, int w_overlap,
                            int scale_size, int scale_size_max)
{
    const int delta = (scale_size_max - scale_size) / 2;
    const int h_avg = h_overlap + scale_size_max - 1;
    const int w_avg = w_overlap + scale_size_max - 1;
    float gauss[64 * 64];
    orc_gen_gauss(scale_size, gauss);
    for (int i = 0; i < 2 * dim * dim; i++) {
        const float* src = in + (size_t)i * h_avg * w_avg;
        float* dst = out + (size_t)i * h_overlap * w_overlap;
#pragma omp parallel for schedule(static)
        for (int x = 0; x < h_overlap; x++)
            for (int y = 0; y < w_overlap; y++) {
                float sum = 0;
                for (int kx = 0; kx < scale_size; kx++)
                    for (int ky = 0; ky < scale_size; ky++)
                        ACC(sum, gauss[kx * scale_size + ky],
                            src[(size_t)(delta + x + kx) * w_avg + delta + y + ky]);
                dst[(size_t)x * w_overlap + y] = sum;
            }
    }
}

/* compute_pixel_similarities_{2x2,3x3}.cl.in:21-103 launched for the two quadrants of
 * cl_wrappers/compute_pixel_similarities_2x2.cpp:51-79 (block size 16, quadrant 2 with a
 * global offset of wsh columns).  Output planes hh*sw+ww, (h_overlap-wsh) x w_overlap. */
static void pixel_similarities(const float* covmat, float* sims, int h_overlap, int w_overlap,
                               int dim, int nlooks, int sw)
{
    const int wsh = (sw - 1) / 2;
    const int h_symm = h_overlap - wsh;
    const int w_symm = w_overlap;
    const int bs = 16;
    const int gh = bs * ((h_overlap - wsh - 1) / bs + 1);
    const int gw = bs * ((w_overlap - wsh - 1) / bs + 1);
    const size_t plane = (size_t)h_overlap * w_overlap;
    const size_t splane = (size_t)h_symm * w_symm;
    for (int q = 0; q < 2; q++) {
        const int hh0 = 0, hh1 = q == 0 ? wsh : wsh + 1;
        const int ww0 = q == 0 ? wsh : 0, ww1 = q == 0 ? sw : wsh;
        const int w_lo = q == 0 ? 0 : wsh;
        const int w_hi = imin(w_lo + gw, w_symm);
        const int h_hi = imin(gh, h_symm);
#pragma omp parallel for schedule(static)
        for (int h = 0; h < h_hi; h++)
            for (int w = w_lo; w < w_hi; w++)
                for (int hh = hh0; hh < hh1; hh++)
                    for (int ww = ww0; ww < ww1; ww++) {
                        const size_t cidx = (size_t)(h + wsh) * w_overlap + w;
                        const long long widx = (long long)(h + hh) * w_overlap + (w - wsh + ww);
                        const size_t oidx = (size_t)(hh * sw + ww) * splane + (size_t)h * w_symm + w;
                        if (widx < 0 || (size_t)widx >= plane) { sims[oidx] = 0; continue; }
                        if (dim == 2) {
                            float p1[4], p2[4];
                            gather2(covmat, cidx, plane, p1);
                            gather2(covmat, (size_t)widx, plane, p2);
                            sims[oidx] = orc_pixel_similarity_2x2(p1, p2, nlooks);
                        } else {
                            float p1[9], p2[9];
                            gather3(covmat, cidx, plane, p1);
                            gather3(covmat, (size_t)widx, plane, p2);
                            sims[oidx] = orc_pixel_similarity_3x3(p1, p2, nlooks);
                        }
                    }
    }
}

/* patch_similarities_row_pass.cl:19-71 then patch_similarities_col_pass.cl:19-70 as launched
 * by cl_wrappers/compute_patch_similarities.cpp:60-117.  Row sums left to right, then column
 * sums top to bottom, each starting from 0; the patch is centred (offset (pmax-p)/2). */
void orc_patch_similarities(const float* pix, float* pat, int h_sim, int w_sim, int sw,
                            int patch_size, int patch_size_max)
{
    const int wsh = (sw - 1) / 2;
    const int nplanes = sw * wsh + wsh;
    const int offset = (patch_size_max - patch_size) / 2;
    const int ori_offset = (patch_size_max - 1) / 2;
    const int h_ori = h_sim - patch_size_max + 1;
    const int w_ori = w_sim - patch_size_max + 1;
    const int h_int = h_sim - 2 * offset;
    const int psh = (patch_size - 1) / 2;
    float* intermed = (float*)malloc(sizeof(float) * (size_t)nplanes * h_int * w_ori);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nplanes; z++) {
        const float* src = pix + (size_t)z * h_sim * w_sim;
        float* mid = intermed + (size_t)z * h_int * w_ori;
        for (int y = 0; y < h_int; y++)
            for (int x = 0; x < w_ori; x++) {
                float sum = 0;
                for (int k = -psh; k <= psh; k++) {
                    const int xi = ori_offset + x + k;
                    const int yi = offset + y;
                    const float v = (xi < 0 || xi >= w_sim || yi >= h_sim)
                                        ? 0.0f : src[(size_t)yi * w_sim + xi];
                    sum += v;
                }
                mid[(size_t)y * w_ori + x] = sum;
            }
        float* dst = pat + (size_t)z * h_ori * w_ori;
        for (int y = 0; y < h_ori; y++)
            for (int x = 0; x < w_ori; x++) {
                float sum = 0;
                for (int k = -psh; k <= psh; k++) {
                    const int yi = psh + y + k;
                    const float v = (yi < 0 || yi >= h_int) ? 0.0f : mid[(size_t)yi * w_ori + x];
                    sum += v;
                }
                dst[(size_t)y * w_ori + x] = sum;
            }
    }
    free(intermed);
}

/* compute_weights.cl:19-56 (NL-SAR): LUT weight, 0 above the trained maximum */
static float lut_weight(float dissim, const orc_stats* st, float h, float c)
{
    if (dissim > st->dissims_max) return 0.0f;
    dissim = fmaxf(dissim, st->dissims_min);
    dissim = fminf(dissim, st->dissims_max);
    const float mapped_idx = (dissim - st->dissims_min) /
                             (st->dissims_max - st->dissims_min) * (float)(st->lut_size - 1);
    const float quantile = st->quantilles[(unsigned int)mapped_idx];
    const float x = st->chi2cdf_inv[(unsigned int)(quantile * (float)(st->lut_size - 1))];
    return expf(-fabsf(x - c) / h);
}

/* copy_symm_weights.cl:19-78: expand the half window to all sw*sw planes using
 * w(c, d) = w_symm(c + d, -d) for the lower half */
void orc_copy_symm_weights(const float* wsymm, float* wfull, int h_ori, int w_ori, int sw)
{
    const int wsh = (sw - 1) / 2;
    const int w_symm = w_ori + 2 * wsh;
    const int h_symm = h_ori + wsh;
    const size_t splane = (size_t)h_symm * w_symm;
    const size_t oplane = (size_t)h_ori * w_ori;
#pragma omp parallel for schedule(static)
    for (int h = 0; h < h_ori; h++)
        for (int w = 0; w < w_ori; w++) {
            for (int i = 0; i < wsh * sw + wsh; i++)
                wfull[(size_t)i * oplane + (size_t)h * w_ori + w] =
                    wsymm[(size_t)i * splane + (size_t)h * w_symm + w + wsh];
            for (int hh = wsh; hh < sw; hh++) {
                const int ww0 = hh == wsh ? wsh + 1 : 0;
                for (int ww = ww0; ww < sw; ww++) {
                    const int hs = h + hh - wsh;
                    const int ws = wsh + w + ww - wsh;
                    const int hhs = sw - hh - 1;
                    const int wws = sw - ww - 1;
                    wfull[(size_t)(hh * sw + ww) * oplane + (size_t)h * w_ori + w] =
                        wsymm[(size_t)(hhs * sw + wws) * splane + (size_t)hs * w_symm + ws];
                }
            }
        }
}

/* compute_nl_statistics.cl:23-104 */
void orc_nl_statistics(const float* covmat_ori, const float* weights, float* int_nl, float* wvar,
                       float* wsums, float* enl, int h_ori, int w_ori, int sw, int pmax, int smax,
                       int dim)
{
    const int h_avg = h_ori + sw + pmax + smax - 3;
    const int w_avg = w_ori + sw + pmax + smax - 3;
    const int psh = (pmax - 1) / 2;
    const int swh = (smax - 1) / 2;
    const int step_diag = dim * 2 + 2;
    const size_t plane = (size_t)h_avg * w_avg;
    const size_t oplane = (size_t)h_ori * w_ori;
#pragma omp parallel for schedule(static)
    for (int x0 = 0; x0 < h_ori; x0++)
        for (int y0 = 0; y0 < w_ori; y0++) {
            float wv[3] = {0, 0, 0}, inl[3] = {0, 0, 0};
            float ws = 0.0f, ws2 = 0.0f;
            for (int x = 0; x < sw; x++)
                for (int y = 0; y < sw; y++) {
                    const float weight = weights[(size_t)(x * sw + y) * oplane + (size_t)x0 * w_ori + y0];
                    ws += weight;
                    ACC(ws2, weight, weight);
                    for (int d = 0; d < dim; d++) {
                        const float I = covmat_ori[(size_t)d * step_diag * plane +
                                                   (size_t)(x0 + x + psh + swh) * w_avg +
                                                   (y0 + y + psh + swh)];
                        const float I2 = I * I;
                        ACC(inl[d], weight, I);
                        ACC(wv[d], weight, I2);
                    }
                }
            const size_t o = (size_t)x0 * w_ori + y0;
            wsums[o] = ws;
            enl[o] = ws * ws / ws2;
            for (int d = 0; d < dim; d++) {
                inl[d] /= ws;
                wv[d] /= ws;
                wvar[d * oplane + o] = wv[d] - inl[d] * inl[d];
                int_nl[d * oplane + o] = inl[d];
            }
        }
}

/* compute_alphas.cl:19-44.  OpenCL max() is taken as fmax (a NaN operand yields the other). */
void orc_alphas(const float* int_nl, const float* wvar, float* alphas, int h_ori, int w_ori,
                int dim, int nlooks)
{
    const size_t n = (size_t)h_ori * w_ori;
    for (size_t i = 0; i < n; i++) {
        float alpha = 0.0f;
        for (int d = 0; d < dim; d++) {
            const float var = wvar[d * n + i];
            const float inl = int_nl[d * n + i];
            alpha = fmaxf(alpha, fmaxf(0.0f, (var - inl * inl / (float)nlooks) / var));
        }
        alphas[i] = alpha;
    }
}

/* enl_nobias, compute_enls_nobias.cl:19-23 (Deledalle 2015, Eq. 20) */
float orc_enl_nobias(float enl, float alpha, float wsum)
{
    const float oma = 1.0f - alpha;
    return enl / (oma * oma + (alpha * alpha + 2.0f * alpha * (1.0f - alpha) / wsum) * enl);
}

/* weighted_means.cl:19-83: each of the 2*dim*dim planes, written at the interior */
void orc_weighted_means(const float* cin, float* cout, const float* weights, const float* alphas,
                        int h_ori, int w_ori, int sw, int pmax, int smax, int dim)
{
    const int h_ov = h_ori + sw + pmax - 2;
    const int w_ov = w_ori + sw + pmax - 2;
    const int h_avg = h_ov + smax - 1;
    const int w_avg = w_ov + smax - 1;
    const int wsh = (sw - 1) / 2, psh = (pmax - 1) / 2, wwh = (smax - 1) / 2;
    const size_t plane = (size_t)h_avg * w_avg;
    const size_t oplane = (size_t)h_ori * w_ori;
    for (int d = 0; d < 2 * dim * dim; d++) {
        const float* src = cin + d * plane;
#pragma omp parallel for schedule(static)
        for (int x0 = 0; x0 < h_ori; x0++)
            for (int y0 = 0; y0 < w_ori; y0++) {
                float cnew = 0.0f, wsum = 0.0f;
                const float alpha = alphas[(size_t)x0 * w_ori + y0];
                for (int x = 0; x < sw; x++)
                    for (int y = 0; y < sw; y++) {
                        const float weight = weights[(size_t)(x * sw + y) * oplane + (size_t)x0 * w_ori + y0];
                        wsum += weight;
                        ACC(cnew, weight, src[(size_t)(x0 + x + wwh + psh) * w_avg + y0 + y + wwh + psh]);
                    }
                const float centre = src[(size_t)(x0 + wsh + wwh + psh) * w_avg + y0 + wsh + wwh + psh];
                cout[d * plane + (size_t)(wsh + psh + wwh + x0) * w_avg + (wsh + psh + wwh + y0)] =
                    (1 - alpha) * cnew / wsum + alpha * centre;
            }
    }
}

/* covmat_decompose.cl:19-39 */
void orc_covmat_decompose(const float* covmat, float* ref, float* phase, float* coh, int h, int w)
{
    const size_t n = (size_t)h * w;
    for (size_t i = 0; i < n; i++) {
        const float am = sqrtf(covmat[i]);
        const float as = sqrtf(covmat[i + 6 * n]);
        const float s = 0.5f * (am + as);
        ref[i] = s * s;
        phase[i] = atan2f(covmat[i + 3 * n], covmat[i + 2 * n]);
        const float im = covmat[i + 3 * n], re = covmat[i + 2 * n];
        coh[i] = sqrtf(im * im + re * re) / (am * as);
    }
}

/* ====================================================================================
 * chi-squared inverse CDF (the reference calls GSL's gsl_cdf_chisq_Pinv, stats.cpp:76).
 * Own implementation: regularised lower incomplete gamma P(a, x) by its power series
 * (x < a + 1) or Lentz's continued fraction for Q (x >= a + 1), inverted with safeguarded
 * Newton steps from the Wilson-Hilferty start.  Accurate to a few ulp of double, far below
 * the float rounding applied to the table entries.
 * ==================================================================================== */
static double gammainc_lower_reg(double a, double x)
{
    if (x <= 0) return 0.0;
    const double lpre = a * log(x) - x - lgamma(a);
    if (x < a + 1.0) {
        double term = 1.0 / a, sum = term, ap = a;
        for (int n = 0; n < 100000; n++) {
            ap += 1.0;
            term *= x / ap;
            sum += term;
            if (fabs(term) < fabs(sum) * 1e-17) break;
        }
        return sum * exp(lpre);
    } else {
        const double tiny = 1e-300;
        double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, hcf = d;
        for (int i = 1; i < 100000; i++) {
            const double an = -i * (i - a);
            b += 2.0;
            d = an * d + b;
            if (fabs(d) < tiny) d = tiny;
            c = b + an / c;
            if (fabs(c) < tiny) c = tiny;
            d = 1.0 / d;
            const double del = d * c;
            hcf *= del;
            if (fabs(del - 1.0) < 1e-17) break;
        }
        return 1.0 - exp(lpre) * hcf;
    }
}

double orc_chi2_cdf(double x, double nu) { return gammainc_lower_reg(0.5 * nu, 0.5 * x); }

double orc_chi2_inv(double p, double nu)
{
    if (p <= 0.0) return 0.0;
    if (p >= 1.0) return INFINITY;
    const double a = 0.5 * nu;
    /* Wilson-Hilferty start */
    const double t = sqrt(2.0 / (9.0 * nu));
    /* rational approximation of the normal quantile (Abramowitz & Stegun 26.2.23) */
    const double pp = p < 0.5 ? p : 1.0 - p;
    const double tt = sqrt(-2.0 * log(pp));
    double z = tt - (2.515517 + 0.802853 * tt + 0.010328 * tt * tt) /
                        (1.0 + 1.432788 * tt + 0.189269 * tt * tt + 0.001308 * tt * tt * tt);
    if (p < 0.5) z = -z;
    double x = nu * pow(1.0 - t * t + z * t, 3.0);
    if (!(x > 0)) x = 0.5 * nu * 0.01 + 1e-3;
    double lo = 0.0, hi = INFINITY;
    for (int it = 0; it < 200; it++) {
        const double f = orc_chi2_cdf(x, nu) - p;
        if (f > 0) hi = x; else lo = x;
        /* density of chi2(nu) at x */
        const double dens = exp((a - 1.0) * log(0.5 * x) - 0.5 * x - lgamma(a)) * 0.5;
        double xn = x - f / dens;
        if (!(xn > lo && xn < hi) || !isfinite(xn)) xn = isfinite(hi) ? 0.5 * (lo + hi) : 2.0 * x;
        if (fabs(xn - x) <= 1e-15 * fabs(x)) { x = xn; break; }
        x = xn;
    }
    return x;
}

/* ====================================================================================
 * NL-SAR statistics: nlsar::stats::stats, stats.cpp:28-80
 * ==================================================================================== */
static int cmp_float(const void* a, const void* b)
{
    const float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}

int orc_stats_build(const float* dissims, int n, int lut_size, float* dmin, float* dmax,
                    float* quantilles, float* chi2cdf_inv)
{
    /* sort, drop NaN and inf (stats.cpp:30-32); NaNs are removed before sorting here, which
     * gives the same set without std::sort's NaN ordering issue */
    float* v = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    int m = 0;
    for (int i = 0; i < n; i++)
        if (!isnan(dissims[i]) && !isinf(dissims[i])) v[m++] = dissims[i];
    qsort(v, (size_t)m, sizeof(float), cmp_float);
    if (m == 0) { free(v); return 0; }
    *dmin = v[0];
    *dmax = v[m - 1];
    /* get_quantilles, stats.cpp:51-65 */
    const float step = (*dmax - *dmin) / (float)lut_size;
    for (int i = 0; i < lut_size; i++) {
        const float dissim = *dmin + (float)i * step;
        /* std::lower_bound: first element not less than dissim */
        int lo = 0, hi = m;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (v[mid] < dissim) lo = mid + 1; else hi = mid;
        }
        quantilles[i] = ((float)lo) / (float)m;
    }
    /* get_chi2cdf_inv, stats.cpp:67-80 */
    const float step_p = 1.0f / (float)lut_size;
    for (int i = 0; i < lut_size - 1; i++) {
        const float y = (float)i * step_p;
        chi2cdf_inv[i] = (float)orc_chi2_inv((double)y, 49.0);
    }
    chi2cdf_inv[lut_size - 1] = 100000.0f;
    free(v);
    return m;
}

/* ====================================================================================
 * tiling -- data/sub_images.cpp:24-69, data/tile_iterator.cpp:21-75,
 *           compute/common/map_filter_tiles.h:37-104 (tasks emulated in spawn order)
 * ==================================================================================== */
static void get_sub_image(const float* img, int height, int width, int nplanes, int h_low,
                          int w_low, int sh, int sw_, float* out)
{
    for (int d = 0; d < nplanes; d++)
        for (int h = 0; h < sh; h++)
            for (int w = 0; w < sw_; w++) {
                const int hi = imin(height - 1, imax(0, h + h_low));
                const int wi = imin(width - 1, imax(0, w + w_low));
                out[(size_t)d * sh * sw_ + (size_t)h * sw_ + w] =
                    img[(size_t)d * height * width + (size_t)hi * width + wi];
            }
}

/* tilecpy of the tile interior with the clamped write of write_sub_image: interior
 * positions beyond the image edge land on the last row / column, last write wins */
static void write_interior(float* img, int height, int width, int nplanes, const float* tile,
                           int h_low, int w_low, int th, int tw, int overlap)
{
    for (int d = 0; d < nplanes; d++)
        for (int h = overlap; h < th - overlap; h++)
            for (int w = overlap; w < tw - overlap; w++) {
                const int hi = imin(height - 1, imax(0, h + h_low));
                const int wi = imin(width - 1, imax(0, w + w_low));
                img[(size_t)d * height * width + (size_t)hi * width + wi] =
                    tile[(size_t)d * th * tw + (size_t)h * tw + w];
            }
}

/* ====================================================================================
 * NL-SAR per tile -- nlsar_filter_sub_image.cpp:146-390, routines/routines.cpp:21-223,
 * best_params.cpp:27-53, best_alpha_copy.cpp:21-38, copy_best_weights.cl:19-36
 * ==================================================================================== */
static const orc_stats* find_stats(const orc_stats* st, int n, int p, int s)
{
    for (int i = 0; i < n; i++)
        if (st[i].patch_size == p && st[i].scale_size == s) return &st[i];
    return NULL;
}

static int params_less(int p1, int s1, int p2, int s2)
{
    /* nlsar::params::operator<, parameters.h:37-40 */
    return (p1 < p2) || (p1 == p2 && s1 < s2);
}

/* covmat_ori: 2*dim*dim planes of th x tw; covmat_filt receives the interior */
/* cond (optional, th x tw, interior written): min over the diagonal of var / I_nl^2 of the
 * selected parameter; alpha = 1 - I_nl^2 / var is ill-conditioned where it is small */
/* maps (optional, 3*nP*nS planes of th x tw, interior written): every key's ENL_nobias, then
 * every key's alpha, then every key's conditioning (as cond), key index
 * scale_index*nP + patch_index */
static int nlsar_tile(const float* covmat_ori, float* covmat_filt, int th, int tw, int dim,
                      int sw, const int* P, int nP, const int* S, int nS, const orc_stats* st,
                      int nst, float hpar, float cpar, long long* ub_pixels, float* cond,
                      float* maps)
{
    int pmax = 0, smax = 0;
    for (int i = 0; i < nP; i++) pmax = imax(pmax, P[i]);
    for (int i = 0; i < nS; i++) smax = imax(smax, S[i]);
    const int psh = (pmax - 1) / 2, wsh = (sw - 1) / 2;
    const int overlap = wsh + psh;
    const int nlooks = 1; /* nlsar_filter_sub_image.cpp:179 */
    const int nplanes_cov = 2 * dim * dim;
    const int h_ov = th - smax + 1, w_ov = tw - smax + 1;
    const int h_ori = h_ov - 2 * overlap, w_ori = w_ov - 2 * overlap;
    if (h_ori <= 0 || w_ori <= 0) return -2;
    const size_t n_ori = (size_t)h_ori * w_ori;
    const size_t tplane = (size_t)th * tw;
    const int nhalf = sw * wsh + wsh;
    const int nfull = sw * sw;

    float* rescaled = (float*)malloc(sizeof(float) * nplanes_cov * tplane);
    memcpy(rescaled, covmat_ori, sizeof(float) * nplanes_cov * tplane);
    orc_covmat_rescale(rescaled, dim, nlooks, th, tw);

    const int nparam = nP * nS;
    float* all_weights = (float*)malloc(sizeof(float) * (size_t)nparam * nfull * n_ori);
    float* enls_nobias = (float*)malloc(sizeof(float) * (size_t)nparam * n_ori);
    float* alphas = (float*)malloc(sizeof(float) * (size_t)nparam * n_ori);
    float* conds = (float*)malloc(sizeof(float) * (size_t)nparam * n_ori);
    int* keyp = (int*)malloc(sizeof(int) * nparam);
    int* keys = (int*)malloc(sizeof(int) * nparam);

    float* avg = (float*)malloc(sizeof(float) * nplanes_cov * (size_t)h_ov * w_ov);
    const int h_pix = h_ov - wsh, w_pix = w_ov;
    float* pix = (float*)malloc(sizeof(float) * (size_t)nhalf * h_pix * w_pix);
    const int h_pat = h_pix - pmax + 1, w_pat = w_pix - pmax + 1;
    float* pat = (float*)malloc(sizeof(float) * (size_t)nhalf * h_pat * w_pat);
    float* int_nl = (float*)malloc(sizeof(float) * dim * n_ori);
    float* wvar = (float*)malloc(sizeof(float) * dim * n_ori);
    float* wsums = (float*)malloc(sizeof(float) * n_ori);
    float* enl = (float*)malloc(sizeof(float) * n_ori);

    int idx = 0;
    for (int si = 0; si < nS; si++) {
        const int s = S[si];
        orc_covmat_spatial_avg(rescaled, avg, dim, h_ov, w_ov, s, smax);
        pixel_similarities(avg, pix, h_ov, w_ov, dim, nlooks, sw);
        for (int pi = 0; pi < nP; pi++, idx++) {
            const int p = P[pi];
            const orc_stats* ps = find_stats(st, nst, p, s);
            keyp[idx] = p;
            keys[idx] = s;
            orc_patch_similarities(pix, pat, h_pix, w_pix, sw, p, pmax);
            /* compute_weights over all half-window planes, in place into pat */
            const size_t npat = (size_t)nhalf * h_pat * w_pat;
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < (long long)npat; i++) pat[i] = lut_weight(pat[i], ps, hpar, cpar);
            float* wfull = all_weights + (size_t)idx * nfull * n_ori;
            orc_copy_symm_weights(pat, wfull, h_ori, w_ori, sw);
            /* centre weight 1, routines.cpp:151-157 */
            for (size_t i = 0; i < n_ori; i++) wfull[(size_t)(sw * wsh + wsh) * n_ori + i] = 1.0f;
            orc_nl_statistics(covmat_ori, wfull, int_nl, wvar, wsums, enl, h_ori, w_ori, sw, pmax,
                              smax, dim);
            float* al = alphas + (size_t)idx * n_ori;
            orc_alphas(int_nl, wvar, al, h_ori, w_ori, dim, nlooks);
            for (size_t i = 0; i < n_ori; i++) {
                double cmin = INFINITY;
                for (int d = 0; d < dim; d++) {
                    const double v = wvar[d * n_ori + i], m = int_nl[d * n_ori + i];
                    const double q = v / (m * m);
                    if (!(q >= cmin)) cmin = q;
                }
                conds[(size_t)idx * n_ori + i] = (float)cmin;
            }
            float* en = enls_nobias + (size_t)idx * n_ori;
            for (size_t i = 0; i < n_ori; i++) en[i] = orc_enl_nobias(enl[i], al[i], wsums[i]);
        }
    }

    /* get_best_params, best_params.cpp:27-53.  best_enl is never updated (stays -1), so a
     * pixel takes every key whose ENL exceeds -1, in map order: the last such key wins.
     * When a mask entry is false the reference's loop runs past the mask (undefined); the
     * per-pixel reading "last key with ENL > -1, else the last key" is used and counted. */
    int* order = (int*)malloc(sizeof(int) * nparam);
    for (int i = 0; i < nparam; i++) order[i] = i;
    for (int i = 1; i < nparam; i++)
        for (int j = i; j > 0 && params_less(keyp[order[j]], keys[order[j]], keyp[order[j - 1]],
                                              keys[order[j - 1]]); j--) {
            const int t = order[j]; order[j] = order[j - 1]; order[j - 1] = t;
        }
    /* The enls_nobias map has an entry, zero-filled, for every key of the statistics
     * collection (nlsar_filter_sub_image.cpp:224-228), so a collection whose last key is not
     * filtered with makes every pixel select that key (0 > -1).  params2idx[key] and
     * best_alpha_copy's mapptr[key] then read default entries: index 0 (the first key of the
     * scale-major loop) and the zero-filled alphas (nlsar_filter_sub_image.cpp:345-353,
     * best_alpha_copy.cpp:26-35). */
    int lastp = st[0].patch_size, lasts = st[0].scale_size;
    for (int k = 1; k < nst; k++)
        if (params_less(lastp, lasts, st[k].patch_size, st[k].scale_size)) {
            lastp = st[k].patch_size;
            lasts = st[k].scale_size;
        }
    int last_in_set = 0;
    for (int k = 0; k < nparam; k++)
        if (keyp[k] == lastp && keys[k] == lasts) last_in_set = 1;
    int* best = (int*)malloc(sizeof(int) * n_ori);
    long long ub = 0;
    for (size_t i = 0; i < n_ori; i++) {
        int sel = -1, all_true = 1;
        for (int k = 0; k < nparam; k++) {
            const float e = enls_nobias[(size_t)order[k] * n_ori + i];
            if (e > -1.0f) sel = order[k]; else all_true = 0;
        }
        if (!all_true) ub++;
        best[i] = !last_in_set ? 0 : sel >= 0 ? sel : order[nparam - 1];
    }
    if (ub_pixels) *ub_pixels += ub;

    /* best_alpha_copy + copy_best_weights */
    float* best_alpha = (float*)malloc(sizeof(float) * n_ori);
    float* best_w = (float*)malloc(sizeof(float) * (size_t)nfull * n_ori);
    for (size_t i = 0; i < n_ori; i++) {
        best_alpha[i] = last_in_set ? alphas[(size_t)best[i] * n_ori + i] : 0.0f;
        for (int k = 0; k < nfull; k++)
            best_w[(size_t)k * n_ori + i] = all_weights[((size_t)best[i] * nfull + k) * n_ori + i];
    }
    orc_weighted_means(covmat_ori, covmat_filt, best_w, best_alpha, h_ori, w_ori, sw, pmax, smax,
                       dim);
    if (cond) {
        const int o = wsh + psh + (smax - 1) / 2;
        for (int y = 0; y < h_ori; y++)
            for (int x = 0; x < w_ori; x++) {
                const size_t i = (size_t)y * w_ori + x;
                cond[(size_t)(y + o) * tw + x + o] = conds[(size_t)best[i] * n_ori + i];
            }
    }
    if (maps) {
        const int o = wsh + psh + (smax - 1) / 2;
        for (int k = 0; k < nparam; k++)
            for (int y = 0; y < h_ori; y++)
                for (int x = 0; x < w_ori; x++) {
                    const size_t i = (size_t)y * w_ori + x;
                    const size_t t = (size_t)(y + o) * tw + x + o;
                    maps[(size_t)k * tplane + t] = enls_nobias[(size_t)k * n_ori + i];
                    maps[(size_t)(nparam + k) * tplane + t] = alphas[(size_t)k * n_ori + i];
                    maps[(size_t)(2 * nparam + k) * tplane + t] = conds[(size_t)k * n_ori + i];
                }
    }

    free(best_w); free(best_alpha); free(best); free(order);
    free(enl); free(wsums); free(wvar); free(int_nl); free(pat); free(pix); free(avg);
    free(keys); free(keyp); free(conds); free(alphas); free(enls_nobias); free(all_weights); free(rescaled);
    return 0;
}

/* parameter checks shared by both NL-SAR overloads: the stats keys must be exactly the
 * patch x scale set (the reference silently uses default entries otherwise) */
static int check_nlsar_args(int height, int width, int dim, int sw, const int* P, int nP,
                            const int* S, int nS, const orc_stats* st, int nst)
{
    if (height <= 0 || width <= 0) return -1;
    if (dim != 2 && dim != 3) return -1;
    if (sw < 3 || sw % 2 == 0 || nP <= 0 || nS <= 0) return -1;
    for (int i = 0; i < nP; i++) if (P[i] < 1 || P[i] % 2 == 0) return -1;
    for (int i = 0; i < nS; i++) if (S[i] < 1 || S[i] % 2 == 0 || S[i] > 63) return -1;
    /* every (patch, scale) pair needs an entry; further entries are allowed (the reference
     * keeps them in its maps, see nlsar_tile) but no key may appear twice */
    if (!st || nst < nP * nS) return -3;
    for (int i = 0; i < nst; i++)
        for (int j = 0; j < i; j++)
            if (st[i].patch_size == st[j].patch_size && st[i].scale_size == st[j].scale_size) return -3;
    for (int i = 0; i < nP; i++)
        for (int j = 0; j < nS; j++) {
            const orc_stats* s = find_stats(st, nst, P[i], S[j]);
            if (!s || s->lut_size < 2) return -3;
        }
    return 0;
}

static int nlsar_overlap(int sw, const int* P, int nP, const int* S, int nS)
{
    int pmax = 0, smax = 0;
    for (int i = 0; i < nP; i++) pmax = imax(pmax, P[i]);
    for (int i = 0; i < nS; i++) smax = imax(smax, S[i]);
    /* get_overlap, nlsar.cpp:41-51 */
    return (pmax - 1) / 2 + (sw - 1) / 2 + (smax - 1) / 2;
}

/* map_filter_tiles over a planar image of nplanes (input planes first); filt_planes are
 * the planes the tile filter fills; serial tile order */
typedef int (*tile_fn)(const float* tile_in, float* tile_out, int th, int tw, void* ctx);

static int map_tiles(const float* img_in, float* img_out, int height, int width, int nin,
                     int nout, int th, int tw, int overlap, tile_fn fn, void* ctx)
{
    float* tin = (float*)malloc(sizeof(float) * (size_t)nin * th * tw);
    float* tout = (float*)calloc((size_t)nout * th * tw, sizeof(float));
    int rc = 0;
    for (int h_low = -overlap; h_low < height && rc == 0; h_low += th - 2 * overlap)
        for (int w_low = -overlap; w_low < width && rc == 0; w_low += tw - 2 * overlap) {
            get_sub_image(img_in, height, width, nin, h_low, w_low, th, tw, tin);
            rc = fn(tin, tout, th, tw, ctx);
            if (rc == 0) write_interior(img_out, height, width, nout, tout, h_low, w_low, th, tw, overlap);
        }
    free(tin);
    free(tout);
    return rc;
}

typedef struct {
    int dim, sw, nP, nS, nst;
    const int *P, *S;
    const orc_stats* st;
    float h, c;
    long long ub;
    int want_cond, want_maps;
} nlsar_ctx;

/* tile output planes: the filter's (3 for InSAR input, 2*dim*dim for covariance input), then
 * the conditioning plane and the 3*nP*nS map planes when requested */
static float* extra_planes(nlsar_ctx* cx, float* tout, size_t n, int nbase, float** maps)
{
    float* cond = cx->want_cond ? tout + (size_t)nbase * n : NULL;
    *maps = cx->want_maps ? tout + (size_t)(nbase + (cx->want_cond ? 1 : 0)) * n : NULL;
    return cond;
}

static int nlsar_insar_tile(const float* tin, float* tout, int th, int tw, void* vctx)
{
    nlsar_ctx* cx = (nlsar_ctx*)vctx;
    const size_t n = (size_t)th * tw;
    float* cov = (float*)malloc(sizeof(float) * 8 * n);
    float* filt = (float*)calloc(8 * n, sizeof(float));
    float* maps;
    float* cond = extra_planes(cx, tout, n, 3, &maps);
    /* data_to_covmat, conversion.cpp:22-62 */
    orc_covmat_create(tin, tin + n, tin + 2 * n, cov, th, tw);
    int rc = nlsar_tile(cov, filt, th, tw, 2, cx->sw, cx->P, cx->nP, cx->S, cx->nS, cx->st,
                        cx->nst, cx->h, cx->c, &cx->ub, cond, maps);
    /* covmat_to_data, conversion.cpp:88-142 */
    if (rc == 0) orc_covmat_decompose(filt, tout, tout + n, tout + 2 * n, th, tw);
    free(filt);
    free(cov);
    return rc;
}

static int nlsar_covmat_tile(const float* tin, float* tout, int th, int tw, void* vctx)
{
    nlsar_ctx* cx = (nlsar_ctx*)vctx;
    const size_t n = (size_t)th * tw;
    float* maps;
    float* cond = extra_planes(cx, tout, n, 2 * cx->dim * cx->dim, &maps);
    return nlsar_tile(tin, tout, th, tw, cx->dim, cx->sw, cx->P, cx->nP, cx->S, cx->nS, cx->st,
                      cx->nst, cx->h, cx->c, &cx->ub, cond, maps);
}

/* the filter over the image, for InSAR input (dim 0: in = a1, a2, phi; filt = ref, phase,
 * coh) or covariance input (dim 2 / 3: 2*dim*dim planes each) */
static int nlsar_image(const float* in, float* filt, float* cond, float* enl_maps,
                       float* alpha_maps, float* cond_maps, int height, int width, int dim, int sw, const int* P,
                       int nP, const int* S, int nS, const orc_stats* st, int nst, float hpar,
                       float cpar, int tile_h, int tile_w, long long* ub_pixels)
{
    int rc = check_nlsar_args(height, width, dim ? dim : 2, sw, P, nP, S, nS, st, nst);
    if (rc) return rc;
    const int overlap = nlsar_overlap(sw, P, nP, S, nS);
    if (tile_h <= 0) tile_h = height + 3 * overlap;
    if (tile_w <= 0) tile_w = width + 3 * overlap;
    if (tile_h <= 2 * overlap || tile_w <= 2 * overlap) return -1;
    const int nin = dim ? 2 * dim * dim : 3, nbase = nin;
    const int want_maps = enl_maps || alpha_maps || cond_maps;
    const int nparam = nP * nS;
    const int nout = nbase + (cond ? 1 : 0) + (want_maps ? 3 * nparam : 0);
    const size_t n = (size_t)height * width;
    float* out = (float*)calloc((size_t)nout * n, sizeof(float));
    memcpy(out, filt, sizeof(float) * nbase * n);
    nlsar_ctx cx = {dim ? dim : 2, sw, nP, nS, nst, P, S, st, hpar, cpar, 0, cond != NULL, want_maps};
    rc = map_tiles(in, out, height, width, nin, nout, tile_h, tile_w, overlap,
                   dim ? nlsar_covmat_tile : nlsar_insar_tile, &cx);
    if (rc == 0) {
        memcpy(filt, out, sizeof(float) * nbase * n);
        const float* extra = out + (size_t)nbase * n;
        if (cond) { memcpy(cond, extra, sizeof(float) * n); extra += n; }
        if (want_maps) {
            if (enl_maps) memcpy(enl_maps, extra, sizeof(float) * nparam * n);
            if (alpha_maps) memcpy(alpha_maps, extra + (size_t)nparam * n, sizeof(float) * nparam * n);
            if (cond_maps) memcpy(cond_maps, extra + (size_t)2 * nparam * n, sizeof(float) * nparam * n);
        }
    }
    if (ub_pixels) *ub_pixels = cx.ub;
    free(out);
    return rc;
}

int orc_nlsar_insar(const float* a1, const float* a2, const float* phi, float* ref_filt,
                    float* phase_filt, float* coh_filt, int height, int width, int sw,
                    const int* P, int nP, const int* S, int nS, const orc_stats* st, int nst,
                    float hpar, float cpar, int tile_h, int tile_w, long long* ub_pixels)
{
    return orc_nlsar_insar_maps(a1, a2, phi, ref_filt, phase_filt, coh_filt, NULL, NULL, NULL,
                                NULL, height, width, sw, P, nP, S, nS, st, nst, hpar, cpar,
                                tile_h, tile_w, ub_pixels);
}

int orc_nlsar_insar_cond(const float* a1, const float* a2, const float* phi, float* ref_filt,
                         float* phase_filt, float* coh_filt, float* cond, int height, int width,
                         int sw, const int* P, int nP, const int* S, int nS, const orc_stats* st,
                         int nst, float hpar, float cpar, int tile_h, int tile_w,
                         long long* ub_pixels)
{
    return orc_nlsar_insar_maps(a1, a2, phi, ref_filt, phase_filt, coh_filt, cond, NULL, NULL,
                                NULL, height, width, sw, P, nP, S, nS, st, nst, hpar, cpar,
                                tile_h, tile_w, ub_pixels);
}

int orc_nlsar_insar_maps(const float* a1, const float* a2, const float* phi, float* ref_filt,
                         float* phase_filt, float* coh_filt, float* cond, float* enl_maps,
                         float* alpha_maps, float* cond_maps, int height, int width, int sw,
                         const int* P, int nP, const int* S, int nS, const orc_stats* st,
                         int nst, float hpar, float cpar, int tile_h, int tile_w,
                         long long* ub_pixels)
{
    if (height <= 0 || width <= 0) return -1;
    const size_t n = (size_t)height * width;
    float* in = (float*)malloc(sizeof(float) * 3 * n);
    float* filt = (float*)malloc(sizeof(float) * 3 * n);
    memcpy(in, a1, n * sizeof(float));
    memcpy(in + n, a2, n * sizeof(float));
    memcpy(in + 2 * n, phi, n * sizeof(float));
    memcpy(filt, ref_filt, n * sizeof(float));
    memcpy(filt + n, phase_filt, n * sizeof(float));
    memcpy(filt + 2 * n, coh_filt, n * sizeof(float));
    int rc = nlsar_image(in, filt, cond, enl_maps, alpha_maps, cond_maps, height, width, 0, sw, P, nP, S, nS,
                         st, nst, hpar, cpar, tile_h, tile_w, ub_pixels);
    if (rc == 0) {
        memcpy(ref_filt, filt, n * sizeof(float));
        memcpy(phase_filt, filt + n, n * sizeof(float));
        memcpy(coh_filt, filt + 2 * n, n * sizeof(float));
    }
    free(filt);
    free(in);
    return rc;
}

int orc_nlsar_covmat(const float* covmat_raw, float* covmat_filt, int height, int width, int dim,
                     int sw, const int* P, int nP, const int* S, int nS, const orc_stats* st,
                     int nst, float hpar, float cpar, int tile_h, int tile_w, long long* ub_pixels)
{
    return orc_nlsar_covmat_maps(covmat_raw, covmat_filt, NULL, NULL, NULL, NULL, height, width, dim,
                                 sw, P, nP, S, nS, st, nst, hpar, cpar, tile_h, tile_w, ub_pixels);
}

/* as orc_nlsar_covmat; cond (optional, height x width) as in orc_nlsar_insar_cond */
int orc_nlsar_covmat_cond(const float* covmat_raw, float* covmat_filt, float* cond, int height,
                          int width, int dim, int sw, const int* P, int nP, const int* S, int nS,
                          const orc_stats* st, int nst, float hpar, float cpar, int tile_h,
                          int tile_w, long long* ub_pixels)
{
    return orc_nlsar_covmat_maps(covmat_raw, covmat_filt, cond, NULL, NULL, NULL, height, width, dim,
                                 sw, P, nP, S, nS, st, nst, hpar, cpar, tile_h, tile_w, ub_pixels);
}

int orc_nlsar_covmat_maps(const float* covmat_raw, float* covmat_filt, float* cond,
                          float* enl_maps, float* alpha_maps, float* cond_maps, int height,
                          int width, int dim, int sw, const int* P, int nP, const int* S, int nS,
                          const orc_stats* st, int nst, float hpar, float cpar, int tile_h,
                          int tile_w, long long* ub_pixels)
{
    if (dim != 2 && dim != 3) return -1;
    return nlsar_image(covmat_raw, covmat_filt, cond, enl_maps, alpha_maps, cond_maps, height, width, dim,
                       sw, P, nP, S, nS, st, nst, hpar, cpar, tile_h, tile_w, ub_pixels);
}

/* ====================================================================================
 * NL-SAR training -- training/get_dissims.cpp:32-117, patches.cpp:32-94,
 * combinations.cpp:26-41, get_stats.cpp:25-59, data.cpp:60-89
 * ==================================================================================== */
long long orc_training_dissims(const float* covmat, int height, int width, int dim,
                               int patch_size, int scale_size, float* out, long long cap)
{
    const int np = 2 * dim * dim;
    const size_t n = (size_t)height * width;
    float* cov = (float*)malloc(sizeof(float) * np * n);
    memcpy(cov, covmat, sizeof(float) * np * n);
    orc_covmat_rescale(cov, dim, 1, height, width);
    const int h_ov = height - scale_size + 1, w_ov = width - scale_size + 1;
    if (h_ov <= 0 || w_ov <= 0) { free(cov); return -1; }
    float* avg = (float*)malloc(sizeof(float) * np * (size_t)h_ov * w_ov);
    orc_covmat_spatial_avg(cov, avg, dim, h_ov, w_ov, scale_size, scale_size);
    /* get_all_patches: non-overlapping p x p tiles from (0,0), edge-clamped, row-major */
    const int nph = (h_ov + patch_size - 1) / patch_size;
    const int npw = (w_ov + patch_size - 1) / patch_size;
    const int npatch = nph * npw;
    const size_t psz = (size_t)patch_size * patch_size;
    float* patches = (float*)malloc(sizeof(float) * np * psz * npatch);
    int k = 0;
    for (int h_low = 0; h_low < h_ov; h_low += patch_size)
        for (int w_low = 0; w_low < w_ov; w_low += patch_size, k++)
            get_sub_image(avg, h_ov, w_ov, np, h_low, w_low, patch_size, patch_size,
                          patches + (size_t)k * np * psz);
    /* get_all_dissim_combs: head = last patch against every earlier one, recursively;
     * dissimilarity() sums the per-pixel host similarity in row-major order */
    long long cnt = 0;
    for (int head = npatch - 1; head >= 1; head--) {
        const float* first = patches + (size_t)head * np * psz;
        for (int j = 0; j < head; j++) {
            const float* second = patches + (size_t)j * np * psz;
            float sum = 0.0f;
            for (size_t i = 0; i < psz; i++) {
                if (dim == 2) {
                    float p1[4], p2[4];
                    gather2(first, i, psz, p1);
                    gather2(second, i, psz, p2);
                    sum += host_similarity_2x2(p1, p2, 1);
                } else {
                    float p1[9], p2[9];
                    gather3(first, i, psz, p1);
                    gather3(second, i, psz, p2);
                    sum += host_similarity_3x3(p1, p2, 1);
                }
            }
            if (cnt < cap) out[cnt] = sum;
            cnt++;
        }
    }
    free(patches);
    free(avg);
    free(cov);
    return cnt;
}

int orc_nlsar_training_covmat(const float* covmat, int height, int width, int dim, const int* P,
                              int nP, const int* S, int nS, orc_stats* out, float* qbuf, float* cbuf)
{
    const int lut_size = 256; /* get_stats.cpp:31 */
    int k = 0;
    /* std::map order: patch-major, then scale */
    int* Ps = (int*)malloc(sizeof(int) * nP);
    int* Ss = (int*)malloc(sizeof(int) * nS);
    memcpy(Ps, P, sizeof(int) * nP);
    memcpy(Ss, S, sizeof(int) * nS);
    for (int i = 1; i < nP; i++) for (int j = i; j > 0 && Ps[j] < Ps[j - 1]; j--) { int t = Ps[j]; Ps[j] = Ps[j - 1]; Ps[j - 1] = t; }
    for (int i = 1; i < nS; i++) for (int j = i; j > 0 && Ss[j] < Ss[j - 1]; j--) { int t = Ss[j]; Ss[j] = Ss[j - 1]; Ss[j - 1] = t; }
    int rc = 0;
    for (int i = 0; i < nP && rc == 0; i++)
        for (int j = 0; j < nS && rc == 0; j++, k++) {
            const int nph = (height - Ss[j] + 1 + Ps[i] - 1) / Ps[i];
            const int npw = (width - Ss[j] + 1 + Ps[i] - 1) / Ps[i];
            const long long npairs = (long long)nph * npw * (nph * npw - 1) / 2;
            if (npairs <= 0) { rc = -1; break; }
            float* d = (float*)malloc(sizeof(float) * (size_t)npairs);
            orc_training_dissims(covmat, height, width, dim, Ps[i], Ss[j], d, npairs);
            out[k].patch_size = Ps[i];
            out[k].scale_size = Ss[j];
            out[k].lut_size = lut_size;
            out[k].quantilles = qbuf + (size_t)k * lut_size;
            out[k].chi2cdf_inv = cbuf + (size_t)k * lut_size;
            if (orc_stats_build(d, (int)npairs, lut_size, &out[k].dissims_min, &out[k].dissims_max,
                                qbuf + (size_t)k * lut_size, cbuf + (size_t)k * lut_size) == 0)
                rc = -1;
            free(d);
        }
    free(Ps);
    free(Ss);
    return rc;
}

int orc_nlsar_training_insar(const float* a1, const float* a2, const float* phi, int height,
                             int width, const int* P, int nP, const int* S, int nS, orc_stats* out,
                             float* qbuf, float* cbuf)
{
    /* covmat_data(insar_data), data.cpp:60-89: host float cos/sin */
    const size_t n = (size_t)height * width;
    float* cov = (float*)malloc(sizeof(float) * 8 * n);
    for (size_t i = 0; i < n; i++) {
        const float re = a1[i] * a2[i] * cosf(phi[i]);
        const float im = a1[i] * a2[i] * sinf(phi[i]);
        cov[i] = a1[i] * a1[i];
        cov[n + i] = 0;
        cov[2 * n + i] = re;
        cov[3 * n + i] = im;
        cov[4 * n + i] = re;
        cov[5 * n + i] = -im;
        cov[6 * n + i] = a2[i] * a2[i];
        cov[7 * n + i] = 0;
    }
    int rc = orc_nlsar_training_covmat(cov, height, width, 2, P, nP, S, nS, out, qbuf, cbuf);
    free(cov);
    return rc;
}

/* ====================================================================================
 * NL-InSAR -- src/compute/nlinsar/
 * ==================================================================================== */

/* pixel_similarity, nlinsar/sim_measures.h:27-40 (device version: float pow/sqrt/asin/log) */
float orc_nlinsar_pixel_similarity(float am1, float as1, float dp1, float am2, float as2, float dp2)
{
    const float s = am1 * am1 + as1 * as1 + am2 * am2 + as2 * as2;
    const float A = s * s;
    const float B = 4.0f * (am1 * am1 * as1 * as1 + am2 * am2 * as2 * as2 +
                            2.0f * am1 * as1 * am2 * as2 * cosf(dp1 - dp2));
    const float C = am1 * as1 * am2 * as2;
    const float f_el = powf(C / B, 1.5f) * ((A + B) / A * sqrtf(B / (A - B)) - asinf(sqrtf(B / A)));
    if ((f_el == 0) || isnan(f_el)) return 0;
    return logf(f_el);
}

/* pixel_kullback_leibler, nlinsar/sim_measures.h:42-49 */
float orc_nlinsar_pixel_kl(float ref1, float dp1, float coh1, float ref2, float dp2, float coh2)
{
    const float cs = cosf(dp1 - dp2);
    const float a1 = (ref2 != 0.0f) ? ref1 / ref2 * (1.0f - coh1 * coh2 * cs) / (1.0f - coh2 * coh2) : 1.0f;
    const float a2 = (ref2 != 0.0f) ? ref2 / ref1 * (1.0f - coh1 * coh2 * cs) / (1.0f - coh1 * coh1) : 1.0f;
    return -4.0f / (float)M_PI * (a1 + a2 - 2.0f);
}

/* h_theo for patch sizes without a tabulated value (index (p - 1) / 2): the reference draws
 * it by a Monte-Carlo quantile seeded from std::random_device (insarsim_simu.cpp:62-82), so
 * the checker takes the value the filter under test used; 0 = not set. */
static float g_h_theo[11];

int orc_nlinsar_set_h_theo(int patch_size, float h_theo)
{
    if (patch_size < 1 || patch_size % 2 == 0 || patch_size > 21) return -1;
    g_h_theo[(patch_size - 1) / 2] = h_theo;
    return 0;
}

/* nlinsar.cpp:65-78 */
int orc_nlinsar_params(int patch_size, float* h_para, float* T_para)
{
    const int patch_area = patch_size * patch_size;
    float h_theo;
    if (patch_size == 3) h_theo = 0.58f;
    else if (patch_size == 7) h_theo = 0.244f;
    else if (patch_size >= 1 && patch_size % 2 == 1 && patch_size <= 21 && g_h_theo[(patch_size - 1) / 2] != 0.0f)
        h_theo = g_h_theo[(patch_size - 1) / 2];
    else return -1;
    *h_para = h_theo * patch_area;
    *T_para = (float)(2.0 / *h_para * M_PI / 4 * patch_area);
    return 0;
}

/* libstdc++ std::priority_queue<weight_el> with weight_el::operator< = "weight greater"
 * (smoothing.cpp:27-36): a binary heap whose top is the smallest weight.  push = append +
 * sift up; pop = move last to a hole sifted down from the root toward the preferred child
 * (second child unless the first one compares greater), then sift the value up.  These
 * are the sift rules of std::push_heap / std::pop_heap that decide ties. */
typedef struct { float weight; int idx; } wel;
static int wel_less(wel a, wel b) { return a.weight > b.weight; }

static void heap_push(wel* hp, int* n, wel v)
{
    int hole = (*n)++;
    int parent = (hole - 1) / 2;
    while (hole > 0 && wel_less(hp[parent], v)) {
        hp[hole] = hp[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    hp[hole] = v;
}

static void heap_pop(wel* hp, int* n)
{
    const int len = *n - 1; /* heap range after removing the top */
    if (len <= 0) { *n = len < 0 ? 0 : len; return; }
    wel value = hp[len];
    hp[len] = hp[0];
    /* __adjust_heap(first, 0, len, value) */
    int hole = 0, child = 0;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (wel_less(hp[child], hp[child - 1])) child--;
        hp[hole] = hp[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        hp[hole] = hp[child - 1];
        hole = child - 1;
    }
    int parent = (hole - 1) / 2;
    while (hole > 0 && wel_less(hp[parent], value)) {
        hp[hole] = hp[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    hp[hole] = value;
    *n = len;
}

/* search_window_smoothing, smoothing.cpp:78-129 */
void orc_search_window_smoothing(const float* ampl_master, float* weights, int h, int w,
                                 int width_overlap, int patch_size, int sw, int lmin)
{
    const int psh = (patch_size - 1) / 2, wsh = (sw - 1) / 2;
    const int overlap = wsh + psh;
    wel hp[1024];
    int n = 0;
    const double as = ampl_master[(size_t)(h + overlap) * width_overlap + (w + overlap)];
    for (int hh = 0; hh < sw; hh++)
        for (int ww = 0; ww < sw; ww++) {
            const float at = ampl_master[(size_t)(h + overlap + hh - wsh) * width_overlap +
                                         (w + overlap + ww - wsh)];
            if (at < 2.0 * as) {
                const int idx = hh * sw + ww;
                const float weight = weights[idx];
                if (n < lmin) {
                    wel e = {weight, idx};
                    heap_push(hp, &n, e);
                } else if (hp[0].weight < weight) {
                    heap_pop(hp, &n);
                    wel e = {weight, idx};
                    heap_push(hp, &n, e);
                }
            }
        }
    int idxs[1024];
    int m = 0;
    float wsum = 0.0f;
    while (n > 0) {
        wsum += hp[0].weight;
        idxs[m++] = hp[0].idx;
        heap_pop(hp, &n);
    }
    wsum /= (float)m;
    for (int i = 0; i < m; i++) weights[idxs[i]] = wsum;
}

typedef struct {
    int sw, p, lmin;
    float h, T;
    long long nsmooth;
    int want_frag;
    int alt;   /* 1: the equivalent forms below; 2: double precision (sensitivity check only) */
} nlinsar_ctx;

/* Algebraically equivalent forms of the similarity and KL terms (sim_measures.h:27-49):
 * cos(dp1 - dp2) = cos1 cos2 + sin1 sin2, pow(q, 1.5) = q sqrt(q), quotients as products with
 * a reciprocal, the KL's per-pixel factors 1 / (ref (1 - coh^2)).  Each differs from the
 * reference's expression by a few ulps (OpenCL 1.2 itself allows up to 4 ulp for cos, 16 for
 * pow, 2.5 for division), so where they move an output noticeably the reference's output is
 * ill-conditioned: its own builtins may move it as much. */
static float alt_similarity(float am1, float as1, float dp1, float am2, float as2, float dp2)
{
    const float cs = fmaf(cosf(dp1), cosf(dp2), sinf(dp1) * sinf(dp2));
    const float s = am1 * am1 + as1 * as1 + am2 * am2 + as2 * as2;
    const float A = s * s;
    const float B = 4.0f * (am1 * am1 * as1 * as1 + am2 * am2 * as2 * as2 + 2.0f * am1 * as1 * am2 * as2 * cs);
    const float C = am1 * as1 * am2 * as2;
    const float q = C * (1.0f / B);
    const float f_el = q * sqrtf(q) * ((A + B) * (1.0f / A) * sqrtf(B * (1.0f / (A - B))) -
                                       asinf(sqrtf(B * (1.0f / A))));
    if ((f_el == 0) || isnan(f_el)) return 0;
    return logf(f_el);
}

/* the same terms evaluated in double precision, rounded to float */
static float dbl_similarity(float am1_, float as1_, float dp1, float am2_, float as2_, float dp2)
{
    const double am1 = am1_, as1 = as1_, am2 = am2_, as2 = as2_;
    const double s = am1 * am1 + as1 * as1 + am2 * am2 + as2 * as2;
    const double A = s * s;
    const double B = 4.0 * (am1 * am1 * as1 * as1 + am2 * am2 * as2 * as2 +
                            2.0 * am1 * as1 * am2 * as2 * cos((double)dp1 - (double)dp2));
    const double C = am1 * as1 * am2 * as2;
    const double f_el = pow(C / B, 1.5) * ((A + B) / A * sqrt(B / (A - B)) - asin(sqrt(B / A)));
    if ((f_el == 0) || isnan(f_el)) return 0;
    return (float)log(f_el);
}

static float dbl_kl(float ref1, float dp1, float coh1, float ref2, float dp2, float coh2)
{
    const double cs = cos((double)dp1 - (double)dp2);
    const double a1 = (ref2 != 0.0f) ? (double)ref1 / ref2 * (1.0 - (double)coh1 * coh2 * cs) / (1.0 - (double)coh2 * coh2) : 1.0;
    const double a2 = (ref2 != 0.0f) ? (double)ref2 / ref1 * (1.0 - (double)coh1 * coh2 * cs) / (1.0 - (double)coh1 * coh1) : 1.0;
    return (float)(-4.0 / M_PI * (a1 + a2 - 2.0));
}

static float alt_kl(float ref1, float dp1, float coh1, float ref2, float dp2, float coh2)
{
    const float r = 1.0f - fmaf(coh1 * cosf(dp1), coh2 * cosf(dp2), (coh1 * sinf(dp1)) * (coh2 * sinf(dp2)));
    const float u1 = 1.0f / (ref1 * (1.0f - coh1 * coh1)), u2 = 1.0f / (ref2 * (1.0f - coh2 * coh2));
    const float a1 = ref1 * r * u2, a2 = ref2 * r * u1;
    return ref2 != 0.0f ? -4.0f / (float)M_PI * (a1 + a2 - 2.0f) : -0.0f;
}

/* Where a last-bit difference of the weights can change the lmin smoothing's kept set
 * (smoothing.cpp:78-129): the lmin-th and (lmin+1)-th largest candidate weights tied or
 * within a relative 1e-4 (2), or a tiny weight (below 1e-30, a denormal or nearly so, with
 * few significant bits) at the kept threshold (4).  Candidates as in the smoothing. */
static int smoothing_fragility(const float* ampl_master, const float* weights, int h, int w,
                               int width_overlap, int patch_size, int sw, int lmin)
{
    const int psh = (patch_size - 1) / 2, wsh = (sw - 1) / 2;
    const int overlap = wsh + psh;
    float c[1024];
    int n = 0, tiny = 0;
    const double as = ampl_master[(size_t)(h + overlap) * width_overlap + (w + overlap)];
    for (int hh = 0; hh < sw; hh++)
        for (int ww = 0; ww < sw; ww++) {
            const float at = ampl_master[(size_t)(h + overlap + hh - wsh) * width_overlap +
                                         (w + overlap + ww - wsh)];
            if (at < 2.0 * as) {
                c[n++] = weights[hh * sw + ww];
                if (c[n - 1] > 0 && c[n - 1] < 1e-30f) tiny = 1;
            }
        }
    if (n <= lmin) return 0;
    qsort(c, (size_t)n, sizeof(float), cmp_float);   /* ascending */
    const float T = c[n - lmin], U = c[n - lmin - 1];
    int flag = 0;
    if ((T == U && T > 0) || (T != U && T - U <= 1e-4f * T)) flag |= 2;
    if (T < 1e-30f && tiny) flag |= 4;
    return flag;
}

static int is_valid(float x) { return !(isnan(x) || isinf(x) || x == 0); }

/* nlinsar_sub_image, nlinsar_sub_image.cpp:22-262, on one tile of 6 planes
 * (a1, a2, phi, ref, phase, coh); writes the updated estimate planes */
static int nlinsar_tile_impl(const float* tin, float* tout, int th, int tw, nlinsar_ctx* cx);

/* the tile, plus (want_frag) the sensitivity plane (5th): how far the equivalent forms, or
 * the same terms in double precision, move an output (relative for reflectivity and
 * coherence, radians for the phase), and flag 8 where that exceeds 1e-4 */
static int nlinsar_tile(const float* tin, float* tout, int th, int tw, void* vctx)
{
    nlinsar_ctx* cx = (nlinsar_ctx*)vctx;
    int rc = nlinsar_tile_impl(tin, tout, th, tw, cx);
    if (rc || !cx->want_frag) return rc;
    const size_t n = (size_t)th * tw;
    float* alt = (float*)calloc(3 * n, sizeof(float));
    for (int mode = 1; mode <= 2 && rc == 0; mode++) {
    nlinsar_ctx ac = *cx;
    ac.want_frag = 0;
    ac.alt = mode;
    rc = nlinsar_tile_impl(tin, alt, th, tw, &ac);
    const int overlap = (cx->sw - 1) / 2 + (cx->p - 1) / 2;
    for (int h = overlap; rc == 0 && h < th - overlap; h++)
        for (int w = overlap; w < tw - overlap; w++) {
            const size_t i = (size_t)h * tw + w;
            const double r0 = tout[i], r1 = alt[i], c0 = tout[2 * n + i], c1 = alt[2 * n + i];
            double dph = fabs((double)tout[n + i] - (double)alt[n + i]);
            if (dph > M_PI) dph = 2 * M_PI - dph;
            const double er = fabs(r1 - r0) / fabs(r0), ec = fabs(c1 - c0) / fabs(c0);
            double sens = er > ec ? er : ec;
            if (dph > sens) sens = dph;
            if (!(sens <= 1e-4)) tout[3 * n + i] = (float)((int)tout[3 * n + i] | 8);
            if (!(sens <= tout[4 * n + i])) tout[4 * n + i] = isnan(sens) ? INFINITY : (float)sens;
        }
    }
    free(alt);
    return rc;
}

static int nlinsar_tile_impl(const float* tin, float* tout, int th, int tw, nlinsar_ctx* cx)
{
    const int sw = cx->sw, p = cx->p;
    const int psh = (p - 1) / 2, wsh = (sw - 1) / 2, overlap = wsh + psh;
    const size_t n = (size_t)th * tw;
    const float *a1 = tin, *a2 = tin + n, *dp = tin + 2 * n;
    const float *ref = tin + 3 * n, *pha = tin + 4 * n, *coh = tin + 5 * n;
    const int h_sim = th - sw + 1, w_sim = tw - sw + 1;
    const int h_ori = th - 2 * overlap, w_ori = tw - 2 * overlap;
    if (h_ori <= 0 || w_ori <= 0) return -2;
    const int nfull = sw * sw;
    const size_t splane = (size_t)h_sim * w_sim;
    const size_t oplane = (size_t)h_ori * w_ori;
    const int border_max = wsh;
    float* sims = (float*)malloc(sizeof(float) * nfull * splane);
    float* kls = (float*)malloc(sizeof(float) * nfull * splane);
    /* 1st pass, precompute_similarities_1st_pass.cl.in:22-75: interior sim positions get
     * offsets hh >= wsh only */
    for (int h = 0; h < h_sim; h++)
        for (int w = 0; w < w_sim; w++) {
            const int hstart = (h >= border_max && w >= border_max && w < w_sim - border_max) ? wsh : 0;
            for (int hh = hstart; hh < sw; hh++)
                for (int ww = 0; ww < sw; ww++) {
                    const size_t ci = (size_t)(h + wsh) * tw + w + wsh;
                    const size_t wi = (size_t)(h + hh) * tw + w + ww;
                    const size_t o = (size_t)(hh * sw + ww) * splane + (size_t)h * w_sim + w;
                    sims[o] = cx->alt == 1 ? alt_similarity(a1[ci], a2[ci], dp[ci], a1[wi], a2[wi], dp[wi])
                            : cx->alt == 2 ? dbl_similarity(a1[ci], a2[ci], dp[ci], a1[wi], a2[wi], dp[wi])
                                           : orc_nlinsar_pixel_similarity(a1[ci], a2[ci], dp[ci], a1[wi], a2[wi], dp[wi]);
                    kls[o] = cx->alt == 1 ? alt_kl(ref[ci], pha[ci], coh[ci], ref[wi], pha[wi], coh[wi])
                           : cx->alt == 2 ? dbl_kl(ref[ci], pha[ci], coh[ci], ref[wi], pha[wi], coh[wi])
                                          : orc_nlinsar_pixel_kl(ref[ci], pha[ci], coh[ci], ref[wi], pha[wi], coh[wi]);
                }
        }
    /* 2nd pass, precompute_similarities_2nd_pass.cl:19-50 (launched only when the tile is
     * large enough, precompute_similarities_2nd_pass.cpp) */
    if (h_sim > border_max && w_sim > 2 * border_max)
        for (int h = border_max; h < h_sim; h++)
            for (int w = border_max; w < w_sim - border_max; w++)
                for (int hh = 0; hh < wsh; hh++)
                    for (int ww = 0; ww < sw; ww++) {
                        const size_t o = (size_t)(hh * sw + ww) * splane + (size_t)h * w_sim + w;
                        const size_t s = (size_t)((sw - 1 - hh) * sw + (sw - 1 - ww)) * splane +
                                         (size_t)(h - (wsh - hh)) * w_sim + (w - (wsh - ww));
                        sims[o] = sims[s];
                        kls[o] = kls[s];
                    }
    /* patch sums, precompute_patch_similarities.cl:19-66: row sums (along w) first, then
     * the rows accumulated top to bottom */
    float* psim = (float*)malloc(sizeof(float) * nfull * oplane);
    float* pkl = (float*)malloc(sizeof(float) * nfull * oplane);
    for (int which = 0; which < 2; which++) {
        const float* in = which == 0 ? sims : kls;
        float* out = which == 0 ? psim : pkl;
        for (int z = 0; z < nfull; z++)
            for (int x = 0; x < h_ori; x++)
                for (int y = 0; y < w_ori; y++) {
                    float sum = 0;
                    for (int kx = 0; kx < p; kx++) {
                        float rs = 0;
                        for (int ky = 0; ky < p; ky++) rs += in[(size_t)z * splane + (size_t)(x + kx) * w_sim + y + ky];
                        if (kx == 0) sum = rs; else sum += rs;
                    }
                    out[(size_t)z * oplane + (size_t)x * w_ori + y] = sum;
                }
    }
    /* weights, nlinsar compute_weights.cl:19-30; centre weight 0 (nlinsar_sub_image.cpp:189-197) */
    float* wts = (float*)malloc(sizeof(float) * nfull * oplane);
    for (size_t i = 0; i < (size_t)nfull * oplane; i++) {
        float weight = cx->alt == 1 ? expf(psim[i] * (1.0f / cx->h) + pkl[i] * (1.0f / cx->T))
                                    : expf(psim[i] / cx->h + pkl[i] / cx->T);
        if (isnan(weight) || isinf(weight)) weight = 0.0f;
        wts[i] = weight;
    }
    for (size_t i = 0; i < oplane; i++) wts[(size_t)(sw * wsh + wsh) * oplane + i] = 0.0f;
    /* compute_number_of_looks.cl:19-42 */
    float* nols = (float*)malloc(sizeof(float) * oplane);
    for (size_t i = 0; i < oplane; i++) {
        float ws = 0.0f, ws2 = 0.0f;
        for (int k = 0; k < nfull; k++) {
            const float wv = wts[(size_t)k * oplane + i];
            ws += wv;
            ACC(ws2, wv, wv);
        }
        nols[i] = ws * ws / ws2;
    }
    /* transpose to [pixel][441] (transpose.cl) */
    float* wt = (float*)malloc(sizeof(float) * nfull * oplane);
    for (size_t i = 0; i < oplane; i++)
        for (int k = 0; k < nfull; k++) wt[i * nfull + k] = wts[(size_t)k * oplane + i];
    /* fragility flags (optional 4th output plane): ENL within a relative 1e-4 of lmin (1),
     * the smoothing's kept set (2, 4; smoothing_fragility), weights near underflow (16: the
     * float ENL, finite, more than 1e-3 from the ENL of the same weights in double, or on the
     * other side of lmin, as happens when the squares of tiny weights lose their bits; or
     * every weight subnormal) */
    if (cx->want_frag)
        for (int h = 0; h < h_ori; h++)
            for (int w = 0; w < w_ori; w++) {
                const float e = nols[(size_t)h * w_ori + w];
                int flag = fabsf(e - (float)cx->lmin) <= 1e-4f * (float)cx->lmin ? 1 : 0;
                const float* wr = wt + ((size_t)h * w_ori + w) * nfull;
                double dws = 0.0, dws2 = 0.0;
                float wmax = 0.0f;
                for (int k = 0; k < nfull; k++) {
                    dws += wr[k];
                    dws2 += (double)wr[k] * wr[k];
                    wmax = fmaxf(wmax, wr[k]);
                }
                const double denl = dws * dws / dws2;
                if ((isfinite(e) && !(fabs(e - denl) <= 1e-3 * denl)) || ((e < cx->lmin) != (denl < cx->lmin)) ||
                    (wmax > 0.0f && wmax < 1e-36f))
                    flag |= 16;
                if (e < cx->lmin)
                    flag |= smoothing_fragility(a1, wt + ((size_t)h * w_ori + w) * nfull, h, w,
                                                w_ori + 2 * overlap, p, sw, cx->lmin);
                tout[3 * n + (size_t)(h + overlap) * tw + (w + overlap)] = (float)flag;
            }
    /* smoothing::run, smoothing.cpp:39-76 */
    for (int h = 0; h < h_ori; h++)
        for (int w = 0; w < w_ori; w++)
            if (nols[(size_t)h * w_ori + w] < cx->lmin) {
                orc_search_window_smoothing(a1, wt + ((size_t)h * w_ori + w) * nfull, h, w,
                                            w_ori + 2 * overlap, p, sw, cx->lmin);
                cx->nsmooth++;
            }
    /* precompute_filter_values.cl:19-39 and compute_insar.cl:48-119 */
    const int h_sws = th - p + 1, w_sws = tw - p + 1;
    float* fa = (float*)malloc(sizeof(float) * (size_t)h_sws * w_sws);
    float* fxr = (float*)malloc(sizeof(float) * (size_t)h_sws * w_sws);
    float* fxi = (float*)malloc(sizeof(float) * (size_t)h_sws * w_sws);
    for (int x = 0; x < h_sws; x++)
        for (int y = 0; y < w_sws; y++) {
            const size_t i = (size_t)(x + psh) * tw + (y + psh);
            const size_t o = (size_t)x * w_sws + y;
            fa[o] = 0.5f * (a1[i] * a1[i] + a2[i] * a2[i]);
            fxr[o] = a1[i] * a2[i] * cosf(-dp[i]);
            fxi[o] = a1[i] * a2[i] * sinf(-dp[i]);
        }
    memcpy(tout, ref, n * sizeof(float));
    memcpy(tout + n, pha, n * sizeof(float));
    memcpy(tout + 2 * n, coh, n * sizeof(float));
    for (int x0 = 0; x0 < h_ori; x0++)
        for (int y0 = 0; y0 < w_ori; y0++) {
            float N = 0.0f, at = 0.0f, xr = 0.0f, xi = 0.0f;
            const float* wrow = wt + ((size_t)x0 * w_ori + y0) * nfull;
            for (int x = 0; x < sw; x++)
                for (int y = 0; y < sw; y++) {
                    const float weight = wrow[x * sw + y];
                    const size_t o = (size_t)(x0 + x) * w_sws + y0 + y;
                    N += weight;
                    ACC(at, weight, fa[o]);
                    ACC(xr, weight, fxr[o]);
                    ACC(xi, weight, fxi[o]);
                }
            const float reflectivity = at / N;
            const float iphase = atan2f(xi, xr);
            const float coherence = sqrtf(xr * xr + xi * xi) / at;
            const size_t idx = (size_t)(x0 + wsh + psh) * tw + (y0 + wsh + psh);
            if (is_valid(reflectivity)) tout[idx] = reflectivity;
            if (is_valid(iphase)) {
                float ph = iphase;
                if (ph > (float)M_PI) ph -= 2.0f * (float)M_PI;
                if (ph < -(float)M_PI) ph += 2.0f * (float)M_PI;
                tout[n + idx] = ph;
            }
            if (is_valid(coherence)) {
                float c = coherence;
                if (c > 1.0f) c = 1.0f;
                if (c < 0.0f) c = 0.0f;
                tout[2 * n + idx] = c;
            }
        }
    free(fxi); free(fxr); free(fa); free(wt); free(nols); free(wts);
    free(pkl); free(psim); free(kls); free(sims);
    return 0;
}

int orc_nlinsar(const float* a1, const float* a2, const float* phi, float* ref_filt,
                float* phase_filt, float* coh_filt, int height, int width, int sw, int p,
                int niter, int lmin, int tile, long long* n_smoothed)
{
    return orc_nlinsar_frag(a1, a2, phi, ref_filt, phase_filt, coh_filt, NULL, NULL, height, width,
                            sw, p, niter, lmin, tile, n_smoothed);
}

int orc_nlinsar_frag(const float* a1, const float* a2, const float* phi, float* ref_filt,
                     float* phase_filt, float* coh_filt, float* frag, float* sens, int height,
                     int width, int sw, int p, int niter, int lmin, int tile,
                     long long* n_smoothed)
{
    if (height <= 0 || width <= 0 || sw < 3 || sw % 2 == 0 || p < 1 || p % 2 == 0 || niter < 0)
        return -1;
    if (sw * sw > 1024) return -1;
    nlinsar_ctx cx = {sw, p, lmin, 0, 0, 0, frag != NULL || sens != NULL, 0};
    const int nplanes = cx.want_frag ? 5 : 3;
    if (orc_nlinsar_params(p, &cx.h, &cx.T)) return -4;
    if (tile <= 0) tile = 80; /* nlinsar.cpp:106 */
    const int overlap = (p - 1) / 2 + (sw - 1) / 2;
    if (tile <= 2 * overlap) return -1;
    const size_t n = (size_t)height * width;
    /* total_image (6 planes) and the per-iteration deep copy, nlinsar.cpp:59-133 */
    float* img = (float*)malloc(sizeof(float) * 6 * n);
    float* tmp = (float*)calloc((size_t)5 * n, sizeof(float));
    memcpy(img, a1, n * sizeof(float));
    memcpy(img + n, a2, n * sizeof(float));
    memcpy(img + 2 * n, phi, n * sizeof(float));
    memcpy(img + 3 * n, ref_filt, n * sizeof(float));
    memcpy(img + 4 * n, phase_filt, n * sizeof(float));
    memcpy(img + 5 * n, coh_filt, n * sizeof(float));
    const int step = tile - 2 * overlap;
    int rc = 0;
    for (int it = 0; it < niter && rc == 0; it++) {
        memcpy(tmp, img + 3 * n, 3 * n * sizeof(float));
        /* tiles are independent within an iteration: run them in parallel, then write
         * back in the serial spawn order */
        int nth = 0, ntw = 0;
        for (int h_low = -overlap; h_low < height; h_low += step) nth++;
        for (int w_low = -overlap; w_low < width; w_low += step) ntw++;
        const int ntiles = nth * ntw;
        float** outs = (float**)calloc((size_t)ntiles, sizeof(float*));
        long long nsm = 0;
        int trc = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : nsm)
        for (int t = 0; t < ntiles; t++) {
            const int h_low = -overlap + (t / ntw) * step;
            const int w_low = -overlap + (t % ntw) * step;
            float* tin = (float*)malloc(sizeof(float) * 6 * (size_t)tile * tile);
            outs[t] = (float*)calloc((size_t)nplanes * tile * tile, sizeof(float));
            get_sub_image(img, height, width, 6, h_low, w_low, tile, tile, tin);
            nlinsar_ctx lc = cx;
            lc.nsmooth = 0;
            if (nlinsar_tile(tin, outs[t], tile, tile, &lc)) trc = -2;
            nsm += lc.nsmooth;
            free(tin);
        }
        cx.nsmooth += nsm;
        rc = trc;
        for (int t = 0; t < ntiles; t++) {
            if (rc == 0) {
                const int h_low = -overlap + (t / ntw) * step;
                const int w_low = -overlap + (t % ntw) * step;
                write_interior(tmp, height, width, nplanes, outs[t], h_low, w_low, tile, tile, overlap);
            }
            free(outs[t]);
        }
        free(outs);
        memcpy(img + 3 * n, tmp, 3 * n * sizeof(float));
    }
    if (rc == 0) {
        memcpy(ref_filt, img + 3 * n, n * sizeof(float));
        memcpy(phase_filt, img + 4 * n, n * sizeof(float));
        memcpy(coh_filt, img + 5 * n, n * sizeof(float));
        if (frag) memcpy(frag, tmp + 3 * n, n * sizeof(float));
        if (sens) memcpy(sens, tmp + 4 * n, n * sizeof(float));
    }
    if (n_smoothed) *n_smoothed = cx.nsmooth;
    free(tmp);
    free(img);
    return rc;
}

/* ====================================================================================
 * boxcar -- boxcar.cpp:35-92 (512 x 512 tiles, overlap (w - 1) / 2), boxcar_kernel.cl:19-69
 * ==================================================================================== */
typedef struct {
    int w;
} boxcar_ctx;

/* The kernel on one tile: output (ox, oy) = (in + r) from the window rows ox - r .. ox + r
 * in row-major order; inputs outside the tile read as 0 (the kernel's local-memory fill);
 * outputs of the first r rows / columns are not written (they stay 0 here and are never
 * copied back). */
static int boxcar_tile(const float* tin, float* tout, int th, int tw, void* vctx)
{
    const int w = ((boxcar_ctx*)vctx)->w;
    const int r = (w - 1) / 2;
    const size_t n = (size_t)th * tw;
    const float* A1 = tin;
    const float* A2 = tin + n;
    const float* PH = tin + 2 * n;
    memset(tout, 0, sizeof(float) * 3 * n);
    const int N = w * w;
#pragma omp parallel for schedule(static)
    for (int ox = r; ox < th; ox++)
        for (int oy = r; oy < tw; oy++) {
            float a = 0, xr = 0, xi = 0;
            for (int kx = 0; kx < w; kx++)
                for (int ky = 0; ky < w; ky++) {
                    const int ix = ox - r + kx, iy = oy - r + ky;
                    const int in = ix < th && iy < tw;
                    const float m = in ? A1[(size_t)ix * tw + iy] : 0.0f;
                    const float sl = in ? A2[(size_t)ix * tw + iy] : 0.0f;
                    const float p = in ? PH[(size_t)ix * tw + iy] : 0.0f;
                    /* pow(x, 2) as x * x; "+=" of a product contracted to an FMA */
                    a += (m * m + sl * sl) / 2;
                    const float aa = m * sl;
                    xr = fmaf(aa, cosf(p), xr);
                    xi = fmaf(aa, sinf(p), xi);
                }
            tout[(size_t)ox * tw + oy] = a / (float)N;
            tout[n + (size_t)ox * tw + oy] = atan2f(xi, xr);
            tout[2 * n + (size_t)ox * tw + oy] = sqrtf(xr * xr + xi * xi) / a;
        }
    return 0;
}

int orc_boxcar(const float* a1, const float* a2, const float* phi, float* ref_filt,
               float* phase_filt, float* coh_filt, int height, int width, int window_width,
               int tile)
{
    if (height <= 0 || width <= 0 || window_width < 1 || window_width % 2 == 0) return -1;
    const int r = (window_width - 1) / 2;
    if (tile <= 0) tile = 512;
    if (tile <= 2 * r) return -1;
    const size_t n = (size_t)height * width;
    float* in = (float*)malloc(sizeof(float) * 3 * n);
    float* out = (float*)calloc(3 * n, sizeof(float));
    memcpy(in, a1, n * sizeof(float));
    memcpy(in + n, a2, n * sizeof(float));
    memcpy(in + 2 * n, phi, n * sizeof(float));
    boxcar_ctx cx = {window_width};
    const int rc = map_tiles(in, out, height, width, 3, 3, tile, tile, r, boxcar_tile, &cx);
    if (rc == 0) {
        memcpy(ref_filt, out, n * sizeof(float));
        memcpy(phase_filt, out + n, n * sizeof(float));
        memcpy(coh_filt, out + 2 * n, n * sizeof(float));
    }
    free(in);
    free(out);
    return rc;
}

/* ====================================================================================
 * Goldstein -- goldstein.cpp:39-106, goldstein_filter_sub_image.cpp:73-187,
 * tile_size.cpp:24-33 (round_up), cl_kernels/{raw_interferogram, patches_unpack,
 * weighted_multiply, patches_pack, slc2real}.cl, cl_wrappers/patches_pack.cpp:48-72
 * ==================================================================================== */

/* patches_unpack.cl:19-42 */
void orc_gs_unpack(const float* packed_re, const float* packed_im, float* unpacked_re,
                   float* unpacked_im, int height_unpacked, int width_unpacked, int patch_size,
                   int overlap)
{
    const int m = patch_size - 2 * overlap;
    const int width_packed = (width_unpacked / patch_size) * m;
    const int height_packed = (height_unpacked / patch_size) * m;
    for (int ty = 0; ty < height_unpacked; ty++)
        for (int tx = 0; tx < width_unpacked; tx++) {
            const int px = tx / patch_size, py = ty / patch_size;
            const int rx = tx % patch_size, ry = ty % patch_size;
            const int ix = imin(width_packed - 1, imax(0, px * m + (rx - overlap)));
            const int iy = imin(height_packed - 1, imax(0, py * m + (ry - overlap)));
            unpacked_re[(size_t)ty * width_unpacked + tx] = packed_re[(size_t)iy * width_packed + ix];
            unpacked_im[(size_t)ty * width_unpacked + tx] = packed_im[(size_t)iy * width_packed + ix];
        }
}

/* patches_pack.cl:19-66 in the four passes of patches_pack.cpp:48-72, offsets (x, y) =
 * (0, 0), (0, P), (P, 0), (P, P); "+=" into the packed buffer as it is.  Patches beyond
 * the unpacked extent (the reference's launch covers one when a direction has a single
 * patch, and reads out of bounds) contribute nothing. */
void orc_gs_pack(const float* unpacked_re, const float* unpacked_im, float* packed_re,
                 float* packed_im, int height_unpacked, int width_unpacked, int patch_size,
                 int overlap)
{
    const int P = patch_size, m = patch_size - 2 * overlap;
    const int width_packed = width_unpacked / P * m;
    const int height_packed = height_unpacked / P * m;
    const int npx = width_unpacked / P, npy = height_unpacked / P;
    const int offs[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
    for (int o = 0; o < 4; o++)
        for (int py = offs[o][1]; py < npy; py += 2)
            for (int px = offs[o][0]; px < npx; px += 2)
                for (int ry = 0; ry < P; ry++)
                    for (int rx = 0; rx < P; rx++) {
                        const int in_x = px * P + rx, in_y = py * P + ry;
                        const int out_x = px * m + (rx - overlap), out_y = py * m + (ry - overlap);
                        float sf = 1.0f;
                        if (rx < 2 * overlap || rx >= P - 2 * overlap) sf *= 0.5f;
                        if (ry < 2 * overlap || ry >= P - 2 * overlap) sf *= 0.5f;
                        if (out_x < overlap || out_x >= width_packed - overlap) sf *= 2.0f;
                        if (out_y < overlap || out_y >= height_packed - overlap) sf *= 2.0f;
                        const float vr = sf * unpacked_re[(size_t)in_y * width_unpacked + in_x];
                        const float vi = sf * unpacked_im[(size_t)in_y * width_unpacked + in_x];
                        if (out_x >= 0 && out_x < width_packed && out_y >= 0 && out_y < height_packed) {
                            packed_re[(size_t)out_y * width_packed + out_x] += vr;
                            packed_im[(size_t)out_y * width_packed + out_x] += vi;
                        }
                    }
}

/* 1-D DFT of n points in double, in place (sign -1 forward, +1 inverse, unscaled) */
static void dft1(double* re, double* im, int n, int sign, const double* c, const double* s)
{
    double tr[256], ti[256];
    for (int k = 0; k < n; k++) {
        double ar = 0, ai = 0;
        for (int j = 0; j < n; j++) {
            const int t = (int)(((long long)j * k) % n);
            ar += re[j] * c[t] - sign * im[j] * s[t];
            ai += im[j] * c[t] + sign * re[j] * s[t];
        }
        tr[k] = ar;
        ti[k] = ai;
    }
    memcpy(re, tr, sizeof(double) * n);
    memcpy(im, ti, sizeof(double) * n);
}

/* 2-D transform of every P x P patch of a planar complex image (goldstein_patch_ft.cpp:24-48
 * with clFFT's default scales: 1 forward, 1/P^2 backward), exact in double, rounded to
 * float */
void orc_gs_patch_ft(float* re, float* im, int height, int width, int P, int inverse)
{
    double c[256], s[256];
    for (int k = 0; k < P; k++) {
        c[k] = cos(2.0 * M_PI * k / P);
        s[k] = sin(2.0 * M_PI * k / P);
    }
    const int sign = inverse ? 1 : -1;
    const double scale = inverse ? 1.0 / ((double)P * P) : 1.0;
    const int npy = height / P, npx = width / P;
#pragma omp parallel for schedule(static)
    for (int q = 0; q < npy * npx; q++) {
        const int py = q / npx, px = q % npx;
        double br[256 * 256 / 4], bi[256 * 256 / 4];   /* P <= 128 */
        double vr[256], vi[256];
        for (int y = 0; y < P; y++)
            for (int x = 0; x < P; x++) {
                const size_t g = (size_t)(py * P + y) * width + px * P + x;
                br[y * P + x] = re[g];
                bi[y * P + x] = im[g];
            }
        for (int y = 0; y < P; y++) dft1(br + y * P, bi + y * P, P, sign, c, s);
        for (int x = 0; x < P; x++) {
            for (int y = 0; y < P; y++) {
                vr[y] = br[y * P + x];
                vi[y] = bi[y * P + x];
            }
            dft1(vr, vi, P, sign, c, s);
            for (int y = 0; y < P; y++) {
                br[y * P + x] = vr[y];
                bi[y * P + x] = vi[y];
            }
        }
        for (int y = 0; y < P; y++)
            for (int x = 0; x < P; x++) {
                const size_t g = (size_t)(py * P + y) * width + px * P + x;
                re[g] = (float)(br[y * P + x] * scale);
                im[g] = (float)(bi[y * P + x] * scale);
            }
    }
}

typedef struct {
    int P, ov;
    float alpha;
} gs_ctx;

/* goldstein::filter_sub_image on one T x T tile; outputs |z| and arg z */
static int goldstein_tile(const float* tin, float* tout, int T, int Tw, void* vctx)
{
    const gs_ctx* cx = (const gs_ctx*)vctx;
    const int P = cx->P, ov = cx->ov, m = P - 2 * ov;
    const size_t n = (size_t)T * Tw;
    const int npy = T / m, npx = Tw / m;
    const int hu = P * npy, wu = P * npx;
    float* zr = (float*)calloc(n, sizeof(float));
    float* zi = (float*)calloc(n, sizeof(float));
    float* ur = (float*)malloc(sizeof(float) * (size_t)hu * wu);
    float* ui = (float*)malloc(sizeof(float) * (size_t)hu * wu);
    /* raw_interferogram.cl:19-35 */
    for (size_t i = 0; i < n; i++) {
        const float a = 0.5f * (tin[i] + tin[n + i]);
        zr[i] = a * cosf(tin[2 * n + i]);
        zi[i] = a * sinf(tin[2 * n + i]);
    }
    orc_gs_unpack(zr, zi, ur, ui, hu, wu, P, ov);
    orc_gs_patch_ft(ur, ui, hu, wu, P, 0);
    /* weighted_multiply.cl:19-36: |Z|^alpha Z */
    for (size_t i = 0; i < (size_t)hu * wu; i++) {
        const float psd = sqrtf(ur[i] * ur[i] + ui[i] * ui[i]);
        ur[i] = powf(psd, cx->alpha) * ur[i];
        ui[i] = powf(psd, cx->alpha) * ui[i];
    }
    orc_gs_patch_ft(ur, ui, hu, wu, P, 1);
    /* the packed buffer still holds the raw interferogram */
    orc_gs_pack(ur, ui, zr, zi, hu, wu, P, ov);
    /* slc2real.cl:19-35 */
    for (size_t i = 0; i < n; i++) {
        tout[i] = fabsf(sqrtf(zr[i] * zr[i] + zi[i] * zi[i]));
        tout[n + i] = atan2f(zi[i], zr[i]);
    }
    free(zr);
    free(zi);
    free(ur);
    free(ui);
    return 0;
}

int orc_goldstein(const float* a1, const float* a2, const float* phi, float* ref_filt,
                  float* phase_filt, int height, int width, int patch_size, int overlap,
                  float alpha)
{
    if (height <= 0 || width <= 0 || patch_size <= 0 || patch_size > 128 || overlap < 0 ||
        patch_size - 2 * overlap <= 0)
        return -1;
    const int m = patch_size - 2 * overlap;
    /* goldstein.cpp:64-67 with tile_size.cpp:30-33 (round_up adds a full multiple); the
     * device-memory cap of tile_size.cpp:35-77 is taken as inactive */
    const int big = imax(height + 2 * overlap, width + 2 * overlap);
    const int T = big - big % m + m;
    const size_t n = (size_t)height * width;
    float* in = (float*)malloc(sizeof(float) * 3 * n);
    float* out = (float*)calloc(2 * n, sizeof(float));
    memcpy(in, a1, n * sizeof(float));
    memcpy(in + n, a2, n * sizeof(float));
    memcpy(in + 2 * n, phi, n * sizeof(float));
    memcpy(out, ref_filt, n * sizeof(float));
    memcpy(out + n, phase_filt, n * sizeof(float));
    gs_ctx cx = {patch_size, overlap, alpha};
    const int rc = map_tiles(in, out, height, width, 3, 2, T, T, overlap, goldstein_tile, &cx);
    if (rc == 0) {
        memcpy(ref_filt, out, n * sizeof(float));
        memcpy(phase_filt, out + n, n * sizeof(float));
    }
    free(in);
    free(out);
    return rc;
}
