/*
 * brute.c -- per-pixel brute-force NL-SAR and NL-InSAR: the independent check of the
 * oracle's composition.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ to pin oracle.c.  The product never links or
 * calls this file.
 *
 * oracle.c restates the reference stage by stage: tiles, half-window similarity planes
 * with their launch quadrants, patch-sum planes, the copy_symm mirror, full 441-plane
 * weights, the tile write-back.  Here every output pixel is computed directly from what
 * those stages amount to, with none of their buffers or index bookkeeping:
 *  - the image is read with clamp-to-edge at every position (sub_images.cpp:37-40);
 *  - the last row / column take the value of the position where the reference's clamped
 *    write-back last writes them (sub_images.cpp:47-69, tile_iterator.cpp:48-55);
 *  - NL-SAR: the dissimilarity of an offset d is the p x p patch sum of delta over the
 *    pixel pairs (q, q + d), each pair evaluated as delta(later raster pixel, earlier one)
 *    (compute_pixel_similarities_2x2.cpp:51-79 with copy_symm_weights.cl:53-74); the
 *    statistics, alpha, ENL, the selection and the weighted mean follow their kernels'
 *    formulas over the full window in plane order;
 *  - NL-InSAR: the same patch sums of the likelihood similarity and KL term, with the
 *    argument order the reference's first / second pass gives each pair, which depends on
 *    where the pair's first pixel lies in its 80 x 80 tile
 *    (precompute_similarities_1st_pass.cl.in:36-42, precompute_similarities_2nd_pass.cl:19-50).
 * It calls only the oracle's single-pair formulas, each pinned by closed forms of its own
 * (orc_pixel_similarity_2x2/3x3, orc_gen_gauss, orc_enl_nobias, orc_nlinsar_pixel_similarity,
 * orc_nlinsar_pixel_kl) and the lmin smoothing of one window (orc_search_window_smoothing,
 * pinned against std::priority_queue).  Summation orders follow the reference, so a correct
 * oracle agrees with this file bit for bit.
 *
 * `flip` (tests only) reverses the argument order of every pair, so a test can show that
 * the bit-exact comparison detects an orientation mistake.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* later raster position of (y1, x1) and (y2, x2)? */
static int is_later(int y1, int x1, int y2, int x2) { return y1 > y2 || (y1 == y2 && x1 > x2); }

/* ------------------------------------------------------------------------------------
 * NL-SAR
 * ------------------------------------------------------------------------------------ */

/* the 2*dim*dim planes of one pixel, clamp-to-edge */
static void cov_at(const float* cov, int H, int W, int np, int y, int x, float* out)
{
    const size_t i = (size_t)clampi(y, 0, H - 1) * W + clampi(x, 0, W - 1);
    for (int k = 0; k < np; k++) out[k] = cov[(size_t)k * H * W + i];
}

/* similarity operands of the rescaled, scale-averaged covariance at (y, x) for scale s:
 * the s x s Gaussian taps centred on (y, x), rows outer, fmaf accumulation
 * (covmat_spatial_avg.cl:19-63; off-diagonals rescaled first, covmat_rescale.cl:19-41) */
static void avg_at(const float* cov, int H, int W, int dim, int s, const float* gauss,
                   float gamma, int y, int x, float* out)
{
    const int np = 2 * dim * dim;
    /* operand planes: dim 2 {00, 01re, 01im, 11}; dim 3 {00, 11, 22, 01, 02, 12 (re, im)} */
    static const int pl2[4] = {0, 2, 3, 6};
    static const int pl3[9] = {0, 8, 16, 2, 3, 4, 5, 10, 11};
    const int* pl = dim == 2 ? pl2 : pl3;
    const int nv = dim == 2 ? 4 : 9;
    float acc[9] = {0};
    float c[18];
    const int ssh = (s - 1) / 2;
    for (int a = 0; a < s; a++)
        for (int b = 0; b < s; b++) {
            cov_at(cov, H, W, np, y - ssh + a, x - ssh + b, c);
            for (int v = 0; v < nv; v++) {
                const int p = pl[v];
                const int r = (p / 2) / dim, cc = (p / 2) % dim;
                const float val = r != cc ? c[p] * gamma : c[p];
                acc[v] = fmaf(gauss[a * s + b], val, acc[v]);
            }
        }
    for (int v = 0; v < nv; v++) out[v] = acc[v];
}

typedef struct {
    int H, W, dim, s, R;   /* grid covers [-R, H+R) x [-R, W+R) */
    int nv;
    float* v;
} avg_grid;

static const float* avg_get(const avg_grid* g, int y, int x)
{
    return g->v + ((size_t)(y + g->R) * (g->W + 2 * g->R) + (x + g->R)) * g->nv;
}

static float pair_delta(const avg_grid* g, int y1, int x1, int y2, int x2, int flip)
{
    /* delta(later, earlier) */
    int later = is_later(y1, x1, y2, x2);
    if (flip) later = !later;
    const float* p = later ? avg_get(g, y1, x1) : avg_get(g, y2, x2);
    const float* q = later ? avg_get(g, y2, x2) : avg_get(g, y1, x1);
    return g->dim == 2 ? orc_pixel_similarity_2x2(p, q, 1) : orc_pixel_similarity_3x3(p, q, 1);
}

static float brute_lut_weight(float d, const orc_stats* st, float h, float c)
{
    /* compute_weights.cl:37-52 */
    if (d > st->dissims_max) return 0.0f;
    d = fmaxf(d, st->dissims_min);
    d = fminf(d, st->dissims_max);
    const float idx = (d - st->dissims_min) / (st->dissims_max - st->dissims_min) * (float)(st->lut_size - 1);
    const float q = st->quantilles[(unsigned int)idx];
    const float x = st->chi2cdf_inv[(unsigned int)(q * (float)(st->lut_size - 1))];
    return expf(-fabsf(x - c) / h);
}

static int key_less(int p1, int s1, int p2, int s2) { return p1 < p2 || (p1 == p2 && s1 < s2); }

/* cov: 2*dim*dim planes of H x W (interlaced re / im per element).  out: same planes.
 * enl_maps / alpha_maps (optional): [nP*nS][H*W], index scale_index*nP + patch_index.
 * The statistics may hold keys beyond P x S: as in the reference, an unused last key of the
 * collection selects the weights of index 0 with alpha 0 (best_params.cpp:27-53 with
 * params2idx's and best_alpha_copy's operator[] defaults). */
int brute_nlsar_covmat(const float* cov, float* out, int H, int W, int dim, int sw, const int* P,
                       int nP, const int* S, int nS, const orc_stats* st, int nst, float h,
                       float c, float* enl_maps, float* alpha_maps, int flip)
{
    if (H <= 0 || W <= 0 || (dim != 2 && dim != 3) || sw < 3 || sw % 2 == 0) return -1;
    const int np = 2 * dim * dim;
    const int wsh = (sw - 1) / 2;
    int pmax = 0, smax = 0;
    for (int i = 0; i < nP; i++) pmax = pmax > P[i] ? pmax : P[i];
    for (int i = 0; i < nS; i++) smax = smax > S[i] ? smax : S[i];
    const int overlap = (pmax - 1) / 2 + wsh + (smax - 1) / 2;
    const int nkey = nP * nS;
    const int nfull = sw * sw;
    /* the keys' statistics; the collection's last key decides the selection */
    const orc_stats** ks = (const orc_stats**)malloc(sizeof(*ks) * nkey);
    for (int si = 0; si < nS; si++)
        for (int pi = 0; pi < nP; pi++) {
            ks[si * nP + pi] = NULL;
            for (int k = 0; k < nst; k++)
                if (st[k].patch_size == P[pi] && st[k].scale_size == S[si]) ks[si * nP + pi] = &st[k];
            if (!ks[si * nP + pi]) { free(ks); return -3; }
        }
    int lastp = st[0].patch_size, lasts = st[0].scale_size;
    for (int k = 1; k < nst; k++)
        if (key_less(lastp, lasts, st[k].patch_size, st[k].scale_size)) {
            lastp = st[k].patch_size;
            lasts = st[k].scale_size;
        }
    int last_in_set = 0;
    for (int i = 0; i < nP; i++)
        for (int j = 0; j < nS; j++)
            if (P[i] == lastp && S[j] == lasts) last_in_set = 1;

    const float LD = 1.0f / dim;
    const float gamma = powf(fminf(LD, 1.0f), 0.33333f);
    /* scale-averaged operands on a grid reaching every position a virtual pixel reads */
    const int R = 2 * overlap + 2;
    avg_grid* grids = (avg_grid*)malloc(sizeof(avg_grid) * nS);
    for (int si = 0; si < nS; si++) {
        float gauss[31 * 31];
        orc_gen_gauss(S[si], gauss);
        avg_grid* g = &grids[si];
        g->H = H; g->W = W; g->dim = dim; g->s = S[si]; g->R = R;
        g->nv = dim == 2 ? 4 : 9;
        const int gh = H + 2 * R, gw = W + 2 * R;
        g->v = (float*)malloc(sizeof(float) * (size_t)gh * gw * g->nv);
#pragma omp parallel for schedule(static)
        for (int y = -R; y < H + R; y++)
            for (int x = -R; x < W + R; x++)
                avg_at(cov, H, W, dim, S[si], gauss, gamma, y, x,
                       g->v + ((size_t)(y + R) * gw + (x + R)) * g->nv);
    }

#pragma omp parallel for schedule(dynamic)
    for (int o = 0; o < H * W; o++) {
        const int oy = o / W, ox = o % W;
        const int vy = oy < H - 1 ? oy : H - 1 + overlap;
        const int vx = ox < W - 1 ? ox : W - 1 + overlap;
        float* wts = (float*)malloc(sizeof(float) * (size_t)nkey * nfull);
        float* alphas = (float*)malloc(sizeof(float) * nkey);
        float* enlnb = (float*)malloc(sizeof(float) * nkey);
        float cc[18];
        for (int si = 0; si < nS; si++)
            for (int pi = 0; pi < nP; pi++) {
                const int key = si * nP + pi;
                const int psh = (P[pi] - 1) / 2;
                float* w = wts + (size_t)key * nfull;
                float ws = 0.0f, ws2 = 0.0f, inl[3] = {0, 0, 0}, wv[3] = {0, 0, 0};
                for (int dy = -wsh; dy <= wsh; dy++)
                    for (int dx = -wsh; dx <= wsh; dx++) {
                        float weight;
                        if (dy == 0 && dx == 0) {
                            weight = 1.0f; /* routines.cpp:151-157 */
                        } else {
                            /* row sums left to right, then the rows top to bottom */
                            float sum = 0.0f;
                            for (int ky = -psh; ky <= psh; ky++) {
                                float rs = 0.0f;
                                for (int kx = -psh; kx <= psh; kx++) {
                                    const int qy = vy + ky, qx = vx + kx;
                                    rs += pair_delta(&grids[si], qy, qx, qy + dy, qx + dx, flip);
                                }
                                sum += rs;
                            }
                            weight = brute_lut_weight(sum, ks[key], h, c);
                        }
                        w[(dy + wsh) * sw + (dx + wsh)] = weight;
                        /* compute_nl_statistics.cl:69-84 on the original covariance */
                        cov_at(cov, H, W, np, vy + dy, vx + dx, cc);
                        ws += weight;
                        ws2 = fmaf(weight, weight, ws2);
                        for (int d = 0; d < dim; d++) {
                            const float I = cc[(2 * dim + 2) * d];
                            inl[d] = fmaf(weight, I, inl[d]);
                            wv[d] = fmaf(weight, I * I, wv[d]);
                        }
                    }
                const float enl = ws * ws / ws2;
                float alpha = 0.0f;
                for (int d = 0; d < dim; d++) {
                    const float m = inl[d] / ws;
                    const float var = wv[d] / ws - m * m;
                    alpha = fmaxf(alpha, fmaxf(0.0f, (var - m * m / 1.0f) / var));
                }
                alphas[key] = alpha;
                enlnb[key] = orc_enl_nobias(enl, alpha, ws);
                if (enl_maps) enl_maps[(size_t)key * H * W + o] = enlnb[key];
                if (alpha_maps) alpha_maps[(size_t)key * H * W + o] = alpha;
            }
        /* selection: the last key in (patch, scale) order whose ENL exceeds -1, else the
         * last key; an unused last key of the collection gives index 0 and alpha 0 */
        int sel = -1;
        float alpha_sel;
        if (last_in_set) {
            int best_any = -1;
            for (int si = 0; si < nS; si++)
                for (int pi = 0; pi < nP; pi++) {
                    const int key = si * nP + pi;
                    if (best_any < 0 || key_less(P[best_any % nP], S[best_any / nP], P[pi], S[si])) best_any = key;
                    if (enlnb[key] > -1.0f &&
                        (sel < 0 || key_less(P[sel % nP], S[sel / nP], P[pi], S[si])))
                        sel = key;
                }
            if (sel < 0) sel = best_any;
            alpha_sel = alphas[sel];
        } else {
            sel = 0;
            alpha_sel = 0.0f;
        }
        /* weighted_means.cl:19-83 over the full window in plane order */
        const float* w = wts + (size_t)sel * nfull;
        float centre[18];
        cov_at(cov, H, W, np, vy, vx, centre);
        for (int pl = 0; pl < np; pl++) {
            float cnew = 0.0f, wsum = 0.0f;
            for (int dy = -wsh; dy <= wsh; dy++)
                for (int dx = -wsh; dx <= wsh; dx++) {
                    const float weight = w[(dy + wsh) * sw + (dx + wsh)];
                    cov_at(cov, H, W, np, vy + dy, vx + dx, cc);
                    wsum += weight;
                    cnew = fmaf(weight, cc[pl], cnew);
                }
            out[(size_t)pl * H * W + o] = (1 - alpha_sel) * cnew / wsum + alpha_sel * centre[pl];
        }
        free(enlnb);
        free(alphas);
        free(wts);
    }
    for (int si = 0; si < nS; si++) free(grids[si].v);
    free(grids);
    free(ks);
    return 0;
}

/* InSAR input: covmat_create.cl:19-56 per pixel, the filter above, covmat_decompose.cl:19-39 */
int brute_nlsar_insar(const float* a1, const float* a2, const float* phi, float* ref, float* pha,
                      float* coh, int H, int W, int sw, const int* P, int nP, const int* S, int nS,
                      const orc_stats* st, int nst, float h, float c, float* enl_maps,
                      float* alpha_maps, int flip)
{
    const size_t n = (size_t)H * W;
    if (H <= 0 || W <= 0) return -1;
    float* cov = (float*)calloc(8 * n, sizeof(float));
    float* out = (float*)calloc(8 * n, sizeof(float));
    for (size_t i = 0; i < n; i++) {
        const float re = a1[i] * a2[i] * cosf(phi[i]);
        const float im = a1[i] * a2[i] * sinf(phi[i]);
        cov[i] = a1[i] * a1[i];
        cov[n + i] = 0.0f;
        cov[2 * n + i] = re;
        cov[3 * n + i] = im;
        cov[4 * n + i] = re;
        cov[5 * n + i] = -im;
        cov[6 * n + i] = a2[i] * a2[i];
        cov[7 * n + i] = 0.0f;
    }
    int rc = brute_nlsar_covmat(cov, out, H, W, 2, sw, P, nP, S, nS, st, nst, h, c, enl_maps,
                                alpha_maps, flip);
    if (rc == 0)
        for (size_t i = 0; i < n; i++) {
            const float am = sqrtf(out[i]);
            const float as = sqrtf(out[6 * n + i]);
            const float s = 0.5f * (am + as);
            const float re = out[2 * n + i], im = out[3 * n + i];
            ref[i] = s * s;
            pha[i] = atan2f(im, re);
            coh[i] = sqrtf(im * im + re * re) / (am * as);
        }
    free(out);
    free(cov);
    return rc;
}

/* ------------------------------------------------------------------------------------
 * NL-InSAR
 * ------------------------------------------------------------------------------------ */

static int valid_value(float x) { return !(isnan(x) || isinf(x) || x == 0); }

/* where the reference's tiling last writes output index o (rows or columns): the tile
 * index along that axis and the virtual image position written */
static void nl_virtual(int o, int n, int step, int overlap, int* tile_idx, int* virt)
{
    if (o < n - 1) {
        *tile_idx = o / step;
        *virt = o;
    } else {
        /* tiles start at -overlap + k*step while that is < n; the last one's last interior
         * position is the last write to index n - 1 */
        const int k = (n - 1 + overlap) / step;
        *tile_idx = k;
        *virt = k * step + step - 1;
    }
}

/* One NL-InSAR iteration (nlinsar.cpp:80-133 over 80 x 80 tiles, nlinsar_sub_image.cpp),
 * per output pixel.  est_* are the iteration's input estimates; out_* receive the new ones.
 * n_smoothed (optional) counts pixels that went through the lmin smoothing. */
int brute_nlinsar_iteration(const float* a1, const float* a2, const float* phi,
                            const float* est_ref, const float* est_pha, const float* est_coh,
                            float* out_ref, float* out_pha, float* out_coh, int H, int W, int sw,
                            int p, int lmin, int tile, long long* n_smoothed, int flip)
{
    float hpar, Tpar;
    if (orc_nlinsar_params(p, &hpar, &Tpar)) return -4;
    const int wsh = (sw - 1) / 2, psh = (p - 1) / 2, overlap = wsh + psh;
    if (tile <= 0) tile = 80;
    const int step = tile - 2 * overlap;
    if (step <= 0 || sw * sw > 1024) return -1;
    const int w_sim = tile - sw + 1;      /* similarity planes of a tile: w_sim x w_sim */
    const int aw = 2 * overlap + 1;       /* local amplitude window of the smoothing */
    long long nsm = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : nsm)
    for (int o = 0; o < H * W; o++) {
        const int oy = o / W, ox = o % W;
        int ty, tx, vy, vx;
        nl_virtual(oy, H, step, overlap, &ty, &vy);
        nl_virtual(ox, W, step, overlap, &tx, &vx);
        const int h_low = -overlap + ty * step, w_low = -overlap + tx * step;
#define CL(arr, y, x) (arr)[(size_t)clampi((y), 0, H - 1) * W + clampi((x), 0, W - 1)]
        float wts[1024];
        float ws = 0.0f, ws2 = 0.0f;
        for (int dy = -wsh; dy <= wsh; dy++)
            for (int dx = -wsh; dx <= wsh; dx++) {
                float weight = 0.0f;
                if (!(dy == 0 && dx == 0)) {
                    /* patch sums (precompute_patch_similarities.cl): row sums along x
                     * inside each patch row, the rows added top to bottom */
                    float S = 0.0f, K = 0.0f;
                    for (int ky = -psh; ky <= psh; ky++) {
                        float rs = 0.0f, rk = 0.0f;
                        for (int kx = -psh; kx <= psh; kx++) {
                            const int qy = vy + ky, qx = vx + kx;
                            const int ry = qy + dy, rx = qx + dx;
                            /* similarity-plane position of q in its tile: interior positions
                             * hold offsets with dy >= 0 themselves and take dy < 0 from the
                             * partner, sim(q + d, q) */
                            const int hs = qy - h_low - wsh, wsp = qx - w_low - wsh;
                            const int interior = hs >= wsh && wsp >= wsh && wsp < w_sim - wsh;
                            int swap = interior && dy < 0;
                            if (flip) swap = !swap;
                            const int y1 = swap ? ry : qy, x1 = swap ? rx : qx;
                            const int y2 = swap ? qy : ry, x2 = swap ? qx : rx;
                            rs += orc_nlinsar_pixel_similarity(CL(a1, y1, x1), CL(a2, y1, x1), CL(phi, y1, x1),
                                                               CL(a1, y2, x2), CL(a2, y2, x2), CL(phi, y2, x2));
                            rk += orc_nlinsar_pixel_kl(CL(est_ref, y1, x1), CL(est_pha, y1, x1), CL(est_coh, y1, x1),
                                                       CL(est_ref, y2, x2), CL(est_pha, y2, x2), CL(est_coh, y2, x2));
                        }
                        if (ky == -psh) { S = rs; K = rk; } else { S += rs; K += rk; }
                    }
                    weight = expf(S / hpar + K / Tpar);   /* compute_weights.cl:19-30 */
                    if (isnan(weight) || isinf(weight)) weight = 0.0f;
                }
                wts[(dy + wsh) * sw + (dx + wsh)] = weight;
                ws += weight;
                ws2 = fmaf(weight, weight, ws2);
            }
        const float enl = ws * ws / ws2;   /* compute_number_of_looks.cl:19-42 */
        if (enl < lmin) {
            float amp[(2 * 32 + 1) * (2 * 32 + 1)];
            for (int y = 0; y < aw; y++)
                for (int x = 0; x < aw; x++) amp[y * aw + x] = CL(a1, vy - overlap + y, vx - overlap + x);
            orc_search_window_smoothing(amp, wts, 0, 0, aw, p, sw, lmin);
            nsm++;
        }
        /* precompute_filter_values.cl:19-39 and compute_insar.cl:48-119 */
        float N = 0.0f, at = 0.0f, xr = 0.0f, xi = 0.0f;
        for (int dy = -wsh; dy <= wsh; dy++)
            for (int dx = -wsh; dx <= wsh; dx++) {
                const float weight = wts[(dy + wsh) * sw + (dx + wsh)];
                const float m = CL(a1, vy + dy, vx + dx), s = CL(a2, vy + dy, vx + dx);
                const float d = CL(phi, vy + dy, vx + dx);
                const float fa = 0.5f * (m * m + s * s);
                const float fr = m * s * cosf(-d), fi = m * s * sinf(-d);
                N += weight;
                at = fmaf(weight, fa, at);
                xr = fmaf(weight, fr, xr);
                xi = fmaf(weight, fi, xi);
            }
        const float reflectivity = at / N;
        const float iphase = atan2f(xi, xr);
        const float coherence = sqrtf(xr * xr + xi * xi) / at;
        out_ref[o] = valid_value(reflectivity) ? reflectivity : est_ref[o];
        if (valid_value(iphase)) {
            float ph = iphase;
            if (ph > (float)M_PI) ph -= 2.0f * (float)M_PI;
            if (ph < -(float)M_PI) ph += 2.0f * (float)M_PI;
            out_pha[o] = ph;
        } else {
            out_pha[o] = est_pha[o];
        }
        if (valid_value(coherence)) {
            float cv = coherence;
            if (cv > 1.0f) cv = 1.0f;
            if (cv < 0.0f) cv = 0.0f;
            out_coh[o] = cv;
        } else {
            out_coh[o] = est_coh[o];
        }
#undef CL
    }
    if (n_smoothed) *n_smoothed = nsm;
    return 0;
}
