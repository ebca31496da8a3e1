// Goldstein phase filter on CDNA4 (src/compute/goldstein/goldstein.cpp:39-106,
// goldstein_filter_sub_image.cpp:73-187 and the kernels of goldstein/cl_kernels).
//
// The reference, per T x T tile (T = round_up(max(H, W) + 2 ov, P - 2 ov), edge-replicated):
//   z = 0.5 (a1 + a2) e^{i phi}                          raw_interferogram.cl
//   patches of P x P at stride m = P - 2 ov, offset -ov     patches_unpack.cl (clamped)
//   Z = FFT2(patch); Z <- |Z|^alpha Z; IFFT2 / P^2        clFFT, weighted_multiply.cl
//   z += sum of weighted patches, four checkerboard passes patches_pack.cl ("+=" onto the
//                                                          raw interferogram still there)
//   ref = |z|, phase = arg z                              slc2real.cl
// then copies the tile interior back with the clamped write of the tile machinery, so the
// last row / column carry the value at the last interior row / column of the last tile.
//
// Here gs_patches fuses the unpack, both 2-D transforms and the weighting of a group of
// patches: the patch is formed straight from the inputs into LDS, each thread transforms one
// row (then one column) in registers (radix-2, P <= 64), and the result, scaled by 1/P^2 and
// the pack weight, is stored once.  gs_pack forms every output pixel from the raw value and
// the (at most four) patch values covering it, added in the reference's pass order.  Only
// the patches that cover an output position are transformed (a second tile per direction,
// which the reference's tiling creates for some sizes, contributes only the last row or
// column).
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_goldstein.h"

#include <algorithm>

namespace dkx {

// threads per block: 128 (several patches per block) up to P = 32, one 64-point patch per
// 64-thread block (its padded tile fills 33 KB of LDS)
constexpr int gs_threads(int P) { return P >= 64 ? 64 : 128; }

struct GsTileDev {
    int h_low, w_low;      // tile origin in image coordinates
    int py0, npy, px0, npx;
    long long u_off;       // first float2 of the tile's unpacked patches in U (row-major,
                           // npy * P rows of npx * P)
    int job0, jobs;
};

struct GsArgs {
    const float* a1;
    const float* a2;
    const float* phi;
    float* ref;
    float* phase;
    float2* U;
    float2* Z;             // raw interferogram of the image
    int H, W, T, P, ov, m, n, step;
    float alpha;
    int ntx, nty, ntiles, jobs;
    GsTileDev tile[4];
};

__device__ __forceinline__ int gs_clamp(int v, int hi) { return min(hi, max(0, v)); }

// raw_interferogram.cl:19-35, once per image pixel (the edge-replicated tiles read it
// through gs_raw)
// image pixel of a launch of ceil(W / 256) blocks per row
__device__ __forceinline__ bool gs_pixel(const GsArgs& a, int& y, int& x)
{
    const int bpr = (a.W + 255) >> 8;
    y = blockIdx.x / bpr;
    x = (blockIdx.x - y * bpr) * 256 + threadIdx.x;
    return x < a.W;
}

__global__ void gs_raw_image(const GsArgs a)
{
    int y, x;
    if (!gs_pixel(a, y, x)) return;
    const long long g = (long long)y * a.W + x;
    const float amp = 0.5f * (gptr(a.a1)[g] + gptr(a.a2)[g]);
    float s, c;
    sincosf(gptr(a.phi)[g], &s, &c);
    gstore2(gptr(a.Z) + g, make_float2(amp * c, amp * s));
}

// the raw interferogram at tile-local packed position (ty, tx) of a tile
__device__ __forceinline__ float2 gs_raw(const GsArgs& a, const GsTileDev& t, int ty, int tx)
{
    const int y = gs_clamp(ty + t.h_low, a.H - 1), x = gs_clamp(tx + t.w_low, a.W - 1);
    return gload2(gptr(a.Z) + (long long)y * a.W + x);
}

__device__ __forceinline__ int gs_tile_of_job(const GsArgs& a, int job)
{
    int t = 0;
    while (t + 1 < a.ntiles && job >= a.tile[t + 1].job0) t++;
    return t;
}

constexpr int gs_rev(int i, int bits)
{
    int r = 0;
    for (int b = 0; b < bits; b++)
        if (i & (1 << b)) r |= 1 << (bits - 1 - b);
    return r;
}

constexpr int gs_log2(int n) { return n <= 1 ? 0 : 1 + gs_log2(n / 2); }

// In-register radix-2 DIT transform of P points: sign -1 forward, +1 inverse (unscaled).
// tw[k] = (cos 2 pi k / P, sin 2 pi k / P).
template <int P, int SIGN>
__device__ __forceinline__ void gs_fft(float2 (&v)[P], const float2* tw)
{
    constexpr int B = gs_log2(P);
#pragma unroll
    for (int i = 0; i < P; i++) {
        const int r = gs_rev(i, B);
        if (i < r) {
            const float2 t = v[i];
            v[i] = v[r];
            v[r] = t;
        }
    }
#pragma unroll
    for (int len = 2; len <= P; len <<= 1) {
        const int half = len >> 1;
#pragma unroll
        for (int j = 0; j < half; j++) {
            const float2 w = tw[j * (P / len)];
            const float wr = w.x, wi = SIGN * w.y;
#pragma unroll
            for (int i = 0; i < P; i += len) {
                const float2 u = v[i + j], x = v[i + j + half];
                const float tr = x.x * wr - x.y * wi;
                const float ti = x.x * wi + x.y * wr;
                v[i + j] = make_float2(u.x + tr, u.y + ti);
                v[i + j + half] = make_float2(u.x - tr, u.y - ti);
            }
        }
    }
}

template <int P>
__global__ __launch_bounds__(gs_threads(P)) void gs_patches(const GsArgs a)
{
    constexpr int NT = gs_threads(P);
    constexpr int NPB = NT / P;            // patches per block
    constexpr int LP = P + 1;              // padded LDS row
    __shared__ float2 sp[NPB * P * LP];
    __shared__ float2 tw[P];
    const int tid = threadIdx.x;
    if (tid < P) {
        float s, c;
        sincospif(2.0f * (float)tid / (float)P, &s, &c);
        tw[tid] = make_float2(c, s);
    }
    const int m = a.m, ov = a.ov, T = a.T;
    // patch origin of each slot: tile-local packed row / column of its first pixel and the
    // tile origin in the image (-1 row origin: no patch)
    __shared__ int s_org[NPB][4];
    // output of each slot: first float2 of the patch in U, and the U row pitch
    __shared__ long long s_ubase[NPB];
    __shared__ int s_upitch[NPB];
    if (tid < NPB) {
        const int job = blockIdx.x * NPB + tid;
        if (job < a.jobs) {
            const GsTileDev& td = a.tile[gs_tile_of_job(a, job)];
            const int local = job - td.job0;
            const int qy = local / td.npx, qx = local - (local / td.npx) * td.npx;
            s_org[tid][0] = (td.py0 + qy) * m - ov;
            s_org[tid][1] = (td.px0 + qx) * m - ov;
            s_org[tid][2] = td.h_low;
            s_org[tid][3] = td.w_low;
            s_upitch[tid] = td.npx * P;
            s_ubase[tid] = td.u_off + (long long)qy * P * (td.npx * P) + (long long)qx * P;
        } else {
            s_org[tid][0] = -(1 << 30);
            s_upitch[tid] = 0;
            s_ubase[tid] = -1;
        }
    }
    __syncthreads();
    // 1. unpack from the raw interferogram (patches_unpack.cl:19-42)
#pragma unroll 8
    for (int e = tid; e < NPB * P * P; e += NT) {
        const int sl = e / (P * P), r = (e / P) % P, c = e % P;
        float2 z = make_float2(0.f, 0.f);
        if (s_org[sl][0] > -(1 << 29)) {
            const int y = gs_clamp(gs_clamp(s_org[sl][0] + r, T - 1) + s_org[sl][2], a.H - 1);
            const int x = gs_clamp(gs_clamp(s_org[sl][1] + c, T - 1) + s_org[sl][3], a.W - 1);
            z = gload2(gptr(a.Z) + (long long)y * a.W + x);
        }
        sp[(sl * P + r) * LP + c] = z;
    }
    __syncthreads();
    const int slot = tid / P, k = tid % P;
    float2* base = sp + slot * P * LP;
    float2 v[P];
    // 2. forward transform of row k
#pragma unroll
    for (int i = 0; i < P; i++) v[i] = base[k * LP + i];
    gs_fft<P, -1>(v, tw);
#pragma unroll
    for (int i = 0; i < P; i++) base[k * LP + i] = v[i];
    __syncthreads();
    // 3. column k: forward, |Z|^alpha Z (weighted_multiply.cl:19-36), inverse
#pragma unroll
    for (int i = 0; i < P; i++) v[i] = base[i * LP + k];
    gs_fft<P, -1>(v, tw);
#pragma unroll
    for (int i = 0; i < P; i++) {
        const float psd = sqrtf(v[i].x * v[i].x + v[i].y * v[i].y);
        // pow(psd, alpha) as exp2(alpha log2 psd); alpha = 0 gives 1 also at psd = 0
        const float w = a.alpha == 0.0f ? 1.0f : exp2f(a.alpha * log2f(psd));
        v[i] = make_float2(w * v[i].x, w * v[i].y);
    }
    gs_fft<P, 1>(v, tw);
#pragma unroll
    for (int i = 0; i < P; i++) base[i * LP + k] = v[i];
    __syncthreads();
    // 4. inverse transform of row k, clFFT's backward scale 1/P^2, the pack weight
    //    (patches_pack.cl:44-56)
#pragma unroll
    for (int i = 0; i < P; i++) v[i] = base[k * LP + i];
    gs_fft<P, 1>(v, tw);
    const int job = blockIdx.x * NPB + slot;
    if (job < a.jobs) {
        const int t = gs_tile_of_job(a, job);
        const GsTileDev& td = a.tile[t];
        const int local = job - td.job0;
        const int py = td.py0 + local / td.npx, px = td.px0 + local % td.npx;
        const float scale = 1.0f / (float)(P * P);
        const int out_y = py * m + (k - ov);
        float sfy = 1.0f;
        if (k < 2 * ov || k >= P - 2 * ov) sfy *= 0.5f;
#pragma unroll
        for (int i = 0; i < P; i++) {
            float sf = sfy;
            if (i < 2 * ov || i >= P - 2 * ov) sf *= 0.5f;
            const int out_x = px * m + (i - ov);
            if (out_x < ov || out_x >= T - ov) sf *= 2.0f;
            if (out_y < ov || out_y >= T - ov) sf *= 2.0f;
            const float2 x = make_float2(v[i].x * scale, v[i].y * scale);
            base[k * LP + i] = make_float2(sf * x.x, sf * x.y);
        }
    }
    __syncthreads();
    // 5. coalesced store of the weighted patches
#pragma unroll 4
    for (int e = tid; e < NPB * P * P; e += NT) {
        const int s = e / (P * P), r = (e / P) % P, c = e % P;
        const long long ub = s_ubase[s];
        if (ub < 0) continue;
        gstore2(gptr(a.U) + ub + (long long)r * s_upitch[s] + c, sp[(s * P + r) * LP + c]);
    }
}

// Which tile writes image row y last, and at which tile row (the clamped copy-back of the
// last tile lands on the last row from its last interior row).
__device__ __forceinline__ void gs_map(int y, int n_img, int ntiles, int step, int ov, int T, int& k, int& h)
{
    if (y == n_img - 1) {
        k = ntiles - 1;
        h = T - ov - 1;
    } else {
        k = y / step;
        h = y + ov - k * step;
    }
}

__global__ void gs_pack(const GsArgs a)
{
    int y, x;
    if (!gs_pixel(a, y, x)) return;
    const long long i = (long long)y * a.W + x;
    int ky, hy, kx, hx;
    gs_map(y, a.H, a.nty, a.step, a.ov, a.T, ky, hy);
    gs_map(x, a.W, a.ntx, a.step, a.ov, a.T, kx, hx);
    // the tile's fields, selected from uniform loads of all (at most four) tiles
    const int t = ky * a.ntx + kx;
#define DKX_GS_SEL(f) (t == 0 ? a.tile[0].f : t == 1 ? a.tile[1].f : t == 2 ? a.tile[2].f : a.tile[3].f)
    const int h_low = DKX_GS_SEL(h_low), w_low = DKX_GS_SEL(w_low);
    const int py0 = DKX_GS_SEL(py0), npy = DKX_GS_SEL(npy), px0 = DKX_GS_SEL(px0), npx = DKX_GS_SEL(npx);
    const long long u_off = DKX_GS_SEL(u_off);
#undef DKX_GS_SEL
    const int P = a.P, m = a.m, ov = a.ov;
    const DKX_GLOBAL float2* U = gptr(a.U);
    // the patches covering (hy, hx): rows pa (rel ra) and pa - 1 (rel ra + m), likewise cols;
    // [0] the even patch, [1] the odd one
    const int pya = (hy + ov) / m, rya = hy + ov - pya * m;
    const int pxa = (hx + ov) / m, rxa = hx + ov - pxa * m;
    const bool ya = pya < a.n, yb = pya >= 1 && rya + m < P;
    const bool xa = pxa < a.n, xb = pxa >= 1 && rxa + m < P;
    const int ea = pya & 1;                 // parity slot of patch row pya
    int py[2], ry[2], px[2], rx[2];
    bool vy[2], vx[2];
    py[ea] = pya; ry[ea] = rya; vy[ea] = ya;
    py[ea ^ 1] = pya - 1; ry[ea ^ 1] = rya + m; vy[ea ^ 1] = yb;
    const int fa = pxa & 1;
    px[fa] = pxa; rx[fa] = rxa; vx[fa] = xa;
    px[fa ^ 1] = pxa - 1; rx[fa ^ 1] = rxa + m; vx[fa ^ 1] = xb;
    const long long wu = (long long)npx * P;
    // all loads first (invalid ones read U[0]), then the additions in the pass order of
    // patches_pack.cpp:48-72: (x, y) parity (0, 0), (0, 1), (1, 0), (1, 1)
    float2 u[4];
    bool ok[4];
#pragma unroll
    for (int o = 0; o < 4; o++) {
        const int ox = o >> 1, oy = o & 1;
        ok[o] = vy[oy] && vx[ox] && py[oy] >= py0 && py[oy] < py0 + npy && px[ox] >= px0 && px[ox] < px0 + npx;
        const long long idx = ok[o] ? u_off + ((long long)(py[oy] - py0) * P + ry[oy]) * wu + (px[ox] - px0) * P + rx[ox]
                                    : 0;
        u[o] = gload2(U + idx);
    }
    const int yy = gs_clamp(hy + h_low, a.H - 1), xx = gs_clamp(hx + w_low, a.W - 1);
    float2 z = gload2(gptr(a.Z) + (long long)yy * a.W + xx);
#pragma unroll
    for (int o = 0; o < 4; o++) {
        z.x = ok[o] ? z.x + u[o].x : z.x;
        z.y = ok[o] ? z.y + u[o].y : z.y;
    }
    gptr(a.ref)[i] = fabsf(sqrtf(z.x * z.x + z.y * z.y));
    gptr(a.phase)[i] = atan2f(z.y, z.x);
}

// Patch rows (cols) of a tile that cover its needed output rows (cols): the interior rows
// that are the final writers of image rows below the last, and for the last tile also its
// last interior row.
static void gs_needed(int n_img, int k, int ntiles, int T, int P, int ov, int m, int n, int& p0, int& np)
{
    const int step = T - 2 * ov;
    int lo = -1, hi = -1;
    const int first = k * step, end = std::min((k + 1) * step, n_img - 1);   // image rows
    if (end > first) {
        lo = first + ov - k * step;
        hi = end - 1 + ov - k * step;
    }
    if (k == ntiles - 1) {
        if (lo < 0) lo = T - ov - 1;
        hi = T - ov - 1;
    }
    if (lo < 0) {
        p0 = 0;
        np = 0;
        return;
    }
    // patch p covers local rows [p m - ov, p m - ov + P)
    int pl = (lo + ov - P + 1 + m - 1) / m;
    if (lo + ov - P + 1 <= 0) pl = 0;
    const int ph = std::min(n - 1, (hi + ov) / m);
    p0 = pl;
    np = std::max(0, ph - pl + 1);
}

int goldstein_plan(const GoldsteinParams& gp, GoldsteinPlan* pl)
{
    if (gp.height <= 0 || gp.width <= 0) {
        set_error("goldstein: empty image");
        return DKX_EINVAL;
    }
    if (gp.overlap < 0 || gp.patch_size - 2 * gp.overlap <= 0) {
        set_error("goldstein: the overlap must be below half the patch size");
        return DKX_EINVAL;
    }
    if (4 * gp.overlap > gp.patch_size) {
        // more than two patches would overlap per direction: the reference's packing passes
        // then add overlapping patches concurrently (patches_pack.cpp:33-47) and its 1/2
        // weights no longer sum to one
        set_error("goldstein: overlaps above a quarter of the patch size are not supported");
        return DKX_EUNSUPPORTED;
    }
    if (gp.patch_size != 16 && gp.patch_size != 32 && gp.patch_size != 64) {
        set_error("goldstein: patch sizes 16, 32 and 64 are supported (the reference's 16 x 16 "
                  "work groups need multiples of 16)");
        return DKX_EUNSUPPORTED;
    }
    const int P = gp.patch_size, ov = gp.overlap, m = P - 2 * ov;
    // goldstein.cpp:64-67, tile_size.cpp:30-33 (round_up adds a full multiple); the
    // device-memory cap of tile_size.cpp:35-77 is taken as inactive
    const long long big = std::max((long long)gp.height, (long long)gp.width) + 2 * ov;
    const long long T = big - big % m + m;
    if (T > (1 << 28)) {
        set_error("goldstein: image too large");
        return DKX_EINVAL;
    }
    pl->P = P;
    pl->ov = ov;
    pl->m = m;
    pl->T = (int)T;
    pl->n = (int)(T / m);
    pl->step = (int)T - 2 * ov;
    pl->nty = (gp.height + ov - 1) / pl->step + 1;
    pl->ntx = (gp.width + ov - 1) / pl->step + 1;
    if (pl->nty > 2 || pl->ntx > 2) {
        set_error("goldstein: internal error, unexpected tiling");
        return DKX_EINVAL;
    }
    long long u = 0;
    int jobs = 0;
    for (int ky = 0; ky < pl->nty; ky++)
        for (int kx = 0; kx < pl->ntx; kx++) {
            GoldsteinPlan::Tile& t = pl->tile[ky * pl->ntx + kx];
            t.h_low = -ov + ky * pl->step;
            t.w_low = -ov + kx * pl->step;
            gs_needed(gp.height, ky, pl->nty, pl->T, P, ov, m, pl->n, t.py0, t.npy);
            gs_needed(gp.width, kx, pl->ntx, pl->T, P, ov, m, pl->n, t.px0, t.npx);
            t.u_off = u;
            t.job0 = jobs;
            t.jobs = t.npy * t.npx;
            u += (long long)t.jobs * P * P;
            jobs += t.jobs;
        }
    pl->jobs = jobs;
    pl->u_bytes = ((size_t)u * sizeof(float2) + 255) / 256 * 256;
    pl->ws_bytes = pl->u_bytes + (size_t)gp.height * gp.width * sizeof(float2);
    return 0;
}

template <int P>
static void gs_launch(const GsArgs& a, hipStream_t stream)
{
    constexpr int NPB = gs_threads(P) / P;
    const unsigned blocks = (unsigned)((a.jobs + NPB - 1) / NPB);
    if (blocks) hipLaunchKernelGGL(gs_patches<P>, dim3(blocks), dim3(gs_threads(P)), 0, stream, a);
}

int goldstein_run(const GoldsteinParams& gp, const GoldsteinPlan& pl, const float* a1,
                  const float* a2, const float* phi, float* ref, float* phase, void* ws,
                  hipStream_t stream)
{
    GsArgs a{};
    a.a1 = a1;
    a.a2 = a2;
    a.phi = phi;
    a.ref = ref;
    a.phase = phase;
    a.U = (float2*)ws;
    a.Z = (float2*)((char*)ws + pl.u_bytes);
    a.H = gp.height;
    a.W = gp.width;
    a.T = pl.T;
    a.P = pl.P;
    a.ov = pl.ov;
    a.m = pl.m;
    a.n = pl.n;
    a.step = pl.step;
    a.alpha = gp.alpha;
    a.ntx = pl.ntx;
    a.nty = pl.nty;
    a.ntiles = pl.ntx * pl.nty;
    a.jobs = pl.jobs;
    for (int t = 0; t < a.ntiles; t++) {
        const GoldsteinPlan::Tile& s = pl.tile[t];
        a.tile[t] = GsTileDev{s.h_low, s.w_low, s.py0, s.npy, s.px0, s.npx, s.u_off, s.job0, s.jobs};
    }
    const dim3 gimg((unsigned)(((gp.width + 255) / 256) * (long long)gp.height));
    hipLaunchKernelGGL(gs_raw_image, gimg, dim3(256), 0, stream, a);
    DKX_HIP_CHECK(hipGetLastError());
    switch (pl.P) {
    case 16: gs_launch<16>(a, stream); break;
    case 32: gs_launch<32>(a, stream); break;
    default: gs_launch<64>(a, stream); break;
    }
    DKX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(gs_pack, gimg, dim3(256), 0, stream, a);
    DKX_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace dkx
