// Boxcar InSAR filter on CDNA4 (src/compute/boxcar/boxcar.cpp:35-92,
// cl_kernels/boxcar_kernel.cl:19-69).
//
// Each output is the w x w window average of the interferogram around its pixel:
// reflectivity sum((a1^2 + a2^2) / 2) / w^2, phase arg(sum a1 a2 e^{i phi}), coherence
// |sum a1 a2 e^{i phi}| / sum((a1^2 + a2^2) / 2).  The reference runs the kernel on 512 x 512
// tiles with an overlap of (w - 1) / 2, replicating edge pixels into the tiles, so:
//  - every pixel except the last row / column sees the edge-replicated image;
//  - the last row (column) receives, through the clamped copy-back of the last tile, the
//    value computed at that tile's last interior row (column), a virtual position at or
//    beyond the edge (boxcar_virtual_last).
// boxcar_tile stages a 16 x 64 output tile plus halo of per-pixel records
// (t = (a1^2 + a2^2)/2, a1*a2, cos phi, sin phi) in LDS once; each thread then sums four
// adjacent outputs in the reference's window order (rows, then columns), holding one window
// row in registers.  boxcar_edge computes the last row and column.
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"

#include <algorithm>

namespace dkx {

constexpr int BOX_TX = 16;                 // threads along x
constexpr int BOX_TY = 16;                 // threads along y (one output row each)
constexpr int BOX_OPT = 4;                 // adjacent outputs per thread along x
constexpr int BOX_BW = BOX_TX * BOX_OPT;   // 64 output columns per block
constexpr int BOX_BH = BOX_TY;             // 16 output rows per block
constexpr int BOX_WMAX = 15;               // the reference's 16-wide blocks leave no output above 15
constexpr int BOX_TILE = 512;              // boxcar.cpp:67

struct BoxArgs {
    const float* a1;
    const float* a2;
    const float* phi;
    float* ref;
    float* phase;
    float* coh;
    int H, W;
    int vh, vw;     // virtual row / column of the last row / column
};

// Per-pixel record; pow(x, 2) of the OpenCL source is x * x.
__device__ __forceinline__ float4 box_record(const BoxArgs& a, long long g)
{
    const float m = gptr(a.a1)[g], s = gptr(a.a2)[g], p = gptr(a.phi)[g];
    float sn, cs;
    sincosf(p, &sn, &cs);
    return make_float4((m * m + s * s) / 2, m * s, cs, sn);
}

__device__ __forceinline__ void box_accumulate(float4 t, float& acc, float& xr, float& xi)
{
    // "a += ...", "x_real += aa * cos(phi)" (contracted), "x_imag += aa * sin(phi)"
    acc += t.x;
    xr = fmaf(t.y, t.z, xr);
    xi = fmaf(t.y, t.w, xi);
}

struct BoxOut {
    float ref, phase, coh;
};

__device__ __forceinline__ BoxOut box_result(float acc, float xr, float xi, float n)
{
    return BoxOut{acc / n, atan2f(xi, xr), sqrtf(xr * xr + xi * xi) / acc};
}

__device__ __forceinline__ void box_store(const BoxArgs& a, long long o, float acc, float xr, float xi, float n)
{
    const BoxOut r = box_result(acc, xr, xi, n);
    gptr(a.ref)[o] = r.ref;
    gptr(a.phase)[o] = r.phase;
    gptr(a.coh)[o] = r.coh;
}

// LDS position of staged column c: one pad record after every four, so that the threads of
// a row, which read four columns apart, fall on distinct banks
__device__ __forceinline__ constexpr int box_skew(int c) { return c + (c >> 2); }

template <int WW>
__global__ __launch_bounds__(BOX_TX * BOX_TY) void boxcar_tile(const BoxArgs a)
{
    constexpr int R = (WW - 1) / 2;
    constexpr int LH = BOX_BH + 2 * R, LW = BOX_BW + 2 * R;
    constexpr int LWP = LW + LW / 4 + 1;   // skewed row pitch
    __shared__ float4 s[LH * LWP];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * BOX_BW, y0 = blockIdx.y * BOX_BH;
    const int H = a.H, W = a.W;
    for (int i = tid; i < LH * LW; i += BOX_TX * BOX_TY) {
        const int ly = i / LW, lx = i - ly * LW;
        const int gy = min(H - 1, max(0, y0 - R + ly));
        const int gx = min(W - 1, max(0, x0 - R + lx));
        s[ly * LWP + box_skew(lx)] = box_record(a, (long long)gy * W + gx);
    }
    __syncthreads();
    const int ty = tid / BOX_TX, tx = tid - ty * BOX_TX;
    const int y = y0 + ty;
    const int xb = x0 + tx * BOX_OPT;
    if (y >= H - 1 || xb >= W - 1) return;   // the last row / column: boxcar_edge
    float acc[BOX_OPT], xr[BOX_OPT], xi[BOX_OPT];
#pragma unroll
    for (int o = 0; o < BOX_OPT; o++) acc[o] = xr[o] = xi[o] = 0.f;
#pragma unroll 1
    for (int kx = 0; kx < WW; kx++) {   // one window row live in registers at a time
        float4 row[BOX_OPT + WW - 1];
        const float4* src = s + (ty + kx) * LWP + box_skew(tx * BOX_OPT);
#pragma unroll
        for (int j = 0; j < BOX_OPT + WW - 1; j++) row[j] = src[j + (j >> 2)];
#pragma unroll
        for (int o = 0; o < BOX_OPT; o++)
#pragma unroll
            for (int ky = 0; ky < WW; ky++) box_accumulate(row[o + ky], acc[o], xr[o], xi[o]);
    }
    const float n = (float)(WW * WW);
    const long long ob = (long long)y * W + xb;
    if (xb + BOX_OPT - 1 < W - 1 && (ob & 3) == 0) {
        BoxOut r[BOX_OPT];
#pragma unroll
        for (int o = 0; o < BOX_OPT; o++) r[o] = box_result(acc[o], xr[o], xi[o], n);
        gstore4(gptr(a.ref) + ob, make_float4(r[0].ref, r[1].ref, r[2].ref, r[3].ref));
        gstore4(gptr(a.phase) + ob, make_float4(r[0].phase, r[1].phase, r[2].phase, r[3].phase));
        gstore4(gptr(a.coh) + ob, make_float4(r[0].coh, r[1].coh, r[2].coh, r[3].coh));
    } else {
#pragma unroll
        for (int o = 0; o < BOX_OPT; o++)
            if (xb + o < W - 1) box_store(a, ob + o, acc[o], xr[o], xi[o], n);
    }
}

// Last row (all columns) and last column (rows 0 .. H-2), at their virtual positions.
template <int WW>
__global__ void boxcar_edge(const BoxArgs a)
{
    constexpr int R = (WW - 1) / 2;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = a.H, W = a.W;
    if (e >= W + H - 1) return;
    const int y = e < W ? H - 1 : e - W;
    const int x = e < W ? e : W - 1;
    const int vy = y == H - 1 ? a.vh : y;
    const int vx = x == W - 1 ? a.vw : x;
    float acc = 0.f, xr = 0.f, xi = 0.f;
    for (int kx = 0; kx < WW; kx++) {
        const int gy = min(H - 1, max(0, vy - R + kx));
        for (int ky = 0; ky < WW; ky++) {
            const int gx = min(W - 1, max(0, vx - R + ky));
            box_accumulate(box_record(a, (long long)gy * W + gx), acc, xr, xi);
        }
    }
    box_store(a, (long long)y * W + x, acc, xr, xi, (float)(WW * WW));
}

// The image row (column) whose value the last row (column) receives: the last interior row
// of the last 512-tile (tile_iterator.cpp:48-64, map_filter_tiles.h:85-91 with the clamped
// copy-back of sub_images.cpp:47-69).  Tiles start at -r and advance by 512 - 2r; the last
// one starts below n; its interior ends at (k_last + 1) * step - 1 >= n - 1.
static int boxcar_virtual_last(int n, int w)
{
    const int r = (w - 1) / 2;
    const int step = BOX_TILE - 2 * r;
    const int k_last = (n + r - 1) / step;
    return (k_last + 1) * step - 1;
}

template <int WW>
static void boxcar_launch(const BoxArgs& a, hipStream_t stream)
{
    const dim3 grid((unsigned)((a.W + BOX_BW - 1) / BOX_BW), (unsigned)((a.H + BOX_BH - 1) / BOX_BH));
    if (a.H > 1 && a.W > 1) hipLaunchKernelGGL(boxcar_tile<WW>, grid, dim3(BOX_TX * BOX_TY), 0, stream, a);
    const int ne = a.W + a.H - 1;
    hipLaunchKernelGGL(boxcar_edge<WW>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, stream, a);
}

static int boxcar_check(int height, int width, int w)
{
    if (height <= 0 || width <= 0) {
        set_error("boxcar: empty image");
        return DKX_EINVAL;
    }
    if (w < 1 || w % 2 == 0) {
        set_error("boxcar: the window width has to be an odd number");
        return DKX_EINVAL;
    }
    if (w > BOX_WMAX) {
        set_error("boxcar: window widths above 15 are not supported (the reference's 16-wide "
                  "work groups have no output pixels left for them)");
        return DKX_EUNSUPPORTED;
    }
    if ((long long)height * width >= (1ll << 40)) {
        set_error("boxcar: image too large");
        return DKX_EINVAL;
    }
    return 0;
}

static int boxcar_enqueue(const float* a1, const float* a2, const float* phi, float* ref,
                          float* pha, float* coh, int height, int width, int w, hipStream_t stream)
{
    BoxArgs a{a1, a2, phi, ref, pha, coh, height, width, boxcar_virtual_last(height, w),
              boxcar_virtual_last(width, w)};
    switch (w) {
    case 1: boxcar_launch<1>(a, stream); break;
    case 3: boxcar_launch<3>(a, stream); break;
    case 5: boxcar_launch<5>(a, stream); break;
    case 7: boxcar_launch<7>(a, stream); break;
    case 9: boxcar_launch<9>(a, stream); break;
    case 11: boxcar_launch<11>(a, stream); break;
    case 13: boxcar_launch<13>(a, stream); break;
    default: boxcar_launch<15>(a, stream); break;
    }
    DKX_HIP_CHECK(hipGetLastError());
    return 0;
}

struct BoxStage {
    void* p = nullptr;
    ~BoxStage() { if (p) (void)hipFree(p); }
};

}  // namespace dkx

using namespace dkx;

extern "C" {

int dkx_boxcar_device(const float* a1, const float* a2, const float* phi, float* ref, float* pha,
                      float* coh, int height, int width, int window_width, void* stream)
{
    return dkx::guarded<int>([&]() -> int {
        int rc = boxcar_check(height, width, window_width);
        if (rc) return rc;
        if (!a1 || !a2 || !phi || !ref || !pha || !coh) { set_error("boxcar: null array argument"); return DKX_EINVAL; }
        return boxcar_enqueue(a1, a2, phi, ref, pha, coh, height, width, window_width, (hipStream_t)stream);
    });
}

int dkx_boxcar(const float* a1, const float* a2, const float* phi, float* ref, float* pha,
               float* coh, int height, int width, int window_width)
{
    return dkx::guarded<int>([&]() -> int {
        int rc = boxcar_check(height, width, window_width);
        if (rc) return rc;
        if (!a1 || !a2 || !phi || !ref || !pha || !coh) { set_error("boxcar: null array argument"); return DKX_EINVAL; }
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
            set_error("boxcar: no HIP device");
            return DKX_EHIP;
        }
        const size_t n = (size_t)height * width;
        BoxStage st;
        DKX_HIP_CHECK(hipMalloc(&st.p, 6 * n * sizeof(float)));
        float* d = (float*)st.p;
        const float* src[3] = {a1, a2, phi};
        for (int k = 0; k < 3; k++) DKX_HIP_CHECK(hipMemcpy(d + k * n, src[k], n * sizeof(float), hipMemcpyHostToDevice));
        rc = boxcar_enqueue(d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, height, width, window_width, nullptr);
        if (rc) return rc;
        float* dst[3] = {ref, pha, coh};
        for (int k = 0; k < 3; k++) DKX_HIP_CHECK(hipMemcpy(dst[k], d + (3 + k) * n, n * sizeof(float), hipMemcpyDeviceToHost));
        return 0;
    });
}

}  // extern "C"
