// NL-SAR training on the device: the dissimilarity statistics of a homogeneous area for every
// (patch size, scale size) pair (src/compute/nlsar/training/nlsar_training.cpp:35-96,
// get_dissims.cpp:32-117, patches.cpp:32-177, combinations.cpp:26-41, get_stats.cpp:25-59).
//
// The reference averages the training covariance on the OpenCL device, cuts it into
// non-overlapping patches and compares every unordered pair of patches on the host, pixel by
// pixel, into one sum per pair; nlsar::stats then sorts the sums into a quantile LUT.  Here:
//  1. train_patches: one thread per patch pixel forms the rescaled, Gaussian-averaged
//     operands of that pixel (the values the dissimilarity reads) and the double log of
//     their determinant;
//  2. train_pairs<pass 0>: one thread per patch pair forms the dissimilarity sum and the
//     block reduces the finite minimum, maximum and count;
//  3. train_pairs<pass 1>: the same sums are binned against the 256 quantile thresholds;
//     the prefix sums of the bins are exactly the std::lower_bound counts of the
//     reference's quantile LUT (stats.cpp:52-62), with no sort and no stored sums.
// Arithmetic follows the reference's host training code, whose sim_measures.h is compiled
// without <math.h>: the logs are ::log(double) and the expression is evaluated in double,
// rounded to float per pixel, then summed in float in pixel order (patches.cpp:96-177).
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_nlsar.h"
#include "dkx_stats.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace dkx {

constexpr int TRAIN_LUT = 256;   // get_stats.cpp:31

struct TrainCounts {
    unsigned int min_key, max_key;    // order-preserving keys of the finite extremes
    unsigned long long finite;
    unsigned long long bins[TRAIN_LUT + 1];
};

__device__ __forceinline__ unsigned int float_key(float f)
{
    const unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

static float key_float(unsigned int k)
{
    const unsigned int u = (k & 0x80000000u) ? (k ^ 0x80000000u) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Operands of one pixel, as sim_measures.h orders them: {a00, a01r, a01i, a11} for dim 2,
// {a00, a11, a22, a01r, a01i, a02r, a02i, a12r, a12i} for dim 3 (interlaced plane indices).
template <int DIM>
__device__ __forceinline__ int train_plane(int k)
{
    if constexpr (DIM == 2) {
        constexpr int pl[4] = {0, 2, 3, 6};
        return pl[k];
    } else {
        constexpr int pl[9] = {0, 8, 16, 2, 3, 4, 5, 10, 11};
        return pl[k];
    }
}

template <int DIM>
__device__ __forceinline__ float train_det(const float* p)
{
    // det_covmat_2x2 / det_covmat_3x3, sim_measures.h:31-34, 79-97 (float operations)
    if constexpr (DIM == 2) {
        return (p[0] * p[3]) - (p[1] * p[1] + p[2] * p[2]);
    } else {
        return p[0] * p[1] * p[2] - p[0] * (p[7] * p[7] + p[8] * p[8]) - p[1] * (p[5] * p[5] + p[6] * p[6]) -
               p[2] * (p[3] * p[3] + p[4] * p[4]) +
               2 * (p[3] * p[7] * p[5] - p[4] * p[8] * p[5] + p[3] * p[8] * p[6] + p[4] * p[7] * p[6]);
    }
}

struct PatchArgs {
    const float* cov;       // 2*dim*dim planes of H x W
    int H, W, ho, wo;       // input, averaged extent (H - s + 1, W - s + 1)
    int p, s, npw, npatch;
    const float* gauss;     // s x s taps
    float gamma;            // off-diagonal rescaling, covmat_rescale.cl:24-25
    float* pat;             // [patch][pixel][operand]
    double* logdet;         // [patch][pixel]
};

// get_dissims.cpp:60-95 (rescale, average with the scale as its own maximum) and
// patches.cpp:32-41 (non-overlapping p x p patches, clamped to the averaged extent)
template <int DIM>
__global__ void train_patches(const PatchArgs a)
{
    constexpr int NK = DIM == 2 ? 4 : 9;
    const int pp = a.p * a.p;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)a.npatch * pp) return;
    const int q = (int)(idx / pp), i = (int)(idx - (long long)q * pp);
    const int ty = q / a.npw, tx = q - ty * a.npw;
    const int y = min(a.ho - 1, ty * a.p + i / a.p), x = min(a.wo - 1, tx * a.p + i % a.p);
    const long long n = (long long)a.H * a.W;
    float v[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int pl = train_plane<DIM>(k);
        const bool off = (pl / 2) / DIM != (pl / 2) % DIM;
        const float* c = a.cov + pl * n;
        float acc = 0.f;
        for (int ky = 0; ky < a.s; ky++)
            for (int kx = 0; kx < a.s; kx++) {
                float val = c[(long long)(y + ky) * a.W + x + kx];
                if (off) val = val * a.gamma;
                acc = fmaf(a.gauss[ky * a.s + kx], val, acc);
            }
        v[k] = acc;
        a.pat[idx * NK + k] = acc;
    }
    a.logdet[idx] = log((double)train_det<DIM>(v));
}

struct PairArgs {
    const float* pat;
    const double* logdet;
    int pp;
    long long npairs;
    double K;               // 2 * dim * log(2.0)
    const float* thr;       // pass 1: TRAIN_LUT thresholds
    TrainCounts* counts;
};

// unordered pair t -> (a, b), a > b (every pair once, combinations.cpp:26-41)
__device__ __forceinline__ void pair_of(long long t, long long& a, long long& b)
{
    long long r = (long long)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
    while (r * (r - 1) / 2 > t) r--;
    while ((r + 1) * r / 2 <= t) r++;
    a = r;
    b = t - r * (r - 1) / 2;
}

template <int DIM>
__device__ __forceinline__ float pair_sum(const PairArgs& g, long long pa, long long pb)
{
    constexpr int NK = DIM == 2 ? 4 : 9;
    const float* p1 = g.pat + pa * g.pp * NK;
    const float* p2 = g.pat + pb * g.pp * NK;
    const double* l1 = g.logdet + pa * g.pp;
    const double* l2 = g.logdet + pb * g.pp;
    float sum = 0.0f;
    for (int i = 0; i < g.pp; i++) {
        float sv[NK];
#pragma unroll
        for (int k = 0; k < NK; k++) sv[k] = p1[i * NK + k] + p2[i * NK + k];
        const double ld = log((double)train_det<DIM>(sv));
        // -1 * (2*dim*log 2 + log nom1 + log nom2 - 2 log det), in double
        float r = (float)(-(((g.K + l1[i]) + l2[i]) - 2.0 * ld));
        if (isnan(r)) r = 0.0f;
        sum += r;
    }
    return sum;
}

template <int DIM, int PASS>
__global__ __launch_bounds__(256) void train_pairs(const PairArgs g)
{
    __shared__ float s_thr[TRAIN_LUT];
    __shared__ unsigned int s_bins[TRAIN_LUT + 1];
    __shared__ unsigned int s_min, s_max;
    __shared__ unsigned long long s_fin;
    const int tid = threadIdx.x;
    if (PASS == 1) {
        for (int i = tid; i < TRAIN_LUT; i += blockDim.x) s_thr[i] = g.thr[i];
        for (int i = tid; i <= TRAIN_LUT; i += blockDim.x) s_bins[i] = 0u;
    }
    if (tid == 0) {
        s_min = 0xffffffffu;
        s_max = 0u;
        s_fin = 0ull;
    }
    __syncthreads();
    unsigned int kmin = 0xffffffffu, kmax = 0u;
    unsigned long long fin = 0ull;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + tid; t < g.npairs; t += stride) {
        long long pa, pb;
        pair_of(t, pa, pb);
        const float v = pair_sum<DIM>(g, pa, pb);
        if (!isfinite(v)) continue;
        if (PASS == 0) {
            const unsigned int k = float_key(v);
            kmin = min(kmin, k);
            kmax = max(kmax, k);
            fin++;
        } else {
            // thresholds at or below v (std::upper_bound over the nondecreasing thresholds)
            int lo = 0, hi = TRAIN_LUT;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_thr[mid] <= v) lo = mid + 1;
                else hi = mid;
            }
            atomicAdd(&s_bins[lo], 1u);
        }
    }
    if (PASS == 0) {
        atomicMin(&s_min, kmin);
        atomicMax(&s_max, kmax);
        atomicAdd(&s_fin, fin);
    }
    __syncthreads();
    if (PASS == 0) {
        if (tid == 0) {
            atomicMin(&g.counts->min_key, s_min);
            atomicMax(&g.counts->max_key, s_max);
            atomicAdd(&g.counts->finite, s_fin);
        }
    } else {
        for (int i = tid; i <= TRAIN_LUT; i += blockDim.x)
            if (s_bins[i]) atomicAdd(&g.counts->bins[i], (unsigned long long)s_bins[i]);
    }
}

__global__ void train_counts_init(TrainCounts* c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        c->min_key = 0xffffffffu;
        c->max_key = 0u;
        c->finite = 0ull;
    }
    if (i <= TRAIN_LUT) c->bins[i] = 0ull;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

// Every (patch, scale) pair in std::map order; the statistics of entry k use
// lut_storage[k*2*256 ...].  Synchronises `stream`.
static int train_device(const float* d_cov, int H, int W, int dim, const int* P, int nP,
                        const int* S, int nS, dkx_nlsar_stats* out, float* lut_storage,
                        hipStream_t stream)
{
    std::vector<int> ps(P, P + nP), ss(S, S + nS);
    std::sort(ps.begin(), ps.end());
    std::sort(ss.begin(), ss.end());
    long long max_elems = 0;
    int max_s = 1;
    for (int p : ps)
        for (int s : ss) {
            const int ho = H - s + 1, wo = W - s + 1;
            if (ho <= 0 || wo <= 0) {
                set_error("nlsar_training: the training area is smaller than scale size " + std::to_string(s));
                return DKX_EINVAL;
            }
            const long long np = (long long)((ho + p - 1) / p) * ((wo + p - 1) / p);
            max_elems = std::max(max_elems, np * p * p);
            max_s = std::max(max_s, s);
        }
    const int nk = dim == 2 ? 4 : 9;
    DevBuf pat, ld, taps, thr, cnt;
    DKX_HIP_CHECK(hipMalloc(&pat.p, sizeof(float) * nk * max_elems));
    DKX_HIP_CHECK(hipMalloc(&ld.p, sizeof(double) * max_elems));
    DKX_HIP_CHECK(hipMalloc(&taps.p, sizeof(float) * max_s * max_s));
    DKX_HIP_CHECK(hipMalloc(&thr.p, sizeof(float) * TRAIN_LUT));
    DKX_HIP_CHECK(hipMalloc(&cnt.p, sizeof(TrainCounts)));
    const double K = (double)(2 * dim) * std::log((double)2.0f);
    const float gamma = std::pow(std::min(1.0f / (float)dim, 1.0f), 0.33333f);
    int k = 0;
    for (int p : ps)
        for (int s : ss) {
            const std::vector<float> g = gen_gauss(s);
            DKX_HIP_CHECK(hipMemcpyAsync(taps.p, g.data(), sizeof(float) * g.size(), hipMemcpyHostToDevice, stream));
            PatchArgs pa;
            pa.cov = d_cov;
            pa.H = H;
            pa.W = W;
            pa.ho = H - s + 1;
            pa.wo = W - s + 1;
            pa.p = p;
            pa.s = s;
            pa.npw = (pa.wo + p - 1) / p;
            pa.npatch = ((pa.ho + p - 1) / p) * pa.npw;
            pa.gauss = (const float*)taps.p;
            pa.gamma = gamma;
            pa.pat = (float*)pat.p;
            pa.logdet = (double*)ld.p;
            const long long elems = (long long)pa.npatch * p * p;
            const dim3 gp((unsigned)((elems + 255) / 256));
            if (dim == 2) hipLaunchKernelGGL(train_patches<2>, gp, dim3(256), 0, stream, pa);
            else hipLaunchKernelGGL(train_patches<3>, gp, dim3(256), 0, stream, pa);
            DKX_HIP_CHECK(hipGetLastError());
            PairArgs ga;
            ga.pat = pa.pat;
            ga.logdet = pa.logdet;
            ga.pp = p * p;
            ga.npairs = (long long)pa.npatch * (pa.npatch - 1) / 2;
            ga.K = K;
            ga.thr = (const float*)thr.p;
            ga.counts = (TrainCounts*)cnt.p;
            hipLaunchKernelGGL(train_counts_init, dim3(2), dim3(256), 0, stream, ga.counts);
            DKX_HIP_CHECK(hipGetLastError());
            const unsigned blocks = (unsigned)std::max<long long>(1, std::min<long long>((ga.npairs + 255) / 256, 8192));
            if (ga.npairs > 0) {
                if (dim == 2) hipLaunchKernelGGL((train_pairs<2, 0>), dim3(blocks), dim3(256), 0, stream, ga);
                else hipLaunchKernelGGL((train_pairs<3, 0>), dim3(blocks), dim3(256), 0, stream, ga);
                DKX_HIP_CHECK(hipGetLastError());
            }
            TrainCounts hc;
            DKX_HIP_CHECK(hipMemcpyAsync(&hc, ga.counts, sizeof(hc), hipMemcpyDeviceToHost, stream));
            DKX_HIP_CHECK(hipStreamSynchronize(stream));
            if (hc.finite == 0) {
                set_error("nlsar_training: the training area yields no finite dissimilarity for patch " +
                          std::to_string(p) + ", scale " + std::to_string(s));
                return DKX_EINVAL;
            }
            dkx_nlsar_stats& o = out[k];
            float* q = lut_storage + (size_t)k * 2 * TRAIN_LUT;
            o.patch_size = p;
            o.scale_size = s;
            o.lut_size = TRAIN_LUT;
            o.quantilles = q;
            o.chi2cdf_inv = q + TRAIN_LUT;
            o.dissims_min = key_float(hc.min_key);
            o.dissims_max = key_float(hc.max_key);
            std::vector<float> th(TRAIN_LUT);
            stats_thresholds(o.dissims_min, o.dissims_max, TRAIN_LUT, th.data());
            DKX_HIP_CHECK(hipMemcpyAsync(thr.p, th.data(), sizeof(float) * TRAIN_LUT, hipMemcpyHostToDevice, stream));
            if (dim == 2) hipLaunchKernelGGL((train_pairs<2, 1>), dim3(blocks), dim3(256), 0, stream, ga);
            else hipLaunchKernelGGL((train_pairs<3, 1>), dim3(blocks), dim3(256), 0, stream, ga);
            DKX_HIP_CHECK(hipGetLastError());
            DKX_HIP_CHECK(hipMemcpyAsync(&hc, ga.counts, sizeof(hc), hipMemcpyDeviceToHost, stream));
            DKX_HIP_CHECK(hipStreamSynchronize(stream));
            unsigned long long below = 0;
            for (int i = 0; i < TRAIN_LUT; i++) {
                below += hc.bins[i];
                q[i] = ((float)below) / (float)hc.finite;
            }
            stats_chi2_lut(TRAIN_LUT, q + TRAIN_LUT);
            k++;
        }
    return 0;
}

static int train_check(int H, int W, int dim, const int* P, int nP, const int* S, int nS,
                       const dkx_nlsar_stats* out, const float* lut_storage)
{
    if (!P || !S || !out || !lut_storage || H <= 0 || W <= 0 || nP <= 0 || nS <= 0) {
        set_error("nlsar_training: invalid arguments");
        return DKX_EINVAL;
    }
    if (dim != 2 && dim != 3) {
        // routines.cpp:77-80
        set_error("currently only 2x2 or 3x3 covariance matrices are supported");
        return DKX_EINVAL;
    }
    for (int i = 0; i < nP; i++)
        if (P[i] < 1 || P[i] % 2 == 0) {
            set_error("nlsar_training: patch and scale sizes must be odd");
            return DKX_EINVAL;
        }
    for (int i = 0; i < nS; i++)
        if (S[i] < 1 || S[i] % 2 == 0 || S[i] > 31) {
            set_error("nlsar_training: patch and scale sizes must be odd (scale sizes up to 31)");
            return DKX_EINVAL;
        }
    return 0;
}

static int have_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        set_error("nlsar_training: no HIP device (training runs on the GPU)");
        return DKX_EHIP;
    }
    return 0;
}

// host covariance planes -> device -> training
static int train_host_planes(const float* cov, int H, int W, int dim, const int* P, int nP,
                             const int* S, int nS, dkx_nlsar_stats* out, float* lut_storage)
{
    DevBuf d;
    const size_t bytes = sizeof(float) * 2 * dim * dim * (size_t)H * W;
    DKX_HIP_CHECK(hipMalloc(&d.p, bytes));
    hipStream_t st;
    DKX_HIP_CHECK(hipStreamCreate(&st));
    int rc = 0;
    if (hipMemcpyAsync(d.p, cov, bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
        set_error("nlsar_training: copy to the device failed");
        rc = DKX_EHIP;
    }
    if (!rc) rc = train_device((const float*)d.p, H, W, dim, P, nP, S, nS, out, lut_storage, st);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    return rc;
}

}  // namespace dkx

using namespace dkx;

extern "C" {

int dkx_nlsar_training_covmat_device(const float* d_covmat, int height, int width, int dim,
                                     const int* P, int nP, const int* S, int nS,
                                     dkx_nlsar_stats* out, float* lut_storage, void* stream)
{
    return dkx::guarded<int>([&]() -> int {
        int rc = train_check(height, width, dim, P, nP, S, nS, out, lut_storage);
        if (rc) return rc;
        if (!d_covmat) { set_error("nlsar_training: null array argument"); return DKX_EINVAL; }
        rc = have_device();
        if (rc) return rc;
        return train_device(d_covmat, height, width, dim, P, nP, S, nS, out, lut_storage, (hipStream_t)stream);
    });
}

int dkx_nlsar_training_covmat(const float* covmat, int height, int width, int dim, const int* P,
                              int nP, const int* S, int nS, dkx_nlsar_stats* out, float* lut_storage)
{
    return dkx::guarded<int>([&]() -> int {
        int rc = train_check(height, width, dim, P, nP, S, nS, out, lut_storage);
        if (rc) return rc;
        if (!covmat) { set_error("nlsar_training: null array argument"); return DKX_EINVAL; }
        rc = have_device();
        if (rc) return rc;
        return train_host_planes(covmat, height, width, dim, P, nP, S, nS, out, lut_storage);
    });
}

int dkx_nlsar_training_insar(const float* a1, const float* a2, const float* phi, int height,
                             int width, const int* P, int nP, const int* S, int nS,
                             dkx_nlsar_stats* out, float* lut_storage)
{
    return dkx::guarded<int>([&]() -> int {
        int rc = train_check(height, width, 2, P, nP, S, nS, out, lut_storage);
        if (rc) return rc;
        if (!a1 || !a2 || !phi) { set_error("nlsar_training: null array argument"); return DKX_EINVAL; }
        rc = have_device();
        if (rc) return rc;
        // covmat_data(insar_data), src/data/data.cpp:60-89: the reference forms the training
        // covariance on the host with float cos / sin, and so does this entry point
        const size_t n = (size_t)height * width;
        std::vector<float> cov(8 * n);
        for (size_t i = 0; i < n; i++) {
            const float re = a1[i] * a2[i] * std::cos(phi[i]);
            const float im = a1[i] * a2[i] * std::sin(phi[i]);
            cov[i] = a1[i] * a1[i];
            cov[n + i] = 0.f;
            cov[2 * n + i] = re;
            cov[3 * n + i] = im;
            cov[4 * n + i] = re;
            cov[5 * n + i] = -im;
            cov[6 * n + i] = a2[i] * a2[i];
            cov[7 * n + i] = 0.f;
        }
        return train_host_planes(cov.data(), height, width, 2, P, nP, S, nS, out, lut_storage);
    });
}

}  // extern "C"
