// The allocation-free NL-SAR entry point: dkx_nlsar_insar_device with its scratch memory (the
// padded covariance and scale-average grids, the undefined-selection flags and count, the
// weight LUTs and Gaussian taps) in one file-scope device array instead of the library's
// workspace, for callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).  The array exists once per device and
// keeps the last call's contents; every kernel that reads a part of it runs after the
// stage that writes that part in the same call, and the undefined-selection count is
// cleared by a kernel, so nothing from an earlier call is read.  Calls that overlap on one
// device share the array and race: the caller orders them.
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_nlsar.h"

#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace dkx {

// 64 MiB: images up to about 1200 x 1200 pixels with a search window of 21, patch sizes up
// to 7 and scale sizes up to 3 (padded grids of 16 + 20 bytes per element, plus one flag
// byte per pixel, plus the constants).
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Per device: the address of its copy of `scratch` (from hipGetSymbolAddress, with that
// device current), the host source of the last constants upload and its completion event.
struct NoallocDevice {
    char* ptr = nullptr;
    std::vector<float> blob;
    hipEvent_t uploaded = nullptr;
};

static std::mutex g_noalloc_mu;
static std::map<int, NoallocDevice> g_noalloc;

// Set-up on the first call on the current device: the array's device address and an event.
static int noalloc_device(NoallocDevice** out)
{
    int dev = 0;
    DKX_HIP_CHECK(hipGetDevice(&dev));
    NoallocDevice& d = g_noalloc[dev];
    if (!d.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("nlsar (allocation-free): the scratch array's address is not "
                                  "available: ") + hipGetErrorString(e));
            return DKX_EHIP;
        }
        DKX_HIP_CHECK(hipEventCreateWithFlags(&d.uploaded, hipEventDisableTiming));
        d.ptr = (char*)p;
    }
    *out = &d;
    return 0;
}

static size_t noalloc_need(const NlsarCall& call)
{
    return (call.ip.bytes + 255) / 256 * 256 + call.cm.const_floats * sizeof(float);
}

}  // namespace dkx

using namespace dkx;

extern "C" {

int dkx_nlsar_insar_device_noalloc(const float* a1, const float* a2, const float* phi, float* ref,
                                   float* pha, float* coh, int height, int width, int sw,
                                   const int* P, int nP, const int* S, int nS,
                                   const dkx_nlsar_stats* stats, int n_stats, float h, float c,
                                   float* enl_maps, float* alpha_maps, unsigned long long* ub,
                                   void* stream)
{
    return dkx::guarded<int>([&]() -> int {
        if (!a1 || !a2 || !phi || !ref || !pha || !coh) { set_error("nlsar: null array argument"); return DKX_EINVAL; }
        NlsarCall call;
        const int diag = (enl_maps || alpha_maps || ub) ? 1 : 0;
        int rc = nlsar_prepare(call, 0, diag, height, width, sw, P, nP, S, nS, stats, n_stats, h, c);
        if (rc) return rc;
        const size_t need = noalloc_need(call);
        if (need > NOALLOC_SCRATCH_BYTES) {
            set_error("nlsar (allocation-free): a " + std::to_string(height) + " x " + std::to_string(width) +
                      " image needs " + std::to_string(need) + " bytes of scratch memory, the entry point "
                      "holds " + std::to_string(NOALLOC_SCRATCH_BYTES) + " (use dkx_nlsar_insar_device)");
            return DKX_ERESOURCE;
        }
        std::vector<float> blob;
        rc = nlsar_build_constants(call.pp, call.cm, blob);
        if (rc) return rc;
        const hipStream_t st = (hipStream_t)stream;
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        NoallocDevice* d = nullptr;
        rc = noalloc_device(&d);
        if (rc) return rc;
        // the constants after the grids; their host source must outlive the asynchronous
        // upload, so it stays with the device until the next call's upload has been issued
        float* d_const = (float*)(d->ptr + (call.ip.bytes + 255) / 256 * 256);
        DKX_HIP_CHECK(hipEventSynchronize(d->uploaded));
        d->blob.swap(blob);
        DKX_HIP_CHECK(hipMemcpyAsync(d_const, d->blob.data(), d->blob.size() * sizeof(float),
                                     hipMemcpyHostToDevice, st));
        DKX_HIP_CHECK(hipEventRecord(d->uploaded, st));
        NlsarImageDesc desc;
        nlsar_image_desc(call.ip, d->ptr, a1, a2, phi, nullptr, ref, pha, coh, nullptr, enl_maps,
                         alpha_maps, &desc);
        rc = nlsar_enqueue(call.pp, call.cm, &desc, 1, nullptr, d_const, st);
        if (rc) return rc;
        if (ub) {
            DKX_HIP_CHECK(hipMemcpyAsync(ub, desc.ub_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
            DKX_HIP_CHECK(hipStreamSynchronize(st));
        }
        return 0;
    });
}

}  // extern "C"
