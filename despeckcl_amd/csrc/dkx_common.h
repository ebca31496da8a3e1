// Shared device/host definitions of the despeckle HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <exception>
#include <new>
#include <string>

namespace dkx {

// Pointers read from descriptors in memory reach the kernels as generic pointers; these casts
// give their loads and stores the global address space (global_load with a scalar base
// instead of flat_load, which also waits on LDS traffic).
#define DKX_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const DKX_GLOBAL T* gptr(const T* p) { return (const DKX_GLOBAL T*)p; }
template <class T>
__device__ __forceinline__ DKX_GLOBAL T* gptr(T* p) { return (DKX_GLOBAL T*)p; }
// vector loads / stores through such pointers (HIP's float4 class has no constructor from
// an address-space-qualified object; the native vector types do)
typedef float dkx_f4 __attribute__((ext_vector_type(4)));
typedef float dkx_f2 __attribute__((ext_vector_type(2)));
#ifdef __HIP__
// Values every lane of a wave holds alike, moved to scalar registers (HIP sources only).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <class T>
__device__ __forceinline__ T* uni(T* p)
{
    const unsigned long long u = reinterpret_cast<unsigned long long>(p);
    const unsigned int lo = __builtin_amdgcn_readfirstlane((unsigned int)u);
    const unsigned int hi = __builtin_amdgcn_readfirstlane((unsigned int)(u >> 32));
    return reinterpret_cast<T*>(((unsigned long long)hi << 32) | lo);
}
#endif

__device__ __forceinline__ float4 gload4(const DKX_GLOBAL void* p)
{
    const dkx_f4 v = *(const DKX_GLOBAL dkx_f4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 gload2(const DKX_GLOBAL void* p)
{
    const dkx_f2 v = *(const DKX_GLOBAL dkx_f2*)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void gstore4(DKX_GLOBAL void* p, float4 v)
{
    *(DKX_GLOBAL dkx_f4*)p = dkx_f4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ void gstore2(DKX_GLOBAL void* p, float2 v)
{
    *(DKX_GLOBAL dkx_f2*)p = dkx_f2{v.x, v.y};
}

// A padded planar grid in image coordinates: element (y, x) of plane k lives at
// base[k * plane + (y + oy) * pitch + (x + ox)], for y in [-oy, h - oy), x in [-ox, w - ox).
// Grids are filled with clamp-to-edge replication of the image, which is how the reference
// pads every tile (src/data/sub_images.cpp:37-40).
struct Grid {
    float* base;
    long long plane;  // elements per plane
    int pitch;        // elements per row
    int oy, ox;       // origin offset (padding before image row/column 0)
    int h, w;         // total rows / columns including padding
    __host__ __device__ inline long long at(int y, int x) const {
        return (long long)(y + oy) * pitch + (x + ox);
    }
};

// Launch mapping of an output-tile kernel: tiles start at virtual image coordinates
// (vy0 + by*TY, vx0 + bx*TX); valid positions are the first ny x nx virtual positions; a
// valid virtual position (vy0+i, vx0+j) is written to output pixel (oy0+i, ox0+j).
// Four mappings (interior, last row, last column, corner) reproduce the reference's
// clamped write-back, where the last image row/column receives the value computed at a
// virtual position beyond the edge (src/data/sub_images.cpp:47-69).
struct Launch {
    int vy0, vx0;
    int ny, nx;
    int oy0, ox0;
};

void set_error(const std::string& msg);
const char* last_error();

// C++ exceptions (host allocation failures, length errors of a malformed file) must not cross
// the C ABI: every entry point runs its body through guarded(), which maps them to
// DKX_ERESOURCE (-2) / DKX_EINVAL (-1) with a message.
template <class R, class F>
R guarded(F&& f) noexcept
{
    try {
        return f();
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return (R)-2;
    } catch (const std::exception& e) {
        set_error(std::string("internal error: ") + e.what());
        return (R)-1;
    } catch (...) {
        set_error("internal error");
        return (R)-1;
    }
}
template <class F>
void guarded_void(F&& f) noexcept
{
    try {
        f();
    } catch (...) {
        set_error("internal error");
    }
}

// Stage timing (off by default): when enabled, the main launch of each fused stage is
// bracketed by HIP events on its stream; dkx_stage_times() reports the durations.
enum StageKind {
    STAGE_NLSAR_SCALE = 1,       // NL-SAR scale pass, not the last
    STAGE_NLSAR_LAST = 2,        // NL-SAR last pass (statistics + weighted mean + outputs)
    STAGE_NLINSAR_WEIGHTS = 11,  // NL-InSAR weights / ENL / estimates
    STAGE_NLINSAR_SMOOTH = 12,   // NL-InSAR lmin smoothing
    STAGE_NLINSAR_EDGE = 13,     // NL-InSAR weights of the last row / column (thin blocks)
    STAGE_NLINSAR_FILL = 14,     // NL-InSAR weights of queued pixels the weights launch did not store
};
// Autotuner records (tune.cpp): the fastest tile variant for filter 0 (NL-SAR; variant 1 =
// every key's statistics, 0 = the selected key's; P, S = patch and scale sizes) or 1
// (NL-InSAR; P = the patch size, no S) at a search window and log2 of the pixel count
// (a batch: of all its images).  tuned_tile returns -1 when nothing is recorded.
int tuned_tile(int filter, int variant, int sw, const int* P, int nP, const int* S, int nS, double pixels);
void record_tuned_tile(int filter, int variant, int sw, const int* P, int nP, const int* S, int nS,
                       double pixels, int tile);

bool timing_on();
void timing_push(int kind, hipEvent_t start, hipEvent_t stop);

}  // namespace dkx

#define DKX_HIP_CHECK(expr)                                                                  \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) {                                                              \
            dkx::set_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " +     \
                           __FILE__ + ":" + std::to_string(__LINE__) + ": " #expr);         \
            return -5;                                                                       \
        }                                                                                    \
    } while (0)
