// Shared device/host definitions of the despeckle HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace dkx {

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

// Stage timing (off by default): when enabled, the main launch of each fused stage is
// bracketed by HIP events on its stream; dkx_stage_times() reports the durations.
enum StageKind {
    STAGE_NLSAR_SCALE = 1,       // NL-SAR scale pass, not the last
    STAGE_NLSAR_LAST = 2,        // NL-SAR last pass (statistics + weighted mean + outputs)
    STAGE_NLINSAR_WEIGHTS = 11,  // NL-InSAR weights / ENL / estimates
    STAGE_NLINSAR_SMOOTH = 12,   // NL-InSAR lmin smoothing
};
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
