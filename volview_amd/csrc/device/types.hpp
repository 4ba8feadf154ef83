// Plain data types shared by the host code and the kernels (no device code,
// so host translation units compile them with a plain C++ compiler).
#pragma once
#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>
#include <stddef.h>
#include <stdint.h>

#include "../../../include/vvr_uniforms.h"

#if defined(__HIPCC__)
#define VVR_HD __host__ __device__ __forceinline__
#else
#define VVR_HD inline
#endif

namespace vvr {

// 4^3 micro-brick layout of a volume of nx x ny x nz voxels.
struct BrickGrid {
    int nx, ny, nz;
    int bx, by, bz;          // bricks per axis, ceil(n / 4)
    uint32_t ystride;        // bx * 64
    uint32_t zstride;        // bx * by * 64
    VVR_HD uint32_t xoff(int x) const { return (uint32_t(x >> 2) << 6) | uint32_t(x & 3); }
    VVR_HD uint32_t yoff(int y) const { return uint32_t(y >> 2) * ystride + (uint32_t(y & 3) << 2); }
    VVR_HD uint32_t zoff(int z) const { return uint32_t(z >> 2) * zstride + (uint32_t(z & 3) << 4); }
    VVR_HD uint32_t at(int x, int y, int z) const { return xoff(x) + yoff(y) + zoff(z); }
    VVR_HD size_t padded() const { return size_t(zstride) * size_t(bz); }
};

inline BrickGrid make_brick_grid(int nx, int ny, int nz) {
    BrickGrid g;
    g.nx = nx, g.ny = ny, g.nz = nz;
    g.bx = (nx + 3) / 4, g.by = (ny + 3) / 4, g.bz = (nz + 3) / 4;
    g.ystride = uint32_t(g.bx) * 64u;
    g.zstride = uint32_t(g.bx) * uint32_t(g.by) * 64u;
    return g;
}

// Everything the render kernels read, by value in the kernel arguments.
struct DevScene {
    BrickGrid g;
    const uint16_t* __restrict__ density;  // unorm16
    const uint8_t* __restrict__ labels;    // label per voxel
    const uint32_t* __restrict__ mask;     // RG16: r low half, g high half
    const uint32_t* __restrict__ normals;  // RGBA8 snorm, x in the low byte
    const float* __restrict__ majorant;    // ceil(dims / 8), x fastest
    int mx, my, mz;
    const float4* __restrict__ lut;     // 4096 transfer-function entries, fp16-rounded
    const float4* __restrict__ preint;  // 4096 cumulative integrals
    const vvr_uniforms* __restrict__ u;
};

}  // namespace vvr
