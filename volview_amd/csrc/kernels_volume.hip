// Volume passes: density upload, normals (normals.comp), label masks
// (mask.comp) and the brick majorant (majorant.comp). They run once per volume
// or per opacity change and are HBM scans; their quantised outputs are
// reproduced bit for bit, so floating-point contraction is off in this file.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <stdexcept>

#include <cmath>

#include "launch.hpp"

namespace vvr {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void brick_coords(const BrickGrid& g, uint32_t o, int& x, int& y, int& z) {
    const uint32_t brick = o >> 6, w = o & 63u;
    const uint32_t bxy = uint32_t(g.bx) * uint32_t(g.by);
    const uint32_t bz = brick / bxy, rem = brick - bz * bxy;
    const uint32_t by = rem / uint32_t(g.bx), bx = rem - by * uint32_t(g.bx);
    x = int(bx * 4u + (w & 3u));
    y = int(by * 4u + ((w >> 2) & 3u));
    z = int(bz * 4u + (w >> 4));
}

// transfer.rs:121-124, in f32 with a correctly rounded division.
__device__ __forceinline__ uint16_t hu_to_unorm16(int16_t hu) {
    float t = (float(hu) - (-1024.0f)) / (3071.0f - (-1024.0f));
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return uint16_t(roundf(t * 65535.0f));
}

__global__ void k_upload_density(const int16_t* __restrict__ hu, uint16_t* __restrict__ out, BrickGrid g) {
    const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= g.padded()) return;
    int x, y, z;
    brick_coords(g, o, x, y, z);
    uint16_t v = 0;
    if (x < g.nx && y < g.ny && z < g.nz) v = hu_to_unorm16(hu[size_t(x) + size_t(g.nx) * (size_t(y) + size_t(g.ny) * z)]);
    out[o] = v;
}

__global__ void k_upload_labels(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, BrickGrid g) {
    const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= g.padded()) return;
    int x, y, z;
    brick_coords(g, o, x, y, z);
    uint8_t v = 0;
    if (x < g.nx && y < g.ny && z < g.nz) v = in[size_t(x) + size_t(g.nx) * (size_t(y) + size_t(g.ny) * z)];
    out[o] = v;
}

template <class T>
__global__ void k_debrick(const T* __restrict__ in, T* __restrict__ out, BrickGrid g) {
    const size_t n = size_t(g.nx) * g.ny * g.nz;
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const int x = int(i % g.nx), y = int((i / g.nx) % g.ny), z = int(i / (size_t(g.nx) * g.ny));
    out[i] = in[g.at(x, y, z)];
}

__device__ __forceinline__ float dtexel(const uint16_t* __restrict__ d, const BrickGrid& g, int x, int y, int z) {
    x = min(max(x, 0), g.nx - 1);
    y = min(max(y, 0), g.ny - 1);
    z = min(max(z, 0), g.nz - 1);
    return float(d[g.at(x, y, z)]) / 65535.0f;
}

__device__ __forceinline__ uint32_t snorm8_bits(float f) {
    f = f < -1.0f ? -1.0f : (f > 1.0f ? 1.0f : f);
    return uint32_t(uint8_t(int8_t(roundf(f * 127.0f))));
}

// normals.comp:14-37: 3x3x3 Sobel, world gradient, outward normal and surfaceness.
// A workgroup covers 2 x 2 x 2 micro-bricks (8^3 voxels). It first stages the
// tile's densities with a one-voxel halo in LDS (10^3 values, each the clamped
// texel as dtexel reads it), then every thread runs the stencil for two voxels
// from LDS with the same arithmetic in the same order. Each wave writes one
// micro-brick's 64 contiguous outputs.
constexpr int kNormTile = 8;                // voxels per axis
constexpr int kNormHalo = kNormTile + 2;  // with the halo
__global__ void __launch_bounds__(kBlock) k_normals(const uint16_t* __restrict__ d, uint32_t* __restrict__ out,
                                                    BrickGrid g, const vvr_uniforms* __restrict__ u) {
    __shared__ float s_d[kNormHalo * kNormHalo * kNormHalo];
    const int x0 = int(blockIdx.x) * kNormTile, y0 = int(blockIdx.y) * kNormTile, z0 = int(blockIdx.z) * kNormTile;
    for (int i = threadIdx.x; i < kNormHalo * kNormHalo * kNormHalo; i += kBlock) {
        const int hx = i % kNormHalo, hy = (i / kNormHalo) % kNormHalo, hz = i / (kNormHalo * kNormHalo);
        s_d[i] = dtexel(d, g, x0 - 1 + hx, y0 - 1 + hy, z0 - 1 + hz);
    }
    __syncthreads();
    const float* m = u->world_to_index;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
        const int mb = (threadIdx.x >> 6) + 4 * half, w = threadIdx.x & 63;  // micro-brick of the tile, voxel in it
        const int mbx = 2 * int(blockIdx.x) + (mb & 1), mby = 2 * int(blockIdx.y) + ((mb >> 1) & 1),
                  mbz = 2 * int(blockIdx.z) + (mb >> 2);
        if (mbx >= g.bx || mby >= g.by || mbz >= g.bz) continue;  // no such micro-brick (whole waves)
        const int x = 4 * mbx + (w & 3), y = 4 * mby + ((w >> 2) & 3), z = 4 * mbz + (w >> 4);
        const uint32_t o = g.at(x, y, z);
        if (x >= g.nx || y >= g.ny || z >= g.nz) {
            out[o] = 0u;
            continue;
        }
        // dtexel(x + i, y + j, z + k) for |i|, |j|, |k| <= 1.
        auto t = [&](int i, int j, int k) {
            return s_d[(x - x0 + 1 + i) + kNormHalo * ((y - y0 + 1 + j) + kNormHalo * (z - z0 + 1 + k))];
        };
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
        for (int a = -1; a <= 1; a++) {
            for (int b = -1; b <= 1; b++) {
                const float wt = float((2 - abs(a)) * (2 - abs(b)));
                gx += wt * (t(1, a, b) - t(-1, a, b));
                gy += wt * (t(a, 1, b) - t(a, -1, b));
                gz += wt * (t(a, b, 1) - t(a, b, -1));
            }
        }
        gx /= 32.0f;
        gy /= 32.0f;
        gz /= 32.0f;
        // transpose(mat3(world_to_index)) * g
        const float wx = m[0] * gx + m[1] * gy + m[2] * gz;
        const float wy = m[4] * gx + m[5] * gy + m[6] * gz;
        const float wz = m[8] * gx + m[9] * gy + m[10] * gz;
        const float len = sqrtf(wx * wx + wy * wy + wz * wz);
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        if (len > 1e-7f) {
            nx = -wx / len;
            ny = -wy / len;
            nz = -wz / len;
        }
        float sness = len * 400.0f;
        sness = sness < 0.0f ? 0.0f : (sness > 1.0f ? 1.0f : sness);
        out[o] = snorm8_bits(nx) | (snorm8_bits(ny) << 8) | (snorm8_bits(nz) << 16) | (snorm8_bits(sness) << 24);
    }
}

__device__ __forceinline__ uint32_t unorm16_bits(float f) {
    f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
    return uint32_t(roundf(f * 65535.0f));
}

__device__ __forceinline__ bool shown(const vvr_uniforms* u, uint32_t l) { return u->label_colors[l][3] > 0.5f; }
__device__ __forceinline__ bool ghosted(const vvr_uniforms* u, uint32_t l) {
    return l != 0u && u->label_colors[l][3] > 1.5f;
}
__device__ __forceinline__ bool solid(const vvr_uniforms* u, uint32_t l, uint32_t mode) {
    if (ghosted(u, l)) return false;
    return mode == 1u ? (l == 0u || shown(u, l)) : (l != 0u && shown(u, l));
}

// mask.comp:32-37
__device__ __forceinline__ float soft_score(bool best, bool second, float margin) {
    if (best == second) return best ? 4.0f : -4.0f;
    return best ? fminf(margin, 4.0f) : -fminf(margin, 4.0f);
}

// Trilinear density at continuous index v (float, as density_at_index).
__device__ __forceinline__ float density_linear(const uint16_t* __restrict__ d, const BrickGrid& g, float vx, float vy,
                                                float vz) {
    const float fx = floorf(vx), fy = floorf(vy), fz = floorf(vz);
    const float ax = vx - fx, ay = vy - fy, az = vz - fz;
    const int x0 = int(fx), y0 = int(fy), z0 = int(fz);
    auto t = [&](int x, int y, int z) { return dtexel(d, g, x, y, z); };
    const float c00 = t(x0, y0, z0) * (1.0f - ax) + t(x0 + 1, y0, z0) * ax;
    const float c10 = t(x0, y0 + 1, z0) * (1.0f - ax) + t(x0 + 1, y0 + 1, z0) * ax;
    const float c01 = t(x0, y0, z0 + 1) * (1.0f - ax) + t(x0 + 1, y0, z0 + 1) * ax;
    const float c11 = t(x0, y0 + 1, z0 + 1) * (1.0f - ax) + t(x0 + 1, y0 + 1, z0 + 1) * ax;
    const float c0 = c00 * (1.0f - ay) + c10 * ay;
    const float c1 = c01 * (1.0f - ay) + c11 * ay;
    return c0 * (1.0f - az) + c1 * az;
}

// The soft grid's mapping (mask.comp:40-60): for grid axis a, the volume axis
// it follows, that axis's extent `can` and the grid-per-voxel scale `to_grid`.
struct SoftAxes {
    int axis[3];
    float can[3], to_grid[3];
};
__device__ __forceinline__ SoftAxes soft_axes(const SoftGrid& soft, const BrickGrid& g) {
    SoftAxes s;
    const int dims[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; a++) {
        s.axis[a] = soft.push[a];
        s.can[a] = float(dims[s.axis[a]]);
        const float den = fmaxf(s.can[a] - 1.0f, 1.0f);
        s.to_grid[a] = (float(soft.dims[a]) - 1.0f) / den;
    }
    return s;
}

// The density at grid node n, trilinear at the node's volume position.
__device__ __forceinline__ float soft_node_density(const uint16_t* __restrict__ d, const BrickGrid& g,
                                                   const SoftGrid& soft, const SoftAxes& s, const int n[3]) {
    float v[3];
    for (int a = 0; a < 3; a++) {
        const float c = float(n[a]) / s.to_grid[a];
        v[s.axis[a]] = ((soft.push[3] >> a) & 1) != 0 ? s.can[a] - 1.0f - c : c;
    }
    return density_linear(d, g, v[0], v[1], v[2]);
}

// Every grid node's density, once, for k_mask_binary's 27-node windows (which
// would otherwise fetch it 27 times per voxel); grid axis 0 fastest.
__global__ void k_soft_node_density(const uint16_t* __restrict__ d, SoftGrid soft, BrickGrid g,
                                    float* __restrict__ there) {
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    const size_t n01 = size_t(soft.dims[0]) * size_t(soft.dims[1]);
    if (i >= n01 * size_t(soft.dims[2])) return;
    const int n[3] = {int(i % size_t(soft.dims[0])), int((i / size_t(soft.dims[0])) % size_t(soft.dims[1])),
                      int(i / n01)};
    there[i] = soft_node_density(d, g, soft, soft_axes(soft, g), n);
}

// Per label: bit 0 solid in the frame's mode, bit 1 ghosted, bit 2 shown.
constexpr uint32_t kLabelSolid = 1u, kLabelGhosted = 2u, kLabelShown = 4u;

// mask.comp:107-138 (binary_masks), with soft_coord / soft_masks (mask.comp:40-100).
// `there` holds every soft-grid node's density (k_soft_node_density), or is
// null and each window fetches it.
__global__ void __launch_bounds__(kBlock) k_mask_binary(const uint16_t* __restrict__ d,
                                                        const uint8_t* __restrict__ labels, SoftGrid soft,
                                                        const float* __restrict__ there_of, int margin,
                                                        uint32_t* __restrict__ out, BrickGrid g,
                                                        const vvr_uniforms* __restrict__ u) {
    __shared__ uint8_t s_label[256];
    const uint32_t mode = uint32_t(u->labels[0]);
    static_assert(kBlock == 256, "one label per thread");
    {
        const uint32_t l = threadIdx.x;
        s_label[l] = uint8_t((solid(u, l, mode) ? kLabelSolid : 0u) | (ghosted(u, l) ? kLabelGhosted : 0u) |
                             (shown(u, l) ? kLabelShown : 0u));
    }
    __syncthreads();
    const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= g.padded()) return;
    int x, y, z;
    brick_coords(g, o, x, y, z);
    if (x >= g.nx || y >= g.ny || z >= g.nz) {
        out[o] = 0u;
        return;
    }
    auto has = [&](uint32_t l, uint32_t bit) { return (s_label[l] & bit) != 0u; };
    const uint32_t label = labels[o];
    float m0 = has(label, kLabelSolid) ? 1.0f : 0.0f;
    float m1 = has(label, kLabelGhosted) ? 1.0f : 0.0f;
    if ((soft.push[3] & 8) != 0 && label < 250u) {
        const int p[3] = {x, y, z};
        const SoftAxes sa = soft_axes(soft, g);
        float xs[3];
        for (int a = 0; a < 3; a++) {
            const float v = float(p[sa.axis[a]]);
            xs[a] = ((soft.push[3] >> a) & 1) != 0 ? sa.can[a] - 1.0f - v : v;
            const float den = fmaxf(sa.can[a] - 1.0f, 1.0f);
            xs[a] = xs[a] * (float(soft.dims[a]) - 1.0f) / den;
        }
        auto node_index = [&](int a, int b, int c) {
            return size_t(a) + size_t(soft.dims[0]) * (size_t(b) + size_t(soft.dims[1]) * c);
        };
        int cx = int(roundf(xs[0])), cy = int(roundf(xs[1])), cz = int(roundf(xs[2]));
        const bool overridden = (soft.nodes[node_index(min(max(cx, 0), soft.dims[0] - 1),
                                                       min(max(cy, 0), soft.dims[1] - 1),
                                                       min(max(cz, 0), soft.dims[2] - 1))] >>
                                 24) != 0u;
        if (!overridden) {
            const float here = dtexel(d, g, x, y, z);
            const float kSigmaCt = 40.0f / 4095.0f;
            float s0 = 0.0f, s1 = 0.0f, weights = 0.0f;
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        const int n[3] = {min(max(cx + dx, 0), soft.dims[0] - 1), min(max(cy + dy, 0), soft.dims[1] - 1),
                                          min(max(cz + dz, 0), soft.dims[2] - 1)};
                        float dd[3];
                        for (int a = 0; a < 3; a++) dd[a] = float(n[a]) - xs[a];
                        const float w_space = expf(-(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]) / (2.0f * 0.6f * 0.6f));
                        const size_t ni = node_index(n[0], n[1], n[2]);
                        const float there = there_of ? there_of[ni] : soft_node_density(d, g, soft, sa, n);
                        const float w_ct = expf(-(here - there) * (here - there) / (2.0f * kSigmaCt * kSigmaCt));
                        const uint32_t t = soft.nodes[ni];
                        const uint32_t best = t & 0xffu, second = (t >> 8) & 0xffu;
                        const float mg = float((t >> 16) & 0xffu) / 16.0f;
                        const float sc0 = soft_score(has(best, kLabelSolid), has(second, kLabelSolid), mg);
                        const float sc1 = soft_score(has(best, kLabelGhosted), has(second, kLabelGhosted), mg);
                        const float w = w_space * w_ct + 1e-6f * w_space;
                        s0 += w * sc0;
                        s1 += w * sc1;
                        weights += w;
                    }
            m0 = 0.5f + (s0 / weights) / (2.0f * 4.0f);
            m1 = 0.5f + (s1 / weights) / (2.0f * 4.0f);
            m0 = m0 < 0.0f ? 0.0f : (m0 > 1.0f ? 1.0f : m0);
            m1 = m1 < 0.0f ? 0.0f : (m1 > 1.0f ? 1.0f : m1);
        }
    }
    if (mode == 1u && margin > 0 && label == 0u) {
        for (int a = -margin; a <= margin && m0 > 0.0f; a++)
            for (int b = -margin; b <= margin && m0 > 0.0f; b++)
                for (int c = -margin; c <= margin; c++) {
                    const int qx = min(max(x + a, 0), g.nx - 1), qy = min(max(y + b, 0), g.ny - 1),
                              qz = min(max(z + c, 0), g.nz - 1);
                    const uint32_t l = labels[g.at(qx, qy, qz)];
                    if (l != 0u && (!has(l, kLabelShown) || has(l, kLabelGhosted))) {
                        m0 = 0.0f;
                        break;
                    }
                }
    }
    out[o] = unorm16_bits(m0) | (unorm16_bits(m1) << 16);
}

// mask.comp:152-170: binomial weights C(2r, r + k) / 4^r along one axis. The
// weights c_j (c_0 = 1, c_j+1 = c_j (2r - j) / (j + 1), each step rounded to
// f32) and 4^r come from the host, computed with the same f32 operations. A
// workgroup covers 4 micro-bricks along x, one per wave, so a lane's voxel
// follows from its indices without divisions.
constexpr int kMaxBlurTaps = 17;  // radius <= 8
struct BlurTaps {
    float c[kMaxBlurTaps];
    float e;
};
__global__ void __launch_bounds__(kBlock) k_mask_blur(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                      int axis, int r, BrickGrid g, BlurTaps taps) {
    const int mbx = int(blockIdx.x) * (kBlock / 64) + int(threadIdx.x >> 6), w = int(threadIdx.x & 63u);
    if (mbx >= g.bx) return;
    const uint32_t o = uint32_t(mbx) * 64u + uint32_t(w) + blockIdx.y * g.ystride + blockIdx.z * g.zstride;
    const int x = 4 * mbx + (w & 3), y = 4 * int(blockIdx.y) + ((w >> 2) & 3), z = 4 * int(blockIdx.z) + (w >> 4);
    if (x >= g.nx || y >= g.ny || z >= g.nz) {
        dst[o] = 0u;
        return;
    }
    float s0 = 0.0f, s1 = 0.0f;
    const int n = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
    for (int j = 0; j <= 2 * r; j++) {
        int q[3] = {x, y, z};
        q[axis] = min(max(q[axis] + j - r, 0), n - 1);
        const uint32_t v = src[g.at(q[0], q[1], q[2])];
        s0 += taps.c[j] * (float(v & 0xffffu) / 65535.0f);
        s1 += taps.c[j] * (float(v >> 16) / 65535.0f);
    }
    dst[o] = unorm16_bits(s0 / taps.e) | (unorm16_bits(s1 / taps.e) << 16);
}

// majorant.comp:26-71. A workgroup covers 4 x 4 x 4 bricks. Every brick's window
// [8b - 2, 8b + 9] (clamped to the volume) is a union of 6^3 aligned 2^3-voxel
// cells, so the group first reduces the 18^3 cells of its bricks and their
// halo into LDS, then each brick reduces its 216 cells. Densities stay integer:
// int(d * 64) of d = c / 65535 equals floor(64 c / 65535) for every c (no
// quotient lies within f32 rounding of a bucket edge), and mask max > 0 iff the
// largest raw mask value is > 0.
constexpr int kTile = 4;                 // bricks per axis per workgroup
constexpr int kCells = 4 * kTile + 2;  // 18 cells per axis: 4 per brick and 1 of halo each side

__global__ void __launch_bounds__(kBlock) k_majorant(const uint16_t* __restrict__ d, const uint32_t* __restrict__ mask,
                                                     float* __restrict__ out, BrickGrid g, int mode, int mx, int my,
                                                     int mz, const vvr_uniforms* __restrict__ u) {
    __shared__ uint32_t s_minmax[kCells * kCells * kCells];  // min | max << 16
    __shared__ uint16_t s_mask[kCells * kCells * kCells];
    const int tbx = blockIdx.x * kTile, tby = blockIdx.y * kTile, tbz = blockIdx.z * kTile;
    // Cell c covers voxels 2c, 2c + 1; the tile's first cell is 4 * tb - 1. A
    // 4^3 micro-brick (64 contiguous voxels) holds 2^3 cells: the tile's cells
    // lie in micro-bricks 2 tb - 1 .. 2 tb + 2 kTile. Eight neighbouring lanes
    // read one micro-brick with one 16-byte load each (density; two of mask),
    // so a wave's loads cover whole lines, and combine their cells by shuffles.
    //   density chunk k (voxels 8k .. 8k + 7): rows dy = 2 (k & 1) + {0, 1} of
    //     dz = k >> 1, i.e. cells cx = 0, 1 of cy = k & 1, cz = k >> 2; lane
    //     k ^ 2 holds the cells' other dz;
    //   mask chunks k and k + 8 (4 voxels each): row dy = k & 3 of dz = k >> 2
    //     and 2 + (k >> 2), i.e. cells cx = 0, 1 of cy = (k >> 1) & 1 and
    //     cz = 0, 1; lanes k ^ 1 and k ^ 4 hold the cells' other rows.
    const int c0x = 4 * tbx - 1, c0y = 4 * tby - 1, c0z = 4 * tbz - 1;
    constexpr int kMb = 2 * kTile + 2;  // micro-bricks per axis
    static_assert((kMb * kMb * kMb * 8) % 64 == 0, "every wave runs the loop's iterations whole");
    for (int t = threadIdx.x; t < kMb * kMb * kMb * 8; t += kBlock) {
        const int m = t >> 3, k = t & 7;
        const int mbx = 2 * tbx - 1 + m % kMb, mby = 2 * tby - 1 + (m / kMb) % kMb, mbz = 2 * tbz - 1 + m / (kMb * kMb);
        uint32_t lo[2] = {0xffffu, 0xffffu}, hi[2] = {0u, 0u}, mm[4] = {0u, 0u, 0u, 0u};  // mm[cx + 2 cz]
        if (mbx >= 0 && mby >= 0 && mbz >= 0 && mbx < g.bx && mby < g.by && mbz < g.bz) {
            const uint32_t base = g.at(4 * mbx, 4 * mby, 4 * mbz);  // the micro-brick's first voxel
            // Voxel v = dx + 4 dy + 16 dz of the micro-brick; inside the volume?
            auto inside = [&](int v) {
                return 4 * mbx + (v & 3) < g.nx && 4 * mby + ((v >> 2) & 3) < g.ny && 4 * mbz + (v >> 4) < g.nz;
            };
            const uint4 w4 = reinterpret_cast<const uint4*>(d + base)[k];
            const uint32_t w[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int v = 8 * k + j;
                if (!inside(v)) continue;
                const uint32_t x = (j & 1) ? (w[j >> 1] >> 16) : (w[j >> 1] & 0xffffu);
                const int cx = (j >> 1) & 1;
                lo[cx] = min(lo[cx], x);
                hi[cx] = max(hi[cx], x);
            }
            if (mode != 0) {
                const uint4* mv = reinterpret_cast<const uint4*>(mask + base);
#pragma unroll
                for (int half = 0; half < 2; half++) {
                    const int q = k + 8 * half;
                    const uint4 m4 = mv[q];
                    const uint32_t mw[4] = {m4.x, m4.y, m4.z, m4.w};
#pragma unroll
                    for (int dx = 0; dx < 4; dx++) {
                        if (!inside(4 * q + dx)) continue;
                        uint32_t& c = mm[(dx >> 1) + 2 * half];
                        c = max(c, max(mw[dx] & 0xffffu, mw[dx] >> 16));
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 2; c++) {
            lo[c] = min(lo[c], uint32_t(__shfl_xor(int(lo[c]), 2)));
            hi[c] = max(hi[c], uint32_t(__shfl_xor(int(hi[c]), 2)));
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            mm[c] = max(mm[c], uint32_t(__shfl_xor(int(mm[c]), 1)));
            mm[c] = max(mm[c], uint32_t(__shfl_xor(int(mm[c]), 4)));
        }
        // Owners write the micro-brick's cells that lie in the tile (all of
        // them but halo halves): lanes 0, 1, 4, 5 the density cells, lanes 0, 2
        // the mask cells.
        auto in_tile = [&](int cx, int cy, int cz, int& i) {
            cx += 2 * mbx - c0x, cy += 2 * mby - c0y, cz += 2 * mbz - c0z;
            i = cx + kCells * (cy + kCells * cz);
            return cx >= 0 && cy >= 0 && cz >= 0 && cx < kCells && cy < kCells && cz < kCells;
        };
        int i;
        if ((k & 2) == 0)
#pragma unroll
            for (int cx = 0; cx < 2; cx++)
                if (in_tile(cx, k & 1, k >> 2, i)) s_minmax[i] = lo[cx] | (hi[cx] << 16);
        if ((k & 5) == 0)
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (in_tile(c & 1, (k >> 1) & 1, c >> 1, i)) s_mask[i] = uint16_t(mm[c]);
    }
    __syncthreads();
    if (threadIdx.x >= kTile * kTile * kTile) return;
    const int lx = threadIdx.x % kTile, ly = (threadIdx.x / kTile) % kTile, lz = threadIdx.x / (kTile * kTile);
    const int bx = tbx + lx, by = tby + ly, bz = tbz + lz;
    if (bx >= mx || by >= my || bz >= mz) return;
    uint32_t lo = 0xffffu, hi = 0u, mm = 0u;
    for (int k = 0; k < 6; k++)
        for (int j = 0; j < 6; j++)
            for (int i = 0; i < 6; i++) {
                const int c = (4 * lx + i) + kCells * ((4 * ly + j) + kCells * (4 * lz + k));
                const uint32_t v = s_minmax[c];
                lo = min(lo, v & 0xffffu);
                hi = max(hi, v >> 16);
                mm = max(mm, uint32_t(s_mask[c]));
            }
    float bound;
    auto tf_max = [&](uint32_t clo, uint32_t chi) {
        const int first = int((64u * clo) / 65535u), last = int((64u * chi) / 65535u);
        float m = 0.0f;
        for (int b = min(first, 63); b <= min(last, 63); b++) m = fmaxf(m, u->tf_bucket_max[b / 4][b % 4]);
        return m;
    };
    if (mode == 0) bound = tf_max(lo, hi);
    else if (mm == 0u) bound = 0.0f;
    else if (mode == 1) bound = tf_max(lo, hi);
    else bound = u->labels[2];
    out[size_t(bx) + size_t(mx) * (size_t(by) + size_t(my) * bz)] = bound;
}

inline unsigned blocks_for(size_t n) { return unsigned((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_upload_density(const int16_t* hu, uint16_t* out, const BrickGrid& g, hipStream_t st) {
    hipLaunchKernelGGL(k_upload_density, dim3(blocks_for(g.padded())), dim3(kBlock), 0, st, hu, out, g);
}

void launch_upload_labels(const uint8_t* labels, uint8_t* out, const BrickGrid& g, hipStream_t st) {
    hipLaunchKernelGGL(k_upload_labels, dim3(blocks_for(g.padded())), dim3(kBlock), 0, st, labels, out, g);
}

void launch_debrick(const void* in, void* out, const BrickGrid& g, int elem_bytes, hipStream_t st) {
    const size_t n = size_t(g.nx) * g.ny * g.nz;
    const dim3 grid(blocks_for(n));
    if (elem_bytes == 1)
        hipLaunchKernelGGL(k_debrick<uint8_t>, grid, dim3(kBlock), 0, st, (const uint8_t*)in, (uint8_t*)out, g);
    else if (elem_bytes == 2)
        hipLaunchKernelGGL(k_debrick<uint16_t>, grid, dim3(kBlock), 0, st, (const uint16_t*)in, (uint16_t*)out, g);
    else
        hipLaunchKernelGGL(k_debrick<uint32_t>, grid, dim3(kBlock), 0, st, (const uint32_t*)in, (uint32_t*)out, g);
}

void launch_normals(const uint16_t* density, uint32_t* normals, const BrickGrid& g, const vvr_uniforms* u,
                    hipStream_t st) {
    const dim3 grid(unsigned(g.bx + 1) / 2, unsigned(g.by + 1) / 2, unsigned(g.bz + 1) / 2);
    hipLaunchKernelGGL(k_normals, grid, dim3(kBlock), 0, st, density, normals, g, u);
}

void launch_mask_binary(const uint16_t* density, const uint8_t* labels, SoftGrid soft, int margin, uint32_t* dst,
                        void* scratch, size_t scratch_bytes, const BrickGrid& g, const vvr_uniforms* u,
                        hipStream_t st) {
    float* there = nullptr;
    const size_t nodes = size_t(soft.dims[0]) * size_t(soft.dims[1]) * size_t(soft.dims[2]);
    if ((soft.push[3] & 8) != 0 && nodes * sizeof(float) <= scratch_bytes) {
        there = static_cast<float*>(scratch);
        hipLaunchKernelGGL(k_soft_node_density, dim3(blocks_for(nodes)), dim3(kBlock), 0, st, density, soft, g, there);
    }
    hipLaunchKernelGGL(k_mask_binary, dim3(blocks_for(g.padded())), dim3(kBlock), 0, st, density, labels, soft,
                       static_cast<const float*>(there), margin, dst, g, u);
}

void launch_mask_blur(const uint32_t* src, uint32_t* dst, int axis, int radius, const BrickGrid& g, hipStream_t st) {
    if (radius < 0 || 2 * radius + 1 > kMaxBlurTaps) throw std::invalid_argument("mask blur radius out of range");
    BlurTaps taps{};
    float c = 1.0f;
    for (int j = 0; j <= 2 * radius; j++) {
        taps.c[j] = c;
        c = c * float(2 * radius - j) / float(j + 1);
    }
    taps.e = std::ldexp(1.0f, 2 * radius);
    const dim3 grid(unsigned(g.bx + kBlock / 64 - 1) / (kBlock / 64), unsigned(g.by), unsigned(g.bz));
    hipLaunchKernelGGL(k_mask_blur, grid, dim3(kBlock), 0, st, src, dst, axis, radius, g, taps);
}

void launch_majorant(const uint16_t* density, const uint32_t* mask, float* out, const BrickGrid& g, int mode,
                     const vvr_uniforms* u, hipStream_t st) {
    const int mx = (g.nx + 7) / 8, my = (g.ny + 7) / 8, mz = (g.nz + 7) / 8;
    const dim3 grid((mx + kTile - 1) / kTile, (my + kTile - 1) / kTile, (mz + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_majorant, grid, dim3(kBlock), 0, st, density, mask, out, g, mode, mx, my, mz, u);
}

}  // namespace vvr
