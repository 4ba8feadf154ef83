// Volume passes: density upload, normals (normals.comp), label masks
// (mask.comp) and the brick majorant (majorant.comp). They run once per volume
// or per opacity change and are HBM scans; their quantised outputs are
// reproduced bit for bit, so floating-point contraction is off in this file.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <type_traits>

#include <cmath>

#include "device/scene.hpp"
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

// The unorm16 density, and its values c / 65535 as floats (what the cubic
// fetch's window reads).
__global__ void k_upload_density(const int16_t* __restrict__ hu, uint16_t* __restrict__ out, BrickGrid g) {
    const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= g.padded()) return;
    int x, y, z;
    brick_coords(g, o, x, y, z);
    uint16_t v = 0;
    if (x < g.nx && y < g.ny && z < g.nz) v = hu_to_unorm16(hu[size_t(x) + size_t(g.nx) * (size_t(y) + size_t(g.ny) * z)]);
    out[o] = v;
}

// The float copy in rows (DevScene::density_f): texel (x, y, z) of the padded
// box -kRowPad .. n - 1 + kRowPad holds the value of the clamped voxel.
__global__ void k_upload_density_rows(const int16_t* __restrict__ hu, float* __restrict__ out, BrickGrid g) {
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    const size_t rx = size_t(g.nx + 2 * kRowPad), ry = size_t(g.ny + 2 * kRowPad);
    if (i >= row_volume_texels(g.nx, g.ny, g.nz)) return;
    const int x = iclamp(int(i % rx) - kRowPad, 0, g.nx - 1), y = iclamp(int((i / rx) % ry) - kRowPad, 0, g.ny - 1),
              z = iclamp(int(i / (rx * ry)) - kRowPad, 0, g.nz - 1);
    out[i] = unorm16_value(float(hu_to_unorm16(hu[size_t(x) + size_t(g.nx) * (size_t(y) + size_t(g.ny) * z)])));
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

// Mask cells (DevScene::mask_cells, cell_labels): one wave per 4^3-cell
// micro-brick of the cell grid cg (n + 1 cells per axis; cell c covers the
// clamped corners c - 1 and c). A cell's bit is set when its corners' largest
// solid code reaches solid_code or their largest ghost code reaches 32768 (0.5);
// its label byte is L when every corner's solid code reaches solid_code and
// every corner has label L, drawn solid in `mode` (label 0 excepted), else 0.
__global__ void __launch_bounds__(64) k_mask_cells(const uint32_t* __restrict__ mask,
                                                   const uint8_t* __restrict__ labels, BrickGrid g, BrickGrid cg,
                                                   uint32_t solid_code, uint32_t mode, const vvr_uniforms* __restrict__ u,
                                                   uint32_t* __restrict__ bits, uint8_t* __restrict__ cell_labels) {
    const uint32_t mb = blockIdx.x;
    const int bx = int(mb % uint32_t(cg.bx)), by = int((mb / uint32_t(cg.bx)) % uint32_t(cg.by));
    const int bz = int(mb / (uint32_t(cg.bx) * uint32_t(cg.by)));
    const int w = int(threadIdx.x);
    const int cx = 4 * bx + (w & 3), cy = 4 * by + ((w >> 2) & 3), cz = 4 * bz + (w >> 4);
    bool on = false;
    uint32_t uniform_label = 0u;
    if (cx <= g.nx && cy <= g.ny && cz <= g.nz) {
        uint32_t mr = 0u, mgh = 0u, lo = 0xffffu, l0 = 0u;
        bool same = true;
        for (int k = 0; k < 8; k++) {
            const int x = min(max(cx - 1 + (k & 1), 0), g.nx - 1);
            const int y = min(max(cy - 1 + ((k >> 1) & 1), 0), g.ny - 1);
            const int z = min(max(cz - 1 + (k >> 2), 0), g.nz - 1);
            const uint32_t at = g.at(x, y, z);
            const uint32_t v = mask[at];
            mr = max(mr, v & 0xffffu);
            lo = min(lo, v & 0xffffu);
            mgh = max(mgh, v >> 16);
            const uint32_t l = labels[at];
            if (k == 0) l0 = l;
            same = same && l == l0;
        }
        on = mr >= solid_code || mgh >= 32768u;
        if (lo >= solid_code && same && l0 != 0u && solid(u, l0, mode)) uniform_label = l0;
    }
    const uint64_t m = __ballot(on);
    if (w < 2) bits[size_t(mb) * 2u + uint32_t(w)] = uint32_t(m >> (32 * w));
    cell_labels[size_t(mb) * 64u + uint32_t(w)] = uint8_t(uniform_label);
}

// ---- majorant (majorant.comp) ----------------------------------------------------------
//
// Brick b (8^3 voxels) bounds the extinction over the voxels 8b - 2 .. 8b + 9 of
// each axis, clamped to the volume (majorant.comp:26-71): the 2^3-voxel cells
// 4b - 1 .. 4b + 4. Along an axis, micro-brick (4 voxels) 2j + 1 covers cells
// 4j + 2 and 4j + 3, both in brick j, the second also in brick j + 1; micro-brick
// 2j covers cells 4j (also in brick j - 1) and 4j + 1. So each micro-brick feeds
// its own brick with all its cells and, per axis, one neighbour with one layer.
//
// k_majorant_cells visits every 4^3 micro-brick (64 contiguous voxels) once:
// eight lanes read it with one 16-byte load each (two of the mask). Each lane
// folds its voxels into the micro-brick's 2^3 cells, the lanes combine by
// shuffles, then
//  - modes off and tint: the cells' density min | (65535 - max) << 16 go to a
//    32-byte record per micro-brick (so that one packed 16-bit minimum combines
//    two cells), and (tint) a byte with bit c set when a mask value (r or g) of
//    cell c is non-zero; k_majorant_bricks reduces each brick's window from a
//    tile of records staged in LDS and writes the bound;
//  - isolate and hybrid: the bound is the label opacity wherever a mask value
//    of the window is non-zero, else 0, so with the grid cleared to 0 the lanes
//    write the opacity straight into the (up to 8) bricks whose windows hold a
//    non-zero cell of the micro-brick (equal values; the order of the stores
//    does not matter). The density is not read.
// Voxels of partial micro-bricks outside the volume do not count.

__global__ void __launch_bounds__(kBlock) k_fill_f32(float* __restrict__ out, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = v;
}

// Micro-bricks per group of eight lanes (their loads are issued before any is
// used). One measured fastest: 83 us on the PT-1 volume against 89 with four,
// where fewer lanes are left to cover the latency, and on 256^3-512^3 volumes
// in mode off (24.8 against 28.0 us at 256^3).
constexpr int kMajRun = 1;

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
// Per-half minimum and maximum of two pairs of unsigned 16-bit values.
__device__ __forceinline__ uint32_t pk_min_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t,
                              __builtin_elementwise_min(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_max_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t,
                              __builtin_elementwise_max(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}

// kWhole: every micro-brick lies wholly inside the volume (each extent a
// multiple of 4), so no voxel needs the inside test and, outside the label
// modes, no lane needs its micro-brick's coordinates.
template <bool kWhole>
__global__ void __launch_bounds__(kBlock) k_majorant_cells(const uint16_t* __restrict__ d,
                                                           const uint32_t* __restrict__ mask,
                                                           uint32_t* __restrict__ cell_mm, uint8_t* __restrict__ cell_nz,
                                                           float* __restrict__ out, BrickGrid g, int mode, int mx,
                                                           int my, int mz, const vvr_uniforms* __restrict__ u,
                                                           bool with_density, bool with_mask) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t m0 = (t >> 3) * kMajRun, k = t & 7u;
    const uint32_t nmb = uint32_t(g.bx) * uint32_t(g.by) * uint32_t(g.bz);
    if (m0 >= nmb) return;  // whole groups of eight lanes, so the shuffles below stay within live groups
    // Loads of the group's micro-bricks m0 .. m0 + kMajRun - 1 (those that exist), stored in this
    // order, 64 voxels each: density chunk k (16 bytes), mask chunks k and k + 8.
    uint4 dw[kMajRun], mw[kMajRun][2];
#pragma unroll
    for (int i = 0; i < kMajRun; i++) {
        const uint32_t m = m0 + i;
        if (m >= nmb) break;
        const size_t base = size_t(m) * 64u;
        if (with_density) dw[i] = reinterpret_cast<const uint4*>(d + base)[k];
        if (with_mask) {
            mw[i][0] = reinterpret_cast<const uint4*>(mask + base)[k];
            mw[i][1] = reinterpret_cast<const uint4*>(mask + base)[k + 8];
        }
    }
#pragma unroll
    for (int i = 0; i < kMajRun; i++) {
        const uint32_t m = m0 + i;
        if (m >= nmb) break;  // uniform over the group
        int mbx = 0, mby = 0, mbz = 0, rx = 4, ry = 4, rz = 4;  // voxels left per axis, >= 1
        if (!kWhole || mode >= 2) {
            mbx = int(m % uint32_t(g.bx));
            mby = int((m / uint32_t(g.bx)) % uint32_t(g.by));
            mbz = int(m / (uint32_t(g.bx) * uint32_t(g.by)));
            if (!kWhole) rx = g.nx - 4 * mbx, ry = g.ny - 4 * mby, rz = g.nz - 4 * mbz;
        }
        // Voxel v = dx + 4 dy + 16 dz of the micro-brick lies inside the volume.
        auto inside = [&](int v) { return kWhole || ((v & 3) < rx && ((v >> 2) & 3) < ry && (v >> 4) < rz); };
        if (with_density) {
            // Chunk k holds voxels 8k .. 8k + 7 as four words of two: dz = k >> 1, dy = 2 (k & 1) + (j >> 1),
            // dx = 2 (j & 1) + {0, 1} for word j, i.e. words 0 and 2 hold cell cx = 0 of cy = k & 1,
            // cz = k >> 2, words 1 and 3 cell cx = 1; lane k ^ 2 holds the same cells' other dz.
            // Voxels outside the volume count as 65535 for the minimum and 0 for the maximum.
            const uint32_t w[4] = {dw[i].x, dw[i].y, dw[i].z, dw[i].w};
            uint32_t wmin[4], wmax[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t keep = (inside(int(8 * k) + 2 * j) ? 0xffffu : 0u) |
                                      (inside(int(8 * k) + 2 * j + 1) ? 0xffff0000u : 0u);
                wmin[j] = w[j] | ~keep;
                wmax[j] = w[j] & keep;
            }
            uint32_t rec[2];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const uint32_t mn = pk_min_u16(wmin[c], wmin[c + 2]), mxv = pk_max_u16(wmax[c], wmax[c + 2]);
                const uint32_t lo = min(mn & 0xffffu, mn >> 16), hi = max(mxv & 0xffffu, mxv >> 16);
                rec[c] = lo | ((0xffffu - hi) << 16);
                rec[c] = pk_min_u16(rec[c], uint32_t(__shfl_xor(int(rec[c]), 2)));
            }
            if ((k & 2u) == 0u) {  // lanes 0, 1, 4, 5: cells 2 (k & 1) + 4 (k >> 2) + {0, 1}
                const uint32_t c0 = 2u * (k & 1u) + 4u * (k >> 2);
                reinterpret_cast<uint2*>(cell_mm + size_t(m) * 8u + c0)[0] = make_uint2(rec[0], rec[1]);
            }
        }
        if (!with_mask) continue;
        // Mask chunks q = k and k + 8 (four texels each): dz = q >> 2, dy = q & 3, i.e. cells
        // cx = dx >> 1 of cy = (q >> 1) & 1, cz = q >> 3.
        uint32_t bits = 0u;
#pragma unroll
        for (int half = 0; half < 2; half++) {
            const int q = int(k) + 8 * half;
            const uint32_t tw[4] = {mw[i][half].x, mw[i][half].y, mw[i][half].z, mw[i][half].w};
#pragma unroll
            for (int dx = 0; dx < 4; dx++)
                if (tw[dx] != 0u && inside(4 * q + dx)) bits |= 1u << ((dx >> 1) + 2 * ((q >> 1) & 1) + 4 * (q >> 3));
        }
        bits |= uint32_t(__shfl_xor(int(bits), 1));
        bits |= uint32_t(__shfl_xor(int(bits), 2));
        bits |= uint32_t(__shfl_xor(int(bits), 4));
        if (mode == 1) {
            if (k == 0u) cell_nz[m] = uint8_t(bits);
            continue;
        }
        // Isolate / hybrid: lane k feeds the brick that is the neighbour along the axes
        // of k's set bits (own brick along the others) with the cells it holds there.
        const int mb[3] = {mbx, mby, mbz};
        int brick[3];
        uint32_t sel = 0xffu;  // cells of the micro-brick in that brick's window
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const bool nb = (k >> a) & 1u;
            const bool odd = mb[a] & 1;
            brick[a] = (mb[a] >> 1) + (nb ? (odd ? 1 : -1) : 0);
            if (nb) {  // one layer: the upper cell of an odd micro-brick, the lower of an even one
                const uint32_t layer = odd ? 1u : 0u;
                uint32_t keep = 0u;
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if (uint32_t((c >> a) & 1) == layer) keep |= 1u << c;
                sel &= keep;
            }
        }
        if ((bits & sel) && brick[0] >= 0 && brick[1] >= 0 && brick[2] >= 0 && brick[0] < mx && brick[1] < my &&
            brick[2] < mz)
            out[size_t(brick[0]) + size_t(mx) * (size_t(brick[1]) + size_t(my) * size_t(brick[2]))] = u->labels[2];
    }
}

// Maxima of the transfer function's 64 buckets over runs of 2^j buckets:
// s[j][b] = max of buckets b .. min(b + 2^j, 64) - 1, so that the maximum over any
// bucket range is that of two runs (exact: max neither rounds nor depends on the
// order). Built by the workgroup's first wave; read after a barrier.
__device__ __forceinline__ void build_run_max(float (&s)[7][64], const vvr_uniforms* __restrict__ u) {
    if (threadIdx.x >= 64) return;
    const int b = int(threadIdx.x);
    float v = u->tf_bucket_max[b / 4][b % 4];
    s[0][b] = v;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        v = fmaxf(v, __shfl(v, min(b + (1 << j), 63)));
        s[j + 1][b] = v;
    }
}

// The transfer function's buckets of the density codes lo .. hi (majorant.comp:
// bucket 64 c / 65535, at most 63); first > last when the window has no voxels
// (lo 65535, hi 0).
__device__ __forceinline__ uint32_t bucket_range(uint32_t lo, uint32_t hi) {
    const uint32_t first = min((64u * lo) / 65535u, 63u), last = min((64u * hi) / 65535u, 63u);
    return first | (last << 8);
}

// The largest bucket maximum over a range from bucket_range, 0 for none.
__device__ __forceinline__ float bucket_range_max(const float (&s)[7][64], uint32_t range) {
    const int first = int(range & 0xffu), last = int(range >> 8);
    if (first > last) return 0.0f;
    const int j = 31 - __clz(last - first + 1);
    return fmaxf(0.0f, fmaxf(s[j][first], s[j][last - (1 << j) + 1]));
}

// Modes off and tint: bounds from the cell records. A workgroup takes a tile of
// 4^3 bricks and stages the records of their 10^3-micro-brick window in LDS.
// Wave q reduces z layer q of the 4^3 micro-bricks of every brick's window (one
// brick per lane), so the cells each layer takes are known when compiling, and
// wave 0 combines the four layers and writes the bounds. (Tiles of 2^3 bricks in
// one-wave workgroups measured slower: 16 against 13 us on the DVR-1 volume.)
constexpr int kMajTile = 4;
constexpr int kMajMb = 2 * kMajTile + 2;  // micro-bricks per axis of a tile's windows
constexpr int kMajThreads = 256;

__global__ void __launch_bounds__(kMajThreads) k_majorant_bricks(const uint32_t* __restrict__ cell_mm,
                                                                 const uint8_t* __restrict__ cell_nz,
                                                                 uint16_t* __restrict__ ranges, float* __restrict__ out,
                                                                 BrickGrid g, int mode, int mx, int my, int mz,
                                                                 const vvr_uniforms* __restrict__ u) {
    __shared__ uint4 s_rec[kMajMb * kMajMb * kMajMb][2];
    __shared__ uint8_t s_nz[kMajMb * kMajMb * kMajMb];
    __shared__ float s_runmax[7][64];
    build_run_max(s_runmax, u);
    const int tbx = blockIdx.x * kMajTile, tby = blockIdx.y * kMajTile, tbz = blockIdx.z * kMajTile;
    const int m0x = 2 * tbx - 1, m0y = 2 * tby - 1, m0z = 2 * tbz - 1;
    // Every load of the staging first, then the LDS stores.
    constexpr int kN = kMajMb * kMajMb * kMajMb * 2, kIter = (kN + kMajThreads - 1) / kMajThreads;
    uint4 v[kIter];
    uint32_t nzv[kIter];
#pragma unroll
    for (int it = 0; it < kIter; it++) {
        const int i = threadIdx.x + it * kMajThreads;
        v[it] = make_uint4(~0u, ~0u, ~0u, ~0u);  // no voxels: min 65535, max 0
        nzv[it] = 0u;
        if (i >= kN) continue;
        const int r = i >> 1, h = i & 1;
        const int mbx = m0x + r % kMajMb, mby = m0y + (r / kMajMb) % kMajMb, mbz = m0z + r / (kMajMb * kMajMb);
        if (mbx >= 0 && mby >= 0 && mbz >= 0 && mbx < g.bx && mby < g.by && mbz < g.bz) {
            const size_t m = size_t(mbx) + size_t(g.bx) * (size_t(mby) + size_t(g.by) * size_t(mbz));
            v[it] = reinterpret_cast<const uint4*>(cell_mm + m * 8u)[h];
            if (mode == 1 && h == 0) nzv[it] = cell_nz[m];
        }
    }
#pragma unroll
    for (int it = 0; it < kIter; it++) {
        const int i = threadIdx.x + it * kMajThreads;
        if (i >= kN) continue;
        s_rec[i >> 1][i & 1] = v[it];
        if ((i & 1) == 0) s_nz[i >> 1] = uint8_t(nzv[it]);
    }
    __syncthreads();
    __shared__ uint32_t s_layer[4][64][2];  // each layer's packed minimum and cell bits, per brick
    const int layer = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)), lb = int(threadIdx.x & 63u);
    const int lx = lb % kMajTile, ly = (lb / kMajTile) % kMajTile, lz = lb / (kMajTile * kMajTile);
    // Cells of a micro-brick inside the window along an axis, by its place 0..3 in
    // the window: the upper cell of the first, both of the middle two, the lower of
    // the last.
    constexpr uint32_t kWin[4] = {2u, 3u, 3u, 1u};
    uint32_t acc = ~0u, nz = 0u;
    auto reduce = [&](auto wz_c) {
        constexpr int wz_i = decltype(wz_c)::value;
        constexpr uint32_t wz = kWin[wz_i];
#pragma unroll
        for (int wy_i = 0; wy_i < 4; wy_i++)
#pragma unroll
            for (int wx_i = 0; wx_i < 4; wx_i++) {
                const uint32_t wx = kWin[wx_i], wy = kWin[wy_i];
                uint32_t sel = 0u;
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if (((wx >> (c & 1)) & 1u) && ((wy >> ((c >> 1) & 1)) & 1u) && ((wz >> (c >> 2)) & 1u))
                        sel |= 1u << c;
                const int r = (2 * lx + wx_i) + kMajMb * ((2 * ly + wy_i) + kMajMb * (2 * lz + wz_i));
                uint32_t w[8] = {~0u, ~0u, ~0u, ~0u, ~0u, ~0u, ~0u, ~0u};
                if (sel & 0x0fu) {
                    const uint4 a = s_rec[r][0];
                    w[0] = a.x, w[1] = a.y, w[2] = a.z, w[3] = a.w;
                }
                if (sel & 0xf0u) {
                    const uint4 b = s_rec[r][1];
                    w[4] = b.x, w[5] = b.y, w[6] = b.z, w[7] = b.w;
                }
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if ((sel >> c) & 1u) acc = pk_min_u16(acc, w[c]);
                if (mode == 1) nz |= uint32_t(s_nz[r]) & sel;
            }
    };
    if (layer == 0)
        reduce(std::integral_constant<int, 0>{});
    else if (layer == 3)
        reduce(std::integral_constant<int, 3>{});
    else if (layer == 1)
        reduce(std::integral_constant<int, 1>{});
    else
        reduce(std::integral_constant<int, 2>{});
    s_layer[layer][lb][0] = acc;
    s_layer[layer][lb][1] = nz;
    __syncthreads();
    const int bx = tbx + lx, by = tby + ly, bz = tbz + lz;
    if (layer != 0 || bx >= mx || by >= my || bz >= mz) return;
#pragma unroll
    for (int q = 1; q < 4; q++) {
        acc = pk_min_u16(acc, s_layer[q][lb][0]);
        nz |= s_layer[q][lb][1];
    }
    const uint32_t range = bucket_range(acc & 0xffffu, 0xffffu - (acc >> 16));
    const size_t brick = size_t(bx) + size_t(mx) * (size_t(by) + size_t(my) * size_t(bz));
    ranges[brick] = uint16_t(range);
    out[brick] = (mode == 1 && nz == 0u) ? 0.0f : bucket_range_max(s_runmax, range);
}

// Mode off after a transfer-function change: each brick's bound from its kept
// bucket range.
__global__ void __launch_bounds__(kBlock) k_majorant_ranges(const uint16_t* __restrict__ ranges,
                                                            float* __restrict__ out, uint32_t bricks,
                                                            const vvr_uniforms* __restrict__ u) {
    __shared__ float s_runmax[7][64];
    build_run_max(s_runmax, u);
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < bricks) out[i] = bucket_range_max(s_runmax, ranges[i]);
}

inline unsigned blocks_for(size_t n) { return unsigned((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_upload_density(const int16_t* hu, uint16_t* out, float* out_f, const BrickGrid& g, hipStream_t st) {
    hipLaunchKernelGGL(k_upload_density, dim3(blocks_for(g.padded())), dim3(kBlock), 0, st, hu, out, g);
    hipLaunchKernelGGL(k_upload_density_rows, dim3(blocks_for(row_volume_texels(g.nx, g.ny, g.nz))), dim3(kBlock), 0,
                       st, hu, out_f, g);
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

void launch_mask_cells(const uint32_t* mask, const uint8_t* labels, const BrickGrid& g, int mode,
                       const vvr_uniforms* u, uint32_t* bits, uint8_t* cell_labels, hipStream_t st) {
    const BrickGrid cg = mask_cell_grid(g);
    const uint32_t solid_code = mode == 3 ? 16384u : 32768u;  // 0.25 (hybrid) or 0.5, as unorm16 codes rounded up
    hipLaunchKernelGGL(k_mask_cells, dim3(unsigned(size_t(cg.bx) * cg.by * cg.bz)), dim3(64), 0, st, mask, labels, g,
                       cg, solid_code, uint32_t(mode), u, bits, cell_labels);
}

void launch_majorant(const uint16_t* density, const uint32_t* mask, float* out, const BrickGrid& g, int mode,
                     const vvr_uniforms* u, const MajorantCells& c, hipStream_t st) {
    const int mx = (g.nx + 7) / 8, my = (g.ny + 7) / 8, mz = (g.nz + 7) / 8;
    const size_t nmb = size_t(g.bx) * g.by * g.bz, bricks = size_t(mx) * my * mz;
    if (mode == 0 && !c.density) {
        hipLaunchKernelGGL(k_majorant_ranges, dim3(blocks_for(bricks)), dim3(kBlock), 0, st, c.ranges, out,
                           uint32_t(bricks), u);
        return;
    }
    if (mode >= 2)
        hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(bricks)), dim3(kBlock), 0, st, out, uint32_t(bricks), 0.0f);
    const bool with_density = mode <= 1 && c.density, with_mask = mode >= 2 || (mode == 1 && c.mask);
    if (with_density || with_mask) {
        const size_t groups = (nmb + kMajRun - 1) / kMajRun;
        const bool whole = g.nx % 4 == 0 && g.ny % 4 == 0 && g.nz % 4 == 0;
        hipLaunchKernelGGL(whole ? k_majorant_cells<true> : k_majorant_cells<false>, dim3(blocks_for(groups * 8)),
                           dim3(kBlock), 0, st, density, mask, c.cells, c.nz, out, g, mode, mx, my, mz, u,
                           with_density, with_mask);
    }
    if (mode <= 1) {
        const dim3 grid((mx + kMajTile - 1) / kMajTile, (my + kMajTile - 1) / kMajTile, (mz + kMajTile - 1) / kMajTile);
        hipLaunchKernelGGL(k_majorant_bricks, grid, dim3(kMajThreads), 0, st, c.cells, c.nz, c.ranges, out, g, mode,
                           mx, my, mz, u);
    }
}

}  // namespace vvr
