// Device-side volume access and the shared fetch library of the ray marcher and
// path tracer: the HIP counterpart of shaders/common.glsl.
//
// Volumes live in HBM as 4^3-voxel micro-bricks (64 voxels, x fastest inside a
// brick, bricks x fastest): a 2^3 trilinear footprint touches 1-2 bricks on
// average and the 4^3 cubic B-spline footprint at most 8 instead of 16 rows of a
// linear layout. Filtering is done in software with the exact texture-unit
// model: weights frac(idx), texel indices clamped to the volume.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/vvr_uniforms.h"
#include "types.hpp"
#include "vec.hpp"

namespace vvr {

__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return min(max(v, lo), hi); }
// Float to int for an index that may lie far outside the volume.
__device__ __forceinline__ int fidx(float f) { return int(fminf(fmaxf(f, -1.0e8f), 1.0e8f)); }

constexpr float kInv65535 = 1.0f / 65535.0f;

// A unorm16 texel's value c / 65535, correctly rounded as a division would give
// it, without the division: the product with the rounded reciprocal is off by
// one ulp for 512 of the 65,536 codes, and one Newton correction of its
// residual (exact in an FMA) fixes every one of them.
__device__ __forceinline__ float unorm16_value(float c) {
    const float p = c * kInv65535;
    return fmaf(fmaf(-p, 65535.0f, c), kInv65535, p);
}
// A snorm8 texel's value max(c / 127, -1), correctly rounded likewise (16 of the
// 256 codes need the correction).
__device__ __forceinline__ float snorm8_value(float c) {
    constexpr float kInv127 = 1.0f / 127.0f;
    const float p = c * kInv127;
    return gmax(fmaf(fmaf(-p, 127.0f, c), kInv127, p), -1.0f);
}

// ---- linear fetches (common.glsl:76-78 and the mask / normal images) -----------

__device__ __forceinline__ float density_at_index(const DevScene& s, V3 idx) {
    const float fx = floorf(idx.x), fy = floorf(idx.y), fz = floorf(idx.z);
    const float ax = idx.x - fx, ay = idx.y - fy, az = idx.z - fz;
    const int x0 = fidx(fx), y0 = fidx(fy), z0 = fidx(fz);
    const BrickGrid& g = s.g;
    const uint32_t X0 = g.xoff(iclamp(x0, 0, g.nx - 1)), X1 = g.xoff(iclamp(x0 + 1, 0, g.nx - 1));
    const uint32_t Y0 = g.yoff(iclamp(y0, 0, g.ny - 1)), Y1 = g.yoff(iclamp(y0 + 1, 0, g.ny - 1));
    const uint32_t Z0 = g.zoff(iclamp(z0, 0, g.nz - 1)), Z1 = g.zoff(iclamp(z0 + 1, 0, g.nz - 1));
    const uint16_t* d = s.density;
    auto v = [&](uint32_t i) { return unorm16_value(float(d[i])); };
    const float c000 = v(X0 + Y0 + Z0), c100 = v(X1 + Y0 + Z0), c010 = v(X0 + Y1 + Z0), c110 = v(X1 + Y1 + Z0);
    const float c001 = v(X0 + Y0 + Z1), c101 = v(X1 + Y0 + Z1), c011 = v(X0 + Y1 + Z1), c111 = v(X1 + Y1 + Z1);
    const float c00 = mixf(c000, c100, ax), c10 = mixf(c010, c110, ax);
    const float c01 = mixf(c001, c101, ax), c11 = mixf(c011, c111, ax);
    return mixf(mixf(c00, c10, ay), mixf(c01, c11, ay), az);
}

struct F2 {
    float r, g;
};

__device__ __forceinline__ F2 mask_at(const DevScene& s, V3 idx) {
    const float fx = floorf(idx.x), fy = floorf(idx.y), fz = floorf(idx.z);
    const float ax = idx.x - fx, ay = idx.y - fy, az = idx.z - fz;
    const int x0 = fidx(fx), y0 = fidx(fy), z0 = fidx(fz);
    const BrickGrid& g = s.g;
    const uint32_t X0 = g.xoff(iclamp(x0, 0, g.nx - 1)), X1 = g.xoff(iclamp(x0 + 1, 0, g.nx - 1));
    const uint32_t Y0 = g.yoff(iclamp(y0, 0, g.ny - 1)), Y1 = g.yoff(iclamp(y0 + 1, 0, g.ny - 1));
    const uint32_t Z0 = g.zoff(iclamp(z0, 0, g.nz - 1)), Z1 = g.zoff(iclamp(z0 + 1, 0, g.nz - 1));
    const uint32_t* m = s.mask;
    const uint32_t t[8] = {m[X0 + Y0 + Z0], m[X1 + Y0 + Z0], m[X0 + Y1 + Z0], m[X1 + Y1 + Z0],
                           m[X0 + Y0 + Z1], m[X1 + Y0 + Z1], m[X0 + Y1 + Z1], m[X1 + Y1 + Z1]};
    F2 out;
#pragma unroll
    for (int ch = 0; ch < 2; ch++) {
        const int sh = 16 * ch;
        auto c = [&](int k) { return unorm16_value(float((t[k] >> sh) & 0xffffu)); };
        const float c00 = mixf(c(0), c(1), ax);
        const float c10 = mixf(c(2), c(3), ax);
        const float c01 = mixf(c(4), c(5), ax);
        const float c11 = mixf(c(6), c(7), ax);
        const float v = mixf(mixf(c00, c10, ay), mixf(c01, c11, ay), az);
        if (ch == 0) out.r = v;
        else out.g = v;
    }
    return out;
}

__device__ __forceinline__ float mask_channel(const DevScene& s, V3 idx, int ch) {
    const F2 m = mask_at(s, idx);
    return ch == 0 ? m.r : m.g;
}

struct F4 {
    float x, y, z, w;
};

__device__ __forceinline__ float snorm8(uint32_t word, int ch) {
    return snorm8_value(float(int8_t((word >> (8 * ch)) & 0xffu)));
}

__device__ __forceinline__ F4 normal_at(const DevScene& s, V3 idx) {
    const float fx = floorf(idx.x), fy = floorf(idx.y), fz = floorf(idx.z);
    const float ax = idx.x - fx, ay = idx.y - fy, az = idx.z - fz;
    const int x0 = fidx(fx), y0 = fidx(fy), z0 = fidx(fz);
    const BrickGrid& g = s.g;
    const uint32_t X0 = g.xoff(iclamp(x0, 0, g.nx - 1)), X1 = g.xoff(iclamp(x0 + 1, 0, g.nx - 1));
    const uint32_t Y0 = g.yoff(iclamp(y0, 0, g.ny - 1)), Y1 = g.yoff(iclamp(y0 + 1, 0, g.ny - 1));
    const uint32_t Z0 = g.zoff(iclamp(z0, 0, g.nz - 1)), Z1 = g.zoff(iclamp(z0 + 1, 0, g.nz - 1));
    const uint32_t* n = s.normals;
    const uint32_t t[8] = {n[X0 + Y0 + Z0], n[X1 + Y0 + Z0], n[X0 + Y1 + Z0], n[X1 + Y1 + Z0],
                           n[X0 + Y0 + Z1], n[X1 + Y0 + Z1], n[X0 + Y1 + Z1], n[X1 + Y1 + Z1]};
    float r[4];
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
        const float c00 = mixf(snorm8(t[0], ch), snorm8(t[1], ch), ax);
        const float c10 = mixf(snorm8(t[2], ch), snorm8(t[3], ch), ax);
        const float c01 = mixf(snorm8(t[4], ch), snorm8(t[5], ch), ax);
        const float c11 = mixf(snorm8(t[6], ch), snorm8(t[7], ch), ax);
        r[ch] = mixf(mixf(c00, c10, ay), mixf(c01, c11, ay), az);
    }
    return F4{r[0], r[1], r[2], r[3]};
}

// ---- cubic B-spline, 8 trilinear taps (common.glsl:84-136) --------------------------
//
// The 8 (or 24 for the gradient) trilinear fetches of Sigg and Hadwiger read only
// the 4^3 voxels floor(idx) - 1 .. floor(idx) + 2; they are loaded once and the
// taps are evaluated from registers in the reference's order of operations.

struct Taps1 {
    float h0, h1, g0, g1, dh0, dh1, dg0, dg1;
};

// x / 6 correctly rounded, as the division gives it, for |x| >= 2^-125 or x = 0,
// without the division: x / 2 is exact there, and (x / 2) / 3 is its product with
// RN(1/3) plus one FMA-exact correction of the residual, which is correctly
// rounded for every finite float dividend (checked exhaustively on the CPU, with
// subnormals, as the kernels run). The cubic weights' numerators below are 0 or
// at least 2^-27 apart from f^3, which keeps the division where it may be tiny.
__device__ __forceinline__ float div6_normal(float x) {
    constexpr float kThird = 1.0f / 3.0f;
    const float h = x * 0.5f;
    const float q = h * kThird;
    return fmaf(fmaf(-q, 3.0f, h), kThird, q);
}
__device__ __forceinline__ float div6(float x) {
    if (fabsf(x) < 0x1p-125f) return x / 6.0f;
    return div6_normal(x);
}

__device__ __forceinline__ Taps1 cubic_taps1(float idx, float i) {
    const float f = idx - i;
    const float f2 = f * f;
    const float f3 = f2 * f;
    const float w0 = div6_normal((1.0f - 3.0f * f) + 3.0f * f2 - f3);
    const float w1 = div6_normal((4.0f - 6.0f * f2) + 3.0f * f3);
    const float w2 = div6_normal(((1.0f + 3.0f * f) + 3.0f * f2) - 3.0f * f3);
    const float w3 = div6(f3);
    const float d0 = (-0.5f * (1.0f - f)) * (1.0f - f);
    const float d1 = (0.5f * f) * (3.0f * f - 4.0f);
    const float d2 = (0.5f * (1.0f - f)) * (3.0f * f + 1.0f);
    const float d3 = 0.5f * f2;
    Taps1 t;
    t.g0 = w0 + w1;
    t.g1 = w2 + w3;
    t.h0 = (i - 1.0f) + w1 / t.g0;
    t.h1 = (i + 1.0f) + w3 / t.g1;
    t.dg0 = d0 + d1;
    t.dg1 = d2 + d3;
    t.dh0 = (i - 1.0f) + d1 / gmin(t.dg0, -1e-7f);
    t.dh1 = (i + 1.0f) + d3 / gmax(t.dg1, 1e-7f);
    return t;
}

// A linear tap inside the 4-voxel window starting at i - 1: lower voxel offset and weight.
struct Tap {
    int o;    // 0..3, window offset of the lower texel
    float a;  // weight of the upper texel
};
__device__ __forceinline__ Tap tap_of(float h, float i) {
    const float fl = floorf(h);
    Tap t;
    t.o = int(fl - (i - 1.0f));  // 0, 1 (value/derivative low taps) or 2, 3 (high taps)
    t.a = h - fl;
    return t;
}

// Linear interpolation of 4 window values at tap t (the upper index only leaves
// the window with weight 0, so it is clamped to it).
__device__ __forceinline__ float lerp4(float v0, float v1, float v2, float v3, Tap t) {
    const float lo = t.o <= 0 ? v0 : (t.o == 1 ? v1 : (t.o == 2 ? v2 : v3));
    const float hi = t.o <= 0 ? v1 : (t.o == 1 ? v2 : v3);
    return mixf(lo, hi, t.a);
}

// Addresses of the 4 clamped voxel rows of the window floor(idx) - 1 .. + 2.
struct WindowAddr {
    uint32_t x[4], y[4], z[4];
};
__device__ __forceinline__ WindowAddr window_addr(const BrickGrid& g, int ix, int iy, int iz) {
    WindowAddr w;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        w.x[k] = g.xoff(iclamp(ix - 1 + k, 0, g.nx - 1));
        w.y[k] = g.yoff(iclamp(iy - 1 + k, 0, g.ny - 1));
        w.z[k] = g.zoff(iclamp(iz - 1 + k, 0, g.nz - 1));
    }
    return w;
}

// One z slice (4 x 4 texel values) of the window, from the unorm16 density or
// from its float copy (the same values, without the conversion: the ray marcher
// reads it; the path tracer's scattered first-bounce fetches run faster on the
// 2-byte texels).
__device__ __forceinline__ void load_slice(const uint16_t* __restrict__ d, const WindowAddr& w, int z, float v[4][4]) {
#pragma unroll
    for (int y = 0; y < 4; y++) {
        const uint32_t base = w.y[y] + w.z[z];
#pragma unroll
        for (int x = 0; x < 4; x++) v[y][x] = unorm16_value(float(d[w.x[x] + base]));
    }
}

// The window of the float copy in rows (DevScene::density_f): each of its 16
// rows is 4 consecutive floats, read with one 16-byte load, when the window's
// x range lies in the padded box (always inside the volume's box and one voxel
// around it); else texel by texel with clamped indices, the same values.
struct RowWindow {
    const float* row[4][4];  // [z][y]: the padded row's texel x = -kRowPad
    bool whole;              // the window's rows are 4 consecutive texels from x[0]
    int x[4];                // the 4 clamped x offsets in the padded row
};
__device__ __forceinline__ RowWindow row_window(const DevScene& s, int ix, int iy, int iz) {
    const BrickGrid& g = s.g;
    const size_t rx = size_t(g.nx + 2 * kRowPad), ry = size_t(g.ny + 2 * kRowPad);
    RowWindow w;
    const int x0 = ix - 1;
    w.whole = x0 >= -kRowPad && x0 + 3 <= g.nx - 1 + kRowPad;
#pragma unroll
    for (int k = 0; k < 4; k++) w.x[k] = iclamp(x0 + k, -kRowPad, g.nx - 1 + kRowPad) + kRowPad;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const size_t zz = size_t(iclamp(iz - 1 + c, -kRowPad, g.nz - 1 + kRowPad) + kRowPad);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const size_t yy = size_t(iclamp(iy - 1 + b, -kRowPad, g.ny - 1 + kRowPad) + kRowPad);
            w.row[c][b] = s.density_f + (zz * ry + yy) * rx;
        }
    }
    return w;
}
__device__ __forceinline__ void load_slice(const RowWindow& w, int z, float v[4][4]) {
    if (w.whole) {
#pragma unroll
        for (int y = 0; y < 4; y++) {
            float4 t;
            __builtin_memcpy(&t, w.row[z][y] + w.x[0], 16);
            v[y][0] = t.x, v[y][1] = t.y, v[y][2] = t.z, v[y][3] = t.w;
        }
    } else {
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int x = 0; x < 4; x++) v[y][x] = w.row[z][y][w.x[x]];
    }
}
// The slice's 4 rows lerped along x at tap t.
__device__ __forceinline__ void slice_x(const float v[4][4], Tap t, float out[4]) {
#pragma unroll
    for (int y = 0; y < 4; y++) out[y] = lerp4(v[y][0], v[y][1], v[y][2], v[y][3], t);
}
// Weight of window slice z in the lerp at tap t: 1 - a at o, a at o + 1.
__device__ __forceinline__ float tap_weight(Tap t, int z) {
    return z == t.o ? 1.0f - t.a : (z == t.o + 1 ? t.a : 0.0f);
}
// acc + c w, rounded like the terms of mixf, so that summing the slices in order
// (zero weights add an exact zero) reproduces mixf of the two weighted ones exactly.
__device__ __forceinline__ float accum_tap(float acc, float c, float w) { return acc + c * w; }

// cubic_fetch's final combination (common.glsl:121) of the 8 trilinear values.
__device__ __forceinline__ float fold8(const float T[2][2][2], float gx0, float gx1, float gy0, float gy1, float gz0,
                                       float gz1) {
    const float y0z0 = gx0 * T[0][0][0] + gx1 * T[1][0][0];
    const float y1z0 = gx0 * T[0][1][0] + gx1 * T[1][1][0];
    const float y0z1 = gx0 * T[0][0][1] + gx1 * T[1][0][1];
    const float y1z1 = gx0 * T[0][1][1] + gx1 * T[1][1][1];
    return gz0 * (gy0 * y0z0 + gy1 * y1z0) + gz1 * (gy0 * y0z1 + gy1 * y1z1);
}

// density_cubic (common.glsl:124-127). The 8 trilinear fetches of cubic_fetch
// (common.glsl:116-122) read only the 4^3 voxels around floor(idx): each lerps
// along x, then y, then z as a texture unit does, here a z slice of the window
// at a time. The value taps' lower texels sit at fixed window offsets: h0 -
// (i - 1) = w1 / (w0 + w1) lies in [0.8, 1] and h1 - (i + 1) = w3 / (w2 + w3) in
// [0, 0.2), so the low tap lerps offsets 0 and 1 and the high tap offsets 2 and
// 3. (Where h0 rounds up to i, the texture unit's tap is offset 1 at weight 0;
// the lerp of offsets 0 and 1 at weight 1 gives the same value, exactly.)
template <bool kFloatTexels = false>
__device__ inline float density_cubic(const DevScene& s, V3 idx) {
    const float ix = floorf(idx.x), iy = floorf(idx.y), iz = floorf(idx.z);
    const Taps1 tx = cubic_taps1(idx.x, ix), ty = cubic_taps1(idx.y, iy), tz = cubic_taps1(idx.z, iz);
    const float ax[2] = {tx.h0 - (ix - 1.0f), tx.h1 - (ix + 1.0f)};
    const float ay[2] = {ty.h0 - (iy - 1.0f), ty.h1 - (iy + 1.0f)};
    const float az[2] = {tz.h0 - (iz - 1.0f), tz.h1 - (iz + 1.0f)};
    WindowAddr w;
    RowWindow rw;
    if (kFloatTexels) rw = row_window(s, fidx(ix), fidx(iy), fidx(iz));
    else w = window_addr(s.g, fidx(ix), fidx(iy), fidx(iz));
    float Y[2][2][4];  // [x tap][y tap][z slice]: the fetches lerped along x and y
#pragma unroll
    for (int z = 0; z < 4; z++) {
        float v[4][4];
        if (kFloatTexels) load_slice(rw, z, v);
        else load_slice(s.density, w, z, v);
#pragma unroll
        for (int a = 0; a < 2; a++) {
            float r[4];
#pragma unroll
            for (int y = 0; y < 4; y++) r[y] = mixf(v[y][2 * a], v[y][2 * a + 1], ax[a]);
#pragma unroll
            for (int b = 0; b < 2; b++) Y[a][b][z] = mixf(r[2 * b], r[2 * b + 1], ay[b]);
        }
    }
    float T[2][2][2];  // [x tap][y tap][z tap]
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 2; c++) T[a][b][c] = mixf(Y[a][b][2 * c], Y[a][b][2 * c + 1], az[c]);
    return fold8(T, tx.g0, tx.g1, ty.g0, ty.g1, tz.g0, tz.g1);
}

// density_gradient_cubic (common.glsl:130-136), per index unit: the 24 trilinear
// fetches in one pass over the window.
__device__ inline V3 density_gradient_cubic(const DevScene& s, V3 idx) {
    const float ix = floorf(idx.x), iy = floorf(idx.y), iz = floorf(idx.z);
    const Taps1 tx = cubic_taps1(idx.x, ix), ty = cubic_taps1(idx.y, iy), tz = cubic_taps1(idx.z, iz);
    const Tap hx[2] = {tap_of(tx.h0, ix), tap_of(tx.h1, ix)}, dx[2] = {tap_of(tx.dh0, ix), tap_of(tx.dh1, ix)};
    const Tap hy[2] = {tap_of(ty.h0, iy), tap_of(ty.h1, iy)}, dy[2] = {tap_of(ty.dh0, iy), tap_of(ty.dh1, iy)};
    const Tap hz[2] = {tap_of(tz.h0, iz), tap_of(tz.h1, iz)}, dz[2] = {tap_of(tz.dh0, iz), tap_of(tz.dh1, iz)};
    const WindowAddr w = window_addr(s.g, fidx(ix), fidx(iy), fidx(iz));
    // [x tap][y tap][z tap] of the fetches of d/dx, d/dy and d/dz.
    float GX[2][2][2] = {}, GY[2][2][2] = {}, GZ[2][2][2] = {};
#pragma unroll
    for (int z = 0; z < 4; z++) {
        float v[4][4], rh[2][4], rd[2][4];
        load_slice(s.density, w, z, v);
#pragma unroll
        for (int a = 0; a < 2; a++) {
            slice_x(v, hx[a], rh[a]);
            slice_x(v, dx[a], rd[a]);
        }
        const float wh[2] = {tap_weight(hz[0], z), tap_weight(hz[1], z)};
        const float wd[2] = {tap_weight(dz[0], z), tap_weight(dz[1], z)};
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const float cx = lerp4(rd[a][0], rd[a][1], rd[a][2], rd[a][3], hy[b]);
                const float cy = lerp4(rh[a][0], rh[a][1], rh[a][2], rh[a][3], dy[b]);
                const float cz = lerp4(rh[a][0], rh[a][1], rh[a][2], rh[a][3], hy[b]);
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    GX[a][b][e] = accum_tap(GX[a][b][e], cx, wh[e]);
                    GY[a][b][e] = accum_tap(GY[a][b][e], cy, wh[e]);
                    GZ[a][b][e] = accum_tap(GZ[a][b][e], cz, wd[e]);
                }
            }
    }
    const float gx = fold8(GX, tx.dg0, tx.dg1, ty.g0, ty.g1, tz.g0, tz.g1);
    const float gy = fold8(GY, tx.g0, tx.g1, ty.dg0, ty.dg1, tz.g0, tz.g1);
    const float gz = fold8(GZ, tx.g0, tx.g1, ty.g0, ty.g1, tz.dg0, tz.dg1);
    return V3(gx, gy, gz);
}

// density_view (common.glsl:160-162).
// (kFloatTexels: the cubic window from the float copy.)
template <bool kFloatTexels = false>
__device__ __forceinline__ float density_view(const DevScene& s, V3 idx) {
    return s.u->env[1] > 0.5f ? density_cubic<kFloatTexels>(s, idx) : density_at_index(s, idx);
}

// ---- transfer function (common.glsl:170-230) ------------------------------------------

struct Sample {
    V3 albedo;
    float extinction, roughness, metalness;
    uint32_t label;
    bool ghost;
    float density, subsurface, coat;
};

__device__ __forceinline__ float4 transfer(const DevScene& s, float density) {
    const float t = density * 4095.0f;
    const float fl = floorf(t);
    const float a = t - fl;
    const int i0 = iclamp(fidx(fl), 0, 4095), i1 = iclamp(fidx(fl) + 1, 0, 4095);
    const float4 p = s.lut[i0], q = s.lut[i1];
    return make_float4(mixf(p.x, q.x, a), mixf(p.y, q.y, a), mixf(p.z, q.z, a), mixf(p.w, q.w, a));
}

__device__ __forceinline__ Sample tf_point(const DevScene& s, float density) {
    const float4 tf = transfer(s, density);
    Sample r;
    r.albedo = V3(tf.x, tf.y, tf.z);
    r.extinction = tf.w;
    r.roughness = s.u->material[0];
    r.metalness = s.u->material[1];
    r.label = 0u;
    r.ghost = false;
    r.density = density;
    r.subsurface = s.u->subsurface[0];
    r.coat = 0.0f;
    return r;
}

__device__ __forceinline__ float4 preint_at(const DevScene& s, float density) {
    const float x = clampf(density, 0.0f, 1.0f) * 4095.0f;
    const int i = min(int(x), 4094);
    const float a = x - float(i);
    const float4 p = s.preint[i], q = s.preint[i + 1];
    return make_float4(mixf(p.x, q.x, a), mixf(p.y, q.y, a), mixf(p.z, q.z, a), mixf(p.w, q.w, a));
}

__device__ inline Sample tf_segment(const DevScene& s, float front, float back) {
    if (fabsf(back - front) < 1e-4f) return tf_point(s, 0.5f * (front + back));
    const float4 pb = preint_at(s, back), pf = preint_at(s, front);
    const float den = back - front;
    const float4 mean = make_float4((pb.x - pf.x) / den, (pb.y - pf.y) / den, (pb.z - pf.z) / den, (pb.w - pf.w) / den);
    Sample r = tf_point(s, back);  // albedo of transfer(back) when the mean extinction vanishes
    if (mean.w > 1e-6f) r.albedo = V3(mean.x, mean.y, mean.z) / mean.w;
    r.extinction = gmax(mean.w, 0.0f);
    r.density = back;
    return r;
}

__device__ inline float preint_half_depth(const DevScene& s, float front, float back) {
    const float target = 0.5f * (preint_at(s, front).w + preint_at(s, back).w);
    float lo = 0.0f, hi = 1.0f;
    for (int i = 0; i < 6; i++) {
        const float mid = 0.5f * (lo + hi);
        const bool before = (preint_at(s, mixf(front, back, mid)).w < target) == (back > front);
        lo = before ? mid : lo;
        hi = before ? hi : mid;
    }
    return 0.5f * (lo + hi);
}

// ---- labels (common.glsl:232-351) --------------------------------------------------------

__device__ __forceinline__ bool label_shown(const vvr_uniforms* u, uint32_t l) { return u->label_colors[l][3] > 0.5f; }
__device__ __forceinline__ bool label_ghosted(const vvr_uniforms* u, uint32_t l) {
    return l != 0u && u->label_colors[l][3] > 1.5f;
}
__device__ __forceinline__ bool label_solid(const vvr_uniforms* u, uint32_t l, uint32_t mode) {
    if (label_ghosted(u, l)) return false;
    return mode == 1u ? (l == 0u || label_shown(u, l)) : (l != 0u && label_shown(u, l));
}

__device__ inline Sample with_label(const DevScene& sc, Sample s, uint32_t label, uint32_t mode, bool ghost) {
    const vvr_uniforms* u = sc.u;
    s.label = label;
    s.ghost = ghost;
    if (label == 0u) return s;
    const float4 color = *reinterpret_cast<const float4*>(u->label_colors[label]);
    const float4 mat = *reinterpret_cast<const float4*>(u->label_materials[label]);
    if (mode == 1u) {
        s.albedo = vmix(s.albedo, V3(color.x, color.y, color.z), u->labels[1]);
    } else {
        s.albedo = V3(color.x, color.y, color.z);
        s.extinction = u->labels[2];
        const float4 d = *reinterpret_cast<const float4*>(u->label_density[label]);
        if (mode == 3u && !ghost && d.w > 0.0f) {
            s.extinction *= smoothstepf(d.x, d.y, s.density);
            const float t = clampf((s.density - d.z) / d.w, -1.5f, 1.5f);
            s.albedo = vclamp(s.albedo * (1.0f + 0.2f * t), 0.0f, 1.0f);
            const float kBoneDensity = (300.0f + 1024.0f) / 4095.0f;
            if (d.z > kBoneDensity) {
                const float thin = 1.0f - smoothstepf(-1.5f, -0.5f, t);
                s.albedo = s.albedo * vmix(V3(1.0f), V3(0.92f, 0.6f, 0.5f), thin);
            }
        }
    }
    s.roughness = mat.x;
    s.metalness = mat.y;
    s.subsurface = mat.w;
    s.coat = mat.z;
    return s;
}

__device__ __forceinline__ uint32_t label_texel(const DevScene& s, int x, int y, int z) {
    return s.labels[s.g.at(x, y, z)];
}

// drawn_label_near (common.glsl:302-324): the nearest voxel, the 8 cell corners,
// then the 12 axis neighbours at +-1 and +-2.
__device__ inline bool drawn_label_near(const DevScene& s, V3 idx, uint32_t mode, bool ghost, uint32_t& label) {
    const int hx = s.g.nx - 1, hy = s.g.ny - 1, hz = s.g.nz - 1;
    const int nx = iclamp(fidx(floorf(idx.x + 0.5f)), 0, hx);
    const int ny = iclamp(fidx(floorf(idx.y + 0.5f)), 0, hy);
    const int nz = iclamp(fidx(floorf(idx.z + 0.5f)), 0, hz);
    const int bx = fidx(floorf(idx.x)), by = fidx(floorf(idx.y)), bz = fidx(floorf(idx.z));
    for (int i = 0; i < 21; i++) {
        int cx, cy, cz;
        if (i == 0) {
            cx = nx, cy = ny, cz = nz;
        } else if (i < 9) {
            cx = bx + ((i - 1) & 1), cy = by + (((i - 1) >> 1) & 1), cz = bz + ((i - 1) >> 2);
        } else {
            const int k = i - 9;
            const int d = ((k / 3) % 2 == 0 ? 1 : -1) * (k < 6 ? 1 : 2);
            const int a = k % 3;
            cx = nx + (a == 0 ? d : 0), cy = ny + (a == 1 ? d : 0), cz = nz + (a == 2 ? d : 0);
        }
        label = label_texel(s, iclamp(cx, 0, hx), iclamp(cy, 0, hy), iclamp(cz, 0, hz));
        if (ghost ? label_ghosted(s.u, label) : label_solid(s.u, label, mode)) return true;
    }
    return false;
}

// Whether a sample at idx may be drawn under a label: the bit of its mask cell
// (DevScene::mask_cells). A clear bit means every corner lies below the
// thresholds by at least 7.6e-6 (a unorm16 step), which the trilinear lerps'
// rounding cannot close, so mask_at would give m.r below the solid and m.g below
// the ghost threshold.
__device__ __forceinline__ uint32_t mask_cell(const DevScene& s, V3 idx) {
    const int cx = iclamp(fidx(floorf(idx.x)), -1, s.g.nx - 1) + 1;
    const int cy = iclamp(fidx(floorf(idx.y)), -1, s.g.ny - 1) + 1;
    const int cz = iclamp(fidx(floorf(idx.z)), -1, s.g.nz - 1) + 1;
    return s.mg.at(cx, cy, cz);
}

__device__ __forceinline__ Sample apply_labels(const DevScene& s, Sample smp, V3 idx) {
    const uint32_t mode = uint32_t(s.u->labels[0]);
    if (mode == 0u) return smp;
    if (s.mask_cells) {  // the branches below, without the mask and label loads
        const uint32_t c = mask_cell(s, idx);
        if (((s.mask_cells[c >> 5] >> (c & 31u)) & 1u) == 0u) {
            smp.extinction = 0.0f;
            return smp;
        }
        const uint32_t l = s.cell_labels[c];
        if (l != 0u) return with_label(s, smp, l, mode, false);
    }
    const F2 m = mask_at(s, idx);
    uint32_t label = 0u;
    if (m.r < (mode == 3u ? 0.25f : 0.5f)) {
        if (m.g >= 0.5f && drawn_label_near(s, idx, mode, true, label)) return with_label(s, smp, label, mode, true);
        smp.extinction = 0.0f;
        return smp;
    }
    if (drawn_label_near(s, idx, mode, false, label)) return with_label(s, smp, label, mode, false);
    if (mode >= 2u) smp.extinction = 0.0f;
    return smp;
}

// ---- clipping and geometry (common.glsl:357-442) ----------------------------------------

__device__ __forceinline__ float clip_depth(const vvr_uniforms* u, V3 idx, int& plane) {
    plane = -1;
    const uint32_t count = u->clip[0];
    if (count == 0u) return -1e30f;
    const bool all = u->clip[1] != 0u;
    float depth = all ? 1e30f : -1e30f;
    for (uint32_t i = 0; i < count; i++) {
        const float d = dot(V3(u->clip_planes[i][0], u->clip_planes[i][1], u->clip_planes[i][2]), idx) - u->clip_planes[i][3];
        if (all ? d < depth : d > depth) {
            depth = d;
            plane = int(i);
        }
    }
    return depth;
}

__device__ __forceinline__ bool is_clipped(const vvr_uniforms* u, V3 idx) {
    int plane;
    return clip_depth(u, idx, plane) > 0.0f;
}

// In isolate mode a drawn sample takes the label's albedo and extinction and an
// undrawn one has none (common.glsl:253-351): the density is never read, so
// its fetch is skipped. Every other mode reads it.
__device__ __forceinline__ bool density_unused(const vvr_uniforms* u) { return uint32_t(u->labels[0]) == 2u; }

__device__ __forceinline__ Sample sample_point(const DevScene& s, V3 idx) {
    const float d = density_unused(s.u) ? 0.0f : density_at_index(s, idx);
    Sample smp = apply_labels(s, tf_point(s, d), idx);
    if (smp.ghost || is_clipped(s.u, idx)) smp.extinction = 0.0f;
    return smp;
}

__device__ __forceinline__ M3 to_index(const vvr_uniforms* u) {
    const float* m = u->world_to_index;
    M3 r;
    r.c0 = V3(m[0], m[1], m[2]);
    r.c1 = V3(m[4], m[5], m[6]);
    r.c2 = V3(m[8], m[9], m[10]);
    return r;
}

__device__ __forceinline__ V3 world_to_index_point(const vvr_uniforms* u, V3 p) {
    const float* m = u->world_to_index;
    return V3(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12], m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]);
}

// mask_gradient (common.glsl:395-404): central differences of the trilinear
// mask at idx +- one voxel per axis. When every shifted position has the same
// fraction as idx (it does unless idx +- 1 rounds in f32), the six 2^3
// footprints cover only 32 distinct voxels, each loaded once; the lerps are the
// same as six mask_at calls.
__device__ inline V3 mask_gradient(const DevScene& s, V3 idx, int ch) {
    const V3 e[3] = {V3(1.f, 0.f, 0.f), V3(0.f, 1.f, 0.f), V3(0.f, 0.f, 1.f)};
    const float f0x = floorf(idx.x), f0y = floorf(idx.y), f0z = floorf(idx.z);
    const float ax = idx.x - f0x, ay = idx.y - f0y, az = idx.z - f0z;
    bool shared = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const V3 p = idx + e[a], m = idx - e[a];
        const float fp = floorf(component(p, a)), fm = floorf(component(m, a));
        const float f0 = component(V3(f0x, f0y, f0z), a), a0 = component(V3(ax, ay, az), a);
        shared = shared && fp == f0 + 1.0f && fm == f0 - 1.0f && component(p, a) - fp == a0 &&
                 component(m, a) - fm == a0;
    }
    if (!shared) {
        const float gx = mask_channel(s, idx + e[0], ch) - mask_channel(s, idx - e[0], ch);
        const float gy = mask_channel(s, idx + e[1], ch) - mask_channel(s, idx - e[1], ch);
        const float gz = mask_channel(s, idx + e[2], ch) - mask_channel(s, idx - e[2], ch);
        return to_index(s.u).mul_t(0.5f * V3(gx, gy, gz));
    }
    const BrickGrid& g = s.g;
    const int x0 = fidx(f0x), y0 = fidx(f0y), z0 = fidx(f0z);
    uint32_t X[4], Y[4], Z[4];  // voxel rows x0 - 1 .. x0 + 2 (clamped), likewise y and z
#pragma unroll
    for (int k = 0; k < 4; k++) {
        X[k] = g.xoff(iclamp(x0 - 1 + k, 0, g.nx - 1));
        Y[k] = g.yoff(iclamp(y0 - 1 + k, 0, g.ny - 1));
        Z[k] = g.zoff(iclamp(z0 - 1 + k, 0, g.nz - 1));
    }
    const int sh = 16 * ch;
    auto t = [&](int i, int j, int k) { return unorm16_value(float((s.mask[X[i] + Y[j] + Z[k]] >> sh) & 0xffffu)); };
    // Trilinear value with lower corner (i, j, k) in window indices: lerp x, y, z.
    auto tri = [&](float c000, float c100, float c010, float c110, float c001, float c101, float c011, float c111) {
        const float c00 = mixf(c000, c100, ax), c10 = mixf(c010, c110, ax);
        const float c01 = mixf(c001, c101, ax), c11 = mixf(c011, c111, ax);
        return mixf(mixf(c00, c10, ay), mixf(c01, c11, ay), az);
    };
    // x differences: rows y, z in {1, 2}, x in 0..3.
    float vx[2][2][4];
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int i = 0; i < 4; i++) vx[k][j][i] = t(i, j + 1, k + 1);
    const float xp = tri(vx[0][0][2], vx[0][0][3], vx[0][1][2], vx[0][1][3], vx[1][0][2], vx[1][0][3], vx[1][1][2],
                         vx[1][1][3]);
    const float xm = tri(vx[0][0][0], vx[0][0][1], vx[0][1][0], vx[0][1][1], vx[1][0][0], vx[1][0][1], vx[1][1][0],
                         vx[1][1][1]);
    // The central 2^3 (x, y, z in {1, 2}) is vx[..][..][1..2]; y differences add y rows 0 and 3.
    float y0r[2][2], y3r[2][2];  // [z][x]
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
            y0r[k][i] = t(i + 1, 0, k + 1);
            y3r[k][i] = t(i + 1, 3, k + 1);
        }
    const float yp = tri(vx[0][1][1], vx[0][1][2], y3r[0][0], y3r[0][1], vx[1][1][1], vx[1][1][2], y3r[1][0], y3r[1][1]);
    const float ym = tri(y0r[0][0], y0r[0][1], vx[0][0][1], vx[0][0][2], y0r[1][0], y0r[1][1], vx[1][0][1], vx[1][0][2]);
    float z0r[2][2], z3r[2][2];  // [y][x]
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
            z0r[j][i] = t(i + 1, j + 1, 0);
            z3r[j][i] = t(i + 1, j + 1, 3);
        }
    const float zp = tri(vx[1][0][1], vx[1][0][2], vx[1][1][1], vx[1][1][2], z3r[0][0], z3r[0][1], z3r[1][0], z3r[1][1]);
    const float zm = tri(z0r[0][0], z0r[0][1], z0r[1][0], z0r[1][1], vx[0][0][1], vx[0][0][2], vx[0][1][1], vx[0][1][2]);
    return to_index(s.u).mul_t(0.5f * V3(xp - xm, yp - ym, zp - zm));
}

__device__ __forceinline__ void intersect_volume(const vvr_uniforms* u, V3 origin, V3 dir, float& tn, float& tf) {
    const V3 lo(-0.5f);
    const V3 hi(u->dims[0] - 0.5f, u->dims[1] - 0.5f, u->dims[2] - 0.5f);
    const V3 inv = V3(1.0f) / dir;
    const V3 t0 = (lo - origin) * inv;
    const V3 t1 = (hi - origin) * inv;
    const V3 tmin = vmin(t0, t1);
    const V3 tmax = vmax(t0, t1);
    tn = gmax(gmax(tmin.x, tmin.y), gmax(tmin.z, 0.0f));
    tf = gmin(gmin(tmax.x, tmax.y), tmax.z);
}

// The majorant grid's brick of idx (clamped to the grid), or -1 when skipping
// is off (every brick then bounds by 1e30).
__device__ __forceinline__ int brick_index(const DevScene& s, V3 idx) {
    if (s.u->path[3] == 0u) return -1;
    const V3 b = vfloor((idx + 0.5f) / 8.0f);
    const int bx = iclamp(fidx(b.x), 0, s.mx - 1), by = iclamp(fidx(b.y), 0, s.my - 1),
              bz = iclamp(fidx(b.z), 0, s.mz - 1);
    return int(uint32_t(bx) + uint32_t(s.mx) * (uint32_t(by) + uint32_t(s.my) * uint32_t(bz)));
}
__device__ __forceinline__ float brick_majorant_at(const DevScene& s, int b) {
    return b < 0 ? 1e30f : s.majorant[b];
}
__device__ __forceinline__ float brick_majorant(const DevScene& s, V3 idx) {
    return brick_majorant_at(s, brick_index(s, idx));
}

// 1 / d_idx per axis, with brick_exit's guard against zero components.
__device__ __forceinline__ V3 brick_exit_inverse(V3 d_idx) {
    return V3(1.0f / (d_idx.x == 0.0f ? 1e-9f : d_idx.x), 1.0f / (d_idx.y == 0.0f ? 1e-9f : d_idx.y),
              1.0f / (d_idx.z == 0.0f ? 1e-9f : d_idx.z));
}

// brick_exit with the divisions by the direction taken as products with its
// inverse (within the 2.5 ulp the reference's GLSL division is allowed).
__device__ __forceinline__ float brick_exit_fast(V3 idx, V3 inv_d) {
    const V3 lo = vfloor((idx + 0.5f) / 8.0f) * 8.0f - 0.5f;
    const V3 hi = lo + 8.0f;
    const V3 t_far = vmax((lo - idx) * inv_d, (hi - idx) * inv_d);
    return gmax(gmin(t_far.x, gmin(t_far.y, t_far.z)), 0.0f);
}

__device__ __forceinline__ float brick_exit(V3 idx, V3 d_idx) {
    const V3 lo = vfloor((idx + 0.5f) / 8.0f) * 8.0f - 0.5f;
    const V3 hi = lo + 8.0f;
    const V3 d(d_idx.x == 0.0f ? 1e-9f : d_idx.x, d_idx.y == 0.0f ? 1e-9f : d_idx.y, d_idx.z == 0.0f ? 1e-9f : d_idx.z);
    const V3 t_far = vmax((lo - idx) / d, (hi - idx) / d);
    return gmax(gmin(t_far.x, gmin(t_far.y, t_far.z)), 0.0f);
}

__device__ __forceinline__ V3 camera_ray(const vvr_uniforms* u, float px, float py) {
    const float nx = px / float(u->flags[2]) * 2.0f - 1.0f;
    const float ny = py / float(u->flags[3]) * 2.0f - 1.0f;
    const V3 f(u->forward[0], u->forward[1], u->forward[2]);
    const V3 r(u->right[0], u->right[1], u->right[2]);
    const V3 up(u->up[0], u->up[1], u->up[2]);
    return normalize((f + nx * r) - ny * up);
}

}  // namespace vvr
