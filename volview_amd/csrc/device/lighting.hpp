// Lighting volumes of the ray marcher and their lookup: the key light's
// transmittance with the diffused subsurface light (shadow.comp), the ambient
// visibility (ao.comp), and lighting_lookup (common.glsl:141-156).
//
// Both volumes are linear arrays, x fastest: the shadow volume RGBA16F at
// ceil(dims / 2) (8 bytes a texel: diffused rgb, direct transmittance), the
// occlusion volume R16F at ceil(dims / 4). They are read as through the
// reference's border sampler (renderer.rs:645-648): a linear filter with exact
// weights, texels outside the image opaque white.
#pragma once
#include <hip/hip_fp16.h>

#include "scene.hpp"

namespace vvr {

__device__ __forceinline__ float half_bits(uint32_t bits) { return __half2float(__ushort_as_half((unsigned short)bits)); }

__device__ __forceinline__ uint2 pack_half4(float a, float b, float c, float d) {
    uint2 w;
    w.x = uint32_t(__half_as_ushort(__float2half_rn(a))) | (uint32_t(__half_as_ushort(__float2half_rn(b))) << 16);
    w.y = uint32_t(__half_as_ushort(__float2half_rn(c))) | (uint32_t(__half_as_ushort(__float2half_rn(d))) << 16);
    return w;
}

// During a sweep, the slice being written and the side the light comes from:
// a texel of that slice is read from the previous slice instead. The shader's
// fetch of the previous slice can give the slice being written a weight of a
// rounding error; this keeps the pass from reading what it writes. axis < 0:
// not sweeping.
struct SweepSlice {
    int axis, slice, sign;
};

__device__ __forceinline__ float4 shadow_texel(const DevScene& s, int x, int y, int z, SweepSlice k) {
    if (k.axis == 0 && x == k.slice) x += k.sign;
    else if (k.axis == 1 && y == k.slice) y += k.sign;
    else if (k.axis == 2 && z == k.slice) z += k.sign;
    if (uint32_t(x) >= uint32_t(s.sdims[0]) || uint32_t(y) >= uint32_t(s.sdims[1]) || uint32_t(z) >= uint32_t(s.sdims[2]))
        return make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    const uint2 w = s.shadow[uint32_t(x) + uint32_t(s.sdims[0]) * (uint32_t(y) + uint32_t(s.sdims[1]) * uint32_t(z))];
    return make_float4(half_bits(w.x & 0xffffu), half_bits(w.x >> 16), half_bits(w.y & 0xffffu), half_bits(w.y >> 16));
}

__device__ __forceinline__ float ao_texel(const DevScene& s, int x, int y, int z) {
    if (uint32_t(x) >= uint32_t(s.adims[0]) || uint32_t(y) >= uint32_t(s.adims[1]) || uint32_t(z) >= uint32_t(s.adims[2]))
        return 1.0f;
    return half_bits(s.ao[uint32_t(x) + uint32_t(s.adims[0]) * (uint32_t(y) + uint32_t(s.adims[1]) * uint32_t(z))]);
}

__device__ __forceinline__ float4 mix4(float4 p, float4 q, float a) {
    return make_float4(mixf(p.x, q.x, a), mixf(p.y, q.y, a), mixf(p.z, q.z, a), mixf(p.w, q.w, a));
}

// Linear fetch of the shadow volume at texel-space position t.
__device__ __forceinline__ float4 shadow_linear(const DevScene& s, V3 t, SweepSlice k) {
    const float fx = floorf(t.x), fy = floorf(t.y), fz = floorf(t.z);
    const float ax = t.x - fx, ay = t.y - fy, az = t.z - fz;
    const int x0 = fidx(fx), y0 = fidx(fy), z0 = fidx(fz);
    const float4 c00 = mix4(shadow_texel(s, x0, y0, z0, k), shadow_texel(s, x0 + 1, y0, z0, k), ax);
    const float4 c10 = mix4(shadow_texel(s, x0, y0 + 1, z0, k), shadow_texel(s, x0 + 1, y0 + 1, z0, k), ax);
    const float4 c01 = mix4(shadow_texel(s, x0, y0, z0 + 1, k), shadow_texel(s, x0 + 1, y0, z0 + 1, k), ax);
    const float4 c11 = mix4(shadow_texel(s, x0, y0 + 1, z0 + 1, k), shadow_texel(s, x0 + 1, y0 + 1, z0 + 1, k), ax);
    return mix4(mix4(c00, c10, ay), mix4(c01, c11, ay), az);
}

__device__ __forceinline__ float ao_linear(const DevScene& s, V3 t) {
    const float fx = floorf(t.x), fy = floorf(t.y), fz = floorf(t.z);
    const float ax = t.x - fx, ay = t.y - fy, az = t.z - fz;
    const int x0 = fidx(fx), y0 = fidx(fy), z0 = fidx(fz);
    const float c00 = mixf(ao_texel(s, x0, y0, z0), ao_texel(s, x0 + 1, y0, z0), ax);
    const float c10 = mixf(ao_texel(s, x0, y0 + 1, z0), ao_texel(s, x0 + 1, y0 + 1, z0), ax);
    const float c01 = mixf(ao_texel(s, x0, y0, z0 + 1), ao_texel(s, x0 + 1, y0, z0 + 1), ax);
    const float c11 = mixf(ao_texel(s, x0, y0 + 1, z0 + 1), ao_texel(s, x0 + 1, y0 + 1, z0 + 1), ax);
    return mixf(mixf(c00, c10, ay), mixf(c01, c11, ay), az);
}

// texture_cubic (common.glsl:141-151): 8 linear fetches at the B-spline taps,
// combined in the reference's order.
template <class T, class F>
__device__ __forceinline__ T cubic8(V3 t, F fetch) {
    const float ix = floorf(t.x), iy = floorf(t.y), iz = floorf(t.z);
    const Taps1 tx = cubic_taps1(t.x, ix), ty = cubic_taps1(t.y, iy), tz = cubic_taps1(t.z, iz);
    auto f = [&](float x, float y, float z) { return fetch(V3(x, y, z)); };
    auto add = [](T a, T b) { return a + b; };
    const T z0 = add(ty.g0 * add(tx.g0 * f(tx.h0, ty.h0, tz.h0), tx.g1 * f(tx.h1, ty.h0, tz.h0)),
                     ty.g1 * add(tx.g0 * f(tx.h0, ty.h1, tz.h0), tx.g1 * f(tx.h1, ty.h1, tz.h0)));
    const T z1 = add(ty.g0 * add(tx.g0 * f(tx.h0, ty.h0, tz.h1), tx.g1 * f(tx.h1, ty.h0, tz.h1)),
                     ty.g1 * add(tx.g0 * f(tx.h0, ty.h1, tz.h1), tx.g1 * f(tx.h1, ty.h1, tz.h1)));
    return add(tz.g0 * z0, tz.g1 * z1);
}

__device__ __forceinline__ float4 operator*(float a, float4 v) { return make_float4(a * v.x, a * v.y, a * v.z, a * v.w); }
__device__ __forceinline__ float4 operator+(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

__device__ __forceinline__ V3 index_to_tex(const vvr_uniforms* u, V3 idx) {  // common.glsl:68-70
    return V3((idx.x + 0.5f) / u->dims[0], (idx.y + 0.5f) / u->dims[1], (idx.z + 0.5f) / u->dims[2]);
}
__device__ __forceinline__ V3 tex_to_index(const vvr_uniforms* u, V3 t) {  // common.glsl:72-74
    return V3(t.x * u->dims[0] - 0.5f, t.y * u->dims[1] - 0.5f, t.z * u->dims[2] - 0.5f);
}

// lighting_lookup (common.glsl:154-156) at texture coordinate tc: B-spline
// filtered with smooth sampling, else one linear fetch.
__device__ __forceinline__ float4 shadow_lookup(const DevScene& s, V3 tc) {
    const V3 t(tc.x * float(s.sdims[0]) - 0.5f, tc.y * float(s.sdims[1]) - 0.5f, tc.z * float(s.sdims[2]) - 0.5f);
    const SweepSlice none{-1, 0, 0};
    if (s.u->env[1] > 0.5f) return cubic8<float4>(t, [&](V3 p) { return shadow_linear(s, p, none); });
    return shadow_linear(s, t, none);
}
__device__ __forceinline__ float ao_lookup(const DevScene& s, V3 tc) {
    const V3 t(tc.x * float(s.adims[0]) - 0.5f, tc.y * float(s.adims[1]) - 0.5f, tc.z * float(s.adims[2]) - 0.5f);
    if (s.u->env[1] > 0.5f) return cubic8<float>(t, [&](V3 p) { return ao_linear(s, p); });
    return ao_linear(s, t);
}

}  // namespace vvr
