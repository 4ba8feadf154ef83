// Lighting volumes of the ray marcher (renderer.rs:1707-1749, 1907-1940):
//
//   k_shadow_slice  shadow.comp: one slice of the shadow volume perpendicular to
//                   the light's dominant axis, from the slice before it on the
//                   lit side (a serial recurrence: one launch per slice, in order
//                   on one stream; a slice is 10^4-10^5 texels, each a handful of
//                   fetches)
//   k_ao            ao.comp: mean transmittance over 16 directions within 15 mm
//
// Both sample the scene with sample_point (common.glsl:385-391): trilinear
// density, labels, clipping.
#include <hip/hip_runtime.h>

#include "device/lighting.hpp"
#include "launch.hpp"

namespace vvr {

namespace {

// shadow.comp:18-68 for texel p of slice p[axis].
__device__ __forceinline__ float4 shadow_pass(const DevScene& s, int px, int py, int pz, SweepSlice k) {
    const vvr_uniforms* u = s.u;
    const V3 sd(float(s.sdims[0]), float(s.sdims[1]), float(s.sdims[2]));
    V3 l = to_index(u).mul(V3(u->light_dir[0], u->light_dir[1], u->light_dir[2]));
    l = V3(l.x * (sd.x / u->dims[0]), l.y * (sd.y / u->dims[1]), l.z * (sd.z / u->dims[2]));
    const float la = k.axis == 0 ? l.x : (k.axis == 1 ? l.y : l.z);
    const float step = 1.0f / gmax(fabsf(la), 1e-6f);
    const V3 here = V3(float(px), float(py), float(pz));
    const V3 prev = here + l * step;
    const V3 mid = (here + prev) * 0.5f + V3(0.5f);
    const Sample m = sample_point(s, tex_to_index(u, V3(mid.x / sd.x, mid.y / sd.y, mid.z / sd.z)));
    const float4 at_prev = shadow_linear(s, prev, k);
    const float direct = at_prev.w * vexp(-m.extinction * step);
    // The diffused light gathered from a disc of the previous slice (radius
    // subsurface.z shadow texels along the slice's two axes).
    const float r = u->subsurface[2];
    const V3 ou(k.axis == 2 ? r : 0.0f, k.axis == 0 ? r : 0.0f, k.axis == 1 ? r : 0.0f);  // (axis + 1) % 3
    const V3 ov(k.axis == 1 ? r : 0.0f, k.axis == 2 ? r : 0.0f, k.axis == 0 ? r : 0.0f);  // (axis + 2) % 3
    V3 gathered = V3(at_prev.x, at_prev.y, at_prev.z) * 0.4f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        V3 offset = (i & 1) == 0 ? ou : ov;
        offset = offset * (i < 2 ? 1.0f : -1.0f);
        const float4 g = shadow_linear(s, prev + offset, k);
        gathered = gathered + V3(g.x, g.y, g.z) * 0.15f;
    }
    const float mx = gmax(gmax(m.albedo.x, gmax(m.albedo.y, m.albedo.z)), 1e-4f);
    const V3 chroma = m.albedo / mx;
    const V3 ext = (V3(1.2f) - chroma) * (m.extinction * u->subsurface[1]);
    return make_float4(gathered.x * vexp(-ext.x * step), gathered.y * vexp(-ext.y * step),
                       gathered.z * vexp(-ext.z * step), direct);
}

// Threads run along the slice's axis with the smaller memory stride (x where
// the slice contains it).
__global__ void __launch_bounds__(64) k_shadow_slice(DevScene s, uint2* __restrict__ out, int axis, int slice,
                                                     int sign) {
    const int ua = (axis + 1) % 3, va = (axis + 2) % 3;
    const int fast = va == 0 ? va : ua;
    const int a = int(blockIdx.x * 64u + threadIdx.x), b = int(blockIdx.y);
    const int p0 = axis == 0 ? slice : (fast == 0 ? a : b);
    const int p1 = axis == 1 ? slice : (fast == 1 ? a : b);
    const int p2 = axis == 2 ? slice : (fast == 2 ? a : b);
    if (p0 >= s.sdims[0] || p1 >= s.sdims[1] || p2 >= s.sdims[2]) return;
    const float4 v = shadow_pass(s, p0, p1, p2, SweepSlice{axis, slice, sign});
    out[uint32_t(p0) + uint32_t(s.sdims[0]) * (uint32_t(p1) + uint32_t(s.sdims[1]) * uint32_t(p2))] =
        pack_half4(v.x, v.y, v.z, v.w);
}

// ao.comp:20-43.
__global__ void __launch_bounds__(64) k_ao(DevScene s, uint16_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    const uint32_t nx = uint32_t(s.adims[0]), ny = uint32_t(s.adims[1]), nz = uint32_t(s.adims[2]);
    if (i >= nx * ny * nz) return;
    const uint32_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    const vvr_uniforms* u = s.u;
    const V3 center = tex_to_index(u, V3((float(x) + 0.5f) / float(nx), (float(y) + 0.5f) / float(ny),
                                         (float(z) + 0.5f) / float(nz)));
    const M3 ti = to_index(u);
    const float dt = 15.0f / 8.0f;
    float visible = 0.0f;
    for (int d = 0; d < 16; d++) {
        const float zc = 1.0f - (float(d) + 0.5f) * (2.0f / 16.0f);
        const float r = sqrtf(1.0f - zc * zc);
        const float phi = float(d) * 2.39996323f;
        const V3 dir = ti.mul(V3(r * cr_cos(phi), r * cr_sin(phi), zc));
        float depth = 0.0f;
        for (int j = 0; j < 8; j++) depth += sample_point(s, center + dir * ((float(j) + 0.5f) * dt)).extinction * dt;
        visible += vexp(-depth);
    }
    out[i] = __half_as_ushort(__float2half_rn(gmin(2.0f * visible / 16.0f, 1.0f)));
}

}  // namespace

void shadow_sweep_axis(const vvr_uniforms& u, int& axis, int& sign) {
    // Renderer::record_shadow_sweep (renderer.rs:1914-1928): the light's
    // dominant axis in shadow-volume index units; the sweep starts on its lit side.
    float l[3];
    for (int a = 0; a < 3; a++) {
        const float* m = u.world_to_index;
        const float v = (m[a] * u.light_dir[0] + m[4 + a] * u.light_dir[1]) + m[8 + a] * u.light_dir[2];
        const float sd = float((int(u.dims[a]) + 1) / 2);
        l[a] = v * (sd / u.dims[a]);
    }
    const float ax = fabsf(l[0]), ay = fabsf(l[1]), az = fabsf(l[2]);
    axis = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    sign = l[axis] >= 0.0f ? 1 : -1;
}

void launch_shadow_sweep(const DevScene& s, uint2* out, const vvr_uniforms& hu, hipStream_t st) {
    int axis, sign;
    shadow_sweep_axis(hu, axis, sign);
    const int ua = (axis + 1) % 3, va = (axis + 2) % 3;
    const int fast = va == 0 ? va : ua, slow = fast == ua ? va : ua;
    const dim3 grid((s.sdims[fast] + 63) / 64, s.sdims[slow]);
    const int n = s.sdims[axis];
    for (int i = 0; i < n; i++) {
        const int slice = sign > 0 ? n - 1 - i : i;
        hipLaunchKernelGGL(k_shadow_slice, grid, dim3(64), 0, st, s, out, axis, slice, sign);
    }
}

void launch_ao(const DevScene& s, uint16_t* out, hipStream_t st) {
    const uint32_t n = uint32_t(s.adims[0]) * uint32_t(s.adims[1]) * uint32_t(s.adims[2]);
    hipLaunchKernelGGL(k_ao, dim3((n + 63) / 64), dim3(64), 0, st, s, out);
}

}  // namespace vvr
