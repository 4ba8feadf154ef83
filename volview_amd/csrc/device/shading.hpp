// Surface description and reflectance of the ray marcher and the path tracer:
// the HIP counterpart of shaders/shading.glsl.
#pragma once
#include "scene.hpp"

namespace vvr {

__device__ __forceinline__ float ggx_distribution(float nh, float a2) {  // shading.glsl:4-7
    const float denom = nh * nh * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * denom * denom);
}

// shading.glsl:11-27
__device__ __forceinline__ V3 brdf(V3 n, V3 v, V3 l, V3 albedo, float roughness, float metalness, V3& diffuse) {
    const V3 h = normalize(v + l);
    const float nv = gmax(dot(n, v), 1e-4f);
    const float nl = gmax(dot(n, l), 1e-4f);
    const float nh = gmax(dot(n, h), 0.0f);
    const float vh = gmax(dot(v, h), 0.0f);
    const float a = gmax(roughness * roughness, 2e-3f);
    const float d = ggx_distribution(nh, a * a);
    const float k = a * 0.5f;
    const float g = (nv / (nv * (1.0f - k) + k)) * (nl / (nl * (1.0f - k) + k));
    const V3 f0 = vmix(V3(0.028f), albedo, metalness);
    const V3 f = f0 + (1.0f - f0) * vpow(1.0f - vh, 5.0f);
    const V3 specular = d * g * f / (4.0f * nv * nl);
    diffuse = (1.0f - f) * (1.0f - metalness) * albedo / kPi;
    return specular;
}

__device__ __forceinline__ float coat_roughness(const Sample& s) { return clampf(0.5f * s.roughness, 0.1f, 0.25f); }

__device__ __forceinline__ float smith(float nv, float nl, float a) {
    const float k = a * 0.5f;
    return (nv / (nv * (1.0f - k) + k)) * (nl / (nl * (1.0f - k) + k));
}

// shading.glsl:45-58
__device__ __forceinline__ V3 brdf_coated(V3 n, V3 v, V3 l, const Sample& s, V3& diffuse) {
    const V3 specular = brdf(n, v, l, s.albedo, s.roughness, s.metalness, diffuse);
    if (s.coat <= 0.0f) return specular;
    const V3 h = normalize(v + l);
    const float nv = gmax(dot(n, v), 1e-4f);
    const float nl = gmax(dot(n, l), 1e-4f);
    const float a = coat_roughness(s) * coat_roughness(s);
    const float f = s.coat * (0.02f + 0.98f * vpow(1.0f - gmax(dot(v, h), 0.0f), 5.0f));
    const float coat = ggx_distribution(gmax(dot(n, h), 0.0f), a * a) * smith(nv, nl, a) * f / (4.0f * nv * nl);
    diffuse = diffuse * (1.0f - f);
    return specular * (1.0f - f) + V3(coat);
}

__device__ __forceinline__ float ign(float px, float py) {  // shading.glsl:61-63
    return fractf(52.9829189f * fractf(px * 0.06711056f + py * 0.00583715f));
}

// shading.glsl:73-144.
//
// `lazy`: the caller uses n only where surfaceness > 0 (a path tracer event
// after the first, pt.comp:454-458). Solid labelled tissue in isolate mode is
// shaded by its mask: n is the mask normal wherever the mask gradient is
// nonzero, and where it vanishes the surfaceness is 0, so the normal volume and
// the B-spline gradient are read only when they can matter.
__device__ inline float surface_frame(const DevScene& sc, V3 idx, const Sample& s, V3 v, int cap_plane, float cut,
                                      float cap_depth, V3& n, float& surfaceness, bool lazy = false) {
    const vvr_uniforms* u = sc.u;
    const uint32_t label_mode = uint32_t(u->labels[0]);
    if (lazy && label_mode == 2u && !s.ghost && s.label != 255u) {
        const V3 g = mask_gradient(sc, idx, 0);
        const float g_len = length(g);
        surfaceness = clampf(g_len * 4.0f, 0.0f, 1.0f);
        if (g_len > 1e-5f) {
            n = (-g) / g_len;
        } else if (surfaceness > 0.0f) {
            const F4 nw = normal_at(sc, idx);
            const V3 nxyz(nw.x, nw.y, nw.z);
            const float n_len = length(nxyz);
            n = n_len > 1e-3f ? nxyz / n_len : V3(0.0f);
            if (u->env[1] > 0.5f && n_len > 1e-3f) {
                const V3 gw = to_index(u).mul_t(density_gradient_cubic(sc, idx));
                const float len = length(gw);
                n = len > 1e-7f ? (-gw) / len : n;
            }
        } else {
            n = V3(0.0f);  // surfaceness 0: never read
        }
        if (cap_plane >= 0 && -cut < cap_depth) {
            n = V3(u->clip_normals[cap_plane][0], u->clip_normals[cap_plane][1], u->clip_normals[cap_plane][2]);
            surfaceness = 1.0f;
        }
        return 1.0f;
    }
    const F4 nw = normal_at(sc, idx);
    const V3 nxyz(nw.x, nw.y, nw.z);
    const float n_len = length(nxyz);
    n = n_len > 1e-3f ? nxyz / n_len : V3(0.0f);
    surfaceness = nw.w;
    const bool skin = s.label == 255u;
    if (u->env[1] > 0.5f && n_len > 1e-3f && (!s.ghost || skin)) {
        const V3 gw = to_index(u).mul_t(density_gradient_cubic(sc, idx));
        const float len = length(gw);
        n = len > 1e-7f ? (-gw) / len : n;
    }
    const bool hollow = u->label_density[s.label][3] < 0.0f;
    if (skin) {
    } else if (!s.ghost && label_mode == 3u && !hollow) {
        const V3 g = mask_gradient(sc, idx, 0);
        const float g_len = length(g);
        const V3 mask_n = g_len > 1e-5f ? (-g) / g_len : n;
        const V3 blended = vmix(mask_n, n, surfaceness);
        n = length(blended) > 1e-3f ? normalize(blended) : mask_n;
        surfaceness = gmax(surfaceness, clampf(g_len * 4.0f, 0.0f, 1.0f));
    } else if (!s.ghost && (label_mode >= 2u || (label_mode == 1u && mask_at(sc, idx).r < 0.98f))) {
        const V3 g = mask_gradient(sc, idx, 0);
        const float g_len = length(g);
        const float mask_surface = clampf(g_len * 4.0f, 0.0f, 1.0f);
        if (label_mode >= 2u || mask_surface > surfaceness) {
            n = g_len > 1e-5f ? (-g) / g_len : n;
            surfaceness = label_mode >= 2u ? mask_surface : gmax(surfaceness, mask_surface);
        }
    }
    if (cap_plane >= 0 && !s.ghost && -cut < cap_depth) {
        n = V3(u->clip_normals[cap_plane][0], u->clip_normals[cap_plane][1], u->clip_normals[cap_plane][2]);
        surfaceness = 1.0f;
    }
    if (s.ghost) {
        const V3 g = mask_gradient(sc, idx, 1);
        const float g_len = length(g);
        if (!skin || n_len <= 1e-3f) n = g_len > 1e-5f ? (-g) / g_len : n;
        surfaceness = clampf(g_len * 4.0f, 0.0f, 1.0f);
        const float edge_on = vpow(1.0f - fabsf(dot(n, v)), 3.0f);
        return surfaceness * mixf(u->labels[3], 1.0f, edge_on);
    }
    return 1.0f;
}

__device__ __forceinline__ bool measured(const vvr_uniforms* u, const Sample& s) {  // shading.glsl:147-149
    return u->labels[0] >= 2.0f && s.label != 0u && !s.ghost && u->label_optics[s.label][3] > 0.0f;
}

__device__ __forceinline__ V3 tissue_sigma(const vvr_uniforms* u, const Sample& s) {  // shading.glsl:161-168
    const float4 o = *reinterpret_cast<const float4*>(u->label_optics[s.label]);
    V3 mu(o.x, o.y, o.z);
    mu = vmix(V3(mu.y), mu, u->env[3]);
    mu = mu / (1.0f - clampf(u->env[2], 0.0f, 0.95f));
    mu = mu * gmin(1.0f, u->tissue[1] / gmax(gmax(mu.x, gmax(mu.y, mu.z)), 1e-6f));
    const float occupancy = clampf(s.extinction / gmax(u->labels[2], 1e-6f), 0.0f, 1.0f);
    return occupancy * mu;
}

// ---- procedural surface texture (shading.glsl:175-342) ----------------------------------

struct I3 {
    int x, y, z;
};

__device__ __forceinline__ V3 hash3(I3 c) {
    uint32_t vx = uint32_t(c.x) * 1664525u + 1013904223u;
    uint32_t vy = uint32_t(c.y) * 1664525u + 1013904223u;
    uint32_t vz = uint32_t(c.z) * 1664525u + 1013904223u;
    vx += vy * vz;
    vy += vz * vx;
    vz += vx * vy;
    vx ^= vx >> 16u;
    vy ^= vy >> 16u;
    vz ^= vz >> 16u;
    vx += vy * vz;
    vy += vz * vx;
    vz += vx * vy;
    const float k = 1.0f / 4294967296.0f;
    return V3(float(vx) * k, float(vy) * k, float(vz) * k);
}

__device__ __forceinline__ float gradient_at(I3 c, V3 offset) { return dot(hash3(c) * 2.0f - 1.0f, offset); }

__device__ inline float gradient_noise(V3 p) {
    const V3 fl = vfloor(p);
    const I3 i{fidx(fl.x), fidx(fl.y), fidx(fl.z)};
    const V3 f = p - fl;
    const V3 w = f * f * f * (f * (f * 6.0f - 15.0f) + 10.0f);
    auto I = [&](int a, int b, int c) { return I3{i.x + a, i.y + b, i.z + c}; };
    auto F = [&](int a, int b, int c) { return f - V3(float(a), float(b), float(c)); };
    const float x00 = mixf(gradient_at(I(0, 0, 0), F(0, 0, 0)), gradient_at(I(1, 0, 0), F(1, 0, 0)), w.x);
    const float x10 = mixf(gradient_at(I(0, 1, 0), F(0, 1, 0)), gradient_at(I(1, 1, 0), F(1, 1, 0)), w.x);
    const float x01 = mixf(gradient_at(I(0, 0, 1), F(0, 0, 1)), gradient_at(I(1, 0, 1), F(1, 0, 1)), w.x);
    const float x11 = mixf(gradient_at(I(0, 1, 1), F(0, 1, 1)), gradient_at(I(1, 1, 1), F(1, 1, 1)), w.x);
    return 1.7f * mixf(mixf(x00, x10, w.y), mixf(x01, x11, w.y), w.z);
}

__device__ inline float fbm(V3 p) {
    return (gradient_noise(p) + 0.5f * gradient_noise(p * 2.03f + 17.1f) + 0.25f * gradient_noise(p * 4.07f + 31.7f)) /
           1.75f;
}

__device__ inline void cellular(V3 p, float& nearest, float& border, I3& cell) {
    const V3 fl = vfloor(p);
    const I3 i{fidx(fl.x), fidx(fl.y), fidx(fl.z)};
    const V3 f = p - fl;
    V3 d1(8.0f), d2(8.0f);
    cell = i;
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                const I3 c{i.x + x, i.y + y, i.z + z};
                const V3 d = V3(float(x), float(y), float(z)) + hash3(c) - f;
                if (dot(d, d) < dot(d1, d1)) {
                    d2 = d1;
                    d1 = d;
                    cell = c;
                } else if (dot(d, d) < dot(d2, d2)) {
                    d2 = d;
                }
            }
    nearest = length(d1);
    border = dot(0.5f * (d1 + d2), normalize(d2 - d1));
}

__device__ inline float lobule_border(V3 q, I3& lobule) {
    const V3 warp(gradient_noise(q * 1.7f), gradient_noise(q * 1.7f + 5.2f), gradient_noise(q * 1.7f + 9.4f));
    float nearest, border;
    cellular(q + 0.2f * warp, nearest, border, lobule);
    return border;
}

__device__ __noinline__ float surface_height(V3 p, float4 params) {
    const V3 q = p / params.w;
    if (params.x > 1.5f) {
        I3 lobule;
        const float bulge = smoothstepf(0.0f, 0.35f, lobule_border(q, lobule));
        return 0.7f * bulge + 0.3f * fbm(q * 3.0f);
    }
    return 0.6f * fbm(q) + 0.4f * gradient_noise(q * 4.0f);
}

// inverse(mat3(world_to_index)) * (idx - translation), by the adjugate.
__device__ inline V3 index_to_world(const vvr_uniforms* u, V3 idx) {
    const M3 m = to_index(u);
    const V3 r0 = cross(m.c1, m.c2), r1 = cross(m.c2, m.c0), r2 = cross(m.c0, m.c1);
    const float det = dot(m.c0, r0);
    const V3 q = idx - V3(u->world_to_index[12], u->world_to_index[13], u->world_to_index[14]);
    return V3(dot(r0, q), dot(r1, q), dot(r2, q)) / det;
}

__device__ __forceinline__ bool textured(const vvr_uniforms* u, const Sample& s) {
    return s.label != 0u && !s.ghost && u->label_surface[s.label][0] > 0.5f;
}

__device__ inline void apply_surface_texture(const vvr_uniforms* u, V3 idx, Sample& s) {
    if (!textured(u, s)) return;
    const float4 params = *reinterpret_cast<const float4*>(u->label_surface[s.label]);
    const float amount = params.y;
    const V3 p = index_to_world(u, idx);
    const V3 q = p / params.w;
    V3 tint;
    float wet, rough;
    if (params.x > 1.5f) {
        const float strength = gmin(1.0f, 5.0f * amount);
        I3 lobule;
        const float border = lobule_border(q, lobule);
        const float septum =
            (1.0f - smoothstepf(0.0f, 0.06f, border)) * smoothstepf(-0.7f, 0.3f, gradient_noise(q * 6.0f));
        const float near_septum = 1.0f - smoothstepf(0.0f, 0.2f, border);
        const float patches = fbm(q * 1.5f);
        const float mottle = fbm(q * 4.0f);
        tint = V3(1.0f + amount * (hash3(lobule).x * 2.0f - 1.0f + 0.5f * mottle));
        tint = tint * vmix(V3(1.0f), V3(0.84f, 0.78f, 0.88f), strength * smoothstepf(-0.1f, 0.4f, -patches));
        tint = tint * vmix(V3(1.0f), V3(0.6f, 0.57f, 0.66f), strength * septum);
        float nearest, unused;
        I3 speck;
        cellular(p / 3.0f, nearest, unused, speck);
        const V3 h = hash3(I3{speck.x + 7919, speck.y + 7919, speck.z + 7919});
        const float speck_mask = (h.x < 0.03f + 0.4f * near_septum ? 1.0f : 0.0f) *
                                 (1.0f - smoothstepf(0.12f, 0.2f, nearest / (0.6f + 0.8f * h.y)));
        tint = tint * mixf(1.0f, 0.2f, strength * speck_mask);
        wet = 1.0f - 0.3f * near_septum;
        rough = 1.0f + 0.2f * mottle;
    } else {
        const float patches = fbm(q);
        const float grain = gradient_noise(q * 4.0f);
        tint = V3(1.0f + amount * (0.7f * patches + 0.3f * grain));
        tint = tint * vmix(V3(1.0f), V3(0.97f, 0.94f, 1.0f), clampf(-patches, 0.0f, 1.0f) * gmin(1.0f, 5.0f * amount));
        wet = clampf(0.8f + 0.4f * patches, 0.4f, 1.2f);
        rough = 1.0f + 0.3f * grain;
    }
    s.albedo = vclamp(s.albedo * tint, 0.0f, 1.0f);
    s.roughness = clampf(s.roughness * rough, 0.02f, 1.0f);
    s.coat = clampf(s.coat * wet, 0.0f, 1.0f);
}

__device__ inline V3 bump_normal(const vvr_uniforms* u, V3 idx, const Sample& s, V3 n) {
    if (!textured(u, s) || u->label_surface[s.label][2] <= 0.0f || dot(n, n) < 0.5f) return n;
    const float4 params = *reinterpret_cast<const float4*>(u->label_surface[s.label]);
    const V3 p = index_to_world(u, idx);
    const float e = 0.02f * params.w;
    const float h = surface_height(p, params);
    const V3 g = V3(surface_height(p + V3(e, 0.0f, 0.0f), params) - h, surface_height(p + V3(0.0f, e, 0.0f), params) - h,
                    surface_height(p + V3(0.0f, 0.0f, e), params) - h) *
                 (params.z / e);
    return normalize(n - (g - dot(g, n) * n));
}

}  // namespace vvr
