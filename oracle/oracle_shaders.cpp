// Test infrastructure only: the CPU oracle of the volume-rendering hot path.
// Nothing in the shipped library includes, links or calls this.
//
// Line-by-line restatement of the reference's compute shaders:
//   shaders/common.glsl, shaders/shading.glsl, shaders/majorant.comp,
//   shaders/normals.comp, shaders/mask.comp, shaders/dvr.comp, shaders/pt.comp,
//   shaders/shadow.comp, shaders/ao.comp, shaders/denoise.comp, shaders/bloom.comp,
//   shaders/tonemap.comp.
// Each function names the lines it follows. Templated on the real type R: the
// FP32 oracle (R = float, built without FMA contraction, so every operation is
// one IEEE-rounded step in source order) and the FP64 oracle (R = double).
//
// Texture-unit model. The reference samples through Vulkan samplers. The oracle
// models an exact texture unit: a linear fetch at continuous voxel index idx
// (texture coordinate (idx + 0.5) / dims) filters with weights frac(idx) between
// texels floor(idx) and floor(idx) + 1, each index clamped to the image
// (CLAMP_TO_EDGE, renderer.rs:637-640); unorm16 texels read c / 65535, snorm8
// texels max(c / 127, -1), fp16 texels their exact value; the 1D LUT is
// addressed at density * 4095 (common.glsl:170-173). Real texture units round the
// sub-texel weight (Vulkan requires >= 4 bits, most have 8); that rounding is
// implementation-defined and not modelled. The lighting volumes are read through
// the border sampler (renderer.rs:645-648, 1983-1984): the same filter, with
// texels outside the image opaque white (1, 1, 1, 1).
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle_api.h"
#include "oracle_math.hpp"

using namespace oracle;

namespace {

template <class R> struct V4 {
    R x, y, z, w;
};

template <class R> struct Ctx {
    const oracle_scene* s;
    const vvr_uniforms* u;
    int dims[3];
    int mdims[3];
    uint32_t rng;

    int sdims[3];  // shadow volume: ceil(dims / 2) (SHADOW_DOWNSAMPLE, renderer.rs:35-36, 608-613)
    int adims[3];  // ambient occlusion: ceil(dims / 4) (AO_DOWNSAMPLE, renderer.rs:39-40, 614-618)
    const uint16_t* shadow_img;
    // During a sweep: the slice being written. The shader's linear fetch of the
    // previous slice may give the slice being written a weight of a rounding
    // error (its position lands a hair off the texel centre); that texel is
    // read from the previous slice instead, so the sweep never reads what the
    // same pass writes.
    int sweep_axis = -1, sweep_slice = 0, sweep_sign = 0;

    Ctx(const oracle_scene* sc, const vvr_uniforms* un) : s(sc), u(un), rng(1) {
        for (int a = 0; a < 3; a++) {
            dims[a] = (int)sc->dims[a];
            mdims[a] = (dims[a] + 7) / 8;
            sdims[a] = (dims[a] + 1) / 2;
            adims[a] = (dims[a] + 3) / 4;
        }
        shadow_img = sc->shadow;
    }

    static int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
    size_t vox(int x, int y, int z) const { return (size_t)x + (size_t)dims[0] * ((size_t)y + (size_t)dims[1] * z); }

    // ---- texel fetches ---------------------------------------------------------
    R density_texel(int x, int y, int z) const {
        x = iclamp(x, 0, dims[0] - 1), y = iclamp(y, 0, dims[1] - 1), z = iclamp(z, 0, dims[2] - 1);
        return R(s->density[vox(x, y, z)]) / R(65535);
    }
    R mask_texel(int x, int y, int z, int ch) const {
        x = iclamp(x, 0, dims[0] - 1), y = iclamp(y, 0, dims[1] - 1), z = iclamp(z, 0, dims[2] - 1);
        return R(s->mask[2 * vox(x, y, z) + ch]) / R(65535);
    }
    R normal_texel(int x, int y, int z, int ch) const {
        x = iclamp(x, 0, dims[0] - 1), y = iclamp(y, 0, dims[1] - 1), z = iclamp(z, 0, dims[2] - 1);
        R v = R(s->normals[4 * vox(x, y, z) + ch]) / R(127);
        return gmax(v, R(-1));
    }
    uint32_t label_texel(V3<int> c) const {
        if (!s->labels) return 0;
        return s->labels[vox(c.x, c.y, c.z)];
    }

    // Trilinear filter at texel-space position `t` of a texel function.
    template <class F> R trilinear(V3<R> t, F texel) const {
        R fx = std::floor(t.x), fy = std::floor(t.y), fz = std::floor(t.z);
        R ax = t.x - fx, ay = t.y - fy, az = t.z - fz;
        int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
        R c00 = mixr(texel(x0, y0, z0), texel(x0 + 1, y0, z0), ax);
        R c10 = mixr(texel(x0, y0 + 1, z0), texel(x0 + 1, y0 + 1, z0), ax);
        R c01 = mixr(texel(x0, y0, z0 + 1), texel(x0 + 1, y0, z0 + 1), ax);
        R c11 = mixr(texel(x0, y0 + 1, z0 + 1), texel(x0 + 1, y0 + 1, z0 + 1), ax);
        return mixr(mixr(c00, c10, ay), mixr(c01, c11, ay), az);
    }

    // common.glsl:76-78
    R density_at_index(V3<R> idx) const {
        return trilinear(idx, [this](int x, int y, int z) { return density_texel(x, y, z); });
    }
    R mask_at(V3<R> idx, int ch) const {
        return trilinear(idx, [this, ch](int x, int y, int z) { return mask_texel(x, y, z, ch); });
    }
    V4<R> normal_at(V3<R> idx) const {
        V4<R> r;
        r.x = trilinear(idx, [this](int x, int y, int z) { return normal_texel(x, y, z, 0); });
        r.y = trilinear(idx, [this](int x, int y, int z) { return normal_texel(x, y, z, 1); });
        r.z = trilinear(idx, [this](int x, int y, int z) { return normal_texel(x, y, z, 2); });
        r.w = trilinear(idx, [this](int x, int y, int z) { return normal_texel(x, y, z, 3); });
        return r;
    }

    // ---- cubic B-spline (common.glsl:84-136) --------------------------------------
    struct Taps {
        V3<R> h0, h1, g0, g1, dh0, dh1, dg0, dg1;
    };
    static Taps cubic_taps(V3<R> idx) {
        V3<R> i = vfloor(idx);
        V3<R> f = idx - i;
        V3<R> f2 = f * f;
        V3<R> f3 = f2 * f;
        V3<R> w0 = ((R(1) - R(3) * f) + R(3) * f2 - f3) / R(6);
        V3<R> w1 = ((R(4) - R(6) * f2) + R(3) * f3) / R(6);
        V3<R> w2 = (((R(1) + R(3) * f) + R(3) * f2) - R(3) * f3) / R(6);
        V3<R> w3 = f3 / R(6);
        V3<R> d0 = (R(-0.5) * (R(1) - f)) * (R(1) - f);
        V3<R> d1 = (R(0.5) * f) * (R(3) * f - R(4));
        V3<R> d2 = (R(0.5) * (R(1) - f)) * (R(3) * f + R(1));
        V3<R> d3 = R(0.5) * f2;
        Taps t;
        t.g0 = w0 + w1;
        t.g1 = w2 + w3;
        t.h0 = (i - R(1)) + w1 / t.g0;
        t.h1 = (i + R(1)) + w3 / t.g1;
        t.dg0 = d0 + d1;
        t.dg1 = d2 + d3;
        t.dh0 = (i - R(1)) + d1 / vmin(t.dg0, V3<R>(R(-1e-7)));
        t.dh1 = (i + R(1)) + d3 / vmax(t.dg1, V3<R>(R(1e-7)));
        return t;
    }
    R cubic_fetch(V3<R> h0, V3<R> h1, V3<R> g0, V3<R> g1) const {
        R y0z0 = g0.x * density_at_index(V3<R>(h0.x, h0.y, h0.z)) + g1.x * density_at_index(V3<R>(h1.x, h0.y, h0.z));
        R y1z0 = g0.x * density_at_index(V3<R>(h0.x, h1.y, h0.z)) + g1.x * density_at_index(V3<R>(h1.x, h1.y, h0.z));
        R y0z1 = g0.x * density_at_index(V3<R>(h0.x, h0.y, h1.z)) + g1.x * density_at_index(V3<R>(h1.x, h0.y, h1.z));
        R y1z1 = g0.x * density_at_index(V3<R>(h0.x, h1.y, h1.z)) + g1.x * density_at_index(V3<R>(h1.x, h1.y, h1.z));
        return g0.z * (g0.y * y0z0 + g1.y * y1z0) + g1.z * (g0.y * y0z1 + g1.y * y1z1);
    }
    R density_cubic(V3<R> idx) const {
        Taps t = cubic_taps(idx);
        return cubic_fetch(t.h0, t.h1, t.g0, t.g1);
    }
    V3<R> density_gradient_cubic(V3<R> idx) const {
        Taps t = cubic_taps(idx);
        return V3<R>(cubic_fetch(V3<R>(t.dh0.x, t.h0.y, t.h0.z), V3<R>(t.dh1.x, t.h1.y, t.h1.z),
                                 V3<R>(t.dg0.x, t.g0.y, t.g0.z), V3<R>(t.dg1.x, t.g1.y, t.g1.z)),
                     cubic_fetch(V3<R>(t.h0.x, t.dh0.y, t.h0.z), V3<R>(t.h1.x, t.dh1.y, t.h1.z),
                                 V3<R>(t.g0.x, t.dg0.y, t.g0.z), V3<R>(t.g1.x, t.dg1.y, t.g1.z)),
                     cubic_fetch(V3<R>(t.h0.x, t.h0.y, t.dh0.z), V3<R>(t.h1.x, t.h1.y, t.dh1.z),
                                 V3<R>(t.g0.x, t.g0.y, t.dg0.z), V3<R>(t.g1.x, t.g1.y, t.dg1.z)));
    }
    // common.glsl:160-162
    R density_view(V3<R> idx) const { return u->env[1] > 0.5f ? density_cubic(idx) : density_at_index(idx); }

    // ---- lighting volumes (bindings 6 and 7) ------------------------------------------
    R shadow_texel(int x, int y, int z, int ch) const {
        int c[3] = {x, y, z};
        if (sweep_axis >= 0 && c[sweep_axis] == sweep_slice) c[sweep_axis] += sweep_sign;
        for (int a = 0; a < 3; a++)
            if (c[a] < 0 || c[a] >= sdims[a]) return R(1);
        size_t i = (size_t)c[0] + (size_t)sdims[0] * ((size_t)c[1] + (size_t)sdims[1] * c[2]);
        return R(f16_to_f32(shadow_img[4 * i + ch]));
    }
    R ao_texel(int x, int y, int z) const {
        if (x < 0 || y < 0 || z < 0 || x >= adims[0] || y >= adims[1] || z >= adims[2]) return R(1);
        return R(f16_to_f32(s->ao[(size_t)x + (size_t)adims[0] * ((size_t)y + (size_t)adims[1] * z)]));
    }
    // Linear fetch of channel ch of the shadow (which = 0) or AO (which = 1) volume at texel position t.
    R light_linear(int which, V3<R> t, int ch) const {
        if (which == 0) return trilinear(t, [this, ch](int x, int y, int z) { return shadow_texel(x, y, z, ch); });
        return trilinear(t, [this](int x, int y, int z) { return ao_texel(x, y, z); });
    }
    // texture_cubic (common.glsl:141-151): 8 linear fetches at the B-spline taps.
    R light_cubic(int which, V3<R> t, int ch) const {
        Taps k = cubic_taps(t);
        auto f = [&](R x, R y, R z) { return light_linear(which, V3<R>(x, y, z), ch); };
        R z0 = k.g0.y * (k.g0.x * f(k.h0.x, k.h0.y, k.h0.z) + k.g1.x * f(k.h1.x, k.h0.y, k.h0.z)) +
               k.g1.y * (k.g0.x * f(k.h0.x, k.h1.y, k.h0.z) + k.g1.x * f(k.h1.x, k.h1.y, k.h0.z));
        R z1 = k.g0.y * (k.g0.x * f(k.h0.x, k.h0.y, k.h1.z) + k.g1.x * f(k.h1.x, k.h0.y, k.h1.z)) +
               k.g1.y * (k.g0.x * f(k.h0.x, k.h1.y, k.h1.z) + k.g1.x * f(k.h1.x, k.h1.y, k.h1.z));
        return k.g0.z * z0 + k.g1.z * z1;
    }
    // lighting_lookup (common.glsl:154-156) at texture coordinate tc: cubic with
    // smooth sampling, else one linear fetch.
    R lighting_lookup(int which, V3<R> tc, int ch) const {
        const int* size = which == 0 ? sdims : adims;
        V3<R> t = V3<R>(tc.x * R(size[0]), tc.y * R(size[1]), tc.z * R(size[2])) - V3<R>(R(0.5));
        return u->env[1] > 0.5f ? light_cubic(which, t, ch) : light_linear(which, t, ch);
    }
    V3<R> index_to_tex(V3<R> idx) const {  // common.glsl:68-70
        return V3<R>((idx.x + R(0.5)) / R(u->dims[0]), (idx.y + R(0.5)) / R(u->dims[1]),
                     (idx.z + R(0.5)) / R(u->dims[2]));
    }
    V3<R> tex_to_index(V3<R> t) const {  // common.glsl:72-74
        return V3<R>(t.x * R(u->dims[0]) - R(0.5), t.y * R(u->dims[1]) - R(0.5), t.z * R(u->dims[2]) - R(0.5));
    }

    // ---- transfer function (common.glsl:170-230) -------------------------------------
    V4<R> transfer(R density) const {
        R t = density * R(4095);
        R fl = std::floor(t);
        R a = t - fl;
        int i0 = iclamp((int)gmax(gmin(fl, R(5000)), R(-5000)), 0, 4095);
        int i1 = iclamp((int)gmax(gmin(fl, R(5000)), R(-5000)) + 1, 0, 4095);
        V4<R> r;
        R* rp = &r.x;
        for (int c = 0; c < 4; c++)
            rp[c] = mixr(R(f16_to_f32(s->lut[4 * i0 + c])), R(f16_to_f32(s->lut[4 * i1 + c])), a);
        return r;
    }
    struct Sample {
        V3<R> albedo;
        R extinction, roughness, metalness;
        uint32_t label;
        bool ghost;
        R density, subsurface, coat;
    };
    Sample tf_point(R density) const {
        V4<R> tf = transfer(density);
        return Sample{V3<R>(tf.x, tf.y, tf.z), tf.w,   R(u->material[0]), R(u->material[1]), 0u, false,
                      density,                 R(u->subsurface[0]), R(0)};
    }
    V4<R> preint_at(R density) const {
        R x = clampr(density, R(0), R(1)) * R(4095);
        int i = std::min((int)x, 4094);
        R a = x - R(i);
        const float* p0 = s->preint + 4 * i;
        const float* p1 = s->preint + 4 * (i + 1);
        return V4<R>{mixr(R(p0[0]), R(p1[0]), a), mixr(R(p0[1]), R(p1[1]), a), mixr(R(p0[2]), R(p1[2]), a),
                     mixr(R(p0[3]), R(p1[3]), a)};
    }
    Sample tf_segment(R front, R back) const {
        if (std::fabs(back - front) < R(1e-4)) return tf_point(R(0.5) * (front + back));
        V4<R> pb = preint_at(back), pf = preint_at(front);
        R den = back - front;
        V4<R> mean{(pb.x - pf.x) / den, (pb.y - pf.y) / den, (pb.z - pf.z) / den, (pb.w - pf.w) / den};
        V3<R> albedo;
        if (mean.w > R(1e-6)) {
            albedo = V3<R>(mean.x, mean.y, mean.z) / mean.w;
        } else {
            V4<R> t = transfer(back);
            albedo = V3<R>(t.x, t.y, t.z);
        }
        return Sample{albedo, gmax(mean.w, R(0)), R(u->material[0]), R(u->material[1]), 0u, false, back,
                      R(u->subsurface[0]), R(0)};
    }
    R preint_half_depth(R front, R back) const {
        R target = R(0.5) * (preint_at(front).w + preint_at(back).w);
        R lo = 0, hi = 1;
        for (int i = 0; i < 6; i++) {
            R mid = R(0.5) * (lo + hi);
            bool before = (preint_at(mixr(front, back, mid)).w < target) == (back > front);
            lo = before ? mid : lo;
            hi = before ? hi : mid;
        }
        return R(0.5) * (lo + hi);
    }

    // ---- labels (common.glsl:232-351) -------------------------------------------------
    bool label_shown(uint32_t l) const { return u->label_colors[l][3] > 0.5f; }
    bool label_ghosted(uint32_t l) const { return l != 0u && u->label_colors[l][3] > 1.5f; }
    bool label_solid(uint32_t l, uint32_t mode) const {
        if (label_ghosted(l)) return false;
        return mode == 1u ? (l == 0u || label_shown(l)) : (l != 0u && label_shown(l));
    }
    Sample with_label(Sample s_, uint32_t label, uint32_t mode, bool ghost) const {
        s_.label = label;
        s_.ghost = ghost;
        if (label == 0u) return s_;
        const float* color = u->label_colors[label];
        const float* mat = u->label_materials[label];
        if (mode == 1u) {
            s_.albedo = vmix(s_.albedo, V3<R>(color[0], color[1], color[2]), R(u->labels[1]));
        } else {
            s_.albedo = V3<R>(color[0], color[1], color[2]);
            s_.extinction = R(u->labels[2]);
            if (mode == 3u && !ghost && u->label_density[label][3] > 0.0f) {
                const float* d = u->label_density[label];
                s_.extinction *= smoothstep(R(d[0]), R(d[1]), s_.density);
                R t = clampr((s_.density - R(d[2])) / R(d[3]), R(-1.5), R(1.5));
                s_.albedo = vclamp(s_.albedo * (R(1) + R(0.2) * t), R(0), R(1));
                const R BONE_DENSITY = R((300.0f + 1024.0f) / 4095.0f);
                if (R(d[2]) > BONE_DENSITY) {
                    R thin = R(1) - smoothstep(R(-1.5), R(-0.5), t);
                    s_.albedo = s_.albedo * vmix(V3<R>(1), V3<R>(R(0.92f), R(0.6f), R(0.5f)), thin);
                }
            }
        }
        s_.roughness = R(mat[0]);
        s_.metalness = R(mat[1]);
        s_.subsurface = R(mat[3]);
        s_.coat = R(mat[2]);
        return s_;
    }
    bool drawn_label_near(V3<R> idx, uint32_t mode, bool ghost, uint32_t& label) const {
        V3<int> hi(dims[0] - 1, dims[1] - 1, dims[2] - 1);
        V3<R> fn = vfloor(idx + R(0.5));
        V3<int> nearest(iclamp((int)fn.x, 0, hi.x), iclamp((int)fn.y, 0, hi.y), iclamp((int)fn.z, 0, hi.z));
        V3<R> fb = vfloor(idx);
        V3<int> base((int)fb.x, (int)fb.y, (int)fb.z);
        for (int i = 0; i < 21; i++) {
            V3<int> c;
            if (i == 0) {
                c = nearest;
            } else if (i < 9) {
                c = V3<int>(base.x + ((i - 1) & 1), base.y + (((i - 1) >> 1) & 1), base.z + ((i - 1) >> 2));
            } else {
                int k = i - 9;
                V3<int> d(0, 0, 0);
                d[k % 3] = (k / 3 % 2 == 0 ? 1 : -1) * (k < 6 ? 1 : 2);
                c = V3<int>(nearest.x + d.x, nearest.y + d.y, nearest.z + d.z);
            }
            c = V3<int>(iclamp(c.x, 0, hi.x), iclamp(c.y, 0, hi.y), iclamp(c.z, 0, hi.z));
            label = label_texel(c);
            if (ghost ? label_ghosted(label) : label_solid(label, mode)) return true;
        }
        return false;
    }
    Sample apply_labels(Sample s_, V3<R> idx) const {
        uint32_t mode = (uint32_t)u->labels[0];
        if (mode == 0u) return s_;
        R mr = mask_at(idx, 0), mg = mask_at(idx, 1);
        uint32_t label = 0;
        if (mr < (mode == 3u ? R(0.25) : R(0.5))) {
            if (mg >= R(0.5) && drawn_label_near(idx, mode, true, label)) return with_label(s_, label, mode, true);
            s_.extinction = 0;
            return s_;
        }
        if (drawn_label_near(idx, mode, false, label)) return with_label(s_, label, mode, false);
        if (mode >= 2u) s_.extinction = 0;
        return s_;
    }

    // ---- clipping, geometry (common.glsl:357-442) --------------------------------------------
    R clip_depth(V3<R> idx, int& plane) const {
        plane = -1;
        uint32_t count = u->clip[0];
        if (count == 0u) return R(-1e30);
        bool all = u->clip[1] != 0u;
        R depth = all ? R(1e30) : R(-1e30);
        for (uint32_t i = 0; i < count; i++) {
            const float* cp = u->clip_planes[i];
            R d = dot(V3<R>(cp[0], cp[1], cp[2]), idx) - R(cp[3]);
            if (all ? d < depth : d > depth) {
                depth = d;
                plane = (int)i;
            }
        }
        return depth;
    }
    bool is_clipped(V3<R> idx) const {
        int plane;
        return clip_depth(idx, plane) > R(0);
    }
    Sample sample_point(V3<R> idx) const {
        Sample s_ = apply_labels(tf_point(density_at_index(idx)), idx);
        if (s_.ghost || is_clipped(idx)) s_.extinction = 0;
        return s_;
    }
    M3<R> to_index() const {
        const float* m = u->world_to_index;
        M3<R> r;
        r.c[0] = V3<R>(m[0], m[1], m[2]);
        r.c[1] = V3<R>(m[4], m[5], m[6]);
        r.c[2] = V3<R>(m[8], m[9], m[10]);
        return r;
    }
    V3<R> world_to_index_point(V3<R> p) const {
        const float* m = u->world_to_index;
        return V3<R>(R(m[0]) * p.x + R(m[4]) * p.y + R(m[8]) * p.z + R(m[12]),
                     R(m[1]) * p.x + R(m[5]) * p.y + R(m[9]) * p.z + R(m[13]),
                     R(m[2]) * p.x + R(m[6]) * p.y + R(m[10]) * p.z + R(m[14]));
    }
    V3<R> mask_gradient(V3<R> idx, int ch) const {
        V3<R> g;
        for (int a = 0; a < 3; a++) {
            V3<R> d(0);
            d[a] = R(1);
            g[a] = mask_at(idx + d, ch) - mask_at(idx - d, ch);
        }
        return to_index().mul_t(R(0.5) * g);
    }
    void intersect_volume(V3<R> origin, V3<R> dir, R& tn, R& tf) const {
        V3<R> lo(R(-0.5));
        V3<R> hi(R(u->dims[0]) - R(0.5), R(u->dims[1]) - R(0.5), R(u->dims[2]) - R(0.5));
        V3<R> inv = V3<R>(R(1)) / dir;
        V3<R> t0 = (lo - origin) * inv;
        V3<R> t1 = (hi - origin) * inv;
        V3<R> tmin = vmin(t0, t1);
        V3<R> tmax = vmax(t0, t1);
        tn = gmax(gmax(tmin.x, tmin.y), gmax(tmin.z, R(0)));
        tf = gmin(gmin(tmax.x, tmax.y), tmax.z);
    }
    R brick_majorant(V3<R> idx) const {
        if (u->path[3] == 0u) return R(1e30);
        V3<R> b = vfloor((idx + R(0.5)) / R(8));
        int bx = iclamp((int)gmax(gmin(b.x, R(1e8)), R(-1e8)), 0, mdims[0] - 1);
        int by = iclamp((int)gmax(gmin(b.y, R(1e8)), R(-1e8)), 0, mdims[1] - 1);
        int bz = iclamp((int)gmax(gmin(b.z, R(1e8)), R(-1e8)), 0, mdims[2] - 1);
        return R(s->majorant[(size_t)bx + (size_t)mdims[0] * ((size_t)by + (size_t)mdims[1] * bz)]);
    }
    static R brick_exit(V3<R> idx, V3<R> d_idx) {
        V3<R> lo = vfloor((idx + R(0.5)) / R(8)) * R(8) - R(0.5);
        V3<R> hi = lo + R(8);
        V3<R> d(d_idx.x == R(0) ? R(1e-9) : d_idx.x, d_idx.y == R(0) ? R(1e-9) : d_idx.y,
                d_idx.z == R(0) ? R(1e-9) : d_idx.z);
        V3<R> t_far = vmax((lo - idx) / d, (hi - idx) / d);
        return gmax(gmin(t_far.x, gmin(t_far.y, t_far.z)), R(0));
    }
    V3<R> camera_ray(R px, R py) const {
        R nx = px / R(u->flags[2]) * R(2) - R(1);
        R ny = py / R(u->flags[3]) * R(2) - R(1);
        V3<R> f(u->forward[0], u->forward[1], u->forward[2]);
        V3<R> r(u->right[0], u->right[1], u->right[2]);
        V3<R> up(u->up[0], u->up[1], u->up[2]);
        return normalize((f + nx * r) - ny * up);
    }

    // ---- shading.glsl ------------------------------------------------------------------
    static R PI() { return R(3.14159265358979); }
    static R ggx_distribution(R nh, R a2) {
        R denom = nh * nh * (a2 - R(1)) + R(1);
        return a2 / (PI() * denom * denom);
    }
    static V3<R> brdf(V3<R> n, V3<R> v, V3<R> l, V3<R> albedo, R roughness, R metalness, V3<R>& diffuse) {
        V3<R> h = normalize(v + l);
        R nv = gmax(dot(n, v), R(1e-4));
        R nl = gmax(dot(n, l), R(1e-4));
        R nh = gmax(dot(n, h), R(0));
        R vh = gmax(dot(v, h), R(0));
        R a = gmax(roughness * roughness, R(2e-3));
        R d = ggx_distribution(nh, a * a);
        R k = a * R(0.5);
        R g = (nv / (nv * (R(1) - k) + k)) * (nl / (nl * (R(1) - k) + k));
        V3<R> f0 = vmix(V3<R>(R(0.028)), albedo, metalness);
        V3<R> f = f0 + (R(1) - f0) * vpow(R(1) - vh, R(5));
        V3<R> specular = d * g * f / (R(4) * nv * nl);
        diffuse = (R(1) - f) * (R(1) - metalness) * albedo / PI();
        return specular;
    }
    static R coat_roughness(const Sample& s_) { return clampr(R(0.5) * s_.roughness, R(0.1), R(0.25)); }
    static R smith(R nv, R nl, R a) {
        R k = a * R(0.5);
        return (nv / (nv * (R(1) - k) + k)) * (nl / (nl * (R(1) - k) + k));
    }
    static V3<R> brdf_coated(V3<R> n, V3<R> v, V3<R> l, const Sample& s_, V3<R>& diffuse) {
        V3<R> specular = brdf(n, v, l, s_.albedo, s_.roughness, s_.metalness, diffuse);
        if (s_.coat <= R(0)) return specular;
        V3<R> h = normalize(v + l);
        R nv = gmax(dot(n, v), R(1e-4));
        R nl = gmax(dot(n, l), R(1e-4));
        R a = coat_roughness(s_) * coat_roughness(s_);
        R f = s_.coat * (R(0.02) + R(0.98) * vpow(R(1) - gmax(dot(v, h), R(0)), R(5)));
        R coat = ggx_distribution(gmax(dot(n, h), R(0)), a * a) * smith(nv, nl, a) * f / (R(4) * nv * nl);
        diffuse = diffuse * (R(1) - f);
        return specular * (R(1) - f) + V3<R>(coat);
    }
    static R ign(R px, R py) {
        return fract(R(52.9829189) * fract(px * R(0.06711056) + py * R(0.00583715)));
    }

    R surface_frame(V3<R> idx, const Sample& s_, V3<R> v, int cap_plane, R cut, R cap_depth, V3<R>& n,
                    R& surfaceness) const {
        uint32_t label_mode = (uint32_t)u->labels[0];
        V4<R> nw = normal_at(idx);
        V3<R> nxyz(nw.x, nw.y, nw.z);
        R n_len = length(nxyz);
        n = n_len > R(1e-3) ? nxyz / n_len : V3<R>(0);
        surfaceness = nw.w;
        bool skin = s_.label == 255u;
        if (u->env[1] > 0.5f && n_len > R(1e-3) && (!s_.ghost || skin)) {
            V3<R> gw = to_index().mul_t(density_gradient_cubic(idx));
            R len = length(gw);
            n = len > R(1e-7) ? (-gw) / len : n;
        }
        bool hollow = u->label_density[s_.label][3] < 0.0f;
        if (skin) {
        } else if (!s_.ghost && label_mode == 3u && !hollow) {
            V3<R> g = mask_gradient(idx, 0);
            R g_len = length(g);
            V3<R> mask_n = g_len > R(1e-5) ? (-g) / g_len : n;
            V3<R> blended = vmix(mask_n, n, surfaceness);
            n = length(blended) > R(1e-3) ? normalize(blended) : mask_n;
            surfaceness = gmax(surfaceness, clampr(g_len * R(4), R(0), R(1)));
        } else if (!s_.ghost && (label_mode >= 2u || (label_mode == 1u && mask_at(idx, 0) < R(0.98)))) {
            V3<R> g = mask_gradient(idx, 0);
            R g_len = length(g);
            R mask_surface = clampr(g_len * R(4), R(0), R(1));
            if (label_mode >= 2u || mask_surface > surfaceness) {
                n = g_len > R(1e-5) ? (-g) / g_len : n;
                surfaceness = label_mode >= 2u ? mask_surface : gmax(surfaceness, mask_surface);
            }
        }
        if (cap_plane >= 0 && !s_.ghost && -cut < cap_depth) {
            const float* cn = u->clip_normals[cap_plane];
            n = V3<R>(cn[0], cn[1], cn[2]);
            surfaceness = R(1);
        }
        if (s_.ghost) {
            V3<R> g = mask_gradient(idx, 1);
            R g_len = length(g);
            if (!skin || n_len <= R(1e-3)) n = g_len > R(1e-5) ? (-g) / g_len : n;
            surfaceness = clampr(g_len * R(4), R(0), R(1));
            R edge_on = vpow(R(1) - std::fabs(dot(n, v)), R(3));
            return surfaceness * mixr(R(u->labels[3]), R(1), edge_on);
        }
        return R(1);
    }
    bool measured(const Sample& s_) const {
        return u->labels[0] >= 2.0f && s_.label != 0u && !s_.ghost && u->label_optics[s_.label][3] > 0.0f;
    }
    V3<R> tissue_sigma(const Sample& s_) const {
        const float* o = u->label_optics[s_.label];
        V3<R> mu(o[0], o[1], o[2]);
        mu = vmix(V3<R>(mu.y), mu, R(u->env[3]));
        mu = mu / (R(1) - clampr(R(u->env[2]), R(0), R(0.95)));
        mu = mu * gmin(R(1), R(u->tissue[1]) / gmax(gmax(mu.x, gmax(mu.y, mu.z)), R(1e-6)));
        R occupancy = clampr(s_.extinction / gmax(R(u->labels[2]), R(1e-6)), R(0), R(1));
        return occupancy * mu;
    }

    // Procedural texture (shading.glsl:175-342).
    static V3<R> hash3(V3<int> c) {
        uint32_t v[3] = {(uint32_t)c.x * 1664525u + 1013904223u, (uint32_t)c.y * 1664525u + 1013904223u,
                         (uint32_t)c.z * 1664525u + 1013904223u};
        v[0] += v[1] * v[2];
        v[1] += v[2] * v[0];
        v[2] += v[0] * v[1];
        for (int i = 0; i < 3; i++) v[i] ^= v[i] >> 16u;
        v[0] += v[1] * v[2];
        v[1] += v[2] * v[0];
        v[2] += v[0] * v[1];
        const R k = R(1.0 / 4294967296.0);
        return V3<R>(R(v[0]) * k, R(v[1]) * k, R(v[2]) * k);
    }
    static R gradient_at(V3<int> corner, V3<R> offset) { return dot(hash3(corner) * R(2) - R(1), offset); }
    static R gradient_noise(V3<R> p) {
        V3<R> fl = vfloor(p);
        V3<int> i((int)fl.x, (int)fl.y, (int)fl.z);
        V3<R> f = p - fl;
        V3<R> w = f * f * f * (f * (f * R(6) - R(15)) + R(10));
        auto I = [&](int a, int b, int c) { return V3<int>(i.x + a, i.y + b, i.z + c); };
        auto F = [&](int a, int b, int c) { return f - V3<R>(R(a), R(b), R(c)); };
        R x00 = mixr(gradient_at(I(0, 0, 0), F(0, 0, 0)), gradient_at(I(1, 0, 0), F(1, 0, 0)), w.x);
        R x10 = mixr(gradient_at(I(0, 1, 0), F(0, 1, 0)), gradient_at(I(1, 1, 0), F(1, 1, 0)), w.x);
        R x01 = mixr(gradient_at(I(0, 0, 1), F(0, 0, 1)), gradient_at(I(1, 0, 1), F(1, 0, 1)), w.x);
        R x11 = mixr(gradient_at(I(0, 1, 1), F(0, 1, 1)), gradient_at(I(1, 1, 1), F(1, 1, 1)), w.x);
        return R(1.7) * mixr(mixr(x00, x10, w.y), mixr(x01, x11, w.y), w.z);
    }
    static R fbm(V3<R> p) {
        return (gradient_noise(p) + R(0.5) * gradient_noise(p * R(2.03) + R(17.1)) +
                R(0.25) * gradient_noise(p * R(4.07) + R(31.7))) /
               R(1.75);
    }
    static void cellular(V3<R> p, R& nearest, R& border, V3<int>& cell) {
        V3<R> fl = vfloor(p);
        V3<int> i((int)fl.x, (int)fl.y, (int)fl.z);
        V3<R> f = p - fl;
        V3<R> d1(R(8)), d2(R(8));
        cell = i;
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    V3<int> c(i.x + x, i.y + y, i.z + z);
                    V3<R> d = V3<R>(R(x), R(y), R(z)) + hash3(c) - f;
                    if (dot(d, d) < dot(d1, d1)) {
                        d2 = d1;
                        d1 = d;
                        cell = c;
                    } else if (dot(d, d) < dot(d2, d2)) {
                        d2 = d;
                    }
                }
        nearest = length(d1);
        border = dot(R(0.5) * (d1 + d2), normalize(d2 - d1));
    }
    static R lobule_border(V3<R> q, V3<int>& lobule) {
        V3<R> warp(gradient_noise(q * R(1.7)), gradient_noise(q * R(1.7) + R(5.2)),
                   gradient_noise(q * R(1.7) + R(9.4)));
        R nearest, border;
        cellular(q + R(0.2) * warp, nearest, border, lobule);
        return border;
    }
    static R surface_height(V3<R> p, const float* params) {
        V3<R> q = p / R(params[3]);
        if (params[0] > 1.5f) {
            V3<int> lobule;
            R bulge = smoothstep(R(0), R(0.35), lobule_border(q, lobule));
            return R(0.7) * bulge + R(0.3) * fbm(q * R(3));
        }
        return R(0.6) * fbm(q) + R(0.4) * gradient_noise(q * R(4));
    }
    V3<R> index_to_world(V3<R> idx) const {
        // inverse(mat3(world_to_index)) * (idx - world_to_index[3].xyz), by the adjugate.
        M3<R> m = to_index();
        V3<R> c0 = m.c[0], c1 = m.c[1], c2 = m.c[2];
        V3<R> r0 = cross(c1, c2), r1 = cross(c2, c0), r2 = cross(c0, c1);
        R det = dot(c0, r0);
        const float* w = u->world_to_index;
        V3<R> q = idx - V3<R>(w[12], w[13], w[14]);
        return V3<R>(dot(r0, q), dot(r1, q), dot(r2, q)) / det;
    }
    bool textured(const Sample& s_) const {
        return s_.label != 0u && !s_.ghost && u->label_surface[s_.label][0] > 0.5f;
    }
    void apply_surface_texture(V3<R> idx, Sample& s_) const {
        if (!textured(s_)) return;
        const float* params = u->label_surface[s_.label];
        R amount = R(params[1]);
        V3<R> p = index_to_world(idx);
        V3<R> q = p / R(params[3]);
        V3<R> tint;
        R wet, rough;
        if (params[0] > 1.5f) {
            R strength = gmin(R(1), R(5) * amount);
            V3<int> lobule;
            R border = lobule_border(q, lobule);
            R septum = (R(1) - smoothstep(R(0), R(0.06), border)) * smoothstep(R(-0.7), R(0.3), gradient_noise(q * R(6)));
            R near_septum = R(1) - smoothstep(R(0), R(0.2), border);
            R patches = fbm(q * R(1.5));
            R mottle = fbm(q * R(4));
            tint = V3<R>(R(1) + amount * (hash3(lobule).x * R(2) - R(1) + R(0.5) * mottle));
            tint = tint * vmix(V3<R>(1), V3<R>(R(0.84), R(0.78), R(0.88)), strength * smoothstep(R(-0.1), R(0.4), -patches));
            tint = tint * vmix(V3<R>(1), V3<R>(R(0.6), R(0.57), R(0.66)), strength * septum);
            R nearest, unused;
            V3<int> speck;
            cellular(p / R(3), nearest, unused, speck);
            V3<R> h = hash3(V3<int>(speck.x + 7919, speck.y + 7919, speck.z + 7919));
            R speck_mask = (h.x < R(0.03) + R(0.4) * near_septum ? R(1) : R(0)) *
                           (R(1) - smoothstep(R(0.12), R(0.2), nearest / (R(0.6) + R(0.8) * h.y)));
            tint = tint * mixr(R(1), R(0.2), strength * speck_mask);
            wet = R(1) - R(0.3) * near_septum;
            rough = R(1) + R(0.2) * mottle;
        } else {
            R patches = fbm(q);
            R grain = gradient_noise(q * R(4));
            tint = V3<R>(R(1) + amount * (R(0.7) * patches + R(0.3) * grain));
            tint = tint * vmix(V3<R>(1), V3<R>(R(0.97), R(0.94), R(1)),
                               clampr(-patches, R(0), R(1)) * gmin(R(1), R(5) * amount));
            wet = clampr(R(0.8) + R(0.4) * patches, R(0.4), R(1.2));
            rough = R(1) + R(0.3) * grain;
        }
        s_.albedo = vclamp(s_.albedo * tint, R(0), R(1));
        s_.roughness = clampr(s_.roughness * rough, R(0.02), R(1));
        s_.coat = clampr(s_.coat * wet, R(0), R(1));
    }
    V3<R> bump_normal(V3<R> idx, const Sample& s_, V3<R> n) const {
        if (!textured(s_) || u->label_surface[s_.label][2] <= 0.0f || dot(n, n) < R(0.5)) return n;
        const float* params = u->label_surface[s_.label];
        V3<R> p = index_to_world(idx);
        R e = R(0.02) * R(params[3]);
        R h = surface_height(p, params);
        V3<R> g = V3<R>(surface_height(p + V3<R>(e, 0, 0), params) - h, surface_height(p + V3<R>(0, e, 0), params) - h,
                        surface_height(p + V3<R>(0, 0, e), params) - h) *
                  (R(params[2]) / e);
        return normalize(n - (g - dot(g, n) * n));
    }

    // ---- dvr.comp -------------------------------------------------------------------------
    Sample labelled(Sample s_, V3<R> idx) const {  // dvr.comp:15-21
        s_ = apply_labels(s_, idx);
        if (measured(s_)) s_.extinction = dot(tissue_sigma(s_), V3<R>(R(1.0 / 3.0)));
        return s_;
    }
    void dvr_pixel(uint32_t px, uint32_t py, float out[4]) const {  // dvr.comp:23-147
        V3<R> dir = camera_ray(R(px) + R(0.5), R(py) + R(0.5));
        V3<R> o_idx = world_to_index_point(V3<R>(u->eye[0], u->eye[1], u->eye[2]));
        V3<R> d_idx = to_index().mul(dir);
        R t0, t1;
        intersect_volume(o_idx, d_idx, t0, t1);
        R step_mm = R(u->dims[3]);
        bool preintegrate = u->lighting[2] != 0u;
        V3<R> l(u->light_dir[0], u->light_dir[1], u->light_dir[2]);
        V3<R> v = -dir;
        V3<R> light(R(u->light_dir[3]));
        V3<R> radiance(0);
        R transmittance = 1;
        bool refined = false;
        if (t0 < t1) {
            R tt = t0 + ign(R(px), R(py)) * step_mm;
            R prev_density = density_view(o_idx + d_idx * (tt - step_mm));
            for (; tt < t1 && transmittance > R(0.01); tt += step_mm) {
                V3<R> idx = o_idx + d_idx * tt;
                if (brick_majorant(idx) <= R(0)) {
                    tt += brick_exit(idx, d_idx);
                    prev_density = -1;
                    continue;
                }
                int cap_plane;
                R cut = clip_depth(idx, cap_plane);
                if (cut > R(0)) {
                    prev_density = -1;
                    continue;
                }
                R density = density_view(idx);
                R front_density = prev_density;
                bool segment = preintegrate && front_density >= R(0);
                Sample s_ = segment ? tf_segment(front_density, density) : tf_point(density);
                prev_density = density;
                s_ = labelled(s_, idx);
                if (s_.extinction <= R(0)) continue;
                R alpha = R(1) - vexp(-s_.extinction * step_mm);
                if (segment && alpha > R(0.2) && transmittance > R(0.5) && s_.label == 0u &&
                    std::fabs(density - front_density) > R(1e-4)) {
                    idx = idx - d_idx * (step_mm * (R(1) - preint_half_depth(front_density, density)));
                } else if (!refined && alpha > R(0.2) && transmittance > R(0.9)) {
                    refined = true;
                    R inside = 0, outside = -step_mm;
                    R target = R(0.5) * s_.extinction;
                    for (int i = 0; i < 5; i++) {
                        R mid = R(0.5) * (inside + outside);
                        V3<R> at = idx + d_idx * mid;
                        bool opaque = labelled(tf_point(density_view(at)), at).extinction >= target;
                        inside = opaque ? mid : inside;
                        outside = opaque ? outside : mid;
                    }
                    idx = idx + d_idx * inside;
                }
                V3<R> n;
                R surfaceness;
                R extinction_scale = surface_frame(idx, s_, v, cap_plane, cut, R(1.5) * step_mm, n, surfaceness);
                apply_surface_texture(idx, s_);
                if (s_.ghost) alpha = R(1) - vexp(-s_.extinction * extinction_scale * step_mm);
                // dvr.comp:107-119: lighting volumes looked up just outside the surface,
                // the diffused light a little below it.
                M3<R> ti = to_index();
                V3<R> shadow_at = index_to_tex(idx + ti.mul(n * (R(1.5) * R(u->light_cells[0]) * surfaceness)));
                V3<R> ao_at = index_to_tex(idx + ti.mul(n * (R(u->light_cells[1]) * surfaceness)));
                R shadow = u->lighting[0] != 0u ? lighting_lookup(0, shadow_at, 3) : R(1);
                V3<R> sss_at = index_to_tex(idx - ti.mul(n * (R(u->subsurface[3]) * surfaceness)));
                V3<R> diffused(R(1));
                if (u->lighting[0] != 0u)
                    diffused = V3<R>(lighting_lookup(0, sss_at, 0), lighting_lookup(0, sss_at, 1),
                                     lighting_lookup(0, sss_at, 2));
                R ao = u->lighting[1] != 0u ? lighting_lookup(1, ao_at, 0) : R(1);
                n = bump_normal(idx, s_, n);
                n = dot(n, v) < R(0) ? -n : n;
                R nl = gmax(dot(n, l), R(0));
                V3<R> diffuse;
                V3<R> specular = brdf_coated(n, v, l, s_, diffuse);
                V3<R> surface = (specular + diffuse) * (R(1) - s_.subsurface) * nl * shadow +
                                diffuse * s_.subsurface * diffused;
                V3<R> volume = s_.albedo / (R(4) * PI()) * vmix(V3<R>(shadow), diffused, s_.subsurface);
                V3<R> direct = vmix(volume, surface, surfaceness) * light;
                // dvr.comp:135-139: a light map's diffuse response where surfaces face it.
                V3<R> ambient_light(u->ambient[0], u->ambient[1], u->ambient[2]);
                if (u->lighting[3] == 2u)
                    ambient_light = vmix(ambient_light, env_map_diffuse(n), surfaceness * (R(1) - s_.subsurface));
                V3<R> ambient = s_.albedo * ambient_light * ao;
                radiance = radiance + transmittance * alpha * (direct + ambient);
                transmittance *= R(1) - alpha;
            }
        }
        out[0] = (float)radiance.x;
        out[1] = (float)radiance.y;
        out[2] = (float)radiance.z;
        out[3] = (float)transmittance;
    }

    // ---- shadow.comp:18-68: texel p of slice p[axis] from the previous slice ----------
    void shadow_texel_pass(const int p[3], int axis, R out[4]) const {
        int u_axis = (axis + 1) % 3, v_axis = (axis + 2) % 3;
        V3<R> sd = V3<R>(R(sdims[0]), R(sdims[1]), R(sdims[2]));
        V3<R> ld(u->light_dir[0], u->light_dir[1], u->light_dir[2]);
        V3<R> l_idx = to_index().mul(ld);
        l_idx = V3<R>(l_idx.x * (sd.x / R(u->dims[0])), l_idx.y * (sd.y / R(u->dims[1])),
                      l_idx.z * (sd.z / R(u->dims[2])));
        R la[3] = {l_idx.x, l_idx.y, l_idx.z};
        R step = R(1) / gmax(std::fabs(la[axis]), R(1e-6));
        V3<R> here = V3<R>(R(p[0]), R(p[1]), R(p[2]));
        V3<R> prev = here + l_idx * step;
        V3<R> mid = (here + prev) * R(0.5) + V3<R>(R(0.5));
        V3<R> mid_tex(mid.x / sd.x, mid.y / sd.y, mid.z / sd.z);
        Sample m = sample_point(tex_to_index(mid_tex));
        R direct = light_linear(0, prev, 3) * vexp(-m.extinction * step);
        R au[3] = {0, 0, 0}, av[3] = {0, 0, 0};
        au[u_axis] = R(u->subsurface[2]);
        av[v_axis] = R(u->subsurface[2]);
        V3<R> axis_u(au[0], au[1], au[2]), axis_v(av[0], av[1], av[2]);
        V3<R> gathered = V3<R>(light_linear(0, prev, 0), light_linear(0, prev, 1), light_linear(0, prev, 2)) * R(0.4);
        for (int i = 0; i < 4; i++) {
            V3<R> offset = (i & 1) == 0 ? axis_u : axis_v;
            offset = offset * (i < 2 ? R(1) : R(-1));
            V3<R> at = prev + offset;
            gathered = gathered +
                       V3<R>(light_linear(0, at, 0), light_linear(0, at, 1), light_linear(0, at, 2)) * R(0.15);
        }
        R mx = gmax(gmax(m.albedo.x, gmax(m.albedo.y, m.albedo.z)), R(1e-4));
        V3<R> chroma = m.albedo / mx;
        V3<R> ind_ext = (V3<R>(R(1.2)) - chroma) * (m.extinction * R(u->subsurface[1]));
        V3<R> indirect(gathered.x * vexp(-ind_ext.x * step), gathered.y * vexp(-ind_ext.y * step),
                       gathered.z * vexp(-ind_ext.z * step));
        out[0] = indirect.x, out[1] = indirect.y, out[2] = indirect.z, out[3] = direct;
    }

    // ---- ao.comp:20-43 ----------------------------------------------------------------------
    R ao_texel_pass(const int p[3]) const {
        V3<R> tc((R(p[0]) + R(0.5)) / R(adims[0]), (R(p[1]) + R(0.5)) / R(adims[1]), (R(p[2]) + R(0.5)) / R(adims[2]));
        V3<R> center = tex_to_index(tc);
        M3<R> ti = to_index();
        R dt = R(15.0) / R(8);
        R visible = 0;
        for (int i = 0; i < 16; i++) {
            R z = R(1) - (R(i) + R(0.5)) * (R(2) / R(16));
            R r = std::sqrt(R(1) - z * z);
            R phi = R(i) * R(2.39996323f);
            V3<R> d = ti.mul(V3<R>(r * cr_cos(phi), r * cr_sin(phi), z));
            R depth = 0;
            for (int j = 0; j < 8; j++) depth += sample_point(center + d * ((R(j) + R(0.5)) * dt)).extinction * dt;
            visible += vexp(-depth);
        }
        return gmin(R(2) * visible / R(16), R(1));
    }

    // ---- pt.comp ----------------------------------------------------------------------------
    uint32_t pcg() {  // pt.comp:44-49
        uint32_t state = rng;
        rng = rng * 747796405u + 2891336453u;
        uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
        return (word >> 22u) ^ word;
    }
    R rand() { return R(pcg() >> 8) * R(1.0 / 16777216.0); }
    static R rng_real() { return R(0); }  // the real type, for generic callers

    // ---- light map (common.glsl:459-485, pt.comp:193-226) --------------------------------
    V3<R> basis_f() const { return normalize(V3<R>(u->forward[0], u->forward[1], u->forward[2])); }
    V3<R> basis_r() const { return normalize(V3<R>(u->right[0], u->right[1], u->right[2])); }
    V3<R> basis_u() const { return normalize(V3<R>(u->up[0], u->up[1], u->up[2])); }
    void env_uv(V3<R> d, R& uu, R& vv) const {
        V3<R> f = basis_f(), r = basis_r(), up = basis_u();
        R phi = cr_atan2(dot(d, r), dot(d, f));
        R theta = cr_acos(clampr(dot(d, up), R(-1), R(1)));
        R x = R(0.5) + (phi + R(u->env[0])) / (R(2) * PI());
        uu = x - std::floor(x);
        vv = theta / PI();
    }
    V3<R> env_direction(R uu, R vv) const {
        R phi = (uu - R(0.5)) * R(2) * PI() - R(u->env[0]);
        R theta = vv * PI();
        V3<R> f = basis_f(), r = basis_r(), up = basis_u();
        return cr_sin(theta) * (cr_sin(phi) * r + cr_cos(phi) * f) + cr_cos(theta) * up;
    }
    // Linear clamp-to-edge lookup of an RGBA16F image (the renderer's sampler).
    V3<R> image_rgb(const uint16_t* img, int w, int h, R uu, R vv) const {
        R tx = uu * R(w) - R(0.5), ty = vv * R(h) - R(0.5);
        R fx = std::floor(tx), fy = std::floor(ty);
        R ax = tx - fx, ay = ty - fy;
        int x0 = (int)fx, y0 = (int)fy;
        R c[3];
        for (int k = 0; k < 3; k++) {
            auto t = [&](int x, int y) {
                return R(f16_to_f32(img[4 * ((size_t)iclamp(y, 0, h - 1) * w + iclamp(x, 0, w - 1)) + k]));
            };
            c[k] = mixr(mixr(t(x0, y0), t(x0 + 1, y0), ax), mixr(t(x0, y0 + 1), t(x0 + 1, y0 + 1), ax), ay);
        }
        return V3<R>(c[0], c[1], c[2]);
    }
    V3<R> ambient_rgb() const { return V3<R>(u->ambient[0], u->ambient[1], u->ambient[2]); }
    V3<R> env_map_radiance(V3<R> d) const {
        R uu, vv;
        env_uv(d, uu, vv);
        return ambient_rgb() * image_rgb(s->env_radiance, (int)s->env_w, (int)s->env_h, uu, vv);
    }
    V3<R> env_map_diffuse(V3<R> n) const {
        R uu, vv;
        env_uv(n, uu, vv);
        return ambient_rgb() * image_rgb(s->env_diffuse, 32, 16, uu, vv);
    }
    R table(int x, int y, int ch) const { return R(s->env_table[2 * ((size_t)y * (s->env_w + 1) + x) + ch]); }
    R env_pdf(V3<R> d) const {
        int w = (int)s->env_w, h = (int)s->env_h;
        R uu, vv;
        env_uv(d, uu, vv);
        int tx = std::min((int)(uu * R(w)), w - 1), ty = std::min((int)(vv * R(h)), h - 1);
        R p = table(tx, ty, 1);
        return p * R(w * h) / (R(2) * PI() * PI() * gmax(cr_sin(vv * PI()), R(1e-6)));
    }
    V3<R> sample_env(R& pdf) {
        int w = (int)s->env_w, h = (int)s->env_h;
        R xi = rand();
        int lo = 0, hi = h - 1;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (table(w, mid, 0) >= xi) hi = mid;
            else lo = mid + 1;
        }
        int y = lo;
        xi = rand();
        lo = 0, hi = w - 1;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (table(mid, y, 0) >= xi) hi = mid;
            else lo = mid + 1;
        }
        R ju = rand();
        R jv = rand();
        R uu = (R(lo) + ju) / R(w), vv = (R(y) + jv) / R(h);
        R p = table(lo, y, 1);
        pdf = p * R(w * h) / (R(2) * PI() * PI() * gmax(cr_sin(vv * PI()), R(1e-6)));
        return env_direction(uu, vv);
    }

    V3<R> environment(V3<R> d) const {  // pt.comp:56-82
        V3<R> amb(u->ambient[0], u->ambient[1], u->ambient[2]);
        if (u->lighting[3] == 2u) return env_map_radiance(d);
        if (u->lighting[3] == 1u) {
            V3<R> f = normalize(V3<R>(u->forward[0], u->forward[1], u->forward[2]));
            V3<R> r = normalize(V3<R>(u->right[0], u->right[1], u->right[2]));
            V3<R> up = normalize(V3<R>(u->up[0], u->up[1], u->up[2]));
            R key_side = dot(V3<R>(u->light_dir[0], u->light_dir[1], u->light_dir[2]), r) >= R(0) ? R(1) : R(-1);
            V3<R> fill = normalize(((-f) - (R(0.9) * key_side) * r) - R(0.1) * up);
            V3<R> rim = normalize((f + R(0.8) * up) - (R(0.4) * key_side) * r);
            R fill_lobe = R(0.5) * vpow(gmax(dot(d, fill), R(0)), R(6));
            R rim_lobe = R(2.5) * vpow(gmax(dot(d, rim), R(0)), R(24));
            V3<R> studio = V3<R>(R(0.02)) + fill_lobe * V3<R>(R(0.95), R(0.98), R(1.05)) +
                           rim_lobe * V3<R>(R(0.9), R(0.95), R(1.1));
            return amb * studio * (R(1) / R(0.106));
        }
        R up = R(0.5) + R(0.5) * d.z;
        V3<R> sky(R(0.92), R(0.96), R(1.04));
        V3<R> ground(R(1.04), R(1.0), R(0.94));
        return amb * vmix(ground * R(0.35), sky, up) / R(0.675);
    }
    static M3<R> basis(V3<R> n) {  // pt.comp:85-92
        R s_ = n.z >= R(0) ? R(1) : R(-1);
        R a = R(-1) / (s_ + n.z);
        R b = n.x * n.y * a;
        M3<R> m;
        m.c[0] = V3<R>(R(1) + s_ * n.x * n.x * a, s_ * b, -s_ * n.x);
        m.c[1] = V3<R>(b, s_ + n.y * n.y * a, -n.y);
        m.c[2] = n;
        return m;
    }
    V3<R> sample_light(const M3<R>& frame) {  // pt.comp:96-102
        R cos_max = cr_cos(R(u->material[3]));
        R cos_t = mixr(cos_max, R(1), rand());
        R sin_t = std::sqrt(gmax(R(0), R(1) - cos_t * cos_t));
        R phi = R(2) * PI() * rand();
        return frame.mul(V3<R>(sin_t * cr_cos(phi), sin_t * cr_sin(phi), cos_t));
    }
    static V3<R> single_scattering_albedo(V3<R> a) {  // pt.comp:119-123
        a = vclamp(a, R(0), R(0.999));
        V3<R> sq = R(9.59217) + R(41.6808) * a + R(17.7126) * a * a;
        V3<R> s_ = R(4.09712) + R(4.20863) * a - V3<R>(std::sqrt(sq.x), std::sqrt(sq.y), std::sqrt(sq.z));
        return R(1) - s_ * s_;
    }
    V3<R> chroma(const Sample& s_) const {  // pt.comp:132-139
        if (s_.ghost) return V3<R>(1);
        const V3<R> TISSUE_CHROMA(R(0.62), R(0.77), R(1.0));
        return vmix(V3<R>(1), TISSUE_CHROMA, clampr(R(2) * s_.subsurface, R(0), R(1)) * R(u->env[3]));
    }
    V3<R> path_sigma(const Sample& s_, R extinction) const {  // pt.comp:143-145
        return measured(s_) ? tissue_sigma(s_) : extinction * chroma(s_);
    }
    V3<R> event_albedo(const Sample& s_) const {  // pt.comp:150-153
        V3<R> reduced = single_scattering_albedo(s_.albedo);
        return measured(s_) ? R(1) - (R(1) - reduced) * (R(1) - clampr(R(u->env[2]), R(0), R(0.95))) : reduced;
    }
    R flight_majorant(V3<R> p) const {  // pt.comp:157-163
        R m = brick_majorant(p);
        if (m > R(0) && u->labels[0] >= 2.0f) m = gmax(m, R(u->tissue[0]));
        return gmin(m, gmax(R(8), gmax(R(u->labels[2]), R(u->tissue[0]))));
    }
    R phase_hg(R cos_t) const {  // pt.comp:167-171
        R g = R(u->env[2]);
        R d = R(1) + g * g - R(2) * g * cos_t;
        return (R(1) - g * g) / (R(4) * PI() * d * std::sqrt(gmax(d, R(1e-8))));
    }
    V3<R> sample_hg(V3<R> dir) {  // pt.comp:173-185
        R g = R(u->env[2]);
        R cos_t;
        if (std::fabs(g) < R(1e-3)) {
            cos_t = R(1) - R(2) * rand();
        } else {
            R q = (R(1) - g * g) / (R(1) - g + R(2) * g * rand());
            cos_t = clampr((R(1) + g * g - q * q) / (R(2) * g), R(-1), R(1));
        }
        R sin_t = std::sqrt(gmax(R(0), R(1) - cos_t * cos_t));
        R phi = R(2) * PI() * rand();
        return basis(dir).mul(V3<R>(sin_t * cr_cos(phi), sin_t * cr_sin(phi), cos_t));
    }
    static R mis(R a, R b) { return a * a / gmax(a * a + b * b, R(1e-20)); }
    static R coat_share(const Sample& s_) { return R(0.5) * s_.coat; }
    static R specular_share(const Sample& s_) { return mixr(R(0.2), R(1), s_.metalness); }
    static R ggx_pdf(V3<R> n, V3<R> v, V3<R> l, R a) {
        V3<R> h = normalize(v + l);
        R nh = gmax(dot(n, h), R(0));
        return ggx_distribution(nh, a * a) * nh / (R(4) * gmax(dot(v, h), R(1e-4)));
    }
    V3<R> sample_ggx(const M3<R>& frame, V3<R> v, R a) {
        R a2 = a * a;
        R phi = R(2) * PI() * rand();
        R xi = rand();
        R cos_t = std::sqrt((R(1) - xi) / (R(1) + (a2 - R(1)) * xi));
        R sin_t = std::sqrt(gmax(R(0), R(1) - cos_t * cos_t));
        V3<R> h = frame.mul(V3<R>(sin_t * cr_cos(phi), sin_t * cr_sin(phi), cos_t));
        return reflect(-v, h);
    }
    // The surface event's next direction (pt.comp:491-502): clear coat, GGX or
    // cosine lobe, chosen in that order with their shares.
    V3<R> sample_bsdf(V3<R> n, V3<R> v, const Sample& s_) {
        M3<R> frame = basis(n);
        if (rand() < coat_share(s_)) return sample_ggx(frame, v, coat_roughness(s_) * coat_roughness(s_));
        if (rand() < specular_share(s_)) return sample_ggx(frame, v, gmax(s_.roughness * s_.roughness, R(2e-3)));
        R r = std::sqrt(rand());
        R phi = R(2) * PI() * rand();
        return frame.mul(V3<R>(r * cr_cos(phi), r * cr_sin(phi), std::sqrt(gmax(R(0), R(1) - r * r))));
    }
    static R bsdf_pdf(V3<R> n, V3<R> v, V3<R> l, const Sample& s_) {
        R nl = dot(n, l);
        if (nl <= R(0)) return R(0);
        R a = gmax(s_.roughness * s_.roughness, R(2e-3));
        R share = specular_share(s_);
        R base = share * ggx_pdf(n, v, l, a) + (R(1) - share) * nl / PI();
        R coat = coat_share(s_);
        return coat * ggx_pdf(n, v, l, coat_roughness(s_) * coat_roughness(s_)) + (R(1) - coat) * base;
    }
    R extinction_at(V3<R> idx, V3<R> dir, bool smooth_density, Sample& s_, int& cap_plane, R& cut) const {
        cut = clip_depth(idx, cap_plane);
        if (cut > R(0)) return R(0);
        s_ = apply_labels(tf_point(smooth_density ? density_view(idx) : density_at_index(idx)), idx);
        if (s_.extinction <= R(0)) return R(0);
        if (s_.ghost) {
            V3<R> n;
            R surfaceness;
            return s_.extinction * surface_frame(idx, s_, -dir, cap_plane, cut, R(0), n, surfaceness);
        }
        return s_.extinction;
    }
    bool free_flight(V3<R> origin, V3<R> dir, bool smooth_density, int hero, V3<R>& spectral, V3<R>& idx,
                     Sample& s_, int& cap_plane, R& cut, R* t_hit = nullptr) {
        V3<R> d_idx = to_index().mul(dir);
        R t0, t1;
        intersect_volume(origin, d_idx, t0, t1);
        R tt = t0;
        for (int i = 0; i < 4096 && tt < t1; i++) {
            V3<R> p = origin + d_idx * tt;
            R majorant = flight_majorant(p);
            R exit = tt + brick_exit(p, d_idx) + R(1e-3);
            if (majorant <= R(0)) {
                tt = exit;
                continue;
            }
            R next = tt - vlog(R(1) - rand()) / majorant;
            if (next >= exit) {
                tt = exit;
                continue;
            }
            tt = next;
            idx = origin + d_idx * tt;
            R extinction = extinction_at(idx, dir, smooth_density, s_, cap_plane, cut);
            if (extinction <= R(0)) continue;
            V3<R> sigma = path_sigma(s_, extinction);
            if (rand() * majorant < sigma[hero]) {
                spectral = spectral * (sigma / sigma[hero]);
                if (t_hit) *t_hit = tt;
                return true;
            }
            spectral = spectral * (vmax(V3<R>(majorant) - sigma, V3<R>(0)) / (majorant - sigma[hero]));
        }
        return false;
    }
    V3<R> transmittance(V3<R> origin, V3<R> dir) {
        V3<R> d_idx = to_index().mul(dir);
        R t0, t1;
        intersect_volume(origin, d_idx, t0, t1);
        R tt = t0;
        V3<R> result(1);
        for (int i = 0; i < 4096 && tt < t1; i++) {
            R peak = gmax(result.x, gmax(result.y, result.z));
            if (peak < R(0.1)) {
                if (rand() * R(0.1) >= peak) return V3<R>(0);
                result = result * (R(0.1) / peak);
            }
            V3<R> p = origin + d_idx * tt;
            R majorant = flight_majorant(p);
            R exit = tt + brick_exit(p, d_idx) + R(1e-3);
            if (majorant <= R(0)) {
                tt = exit;
                continue;
            }
            R next = tt - vlog(R(1) - rand()) / majorant;
            if (next >= exit) {
                tt = exit;
                continue;
            }
            tt = next;
            Sample m = sample_point(origin + d_idx * tt);
            result = result * vmax(R(1) - path_sigma(m, m.extinction) / majorant, V3<R>(0));
        }
        return result;
    }

    void seed_pixel(uint32_t px, uint32_t py) {  // pt.comp:378-379
        rng = (px * 1973u + py * 9277u + u->path[0] * 26699u) | 1u;
        pcg();
    }
    void pt_pixel(uint32_t px, uint32_t py, float acc_out[4], float alb_out[4], float nrm_out[4]) {
        seed_pixel(px, py);
        uint32_t samples = u->path[2];
        uint32_t max_bounces = u->path[1];
        uint32_t volume_limit = (uint32_t)(float(32u) / (1.0f - clampr(u->env[2], 0.0f, 0.95f)));
        R step_mm = R(u->dims[3]);
        V3<R> ldir(u->light_dir[0], u->light_dir[1], u->light_dir[2]);
        M3<R> light_frame = basis(ldir);
        V3<R> light(R(u->light_dir[3]));
        R cos_light = cr_cos(R(u->material[3]));
        bool area_light = u->material[3] > 1e-4f;
        R light_pdf = area_light ? R(1) / (R(2) * PI() * (R(1) - cos_light)) : R(0);
        M3<R> ti = to_index();
        V3<R> eye_idx = world_to_index_point(V3<R>(u->eye[0], u->eye[1], u->eye[2]));
        bool env_map = u->lighting[3] == 2u;

        V3<R> sum(0);
        R background = 0;
        R albedo_sum[4] = {0, 0, 0, 0};
        R normal_sum[4] = {0, 0, 0, 0};
        for (uint32_t i = 0; i < samples; i++) {
            R jx = rand();
            R jy = rand();
            V3<R> dir = camera_ray(R(px) + jx, R(py) + jy);
            V3<R> origin = eye_idx;
            if (u->lens[0] > 0.0f) {
                V3<R> eye(u->eye[0], u->eye[1], u->eye[2]);
                V3<R> fwd = normalize(V3<R>(u->forward[0], u->forward[1], u->forward[2]));
                V3<R> focal = eye + dir * (R(u->lens[1]) / dot(dir, fwd));
                R r = R(u->lens[0]) * std::sqrt(rand());
                R phi = R(2) * PI() * rand();
                V3<R> rr = normalize(V3<R>(u->right[0], u->right[1], u->right[2]));
                V3<R> uu = normalize(V3<R>(u->up[0], u->up[1], u->up[2]));
                eye = eye + r * (cr_cos(phi) * rr + cr_sin(phi) * uu);
                dir = normalize(focal - eye);
                origin = world_to_index_point(eye);
            }
            V3<R> throughput(1);
            V3<R> radiance(0);
            int hero = std::min((int)(rand() * R(3)), 2);
            V3<R> spectral(1);
            R last_pdf = 0;
            for (uint32_t bounce = 0, surfaces = 0, volumes = 0; surfaces <= max_bounces && volumes <= volume_limit;
                 bounce++) {
                V3<R> idx;
                Sample s_{};
                int cap_plane = -1;
                R cut = 0;
                if (!free_flight(origin, dir, bounce == 0u, hero, spectral, idx, s_, cap_plane, cut)) {
                    if (bounce == 0u) {
                        background += R(1);
                        albedo_sum[0] += R(1), albedo_sum[1] += R(1), albedo_sum[2] += R(1);
                        normal_sum[3] += R(1e4);
                    } else {
                        R weight = env_map ? mis(last_pdf, env_pdf(dir)) : R(1);
                        V3<R> path_weight = throughput * spectral / dot(spectral, V3<R>(R(1.0 / 3.0)));
                        radiance = radiance + path_weight * environment(dir) * weight;
                        if (area_light && dot(dir, ldir) >= cos_light)
                            radiance = radiance + path_weight * light * light_pdf * mis(last_pdf, light_pdf);
                    }
                    break;
                }
                V3<R> path_weight = throughput * spectral / dot(spectral, V3<R>(R(1.0 / 3.0)));
                V3<R> v = -dir;
                V3<R> n;
                R surfaceness;
                surface_frame(idx, s_, v, cap_plane, cut, R(1.5) * step_mm, n, surfaceness);
                apply_surface_texture(idx, s_);
                R xi_surface = rand();
                bool surface_event = xi_surface < surfaceness * (R(1) - s_.subsurface) && dot(n, n) > R(0.5);
                if (surface_event || bounce == 0u) n = bump_normal(idx, s_, n);
                if (bounce == 0u) {
                    albedo_sum[0] += s_.albedo.x, albedo_sum[1] += s_.albedo.y, albedo_sum[2] += s_.albedo.z;
                    V3<R> nn = dot(n, v) < R(0) ? -n : n;
                    normal_sum[0] += nn.x, normal_sum[1] += nn.y, normal_sum[2] += nn.z;
                    normal_sum[3] += length(idx - eye_idx);
                }
                V3<R> light_dir = sample_light(light_frame);
                V3<R> light_t = transmittance(idx, light_dir);
                // pt.comp:467-476: a light-map sample and its transmittance.
                R map_pdf = 0;
                V3<R> map_dir(0), map_light(0);
                if (env_map) {
                    map_dir = sample_env(map_pdf);
                    map_light = env_map_radiance(map_dir) / gmax(map_pdf, R(1e-8));
                    if (map_pdf > R(0)) map_light = map_light * transmittance(idx, map_dir);
                }
                if (surface_event) {
                    n = dot(n, v) < R(0) ? -n : n;
                    V3<R> diffuse;
                    V3<R> specular = brdf_coated(n, v, light_dir, s_, diffuse);
                    R light_weight = area_light ? mis(light_pdf, bsdf_pdf(n, v, light_dir, s_)) : R(1);
                    radiance = radiance + path_weight * (diffuse + specular) * gmax(dot(n, light_dir), R(0)) * light *
                                              light_t * light_weight;
                    if (map_pdf > R(0) && dot(n, map_dir) > R(0)) {  // pt.comp:485-489
                        specular = brdf_coated(n, v, map_dir, s_, diffuse);
                        radiance = radiance + path_weight * (diffuse + specular) * dot(n, map_dir) * map_light *
                                                  mis(map_pdf, bsdf_pdf(n, v, map_dir, s_));
                    }
                    V3<R> l = sample_bsdf(n, v, s_);
                    R nl = dot(n, l);
                    if (nl <= R(0)) break;
                    R pdf = bsdf_pdf(n, v, l, s_);
                    specular = brdf_coated(n, v, l, s_, diffuse);
                    throughput = throughput * (((diffuse + specular) * nl) / gmax(pdf, R(1e-6)));
                    last_pdf = pdf;
                    surfaces++;
                    origin = idx + ti.mul(n * step_mm);
                    dir = l;
                } else {
                    V3<R> albedo = event_albedo(s_);
                    R light_phase = phase_hg(dot(dir, light_dir));
                    radiance = radiance + path_weight * albedo * light_phase * light * light_t *
                                              (area_light ? mis(light_pdf, light_phase) : R(1));
                    if (map_pdf > R(0)) {  // pt.comp:524-527
                        R phase = phase_hg(dot(dir, map_dir));
                        radiance = radiance + path_weight * albedo * phase * map_light * mis(map_pdf, phase);
                    }
                    throughput = throughput * albedo;
                    volumes++;
                    origin = idx;
                    V3<R> next = sample_hg(dir);
                    last_pdf = phase_hg(dot(dir, next));
                    dir = next;
                }
                if (bounce >= 2u) {
                    V3<R> w = throughput * spectral / dot(spectral, V3<R>(R(1.0 / 3.0)));
                    R survive = clampr(gmax(w.x, gmax(w.y, w.z)), R(0.05), R(1));
                    if (rand() > survive) break;
                    throughput = throughput / survive;
                }
            }
            R peak = gmax(radiance.x, gmax(radiance.y, radiance.z));
            radiance = peak > R(20) ? radiance * (R(20) / peak) : radiance;
            sum = sum + radiance;
            R luminance = dot(radiance, V3<R>(R(0.2126), R(0.7152), R(0.0722)));
            albedo_sum[3] += luminance * luminance;
        }
        acc_out[0] = (float)sum.x, acc_out[1] = (float)sum.y, acc_out[2] = (float)sum.z, acc_out[3] = (float)background;
        for (int c = 0; c < 4; c++) {
            alb_out[c] = (float)albedo_sum[c];
            nrm_out[c] = (float)normal_sum[c];
        }
    }
};

// ---- majorant.comp (float, as the shader) ----------------------------------------------------
void majorant_impl(const oracle_scene* s, const vvr_uniforms* u, float* out) {
    int dims[3] = {(int)s->dims[0], (int)s->dims[1], (int)s->dims[2]};
    int grid[3] = {(dims[0] + 7) / 8, (dims[1] + 7) / 8, (dims[2] + 7) / 8};
    uint32_t mode = (uint32_t)u->labels[0];
    auto tf_max = [&](float lo, float hi) {
        int first = std::min(std::max((int)(lo * 64.0f), 0), 63);
        int last = std::min(std::max((int)(hi * 64.0f), 0), 63);
        float m = 0.0f;
        for (int b = first; b <= last; b++) m = gmax(m, u->tf_bucket_max[b / 4][b % 4]);
        return m;
    };
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int bz = 0; bz < grid[2]; bz++)
        for (int by = 0; by < grid[1]; by++)
            for (int bx = 0; bx < grid[0]; bx++) {
                int b[3] = {bx, by, bz};
                int lo[3], hi[3];
                for (int a = 0; a < 3; a++) {
                    lo[a] = std::max(b[a] * 8 - 2, 0);
                    hi[a] = std::min(b[a] * 8 + 8 + 1, dims[a] - 1);
                }
                float d_min = 1.0f, d_max = 0.0f, mask_max = 0.0f;
                for (int z = lo[2]; z <= hi[2]; z++)
                    for (int y = lo[1]; y <= hi[1]; y++)
                        for (int x = lo[0]; x <= hi[0]; x++) {
                            size_t i = (size_t)x + (size_t)dims[0] * ((size_t)y + (size_t)dims[1] * z);
                            float d = (float)s->density[i] / 65535.0f;
                            d_min = gmin(d_min, d);
                            d_max = gmax(d_max, d);
                            if (mode != 0u) {
                                float mr = (float)s->mask[2 * i] / 65535.0f;
                                float mg = (float)s->mask[2 * i + 1] / 65535.0f;
                                mask_max = gmax(mask_max, gmax(mr, mg));
                            }
                        }
                float bound;
                if (mode == 0u) {
                    bound = tf_max(d_min, d_max);
                } else if (mask_max <= 0.0f) {
                    bound = 0.0f;
                } else if (mode == 1u) {
                    bound = tf_max(d_min, d_max);
                } else {
                    bound = u->labels[2];
                }
                out[(size_t)bx + (size_t)grid[0] * ((size_t)by + (size_t)grid[1] * bz)] = bound;
            }
}

// Vulkan float -> snorm8 / unorm16 conversion (round to nearest).
int8_t to_snorm8(float f) {
    f = clampr(f, -1.0f, 1.0f);
    return (int8_t)std::lround(f * 127.0f);
}
uint16_t to_unorm16(float f) {
    f = clampr(f, 0.0f, 1.0f);
    return (uint16_t)std::lround(f * 65535.0f);
}

// ---- normals.comp ------------------------------------------------------------------------------
void normals_impl(const oracle_scene* s, const vvr_uniforms* u, int8_t* out) {
    Ctx<float> c(s, u);
    int nx = c.dims[0], ny = c.dims[1], nz = c.dims[2];
    M3<float> ti = c.to_index();
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                V3<float> g(0);
                for (int a = -1; a <= 1; a++)
                    for (int b = -1; b <= 1; b++) {
                        float w = (float)((2 - std::abs(a)) * (2 - std::abs(b)));
                        g.x += w * (c.density_texel(x + 1, y + a, z + b) - c.density_texel(x - 1, y + a, z + b));
                        g.y += w * (c.density_texel(x + a, y + 1, z + b) - c.density_texel(x + a, y - 1, z + b));
                        g.z += w * (c.density_texel(x + a, y + b, z + 1) - c.density_texel(x + a, y + b, z - 1));
                    }
                g = g / 32.0f;
                V3<float> gw = ti.mul_t(g);
                float len = length(gw);
                V3<float> n = len > 1e-7f ? (-gw) / len : V3<float>(0);
                size_t i = c.vox(x, y, z);
                out[4 * i + 0] = to_snorm8(n.x);
                out[4 * i + 1] = to_snorm8(n.y);
                out[4 * i + 2] = to_snorm8(n.z);
                out[4 * i + 3] = to_snorm8(clampr(len * 400.0f, 0.0f, 1.0f));
            }
}

// ---- mask.comp ----------------------------------------------------------------------------------
struct MaskPass {
    const oracle_scene* s;
    const vvr_uniforms* u;
    const uint8_t* soft;
    int grid[3];
    int push[4];
    int margin;
    Ctx<float> c;

    MaskPass(const oracle_scene* s_, const vvr_uniforms* u_, const uint8_t* soft_, const uint32_t* sd, const int32_t* sp,
             int m)
        : s(s_), u(u_), soft(soft_), margin(m), c(s_, u_) {
        for (int a = 0; a < 3; a++) grid[a] = soft_ ? (int)sd[a] : 1;
        for (int a = 0; a < 4; a++) push[a] = sp ? sp[a] : 0;
    }
    static float soft_score(bool best, bool second, float margin) {  // mask.comp:32-37
        const float CERTAIN = 4.0f;
        if (best == second) return best ? CERTAIN : -CERTAIN;
        return best ? gmin(margin, CERTAIN) : -gmin(margin, CERTAIN);
    }
    const uint8_t* node(int x, int y, int z) const {
        return soft + 4 * ((size_t)x + (size_t)grid[0] * ((size_t)y + (size_t)grid[1] * z));
    }
    V3<float> soft_coord(V3<int> p) const {  // mask.comp:40-50
        int axis[3] = {push[0], push[1], push[2]};
        float can[3] = {(float)c.dims[axis[0]], (float)c.dims[axis[1]], (float)c.dims[axis[2]]};
        V3<float> x;
        for (int a = 0; a < 3; a++) {
            float v = (float)p[axis[a]];
            x[a] = ((push[3] >> a) & 1) != 0 ? can[a] - 1.0f - v : v;
        }
        V3<float> num = x * V3<float>((float)grid[0] - 1.0f, (float)grid[1] - 1.0f, (float)grid[2] - 1.0f);
        return num / vmax(V3<float>(can[0] - 1.0f, can[1] - 1.0f, can[2] - 1.0f), V3<float>(1.0f));
    }
    static int rnd(float f) { return (int)std::round(f); }
    bool soft_overridden(V3<float> x) const {  // mask.comp:54-56
        int ix = Ctx<float>::iclamp(rnd(x.x), 0, grid[0] - 1), iy = Ctx<float>::iclamp(rnd(x.y), 0, grid[1] - 1),
            iz = Ctx<float>::iclamp(rnd(x.z), 0, grid[2] - 1);
        return node(ix, iy, iz)[3] != 0;
    }
    void soft_masks(V3<int> p, V3<float> x, float out[2]) const {  // mask.comp:62-100
        int axis[3] = {push[0], push[1], push[2]};
        float can[3] = {(float)c.dims[axis[0]], (float)c.dims[axis[1]], (float)c.dims[axis[2]]};
        V3<float> to_grid =
            V3<float>((float)grid[0] - 1.0f, (float)grid[1] - 1.0f, (float)grid[2] - 1.0f) /
            vmax(V3<float>(can[0] - 1.0f, can[1] - 1.0f, can[2] - 1.0f), V3<float>(1.0f));
        uint32_t mode = (uint32_t)u->labels[0];
        float here = c.density_texel(p.x, p.y, p.z);
        V3<int> center(rnd(x.x), rnd(x.y), rnd(x.z));
        float sum[2] = {0, 0};
        float weights = 0;
        const float SIGMA_CT = 40.0f / 4095.0f;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    V3<int> nd(Ctx<float>::iclamp(center.x + dx, 0, grid[0] - 1),
                               Ctx<float>::iclamp(center.y + dy, 0, grid[1] - 1),
                               Ctx<float>::iclamp(center.z + dz, 0, grid[2] - 1));
                    V3<float> d = V3<float>((float)nd.x, (float)nd.y, (float)nd.z) - x;
                    float w_space = std::exp(-dot(d, d) / (2.0f * 0.6f * 0.6f));
                    V3<float> cc = V3<float>((float)nd.x, (float)nd.y, (float)nd.z) / to_grid;
                    V3<float> v;
                    for (int a = 0; a < 3; a++) v[axis[a]] = ((push[3] >> a) & 1) != 0 ? can[a] - 1.0f - cc[a] : cc[a];
                    float there = c.density_at_index(v);
                    float w_ct = std::exp(-(here - there) * (here - there) / (2.0f * SIGMA_CT * SIGMA_CT));
                    const uint8_t* t = node(nd.x, nd.y, nd.z);
                    float margin_ = (float)t[2] / 16.0f;
                    float sc0 = soft_score(c.label_solid(t[0], mode), c.label_solid(t[1], mode), margin_);
                    float sc1 = soft_score(c.label_ghosted(t[0]), c.label_ghosted(t[1]), margin_);
                    float w = w_space * w_ct + 1e-6f * w_space;
                    sum[0] += w * sc0;
                    sum[1] += w * sc1;
                    weights += w;
                }
        for (int k = 0; k < 2; k++) out[k] = clampr(0.5f + (sum[k] / weights) / (2.0f * 4.0f), 0.0f, 1.0f);
    }
    bool label_removed(uint32_t l) const { return l != 0u && (!c.label_shown(l) || c.label_ghosted(l)); }
    void binary(V3<int> p, float m[2]) const {  // mask.comp:107-138
        uint32_t mode = (uint32_t)u->labels[0];
        uint32_t label = c.label_texel(p);
        m[0] = c.label_solid(label, mode) ? 1.0f : 0.0f;
        m[1] = c.label_ghosted(label) ? 1.0f : 0.0f;
        if ((push[3] & 8) != 0 && label < 250u && soft) {
            V3<float> x = soft_coord(p);
            if (!soft_overridden(x)) soft_masks(p, x, m);
        }
        if (mode == 1u && margin > 0 && label == 0u) {
            for (int a = -margin; a <= margin && m[0] > 0.0f; a++)
                for (int b = -margin; b <= margin && m[0] > 0.0f; b++)
                    for (int cc = -margin; cc <= margin; cc++) {
                        V3<int> q(Ctx<float>::iclamp(p.x + a, 0, c.dims[0] - 1),
                                  Ctx<float>::iclamp(p.y + b, 0, c.dims[1] - 1),
                                  Ctx<float>::iclamp(p.z + cc, 0, c.dims[2] - 1));
                        if (label_removed(c.label_texel(q))) {
                            m[0] = 0.0f;
                            break;
                        }
                    }
        }
    }
};

}  // namespace

// ============================== C entry points ===============================================

namespace {

// The sweep's axis and direction (renderer.rs:1914-1928): the light's dominant
// axis in shadow-volume index units, slices from the lit side.
template <class R> void shadow_sweep(const oracle_scene* s, const vvr_uniforms* u, uint16_t* out) {
    Ctx<R> c(s, u);
    c.shadow_img = out;
    const int* sd = c.sdims;
    float l[3];
    for (int a = 0; a < 3; a++) {
        const float* m = u->world_to_index;
        float v = (m[a] * u->light_dir[0] + m[4 + a] * u->light_dir[1]) + m[8 + a] * u->light_dir[2];
        l[a] = v * ((float)sd[a] / u->dims[a]);
    }
    float ax = std::fabs(l[0]), ay = std::fabs(l[1]), az = std::fabs(l[2]);
    int axis = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int sign = l[axis] >= 0.0f ? 1 : -1;
    int ua = (axis + 1) % 3, va = (axis + 2) % 3, n = sd[axis];
    c.sweep_axis = axis, c.sweep_sign = sign;
    for (int i = 0; i < n; i++) {
        int slice = sign > 0 ? n - 1 - i : i;
        c.sweep_slice = slice;
#pragma omp parallel for schedule(static)
        for (int b = 0; b < sd[va]; b++)
            for (int a = 0; a < sd[ua]; a++) {
                int p[3];
                p[axis] = slice, p[ua] = a, p[va] = b;
                R v[4];
                c.shadow_texel_pass(p, axis, v);
                size_t idx = (size_t)p[0] + (size_t)sd[0] * ((size_t)p[1] + (size_t)sd[1] * p[2]);
                for (int ch = 0; ch < 4; ch++) out[4 * idx + ch] = f32_to_f16((float)v[ch]);
            }
    }
}

// denoise.comp: the edge-avoiding a-trous filter of the progressive path
// tracer (pass 0 prepare, passes 1..levels filter at steps 1, 2, 4, ..., the
// finish into the HDR image), as Renderer::record_denoise runs it
// (renderer.rs:1868-1904). Images are width x height float4, row-major.
template <class R> struct Denoise {
    const float *accum, *alb, *nrm;
    int w, h;
    R n;
    static R lum(V3<R> c) { return dot(c, V3<R>(R(0.2126f), R(0.7152f), R(0.0722f))); }
    V3<R> rgb(const float* img, int x, int y) const {
        const float* p = img + 4 * ((size_t)y * w + x);
        return V3<R>(p[0], p[1], p[2]);
    }
    V3<R> mean_albedo(int x, int y) const {  // denoise.comp:37-39
        return vmax(rgb(alb, x, y) / n, V3<R>(R(0.02f)));
    }
    void mean_normal_depth(int x, int y, V3<R>& nd, R& depth) const {  // denoise.comp:41-45
        const float* p = nrm + 4 * ((size_t)y * w + x);
        V3<R> g = V3<R>(p[0], p[1], p[2]) / n;
        depth = R(p[3]) / n;
        R len = length(g);
        nd = len > R(1e-3f) ? g / len : V3<R>(R(0));
    }
    int cl(int v, int hi) const { return v < 0 ? 0 : (v > hi ? hi : v); }

    // denoise.comp:63-80
    void prepare(int x, int y, R out[4]) const {
        V3<R> radiance = rgb(accum, x, y) / n;
        V3<R> irr = radiance / mean_albedo(x, y);
        R l = lum(radiance);
        R temporal = gmax(R(alb[4 * ((size_t)y * w + x) + 3]) / n - l * l, R(0)) / n;
        R s1 = 0, s2 = 0;
        for (int dy = -2; dy <= 2; dy++)
            for (int dx = -2; dx <= 2; dx++) {
                R lq = lum(rgb(accum, cl(x + dx, w - 1), cl(y + dy, h - 1)) / n);
                s1 += lq;
                s2 += lq * lq;
            }
        R spatial = gmax(s2 / R(25) - (s1 / R(25)) * (s1 / R(25)), R(0));
        R variance = gmax(temporal, spatial * gmin(R(1), R(4) / n));
        R a = lum(mean_albedo(x, y));
        out[0] = irr.x, out[1] = irr.y, out[2] = irr.z, out[3] = variance / (a * a);
    }

    // denoise.comp:91-139: one level at `step` from `in`.
    void filter(const std::vector<R>& in, int x, int y, int step, R out[4]) const {
        auto at = [&](int qx, int qy, int c) { return in[4 * ((size_t)qy * w + qx) + c]; };
        V3<R> cg;
        R cd;
        mean_normal_depth(x, y, cg, cd);
        V3<R> albedo = mean_albedo(x, y);
        V3<R> center(at(x, y, 0), at(x, y, 1), at(x, y, 2));
        R l_center = lum(center);
        R bv = 0;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++)
                bv += at(cl(x + dx, w - 1), cl(y + dy, h - 1), 3) * R((2 - std::abs(dx)) * (2 - std::abs(dy))) / R(16);
        R sigma_l = R(4) * std::sqrt(gmax(bv, R(0))) + R(1e-4f);
        const R kernel[3] = {R(3.0 / 8.0), R(1.0 / 4.0), R(1.0 / 16.0)};
        V3<R> sum(R(0));
        R var = 0, weights = 0;
        for (int dy = -2; dy <= 2; dy++)
            for (int dx = -2; dx <= 2; dx++) {
                int qx = x + dx * step, qy = y + dy * step;
                if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
                V3<R> c(at(qx, qy, 0), at(qx, qy, 1), at(qx, qy, 2));
                V3<R> g;
                R gd;
                mean_normal_depth(qx, qy, g, gd);
                R w_normal = std::pow(gmax(dot(cg, g), R(0)), R(64));
                if (dot(cg, cg) < R(0.5) || dot(g, g) < R(0.5))
                    w_normal = R(dot(cg, cg) < R(0.5) && dot(g, g) < R(0.5) ? 1 : 0);
                R w_depth = std::exp(-std::fabs(cd - gd) / (R(0.02f) * cd + R(1e-3f)));
                V3<R> da = albedo - mean_albedo(qx, qy);
                R w_albedo = std::exp(-dot(da, da) / R(0.01f));
                R w_lum = std::exp(-std::fabs(l_center - lum(c)) / sigma_l);
                R wt = kernel[std::abs(dx)] * kernel[std::abs(dy)] * w_normal * w_depth * w_albedo * w_lum;
                sum = sum + c * wt;
                var += wt * wt * at(qx, qy, 3);
                weights += wt;
            }
        if (weights > R(0)) {
            out[0] = sum.x / weights, out[1] = sum.y / weights, out[2] = sum.z / weights;
            out[3] = var / (weights * weights);
        } else {
            for (int c = 0; c < 4; c++) out[c] = at(x, y, c);
        }
    }

    void run(int levels, float* hdr, float* last) const {
        const size_t np = (size_t)w * h;
        std::vector<R> a(4 * np), b(4 * np);
#pragma omp parallel for schedule(static)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) prepare(x, y, &a[4 * ((size_t)y * w + x)]);
        for (int level = 0; level < levels; level++) {
#pragma omp parallel for schedule(static)
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) filter(a, x, y, 1 << level, &b[4 * ((size_t)y * w + x)]);
            std::swap(a, b);
        }
        for (size_t i = 0; i < np; i++) {  // denoise.comp:84-89
            int x = int(i % w), y = int(i / w);
            V3<R> irr(a[4 * i], a[4 * i + 1], a[4 * i + 2]);
            V3<R> o = irr * mean_albedo(x, y);
            hdr[4 * i] = (float)o.x, hdr[4 * i + 1] = (float)o.y, hdr[4 * i + 2] = (float)o.z;
            hdr[4 * i + 3] = (float)(R(accum[4 * i + 3]) / n);
            if (last)
                for (int c = 0; c < 4; c++) last[4 * i + c] = (float)a[4 * i + c];
        }
    }
};

// ---- bloom.comp and tonemap.comp: camera finishing of the HDR image ----------------
// Images of the finishing passes are RGBA16F (renderer.rs:1373-1398): every store
// rounds through binary16. Row-major, 4 floats a texel.
struct Img {
    int w = 0, h = 0;
    std::vector<float> v;
    Img(int w_, int h_) : w(w_), h(h_), v(4 * (size_t)w_ * h_, 0.0f) {}
    float* at(int x, int y) { return &v[4 * ((size_t)y * w + x)]; }
    const float* at(int x, int y) const { return &v[4 * ((size_t)y * w + x)]; }
    void store(int x, int y, const float c[4]) {
        for (int i = 0; i < 4; i++) at(x, y)[i] = f16_to_f32(f32_to_f16(c[i]));
    }
};

inline float lum_f(float r, float g, float b) { return (r * 0.2126f + g * 0.7152f) + b * 0.0722f; }
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// 9-tap Gaussian along `ax` (bloom.comp:46-54).
void blur(const Img& src, Img& dst, int ax) {
    const float w[5] = {0.2270f, 0.1945f, 0.1216f, 0.0540f, 0.0162f};
    for (int y = 0; y < src.h; y++)
        for (int x = 0; x < src.w; x++) {
            float sum[4];
            for (int c = 0; c < 4; c++) sum[c] = src.at(x, y)[c] * w[0];
            for (int i = 1; i < 5; i++) {
                int px = ax == 0 ? clampi(x + i, 0, src.w - 1) : x, py = ax == 1 ? clampi(y + i, 0, src.h - 1) : y;
                int mx = ax == 0 ? clampi(x - i, 0, src.w - 1) : x, my = ax == 1 ? clampi(y - i, 0, src.h - 1) : y;
                for (int c = 0; c < 4; c++) sum[c] += w[i] * (src.at(px, py)[c] + src.at(mx, my)[c]);
            }
            dst.store(x, y, sum);
        }
}

// Linear, clamp-to-edge lookup of an RGBA16F image at uv (the renderer's sampler).
void bilinear(const Img& im, float u, float v, float out[4]) {
    float tx = u * (float)im.w - 0.5f, ty = v * (float)im.h - 0.5f;
    float fx = std::floor(tx), fy = std::floor(ty);
    float ax = tx - fx, ay = ty - fy;
    int x0 = (int)fx, y0 = (int)fy;
    for (int c = 0; c < 4; c++) {
        auto t = [&](int x, int y) { return im.at(clampi(x, 0, im.w - 1), clampi(y, 0, im.h - 1))[c]; };
        float c0 = mixr(t(x0, y0), t(x0 + 1, y0), ax), c1 = mixr(t(x0, y0 + 1), t(x0 + 1, y0 + 1), ax);
        out[c] = mixr(c0, c1, ay);
    }
}

inline float hash01(uint32_t v) {  // tonemap.comp:13-17
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (float)((word >> 22u) ^ word) * (1.0f / 4294967296.0f);
}

inline float aces1(float x) {  // tonemap.comp:20-22
    float v = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
    return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

void agx3(float c[3]) {  // tonemap.comp:25-42
    const float inset[3][3] = {{0.842479062253094f, 0.0423282422610123f, 0.0423756549057051f},
                               {0.0784335999999992f, 0.878468636469772f, 0.0784336f},
                               {0.0792237451477643f, 0.0791661274605434f, 0.879142973793104f}};
    const float outset[3][3] = {{1.19687900512017f, -0.0528968517574562f, -0.0529716355144438f},
                                {-0.0980208811401368f, 1.15190312990417f, -0.0980434501171241f},
                                {-0.0990297440797205f, -0.0989611768448433f, 1.15107367264116f}};
    const float min_ev = -12.47393f, max_ev = 4.026069f;
    float t[3];
    for (int i = 0; i < 3; i++) t[i] = (inset[0][i] * c[0] + inset[1][i] * c[1]) + inset[2][i] * c[2];
    for (int i = 0; i < 3; i++) {
        float l = std::log2(gmax(t[i], 1e-10f));
        l = l < min_ev ? min_ev : (l > max_ev ? max_ev : l);
        l = (l - min_ev) / (max_ev - min_ev);
        float l2 = l * l, l4 = l2 * l2;
        t[i] = (((((15.5f * l4 * l2 - 40.14f * l4 * l) + 31.96f * l4) - 6.868f * l2 * l) + 0.4298f * l2) + 0.1191f * l) -
               0.00232f;
    }
    for (int i = 0; i < 3; i++) {
        float o = (outset[0][i] * t[0] + outset[1][i] * t[1]) + outset[2][i] * t[2];
        o = o < 0.0f ? 0.0f : (o > 1.0f ? 1.0f : o);
        c[i] = std::pow(o, 2.2f);
    }
}

inline float srgb1(float c) {  // common.glsl:444-447
    c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
    return c < 0.0031308f ? 12.92f * c : 1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f;
}

inline uint8_t unorm8(float c) {  // RGBA8_UNORM store: nearest
    c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
    return (uint8_t)std::floor(c * 255.0f + 0.5f);
}

}  // namespace

extern "C" {

uint16_t oracle_f32_to_f16(float f) { return f32_to_f16(f); }
float oracle_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

void oracle_majorant(const oracle_scene* s, const vvr_uniforms* u, float* out) { majorant_impl(s, u, out); }

void oracle_normals(const oracle_scene* s, const vvr_uniforms* u, int8_t* out) { normals_impl(s, u, out); }

void oracle_mask_binary(const oracle_scene* s, const vvr_uniforms* u, const uint8_t* soft, const uint32_t soft_dims[3],
                        const int32_t soft_push[4], int32_t margin, float* out) {
    MaskPass mp(s, u, soft, soft_dims, soft_push, margin);
    int nx = (int)s->dims[0], ny = (int)s->dims[1], nz = (int)s->dims[2];
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                float m[2];
                mp.binary(V3<int>(x, y, z), m);
                size_t i = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                out[2 * i] = m[0];
                out[2 * i + 1] = m[1];
            }
}

void oracle_mask(const oracle_scene* s, const vvr_uniforms* u, const uint8_t* soft, const uint32_t soft_dims[3],
                 const int32_t soft_push[4], int32_t margin, int32_t radius, uint16_t* out) {
    int nx = (int)s->dims[0], ny = (int)s->dims[1], nz = (int)s->dims[2];
    size_t n = (size_t)nx * ny * nz;
    std::vector<float> bin(2 * n);
    oracle_mask_binary(s, u, soft, soft_dims, soft_push, margin, bin.data());
    std::vector<uint16_t> B(2 * n), A(2 * n);
    for (size_t i = 0; i < 2 * n; i++) B[i] = to_unorm16(bin[i]);  // pass 0 -> image B
    // Passes 1-3 (mask.comp:152-170): x B->A, y A->B, z B->A.
    for (int pass = 1; pass <= 3; pass++) {
        const std::vector<uint16_t>& src = pass == 2 ? A : B;
        std::vector<uint16_t>& dst = pass == 2 ? B : A;
        int axis = pass - 1;
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    int p[3] = {x, y, z};
                    int dm[3] = {nx, ny, nz};
                    float sum[2] = {0, 0};
                    float cf = 1.0f;
                    for (int j = 0; j <= 2 * radius; j++) {
                        int q[3] = {p[0], p[1], p[2]};
                        q[axis] = std::min(std::max(p[axis] + j - radius, 0), dm[axis] - 1);
                        size_t qi = (size_t)q[0] + (size_t)nx * ((size_t)q[1] + (size_t)ny * q[2]);
                        sum[0] += cf * ((float)src[2 * qi] / 65535.0f);
                        sum[1] += cf * ((float)src[2 * qi + 1] / 65535.0f);
                        cf = cf * (float)(2 * radius - j) / (float)(j + 1);
                    }
                    float e = std::exp2((float)(2 * radius));
                    size_t i = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                    dst[2 * i] = to_unorm16(sum[0] / e);
                    dst[2 * i + 1] = to_unorm16(sum[1] / e);
                }
    }
    std::memcpy(out, A.data(), 2 * n * sizeof(uint16_t));
}

void oracle_dvr(const oracle_scene* s, const vvr_uniforms* u, const uint32_t* pixels, size_t n, int precision,
                float* rgba) {
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < (long)n; i++) {
        if (precision == 64) {
            Ctx<double> c(s, u);
            c.dvr_pixel(pixels[2 * i], pixels[2 * i + 1], rgba + 4 * i);
        } else {
            Ctx<float> c(s, u);
            c.dvr_pixel(pixels[2 * i], pixels[2 * i + 1], rgba + 4 * i);
        }
    }
}

void oracle_shadow(const oracle_scene* s, const vvr_uniforms* u, int precision, uint16_t* out) {
    if (precision == 64) shadow_sweep<double>(s, u, out);
    else shadow_sweep<float>(s, u, out);
}

void oracle_ao(const oracle_scene* s, const vvr_uniforms* u, int precision, uint16_t* out) {
    Ctx<float> c(s, u);
    const int* ad = c.adims;
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < ad[2]; z++)
        for (int y = 0; y < ad[1]; y++)
            for (int x = 0; x < ad[0]; x++) {
                int p[3] = {x, y, z};
                float v = precision == 64 ? (float)Ctx<double>(s, u).ao_texel_pass(p) : Ctx<float>(s, u).ao_texel_pass(p);
                out[(size_t)x + (size_t)ad[0] * ((size_t)y + (size_t)ad[1] * z)] = f32_to_f16(v);
            }
}

void oracle_lighting_lookup(const oracle_scene* s, const vvr_uniforms* u, int which, const float tc[3], int precision,
                            float out[4]) {
    for (int ch = 0; ch < (which == 0 ? 4 : 1); ch++) {
        if (precision == 64) {
            Ctx<double> c(s, u);
            out[ch] = (float)c.lighting_lookup(which, V3<double>(tc[0], tc[1], tc[2]), ch);
        } else {
            Ctx<float> c(s, u);
            out[ch] = c.lighting_lookup(which, V3<float>(tc[0], tc[1], tc[2]), ch);
        }
    }
}

void oracle_denoise(const float* accum, const float* albedo, const float* normal, uint32_t width, uint32_t height,
                    float samples, int levels, int precision, float* hdr, float* last) {
    if (precision == 64) {
        Denoise<double> d{accum, albedo, normal, (int)width, (int)height, (double)samples};
        d.run(levels, hdr, last);
    } else {
        Denoise<float> d{accum, albedo, normal, (int)width, (int)height, samples};
        d.run(levels, hdr, last);
    }
}

float oracle_linear_to_srgb(float c) { return srgb1(c); }
float oracle_aces(float x) { return aces1(x); }
void oracle_agx(const float in[3], float out[3]) {
    for (int i = 0; i < 3; i++) out[i] = in[i];
    agx3(out);
}
uint32_t oracle_hash01_bits(uint32_t v) {
    float f = hash01(v);
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

void oracle_post(const oracle_scene* s, const vvr_uniforms* u, const float* hdr, uint32_t width, uint32_t height,
                 float* exposure, uint8_t* out, float* bloom_a, float* bloom_c) {
    const int w = (int)width, h = (int)height;
    const int qw = (w + 3) / 4, qh = (h + 3) / 4, ew = (qw + 1) / 2, eh = (qh + 1) / 2;
    Img A(qw, qh), B(qw, qh), C(ew, eh), D(ew, eh);
    auto H = [&](int x, int y) { return hdr + 4 * ((size_t)y * w + x); };
    // Pass 0 (bloom.comp:104-122): quarter-resolution Karis average and coverage.
    for (int y = 0; y < qh; y++)
        for (int x = 0; x < qw; x++) {
            float sr = 0, sg = 0, sb = 0, weights = 0, coverage = 0;
            for (int j = 0; j < 4; j++)
                for (int i = 0; i < 4; i++) {
                    const float* p = H(gmin(x * 4 + i, w - 1), gmin(y * 4 + j, h - 1));
                    float wt = 1.0f / (1.0f + lum_f(p[0], p[1], p[2]));
                    sr += p[0] * wt, sg += p[1] * wt, sb += p[2] * wt;
                    weights += wt;
                    coverage += 1.0f - p[3];
                }
            float c[4] = {sr / weights, sg / weights, sb / weights, coverage / 16.0f};
            A.store(x, y, c);
        }
    // Pass 1 (bloom.comp:60-98): 64 lanes sum strided texels, then a tree.
    {
        float ls[64], ws[64];
        for (int t = 0; t < 64; t++) {
            float sum = 0, wsum = 0;
            for (int i = t; i < qw * qh; i += 64) {
                const float* c = A.at(i % qw, i / qw);
                sum += c[3] * std::log(gmax(lum_f(c[0], c[1], c[2]), 1e-4f));
                wsum += c[3];
            }
            ls[t] = sum, ws[t] = wsum;
        }
        for (int stride = 32; stride > 0; stride >>= 1)
            for (int t = 0; t < stride; t++) ls[t] += ls[t + stride], ws[t] += ws[t + stride];
        float target = 1.0f;
        if (ws[0] > 1.0f) {
            target = 0.24f / std::exp(ls[0] / ws[0]);
            target = target < 0.25f ? 0.25f : (target > 4.0f ? 4.0f : target);
        }
        float prev = *exposure;
        *exposure = prev <= 0.0f ? target : std::exp(mixr(std::log(prev), std::log(target), 0.15f));
    }
    blur(A, B, 0);  // pass 2
    blur(B, A, 1);  // pass 3
    for (int y = 0; y < eh; y++)  // pass 4: half-resolution copy of A
        for (int x = 0; x < ew; x++) {
            float sum[4] = {0, 0, 0, 0};
            for (int i = 0; i < 4; i++) {
                const float* c = A.at(gmin(x * 2 + (i & 1), qw - 1), gmin(y * 2 + (i >> 1), qh - 1));
                for (int k = 0; k < 4; k++) sum[k] += c[k];
            }
            for (int k = 0; k < 4; k++) sum[k] *= 0.25f;
            C.store(x, y, sum);
        }
    blur(C, D, 0);  // pass 5
    blur(D, C, 1);  // pass 6
    if (bloom_a) std::memcpy(bloom_a, A.v.data(), A.v.size() * 4);
    if (bloom_c) std::memcpy(bloom_c, C.v.data(), C.v.size() * 4);
    // tonemap.comp:48-85
    Ctx<float> cx(s, u);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const float* hp = H(x, y);
            float ex = u->ambient[3];
            if (u->finish[3] > 0.5f) ex *= *exposure;
            float c[3] = {hp[0], hp[1], hp[2]};
            float uu = ((float)x + 0.5f) / (float)w, vv = ((float)y + 0.5f) / (float)h;
            if (u->finish[0] > 0.0f) {
                float a4[4], c4[4];
                bilinear(A, uu, vv, a4);
                bilinear(C, uu, vv, c4);
                for (int k = 0; k < 3; k++) c[k] = mixr(c[k], 0.6f * a4[k] + 0.4f * c4[k], u->finish[0]);
            }
            float dx = uu * 2.0f - 1.0f, dy = vv * 2.0f - 1.0f;
            float r = std::sqrt(dx * dx + dy * dy) * 0.7071f;
            float vignette = 1.0f - u->finish[1] * std::pow(r, 2.5f);
            for (int k = 0; k < 3; k++) c[k] *= ex * vignette;
            uint32_t op = (uint32_t)u->material[2];
            if (op == 1u)
                for (int k = 0; k < 3; k++) c[k] = aces1(c[k]);
            else if (op == 2u)
                agx3(c);
            else
                for (int k = 0; k < 3; k++) c[k] = c[k] < 0.0f ? 0.0f : (c[k] > 1.0f ? 1.0f : c[k]);
            if (u->lighting[3] < 1u) {  // background (common.glsl:449-455), display-referred
                V3<float> d = cx.camera_ray((float)x + 0.5f, (float)y + 0.5f);
                float t = 0.5f + 0.5f * -d.z;
                const float lo[3] = {0.02f, 0.022f, 0.028f}, hi[3] = {0.08f, 0.085f, 0.1f};
                for (int k = 0; k < 3; k++) c[k] += hp[3] * mixr(lo[k], hi[k], t);
            }
            if (u->flags[1] != 0u)
                for (int k = 0; k < 3; k++) c[k] = srgb1(c[k]);
            if (u->finish[2] > 0.0f) {
                uint32_t px = (uint32_t)x, py = (uint32_t)y;
                float n = hash01(px + 4099u * py) + hash01(px * 7919u + py + 1u) - 1.0f;
                float yl = lum_f(c[0], c[1], c[2]);
                yl = yl < 0.0f ? 0.0f : (yl > 1.0f ? 1.0f : yl);
                for (int k = 0; k < 3; k++) {
                    float v = c[k] + n * u->finish[2] * (0.3f + 2.8f * yl * (1.0f - yl));
                    c[k] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
                }
            }
            uint8_t* o = out + 4 * ((size_t)y * w + x);
            o[0] = unorm8(c[0]), o[1] = unorm8(c[1]), o[2] = unorm8(c[2]), o[3] = 255;
        }
}

void oracle_sample_env(const oracle_scene* s, const vvr_uniforms* u, uint32_t seed, uint32_t n, float* dirs,
                       float* pdfs) {
    Ctx<float> c(s, u);
    c.rng = seed;
    for (uint32_t i = 0; i < n; i++) {
        float pdf;
        V3<float> d = c.sample_env(pdf);
        dirs[3 * i] = d.x, dirs[3 * i + 1] = d.y, dirs[3 * i + 2] = d.z;
        pdfs[i] = pdf;
    }
}

float oracle_env_pdf(const oracle_scene* s, const vvr_uniforms* u, const float d[3]) {
    Ctx<float> c(s, u);
    return c.env_pdf(V3<float>(d[0], d[1], d[2]));
}

void oracle_pt(const oracle_scene* s, const vvr_uniforms* u, const uint32_t* pixels, size_t n, int precision,
               float* accum, float* albedo, float* normal) {
#pragma omp parallel for schedule(dynamic, 4)
    for (long i = 0; i < (long)n; i++) {
        if (precision == 64) {
            Ctx<double> c(s, u);
            c.pt_pixel(pixels[2 * i], pixels[2 * i + 1], accum + 4 * i, albedo + 4 * i, normal + 4 * i);
        } else {
            Ctx<float> c(s, u);
            c.pt_pixel(pixels[2 * i], pixels[2 * i + 1], accum + 4 * i, albedo + 4 * i, normal + 4 * i);
        }
    }
}

float oracle_density_cubic(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], int precision) {
    if (precision == 64) {
        Ctx<double> c(s, u);
        return (float)c.density_cubic(V3<double>(idx[0], idx[1], idx[2]));
    }
    Ctx<float> c(s, u);
    return c.density_cubic(V3<float>(idx[0], idx[1], idx[2]));
}

void oracle_density_gradient_cubic(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], int precision,
                                   float out[3]) {
    if (precision == 64) {
        Ctx<double> c(s, u);
        V3<double> g = c.density_gradient_cubic(V3<double>(idx[0], idx[1], idx[2]));
        out[0] = (float)g.x, out[1] = (float)g.y, out[2] = (float)g.z;
        return;
    }
    Ctx<float> c(s, u);
    V3<float> g = c.density_gradient_cubic(V3<float>(idx[0], idx[1], idx[2]));
    out[0] = g.x, out[1] = g.y, out[2] = g.z;
}

float oracle_density_linear(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], int precision) {
    if (precision == 64) {
        Ctx<double> c(s, u);
        return (float)c.density_at_index(V3<double>(idx[0], idx[1], idx[2]));
    }
    Ctx<float> c(s, u);
    return c.density_at_index(V3<float>(idx[0], idx[1], idx[2]));
}

uint32_t oracle_pcg_sequence(uint32_t seed, uint32_t n, uint32_t* out) {
    oracle_scene sc{};
    vvr_uniforms un{};
    Ctx<float> c(&sc, &un);
    c.rng = seed;
    for (uint32_t i = 0; i < n; i++) out[i] = c.pcg();
    return c.rng;
}

void oracle_rand_sequence(uint32_t seed, uint32_t n, float* out) {
    oracle_scene sc{};
    vvr_uniforms un{};
    Ctx<float> c(&sc, &un);
    c.rng = seed;
    for (uint32_t i = 0; i < n; i++) out[i] = c.rand();
}

float oracle_phase_hg(float g, float cos_t) {
    oracle_scene sc{};
    vvr_uniforms un{};
    un.env[2] = g;
    Ctx<double> c(&sc, &un);
    return (float)c.phase_hg(cos_t);
}

void oracle_sample_hg(float g, uint32_t seed, uint32_t n, float* cos_out) {
    oracle_scene sc{};
    vvr_uniforms un{};
    un.env[2] = g;
    Ctx<float> c(&sc, &un);
    c.rng = seed;
    for (uint32_t i = 0; i < n; i++) {
        V3<float> d = c.sample_hg(V3<float>(0, 0, 1));
        cos_out[i] = d.z;
    }
}

float oracle_ggx_distribution(float nh, float a2) { return (float)Ctx<double>::ggx_distribution(nh, a2); }

void oracle_single_scattering_albedo(const float a[3], float out[3]) {
    V3<float> r = Ctx<float>::single_scattering_albedo(V3<float>(a[0], a[1], a[2]));
    out[0] = r.x, out[1] = r.y, out[2] = r.z;
}

int oracle_free_flight(const oracle_scene* s, const vvr_uniforms* u, const float origin_idx[3], const float dir[3],
                       uint32_t seed, float* t_out, float spectral[3]) {
    Ctx<float> c(s, u);
    c.rng = seed;
    V3<float> sp(1), idx;
    Ctx<float>::Sample smp{};
    int cap;
    float cut, t = -1;
    bool hit = c.free_flight(V3<float>(origin_idx[0], origin_idx[1], origin_idx[2]), V3<float>(dir[0], dir[1], dir[2]),
                             false, 1, sp, idx, smp, cap, cut, &t);
    *t_out = t;
    spectral[0] = sp.x, spectral[1] = sp.y, spectral[2] = sp.z;
    return hit ? 1 : 0;
}

void oracle_transmittance(const oracle_scene* s, const vvr_uniforms* u, const float origin_idx[3], const float dir[3],
                          uint32_t seed, float out[3]) {
    Ctx<float> c(s, u);
    c.rng = seed;
    V3<float> r = c.transmittance(V3<float>(origin_idx[0], origin_idx[1], origin_idx[2]), V3<float>(dir[0], dir[1], dir[2]));
    out[0] = r.x, out[1] = r.y, out[2] = r.z;
}

// ---- building blocks of the shading, label, clipping and sampling code -------------
}  // extern "C"

namespace {
template <class R> V3<R> v3(const float* a) { return V3<R>(R(a[0]), R(a[1]), R(a[2])); }
template <class R> void put3(V3<R> v, float* out) { out[0] = float(v.x), out[1] = float(v.y), out[2] = float(v.z); }
template <class R> typename Ctx<R>::Sample material(const float albedo[3], float roughness, float metalness, float coat) {
    typename Ctx<R>::Sample s{};
    s.albedo = v3<R>(albedo);
    s.roughness = R(roughness), s.metalness = R(metalness), s.coat = R(coat);
    return s;
}
template <class R>
void brdf_t(const float n[3], const float v[3], const float l[3], const float albedo[3], float roughness,
            float metalness, float coat, float spec[3], float diffuse[3]) {
    V3<R> d;
    V3<R> sp = Ctx<R>::brdf_coated(v3<R>(n), v3<R>(v), v3<R>(l), material<R>(albedo, roughness, metalness, coat), d);
    put3(sp, spec), put3(d, diffuse);
}
}  // namespace

extern "C" {

void oracle_brdf(const float n[3], const float v[3], const float l[3], const float albedo[3], float roughness,
                 float metalness, float coat, int precision, float spec[3], float diffuse[3]) {
    if (precision == 64) brdf_t<double>(n, v, l, albedo, roughness, metalness, coat, spec, diffuse);
    else brdf_t<float>(n, v, l, albedo, roughness, metalness, coat, spec, diffuse);
}

double oracle_smith(double nv, double nl, double a) { return Ctx<double>::smith(nv, nl, a); }

float oracle_bsdf_pdf(const float n[3], const float v[3], const float l[3], float roughness, float metalness,
                      float coat) {
    const float alb[3] = {1, 1, 1};
    return float(Ctx<double>::bsdf_pdf(v3<double>(n), v3<double>(v), v3<double>(l),
                                       material<double>(alb, roughness, metalness, coat)));
}

void oracle_bsdf_sample(const float n[3], const float v[3], float roughness, float metalness, float coat,
                        uint32_t seed, uint32_t count, float* dirs) {
    oracle_scene sc{};
    vvr_uniforms un{};
    Ctx<double> c(&sc, &un);
    c.rng = seed;
    const float alb[3] = {1, 1, 1};
    auto s_ = material<double>(alb, roughness, metalness, coat);
    for (uint32_t i = 0; i < count; i++) put3(c.sample_bsdf(v3<double>(n), v3<double>(v), s_), dirs + 3 * i);
}

void oracle_sample_light(const vvr_uniforms* u, const float dir[3], uint32_t seed, uint32_t count, float* dirs) {
    oracle_scene sc{};
    Ctx<double> c(&sc, u);
    c.rng = seed;
    auto frame = Ctx<double>::basis(v3<double>(dir));
    for (uint32_t i = 0; i < count; i++) put3(c.sample_light(frame), dirs + 3 * i);
}

float oracle_mis(float a, float b) { return float(Ctx<double>::mis(a, b)); }

void oracle_hash3(const int32_t c[3], float out[3]) {
    put3(Ctx<float>::hash3(V3<int>(c[0], c[1], c[2])), out);
}

float oracle_gradient_noise(const float p[3], int precision) {
    return precision == 64 ? float(Ctx<double>::gradient_noise(v3<double>(p))) : Ctx<float>::gradient_noise(v3<float>(p));
}

float oracle_fbm(const float p[3], int precision) {
    return precision == 64 ? float(Ctx<double>::fbm(v3<double>(p))) : Ctx<float>::fbm(v3<float>(p));
}

void oracle_cellular(const float p[3], int precision, float* nearest, float* border, int32_t cell[3]) {
    V3<int> c;
    if (precision == 64) {
        double nd, bd;
        Ctx<double>::cellular(v3<double>(p), nd, bd, c);
        *nearest = float(nd), *border = float(bd);
    } else {
        Ctx<float>::cellular(v3<float>(p), *nearest, *border, c);
    }
    cell[0] = c.x, cell[1] = c.y, cell[2] = c.z;
}

float oracle_surface_height(const float p[3], const float params[4], int precision) {
    return precision == 64 ? float(Ctx<double>::surface_height(v3<double>(p), params))
                           : Ctx<float>::surface_height(v3<float>(p), params);
}

void oracle_index_to_world(const vvr_uniforms* u, const float idx[3], float out[3]) {
    oracle_scene sc{};
    Ctx<double> c(&sc, u);
    put3(c.index_to_world(v3<double>(idx)), out);
}

void oracle_surface_texture(const vvr_uniforms* u, uint32_t label, const float idx[3], const float albedo[3],
                            float roughness, float coat, float out_albedo[3], float* out_roughness, float* out_coat) {
    oracle_scene sc{};
    Ctx<double> c(&sc, u);
    auto s_ = material<double>(albedo, roughness, 0.0f, coat);
    s_.label = label;
    c.apply_surface_texture(v3<double>(idx), s_);
    put3(s_.albedo, out_albedo);
    *out_roughness = float(s_.roughness), *out_coat = float(s_.coat);
}

void oracle_bump_normal(const vvr_uniforms* u, uint32_t label, const float idx[3], const float n[3], float out[3]) {
    oracle_scene sc{};
    Ctx<double> c(&sc, u);
    const float alb[3] = {1, 1, 1};
    auto s_ = material<double>(alb, 0.5f, 0.0f, 0.0f);
    s_.label = label;
    put3(c.bump_normal(v3<double>(idx), s_, v3<double>(n)), out);
}

int oracle_drawn_label_near(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], int ghost,
                            uint32_t* label) {
    Ctx<float> c(s, u);
    return c.drawn_label_near(v3<float>(idx), uint32_t(u->labels[0]), ghost != 0, *label) ? 1 : 0;
}

float oracle_with_label(const oracle_scene* s, const vvr_uniforms* u, float density, uint32_t label, int ghost,
                        float albedo[3]) {
    Ctx<double> c(s, u);
    auto s_ = c.with_label(c.tf_point(density), label, uint32_t(u->labels[0]), ghost != 0);
    put3(s_.albedo, albedo);
    return float(s_.extinction);
}

int oracle_apply_labels(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], float* extinction,
                        int* ghost) {
    Ctx<float> c(s, u);
    auto s_ = c.apply_labels(c.tf_point(0.5f), v3<float>(idx));
    *extinction = s_.extinction;
    *ghost = s_.ghost ? 1 : 0;
    return int(s_.label);
}

float oracle_clip_depth(const vvr_uniforms* u, const float idx[3], int* plane) {
    oracle_scene sc{};
    Ctx<double> c(&sc, u);
    return float(c.clip_depth(v3<double>(idx), *plane));
}

void oracle_tf_segment(const oracle_scene* s, const vvr_uniforms* u, float front, float back, int precision,
                       float out[4]) {
    if (precision == 64) {
        Ctx<double> c(s, u);
        auto t = c.tf_segment(front, back);
        put3(t.albedo, out), out[3] = float(t.extinction);
    } else {
        Ctx<float> c(s, u);
        auto t = c.tf_segment(front, back);
        put3(t.albedo, out), out[3] = t.extinction;
    }
}

float oracle_preint_half_depth(const oracle_scene* s, const vvr_uniforms* u, float front, float back, int precision) {
    if (precision == 64) return float(Ctx<double>(s, u).preint_half_depth(front, back));
    return Ctx<float>(s, u).preint_half_depth(front, back);
}

float oracle_surface_frame(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], uint32_t label,
                           int ghost, const float v[3], int cap_plane, float cut, float cap_depth, int precision,
                           float n[3], float* surfaceness) {
    auto run = [&](auto c) {
        using R = decltype(c.rng_real());
        typename decltype(c)::Sample s_{};
        s_.label = label, s_.ghost = ghost != 0;
        V3<R> nn;
        R sf;
        R scale = c.surface_frame(v3<R>(idx), s_, v3<R>(v), cap_plane, R(cut), R(cap_depth), nn, sf);
        put3(nn, n), *surfaceness = float(sf);
        return float(scale);
    };
    return precision == 64 ? run(Ctx<double>(s, u)) : run(Ctx<float>(s, u));
}

void oracle_vlog_vexp(const float* x, size_t n, float* log_out, float* exp_out) {
    for (size_t i = 0; i < n; i++) log_out[i] = vlog(x[i]), exp_out[i] = vexp(x[i]);
}

void oracle_vpow(const float* x, const float* y, size_t n, float* out) {
    for (size_t i = 0; i < n; i++) out[i] = vpow(x[i], y[i]);
}

// The kernels' division-free x / 6 (volview_amd/csrc/device/scene.hpp), restated
// here only to check it against the division this oracle performs.
static float kernel_div6_normal(float x) {
    const float third = 1.0f / 3.0f;
    const float h = x * 0.5f;
    const float q = h * third;
    return std::fma(std::fma(-q, 3.0f, h), third, q);
}
static float kernel_div6(float x) { return std::fabs(x) < 0x1p-125f ? x / 6.0f : kernel_div6_normal(x); }

uint64_t oracle_div6_mismatches(uint32_t first, uint32_t last, uint32_t stride, int mode) {
    uint64_t bad = 0;
    auto differ = [](float a, float b) { return f32_bits(a) != f32_bits(b); };
    for (uint64_t b = first; b <= last; b += stride) {
        const float x = f32_from_bits(uint32_t(b));
        if (mode == 0) {
            bad += differ(kernel_div6(x), x / 6.0f);
        } else if (mode == 2) {  // the product with RN(1/6) alone, without the correction
            bad += differ(x * (1.0f / 6.0f), x / 6.0f);
        } else {
            const float f = x, f2 = f * f, f3 = f2 * f;
            const float n0 = (1.0f - 3.0f * f) + 3.0f * f2 - f3, n1 = (4.0f - 6.0f * f2) + 3.0f * f3,
                        n2 = ((1.0f + 3.0f * f) + 3.0f * f2) - 3.0f * f3;
            bad += differ(kernel_div6_normal(n0), n0 / 6.0f) || differ(kernel_div6_normal(n1), n1 / 6.0f) ||
                   differ(kernel_div6_normal(n2), n2 / 6.0f) || differ(kernel_div6(f3), f3 / 6.0f);
        }
    }
    return bad;
}

void oracle_tissue_sigma(const vvr_uniforms* u, uint32_t label, float extinction, float out[3]) {
    oracle_scene sc{};
    Ctx<double> c(&sc, u);
    Ctx<double>::Sample s_{};
    s_.label = label, s_.extinction = extinction;
    put3(c.tissue_sigma(s_), out);
}

void oracle_pixel_stream(const vvr_uniforms* u, uint32_t px, uint32_t py, uint32_t n, uint32_t* out) {
    oracle_scene sc{};
    Ctx<float> c(&sc, u);
    c.seed_pixel(px, py);
    for (uint32_t i = 0; i < n; i++) out[i] = c.pcg();
}

}  // extern "C"
