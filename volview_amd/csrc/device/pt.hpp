// Transport of the path tracer: the HIP counterpart of shaders/pt.comp:41-371.
// A path's random numbers come from one PCG stream (pt.comp:41-53) seeded per
// (pixel, sample index) exactly as the reference seeds a frame of one path.
#pragma once
#include <hip/hip_fp16.h>

#include "shading.hpp"

namespace vvr {

constexpr float kMaxExtinction = 8.0f;     // pt.comp:34
constexpr uint32_t kMaxVolumeEvents = 32;  // pt.comp:112

struct Rng {
    uint32_t state;
    __device__ __forceinline__ uint32_t next() {  // pt.comp:44-49
        const uint32_t s = state;
        state = state * 747796405u + 2891336453u;
        const uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
        return (word >> 22u) ^ word;
    }
    __device__ __forceinline__ float rand() { return float(next() >> 8) * (1.0f / 16777216.0f); }  // pt.comp:51-53
};

// pt.comp:378-379 for a frame whose first path index is `path_index`.
__device__ __forceinline__ Rng seed_rng(uint32_t px, uint32_t py, uint32_t path_index) {
    Rng r{(px * 1973u + py * 9277u + path_index * 26699u) | 1u};
    r.next();
    return r;
}

// ---- light map (common.glsl:457-485, pt.comp:192-226) ------------------------------
// Equirectangular, fixed to the camera like the key light: the centre column
// ahead, the top row above, turned by env.x.

__device__ __forceinline__ void env_uv(const vvr_uniforms* u, V3 d, float& uu, float& vv) {
    const V3 f = normalize(V3(u->forward[0], u->forward[1], u->forward[2]));
    const V3 r = normalize(V3(u->right[0], u->right[1], u->right[2]));
    const V3 up = normalize(V3(u->up[0], u->up[1], u->up[2]));
    const float phi = cr_atan2(dot(d, r), dot(d, f));
    const float theta = cr_acos(clampf(dot(d, up), -1.0f, 1.0f));
    const float x = 0.5f + (phi + u->env[0]) / (2.0f * kPi);
    uu = x - floorf(x);
    vv = theta / kPi;
}

__device__ __forceinline__ V3 env_direction(const vvr_uniforms* u, float uu, float vv) {
    const float phi = (uu - 0.5f) * 2.0f * kPi - u->env[0];
    const float theta = vv * kPi;
    const V3 f = normalize(V3(u->forward[0], u->forward[1], u->forward[2]));
    const V3 r = normalize(V3(u->right[0], u->right[1], u->right[2]));
    const V3 up = normalize(V3(u->up[0], u->up[1], u->up[2]));
    return cr_sin(theta) * (cr_sin(phi) * r + cr_cos(phi) * f) + cr_cos(theta) * up;
}

__device__ __forceinline__ V3 env_texel(const uint2* img, int w, int x, int y) {
    const uint2 t = img[y * w + x];
    return V3(__half2float(__ushort_as_half((unsigned short)(t.x & 0xffffu))),
              __half2float(__ushort_as_half((unsigned short)(t.x >> 16))),
              __half2float(__ushort_as_half((unsigned short)(t.y & 0xffffu))));
}

// The renderer's linear clamp-to-edge sampler on an RGBA16F image at (uu, vv).
__device__ __forceinline__ V3 env_image(const uint2* img, int w, int h, float uu, float vv) {
    const float tx = uu * float(w) - 0.5f, ty = vv * float(h) - 0.5f;
    const float fx = floorf(tx), fy = floorf(ty);
    const float ax = tx - fx, ay = ty - fy;
    const int x0 = min(max(int(fx), 0), w - 1), x1 = min(max(int(fx) + 1, 0), w - 1);
    const int y0 = min(max(int(fy), 0), h - 1), y1 = min(max(int(fy) + 1, 0), h - 1);
    return vmix(vmix(env_texel(img, w, x0, y0), env_texel(img, w, x1, y0), ax),
                vmix(env_texel(img, w, x0, y1), env_texel(img, w, x1, y1), ax), ay);
}

__device__ __forceinline__ V3 env_map_radiance(const DevScene& s, V3 d) {  // common.glsl:477-480
    float uu, vv;
    env_uv(s.u, d, uu, vv);
    return V3(s.u->ambient[0], s.u->ambient[1], s.u->ambient[2]) * env_image(s.env_rad, s.env_w, s.env_h, uu, vv);
}

__device__ __forceinline__ V3 env_map_diffuse(const DevScene& s, V3 n) {  // common.glsl:482-485
    float uu, vv;
    env_uv(s.u, n, uu, vv);
    return V3(s.u->ambient[0], s.u->ambient[1], s.u->ambient[2]) * env_image(s.env_dif, 32, 16, uu, vv);
}

// Solid-angle pdf of drawing direction d from the light map (pt.comp:193-199).
__device__ __forceinline__ float env_pdf(const DevScene& s, V3 d) {
    float uu, vv;
    env_uv(s.u, d, uu, vv);
    const int tx = min(int(uu * float(s.env_w)), s.env_w - 1), ty = min(int(vv * float(s.env_h)), s.env_h - 1);
    const float p = s.env_table[ty * (s.env_w + 1) + tx].y;
    return p * float(s.env_w * s.env_h) / (2.0f * kPi * kPi * gmax(cr_sin(vv * kPi), 1e-6f));
}

// A direction drawn proportionally to the map's luminance: a row by binary
// search of the marginal CDF, then a column of the row's CDF (pt.comp:203-226).
__device__ inline V3 sample_env(const DevScene& s, Rng& rng, float& pdf) {
    const int w = s.env_w, h = s.env_h;
    float xi = rng.rand();
    int lo = 0, hi = h - 1;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (s.env_table[mid * (w + 1) + w].x >= xi) hi = mid;
        else lo = mid + 1;
    }
    const int y = lo;
    xi = rng.rand();
    lo = 0, hi = w - 1;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (s.env_table[y * (w + 1) + mid].x >= xi) hi = mid;
        else lo = mid + 1;
    }
    const float ju = rng.rand();
    const float jv = rng.rand();
    const float uu = (float(lo) + ju) / float(w), vv = (float(y) + jv) / float(h);
    const float p = s.env_table[y * (w + 1) + lo].y;
    pdf = p * float(w * h) / (2.0f * kPi * kPi * gmax(cr_sin(vv * kPi), 1e-6f));
    return env_direction(s.u, uu, vv);
}

// pt.comp:56-82: the light map, the studio or the dome.
__device__ inline V3 environment(const DevScene& s, V3 d) {
    const vvr_uniforms* u = s.u;
    if (u->lighting[3] == 2u) return env_map_radiance(s, d);
    const V3 amb(u->ambient[0], u->ambient[1], u->ambient[2]);
    if (u->lighting[3] == 1u) {
        const V3 f = normalize(V3(u->forward[0], u->forward[1], u->forward[2]));
        const V3 r = normalize(V3(u->right[0], u->right[1], u->right[2]));
        const V3 up = normalize(V3(u->up[0], u->up[1], u->up[2]));
        const float key_side = dot(V3(u->light_dir[0], u->light_dir[1], u->light_dir[2]), r) >= 0.0f ? 1.0f : -1.0f;
        const V3 fill = normalize(((-f) - (0.9f * key_side) * r) - 0.1f * up);
        const V3 rim = normalize((f + 0.8f * up) - (0.4f * key_side) * r);
        const float fill_lobe = 0.5f * vpow(gmax(dot(d, fill), 0.0f), 6.0f);
        const float rim_lobe = 2.5f * vpow(gmax(dot(d, rim), 0.0f), 24.0f);
        const V3 studio = V3(0.02f) + fill_lobe * V3(0.95f, 0.98f, 1.05f) + rim_lobe * V3(0.9f, 0.95f, 1.1f);
        return amb * studio * (1.0f / 0.106f);
    }
    const float up = 0.5f + 0.5f * d.z;
    return amb * vmix(V3(1.04f, 1.0f, 0.94f) * 0.35f, V3(0.92f, 0.96f, 1.04f), up) / 0.675f;
}

__device__ __forceinline__ M3 basis(V3 n) {  // pt.comp:85-92
    const float s = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (s + n.z);
    const float b = n.x * n.y * a;
    M3 m;
    m.c0 = V3(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
    m.c1 = V3(b, s + n.y * n.y * a, -n.y);
    m.c2 = n;
    return m;
}

__device__ __forceinline__ V3 sample_light(const vvr_uniforms* u, const M3& frame, Rng& rng) {  // pt.comp:96-102
    const float cos_max = cr_cos(u->material[3]);
    const float cos_t = mixf(cos_max, 1.0f, rng.rand());
    const float sin_t = sqrtf(gmax(0.0f, 1.0f - cos_t * cos_t));
    const float phi = 2.0f * kPi * rng.rand();
    return frame.mul(V3(sin_t * cr_cos(phi), sin_t * cr_sin(phi), cos_t));
}

__device__ __forceinline__ V3 single_scattering_albedo(V3 a) {  // pt.comp:119-123
    a = vclamp(a, 0.0f, 0.999f);
    const V3 sq = 9.59217f + 41.6808f * a + 17.7126f * a * a;
    const V3 s = 4.09712f + 4.20863f * a - V3(sqrtf(sq.x), sqrtf(sq.y), sqrtf(sq.z));
    return 1.0f - s * s;
}

__device__ __forceinline__ V3 chroma(const vvr_uniforms* u, const Sample& s) {  // pt.comp:132-139
    if (s.ghost) return V3(1.0f);
    return vmix(V3(1.0f), V3(0.62f, 0.77f, 1.0f), clampf(2.0f * s.subsurface, 0.0f, 1.0f) * u->env[3]);
}

__device__ __forceinline__ V3 path_sigma(const vvr_uniforms* u, const Sample& s, float extinction) {  // :143-145
    return measured(u, s) ? tissue_sigma(u, s) : extinction * chroma(u, s);
}

__device__ __forceinline__ V3 event_albedo(const vvr_uniforms* u, const Sample& s) {  // pt.comp:150-153
    const V3 reduced = single_scattering_albedo(s.albedo);
    return measured(u, s) ? 1.0f - (1.0f - reduced) * (1.0f - clampf(u->env[2], 0.0f, 0.95f)) : reduced;
}

__device__ __forceinline__ float flight_majorant(const DevScene& sc, V3 p) {  // pt.comp:157-163
    const vvr_uniforms* u = sc.u;
    float m = brick_majorant(sc, p);
    if (m > 0.0f && u->labels[0] >= 2.0f) m = gmax(m, u->tissue[0]);
    return gmin(m, gmax(kMaxExtinction, gmax(u->labels[2], u->tissue[0])));
}

__device__ __forceinline__ float phase_hg(const vvr_uniforms* u, float cos_t) {  // pt.comp:167-171
    const float g = u->env[2];
    const float d = 1.0f + g * g - 2.0f * g * cos_t;
    return (1.0f - g * g) / (4.0f * kPi * d * sqrtf(gmax(d, 1e-8f)));
}

__device__ __forceinline__ V3 sample_hg(const vvr_uniforms* u, V3 dir, Rng& rng) {  // pt.comp:173-185
    const float g = u->env[2];
    float cos_t;
    if (fabsf(g) < 1e-3f) {
        cos_t = 1.0f - 2.0f * rng.rand();
    } else {
        const float q = (1.0f - g * g) / (1.0f - g + 2.0f * g * rng.rand());
        cos_t = clampf((1.0f + g * g - q * q) / (2.0f * g), -1.0f, 1.0f);
    }
    const float sin_t = sqrtf(gmax(0.0f, 1.0f - cos_t * cos_t));
    const float phi = 2.0f * kPi * rng.rand();
    return basis(dir).mul(V3(sin_t * cr_cos(phi), sin_t * cr_sin(phi), cos_t));
}

__device__ __forceinline__ float mis(float a, float b) { return a * a / gmax(a * a + b * b, 1e-20f); }
__device__ __forceinline__ float coat_share(const Sample& s) { return 0.5f * s.coat; }
__device__ __forceinline__ float specular_share(const Sample& s) { return mixf(0.2f, 1.0f, s.metalness); }

__device__ __forceinline__ float ggx_pdf(V3 n, V3 v, V3 l, float a) {  // pt.comp:246-250
    const V3 h = normalize(v + l);
    const float nh = gmax(dot(n, h), 0.0f);
    return ggx_distribution(nh, a * a) * nh / (4.0f * gmax(dot(v, h), 1e-4f));
}

__device__ __forceinline__ V3 sample_ggx(const M3& frame, V3 v, float a, Rng& rng) {  // pt.comp:252-260
    const float a2 = a * a;
    const float phi = 2.0f * kPi * rng.rand();
    const float xi = rng.rand();
    const float cos_t = sqrtf((1.0f - xi) / (1.0f + (a2 - 1.0f) * xi));
    const float sin_t = sqrtf(gmax(0.0f, 1.0f - cos_t * cos_t));
    const V3 h = frame.mul(V3(sin_t * cr_cos(phi), sin_t * cr_sin(phi), cos_t));
    return reflect(-v, h);
}

__device__ __forceinline__ float bsdf_pdf(V3 n, V3 v, V3 l, const Sample& s) {  // pt.comp:263-273
    const float nl = dot(n, l);
    if (nl <= 0.0f) return 0.0f;
    const float a = gmax(s.roughness * s.roughness, 2e-3f);
    const float share = specular_share(s);
    const float base = share * ggx_pdf(n, v, l, a) + (1.0f - share) * nl / kPi;
    const float coat = coat_share(s);
    return coat * ggx_pdf(n, v, l, coat_roughness(s) * coat_roughness(s)) + (1.0f - coat) * base;
}

// pt.comp:277-292. kGhosts false: for frames whose labels cannot be drawn as
// ghosts (no label colour's alpha above 1.5), which never take the ghost branch;
// without its surface frame the free-flight kernel needs 222 instead of 256+
// registers and no spills. kDensity false: for frames in isolate mode, which
// never read the density (density_unused); without its fetches the kernel runs
// at three waves per SIMD.
template <bool kGhosts = true, bool kDensity = true>
__device__ inline float extinction_at(const DevScene& sc, V3 idx, V3 dir, bool smooth_density, Sample& s, int& cap_plane,
                                      float& cut) {
    cut = clip_depth(sc.u, idx, cap_plane);
    if (cut > 0.0f) return 0.0f;
    const float d = (!kDensity || density_unused(sc.u))
                        ? 0.0f
                        : (smooth_density ? density_view(sc, idx) : density_at_index(sc, idx));
    s = apply_labels(sc, tf_point(sc, d), idx);
    if (s.extinction <= 0.0f) return 0.0f;
    if (kGhosts && s.ghost) {
        V3 n;
        float surfaceness;
        return s.extinction * surface_frame(sc, idx, s, -dir, cap_plane, cut, 0.0f, n, surfaceness);
    }
    return s.extinction;
}

// The loop state of delta and ratio tracking (pt.comp:300-371): the ray in
// index space, the current distance and far end, and the iteration count (at
// most 4096).
struct Flight {
    V3 origin, d_idx;
    V3 inv_d;  // brick_exit_inverse(d_idx)
    float tt, t1;
    int i;
};

__device__ __forceinline__ Flight flight_begin(const vvr_uniforms* u, V3 origin, V3 dir) {
    Flight f;
    f.origin = origin;
    f.d_idx = to_index(u).mul(dir);
    f.inv_d = brick_exit_inverse(f.d_idx);
    float t0;
    intersect_volume(u, origin, f.d_idx, t0, f.t1);
    f.tt = t0;
    f.i = 0;
    return f;
}

// One iteration of the tracking loop up to the density sample (pt.comp:302-318,
// 351-362): the majorant of the brick at f.tt, then either the brick exit (an
// empty brick, or a tentative distance past the exit: false) or a tentative
// collision at the new f.tt (true, `majorant` set).
__device__ __forceinline__ bool flight_step(const DevScene& sc, Flight& f, Rng& rng, float& majorant) {
    const V3 p = f.origin + f.d_idx * f.tt;
    majorant = flight_majorant(sc, p);
    const float exit = f.tt + brick_exit_fast(p, f.inv_d) + 1e-3f;
    if (majorant <= 0.0f) {
        f.tt = exit;
        return false;
    }
    const float next = f.tt - vlog(1.0f - rng.rand()) / majorant;
    if (next >= exit) {
        f.tt = exit;
        return false;
    }
    f.tt = next;
    return true;
}

// How a bounded run of a tracking loop ended.
enum Flow : int { kFlowExit = 0, kFlowHit = 1, kFlowSuspended = 2 };

// Delta tracking against the brick bounds (pt.comp:300-334) from the loop state
// `f`, stopping before iteration `stop` (kFlowSuspended; `f`, `spectral` and
// `rng` then resume it exactly where it stopped).
template <bool kGhosts = true, bool kDensity = true>
__device__ inline Flow free_flight_steps(const DevScene& sc, Flight& f, V3 dir, bool smooth_density, int hero,
                                         V3& spectral, V3& idx, Sample& s, int& cap_plane, float& cut, Rng& rng,
                                         int stop) {
    const vvr_uniforms* u = sc.u;
    const int end = min(stop, 4096);
    for (; f.i < end && f.tt < f.t1; f.i++) {
        float majorant;
        if (!flight_step(sc, f, rng, majorant)) continue;
        idx = f.origin + f.d_idx * f.tt;
        const float extinction = extinction_at<kGhosts, kDensity>(sc, idx, dir, smooth_density, s, cap_plane, cut);
        if (extinction <= 0.0f) continue;
        const V3 sigma = path_sigma(u, s, extinction);
        const float sh = component(sigma, hero);
        if (rng.rand() * majorant < sh) {
            spectral = spectral * (sigma / sh);
            return kFlowHit;
        }
        spectral = spectral * (vmax(V3(majorant) - sigma, V3(0.0f)) / (majorant - sh));
    }
    return f.i < 4096 && f.tt < f.t1 ? kFlowSuspended : kFlowExit;
}

__device__ inline bool free_flight(const DevScene& sc, V3 origin, V3 dir, bool smooth_density, int hero, V3& spectral,
                                   V3& idx, Sample& s, int& cap_plane, float& cut, Rng& rng) {
    Flight f = flight_begin(sc.u, origin, dir);
    return free_flight_steps(sc, f, dir, smooth_density, hero, spectral, idx, s, cap_plane, cut, rng, 4096) ==
           kFlowHit;
}

// Ratio tracking with Russian roulette (pt.comp:339-371) from the loop state
// `f` and running `result`, stopping before iteration `stop` (false; resumable
// as free_flight_steps). True once the transmittance is final.
__device__ __forceinline__ bool transmittance_steps(const DevScene& sc, Flight& f, V3& result, Rng& rng, int stop) {
    const vvr_uniforms* u = sc.u;
    const int end = min(stop, 4096);
    for (; f.i < end && f.tt < f.t1; f.i++) {
        const float peak = gmax(result.x, gmax(result.y, result.z));
        if (peak < 0.1f) {
            if (rng.rand() * 0.1f >= peak) {
                result = V3(0.0f);
                return true;
            }
            result = result * (0.1f / peak);
        }
        float majorant;
        if (!flight_step(sc, f, rng, majorant)) continue;
        const Sample m = sample_point(sc, f.origin + f.d_idx * f.tt);
        result = result * vmax(1.0f - path_sigma(u, m, m.extinction) / majorant, V3(0.0f));
    }
    return !(f.i < 4096 && f.tt < f.t1);
}

__device__ __forceinline__ V3 transmittance(const DevScene& sc, V3 origin, V3 dir, Rng& rng) {
    Flight f = flight_begin(sc.u, origin, dir);
    V3 result(1.0f);
    transmittance_steps(sc, f, result, rng, 4096);
    return result;
}

// Per-frame constants of the path tracer (pt.comp:381-397).
struct FrameConsts {
    uint32_t max_bounces, volume_limit;
    float step_mm;
    V3 ldir;
    M3 light_frame;
    V3 light;
    float cos_light, light_pdf;
    bool area_light;
    bool env_map;  // Environment::Map with a light map (pt.comp:395)
    M3 ti;
    V3 eye_idx;
};

__device__ __forceinline__ FrameConsts frame_consts(const vvr_uniforms* u) {
    FrameConsts c;
    c.max_bounces = u->path[1];
    c.volume_limit = uint32_t(float(kMaxVolumeEvents) / (1.0f - clampf(u->env[2], 0.0f, 0.95f)));
    c.step_mm = u->dims[3];
    c.ldir = V3(u->light_dir[0], u->light_dir[1], u->light_dir[2]);
    c.light_frame = basis(c.ldir);
    c.light = V3(u->light_dir[3]);
    c.cos_light = cr_cos(u->material[3]);
    c.area_light = u->material[3] > 1e-4f;
    c.light_pdf = c.area_light ? 1.0f / (2.0f * kPi * (1.0f - c.cos_light)) : 0.0f;
    c.env_map = u->lighting[3] == 2u;
    c.ti = to_index(u);
    c.eye_idx = world_to_index_point(u, V3(u->eye[0], u->eye[1], u->eye[2]));
    return c;
}

// One path in flight: the locals of pt.comp's sample loop (pt.comp:404-545),
// resumable between scattering events so a persistent lane can interleave them.
struct PathState {
    V3 origin, dir, throughput, spectral, radiance;
    float last_pdf;
    uint32_t bounce, surfaces, volumes;
    int hero;
    Rng rng;
    // What the path adds to the guide images (first event or background).
    float background;
    V3 albedo, normal;
    float depth;
};

// pt.comp:404-423: camera ray (thin lens), hero channel.
__device__ inline void path_begin(const DevScene& sc, const FrameConsts& fc, uint32_t px, uint32_t py,
                                  uint32_t path_index, PathState& st) {
    const vvr_uniforms* u = sc.u;
    st.rng = seed_rng(px, py, path_index);
    const float jx = st.rng.rand();
    const float jy = st.rng.rand();
    V3 dir = camera_ray(u, float(px) + jx, float(py) + jy);
    V3 origin = fc.eye_idx;
    if (u->lens[0] > 0.0f) {  // thin lens (pt.comp:406-416)
        V3 eye(u->eye[0], u->eye[1], u->eye[2]);
        const V3 fwd = normalize(V3(u->forward[0], u->forward[1], u->forward[2]));
        const V3 focal = eye + dir * (u->lens[1] / dot(dir, fwd));
        const float r = u->lens[0] * sqrtf(st.rng.rand());
        const float phi = 2.0f * kPi * st.rng.rand();
        const V3 rr = normalize(V3(u->right[0], u->right[1], u->right[2]));
        const V3 uu = normalize(V3(u->up[0], u->up[1], u->up[2]));
        eye = eye + r * (cr_cos(phi) * rr + cr_sin(phi) * uu);
        dir = normalize(focal - eye);
        origin = world_to_index_point(u, eye);
    }
    st.origin = origin;
    st.dir = dir;
    st.throughput = V3(1.0f);
    st.radiance = V3(0.0f);
    st.hero = min(int(st.rng.rand() * 3.0f), 2);
    st.spectral = V3(1.0f);
    st.last_pdf = 0.0f;
    st.bounce = st.surfaces = st.volumes = 0u;
    st.background = 0.0f;
    st.albedo = V3(0.0f);
    st.normal = V3(0.0f);
    st.depth = 0.0f;
}

// A scattering event between the stages of one bounce (pt.comp:429-534).
struct EventState {
    V3 idx;          // event position (index space)
    Sample s;        // the material there (after the surface texture)
    int cap_plane;
    float cut;
    V3 n;            // shading normal (valid where used)
    bool surface_event;
    V3 light_dir;    // key-light sample
    V3 light_t;      // its transmittance
    V3 map_dir;      // light-map sample (pt.comp:467-476)
    float map_pdf;   // 0: none
    V3 map_light;    // its radiance / pdf x transmittance
};

// pt.comp:427-428: whether the bounce loop runs another iteration.
__device__ __forceinline__ bool path_continues(const FrameConsts& fc, const PathState& st) {
    return st.surfaces <= fc.max_bounces && st.volumes <= fc.volume_limit;
}

// pt.comp:434-447: what a path that left the volume adds: the background guides
// on bounce 0, else the environment (MIS-weighted against the light-map
// samples) and the key light seen along the ray.
__device__ __forceinline__ void path_escape(const DevScene& sc, const FrameConsts& fc, PathState& st) {
    if (st.bounce == 0u) {
        st.background = 1.0f;
        st.albedo = V3(1.0f);
        st.depth = 1e4f;
    } else {
        const float weight = fc.env_map ? mis(st.last_pdf, env_pdf(sc, st.dir)) : 1.0f;
        const V3 path_weight = st.throughput * st.spectral / dot(st.spectral, V3(1.0f / 3.0f));
        st.radiance = st.radiance + path_weight * environment(sc, st.dir) * weight;
        if (fc.area_light && dot(st.dir, fc.ldir) >= fc.cos_light)
            st.radiance = st.radiance + path_weight * fc.light * fc.light_pdf * mis(st.last_pdf, fc.light_pdf);
    }
}

// pt.comp:433-447: free flight; on escape the environment and the key light
// (bounce > 0) or the background guides (bounce 0). False once the path ended.
__device__ __forceinline__ bool stage_extend(const DevScene& sc, const FrameConsts& fc, PathState& st, EventState& ev) {
    ev.cap_plane = -1;
    ev.cut = 0.0f;
    if (free_flight(sc, st.origin, st.dir, st.bounce == 0u, st.hero, st.spectral, ev.idx, ev.s, ev.cap_plane, ev.cut,
                    st.rng))
        return true;
    path_escape(sc, fc, st);
    return false;
}

// stage_extend as a bounded run from the flight state `f` (flight_begin of the
// path's ray, or a suspended run's state): at most up to iteration `stop`.
template <bool kGhosts = true, bool kDensity = true>
__device__ __forceinline__ Flow stage_extend_steps(const DevScene& sc, const FrameConsts& fc, PathState& st,
                                                   EventState& ev, Flight& f, int stop) {
    ev.cap_plane = -1;
    ev.cut = 0.0f;
    const Flow flow = free_flight_steps<kGhosts, kDensity>(sc, f, st.dir, st.bounce == 0u, st.hero, st.spectral, ev.idx, ev.s,
                                                 ev.cap_plane, ev.cut, st.rng, stop);
    if (flow == kFlowExit) path_escape(sc, fc, st);
    return flow;
}

// pt.comp:448-466 up to the shadow ray: shading frame, surface texture, the
// surface-or-volume decision, first-event guides and the key-light sample.
__device__ __forceinline__ void stage_shade(const DevScene& sc, const FrameConsts& fc, PathState& st, EventState& ev) {
    const vvr_uniforms* u = sc.u;
    const V3 v = -st.dir;
    float surfaceness;
    surface_frame(sc, ev.idx, ev.s, v, ev.cap_plane, ev.cut, 1.5f * fc.step_mm, ev.n, surfaceness, st.bounce > 0u);
    apply_surface_texture(u, ev.idx, ev.s);
    const float xi_surface = st.rng.rand();
    ev.surface_event = xi_surface < surfaceness * (1.0f - ev.s.subsurface) && dot(ev.n, ev.n) > 0.5f;
    if (ev.surface_event || st.bounce == 0u) ev.n = bump_normal(u, ev.idx, ev.s, ev.n);
    if (st.bounce == 0u) {
        st.albedo = ev.s.albedo;
        st.normal = dot(ev.n, v) < 0.0f ? -ev.n : ev.n;
        st.depth = length(ev.idx - fc.eye_idx);
    }
    ev.light_dir = sample_light(u, fc.light_frame, st.rng);
}

// pt.comp:466: the shadow ray.
__device__ __forceinline__ void stage_shadow(const DevScene& sc, PathState& st, EventState& ev) {
    ev.light_t = transmittance(sc, ev.idx, ev.light_dir, st.rng);
}

// pt.comp:467-476: with a light map, a light-map sample and its shadow ray
// (after the key light's, in the stream's order). A stage of its own so frames
// without a map do not carry its registers.
__device__ __forceinline__ void stage_shadow_map(const DevScene& sc, PathState& st, EventState& ev) {
    ev.map_pdf = 0.0f;
    ev.map_dir = ev.map_light = V3(0.0f);
    if (sc.u->lighting[3] != 2u) return;
    ev.map_dir = sample_env(sc, st.rng, ev.map_pdf);
    ev.map_light = env_map_radiance(sc, ev.map_dir) / gmax(ev.map_pdf, 1e-8f);
    if (ev.map_pdf > 0.0f) ev.map_light = ev.map_light * transmittance(sc, ev.idx, ev.map_dir, st.rng);
}

// pt.comp:478-544: next-event estimate, the next direction, Russian roulette.
// False once the path ended.
__device__ __forceinline__ bool stage_scatter(const DevScene& sc, const FrameConsts& fc, PathState& st,
                                              EventState& ev) {
    const vvr_uniforms* u = sc.u;
    const V3 path_weight = st.throughput * st.spectral / dot(st.spectral, V3(1.0f / 3.0f));
    const V3 v = -st.dir;
    const Sample& s = ev.s;
    if (ev.surface_event) {
        const V3 n = dot(ev.n, v) < 0.0f ? -ev.n : ev.n;
        V3 diffuse;
        V3 specular = brdf_coated(n, v, ev.light_dir, s, diffuse);
        const float light_weight = fc.area_light ? mis(fc.light_pdf, bsdf_pdf(n, v, ev.light_dir, s)) : 1.0f;
        st.radiance = st.radiance + path_weight * (diffuse + specular) * gmax(dot(n, ev.light_dir), 0.0f) * fc.light *
                                        ev.light_t * light_weight;
        if (ev.map_pdf > 0.0f && dot(n, ev.map_dir) > 0.0f) {  // pt.comp:485-489
            specular = brdf_coated(n, v, ev.map_dir, s, diffuse);
            st.radiance = st.radiance + path_weight * (diffuse + specular) * dot(n, ev.map_dir) * ev.map_light *
                                            mis(ev.map_pdf, bsdf_pdf(n, v, ev.map_dir, s));
        }
        const M3 frame = basis(n);
        V3 l;
        if (st.rng.rand() < coat_share(s)) {
            l = sample_ggx(frame, v, coat_roughness(s) * coat_roughness(s), st.rng);
        } else if (st.rng.rand() < specular_share(s)) {
            l = sample_ggx(frame, v, gmax(s.roughness * s.roughness, 2e-3f), st.rng);
        } else {
            const float r = sqrtf(st.rng.rand());
            const float phi = 2.0f * kPi * st.rng.rand();
            l = frame.mul(V3(r * cr_cos(phi), r * cr_sin(phi), sqrtf(gmax(0.0f, 1.0f - r * r))));
        }
        const float nl = dot(n, l);
        if (nl <= 0.0f) return false;
        const float pdf = bsdf_pdf(n, v, l, s);
        specular = brdf_coated(n, v, l, s, diffuse);
        st.throughput = st.throughput * (((diffuse + specular) * nl) / gmax(pdf, 1e-6f));
        st.last_pdf = pdf;
        st.surfaces++;
        st.origin = ev.idx + fc.ti.mul(n * fc.step_mm);
        st.dir = l;
    } else {
        const V3 albedo = event_albedo(u, s);
        const float light_phase = phase_hg(u, dot(st.dir, ev.light_dir));
        st.radiance = st.radiance + path_weight * albedo * light_phase * fc.light * ev.light_t *
                                        (fc.area_light ? mis(fc.light_pdf, light_phase) : 1.0f);
        if (ev.map_pdf > 0.0f) {  // pt.comp:524-527
            const float phase = phase_hg(u, dot(st.dir, ev.map_dir));
            st.radiance = st.radiance + path_weight * albedo * phase * ev.map_light * mis(ev.map_pdf, phase);
        }
        st.throughput = st.throughput * albedo;
        st.volumes++;
        st.origin = ev.idx;
        const V3 next = sample_hg(u, st.dir, st.rng);
        st.last_pdf = phase_hg(u, dot(st.dir, next));
        st.dir = next;
    }
    if (st.bounce >= 2u) {  // Russian roulette (pt.comp:536-544)
        const V3 w = st.throughput * st.spectral / dot(st.spectral, V3(1.0f / 3.0f));
        const float survive = clampf(gmax(w.x, gmax(w.y, w.z)), 0.05f, 1.0f);
        if (st.rng.rand() > survive) return false;
        st.throughput = st.throughput / survive;
    }
    st.bounce++;
    return true;
}

// Firefly clamp (pt.comp:546-548): the radiance the path adds.
__device__ __forceinline__ V3 path_end(const PathState& st) {
    const float peak = gmax(st.radiance.x, gmax(st.radiance.y, st.radiance.z));
    return peak > 20.0f ? st.radiance * (20.0f / peak) : st.radiance;
}

}  // namespace vvr
