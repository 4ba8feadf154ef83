// Test infrastructure only: the CPU oracle of the volume-rendering hot path.
// Nothing in the shipped library includes, links or calls this.
//
// Restatement of the host-side arithmetic that feeds the kernels: transfer
// functions (crates/render/src/transfer.rs), the orbit camera
// (crates/render/src/camera.rs), volume geometry (crates/volio/src/volume.rs)
// and the per-frame uniform block (crates/render/src/renderer.rs:1499-1660).
// All in f32 like the Rust host, except where the reference uses f64.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../include/vvr.h"
#include "../include/vvr_uniforms.h"
#include "oracle_api.h"
#include "oracle_math.hpp"

using namespace oracle;
using V = V3<float>;

static const float HU_MIN = -1024.0f;  // transfer.rs:2
static const float HU_MAX = 3071.0f;   // transfer.rs:3
static const int LUT_SIZE = 4096;      // transfer.rs:4

// Rust f32::clamp (panics aside): below min -> min, above max -> max.
static float rclamp(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

// transfer.rs:121-124
extern "C" uint16_t oracle_hu_to_unorm16(int16_t hu) {
    float t = rclamp(((float)hu - HU_MIN) / (HU_MAX - HU_MIN), 0.0f, 1.0f);
    return (uint16_t)std::round(t * 65535.0f);  // f32::round: half away from zero
}

// transfer.rs:72-86
extern "C" void oracle_tf_evaluate(const vvr_control_point* pts, uint32_t n, float hu, float out[4]) {
    auto rgba = [&](const vvr_control_point& p, float* o) {
        o[0] = p.color[0];
        o[1] = p.color[1];
        o[2] = p.color[2];
        o[3] = p.extinction;
    };
    if (hu <= pts[0].hu) {
        rgba(pts[0], out);
        return;
    }
    for (uint32_t w = 0; w + 1 < n; w++) {
        if (hu <= pts[w + 1].hu) {
            float t = (hu - pts[w].hu) / std::max(pts[w + 1].hu - pts[w].hu, 1e-6f);
            float a[4], b[4];
            rgba(pts[w], a);
            rgba(pts[w + 1], b);
            for (int i = 0; i < 4; i++) out[i] = a[i] + (b[i] - a[i]) * t;
            return;
        }
    }
    rgba(pts[n - 1], out);
}

// transfer.rs:90-97
extern "C" void oracle_tf_lut(const vvr_control_point* pts, uint32_t n, float* lut) {
    for (int i = 0; i < LUT_SIZE; i++) {
        float t = (float)i / (float)(LUT_SIZE - 1);
        oracle_tf_evaluate(pts, n, HU_MIN + t * (HU_MAX - HU_MIN), lut + 4 * i);
    }
}

// renderer.rs:2177-2195: the uploaded RGBA16F LUT.
extern "C" void oracle_tf_lut_f16(const vvr_control_point* pts, uint32_t n, uint16_t* out) {
    float lut[LUT_SIZE * 4];
    oracle_tf_lut(pts, n, lut);
    for (int i = 0; i < LUT_SIZE * 4; i++) out[i] = f32_to_f16(lut[i]);
}

// transfer.rs:103-117 (trapezoid rule in f64).
extern "C" void oracle_tf_cumulative(const vvr_control_point* pts, uint32_t n, float* out) {
    float lut[LUT_SIZE * 4];
    oracle_tf_lut(pts, n, lut);
    double dx = 1.0 / (double)(LUT_SIZE - 1);
    double sum[4] = {0, 0, 0, 0};
    for (int c = 0; c < 4; c++) out[c] = 0.0f;
    for (int i = 1; i < LUT_SIZE; i++) {
        const float* e0 = lut + 4 * (i - 1);
        const float* e1 = lut + 4 * i;
        double a[4] = {(double)(e0[0] * e0[3]), (double)(e0[1] * e0[3]), (double)(e0[2] * e0[3]), (double)e0[3]};
        double b[4] = {(double)(e1[0] * e1[3]), (double)(e1[1] * e1[3]), (double)(e1[2] * e1[3]), (double)e1[3]};
        for (int c = 0; c < 4; c++) {
            sum[c] += 0.5 * (a[c] + b[c]) * dx;
            out[4 * i + c] = (float)sum[c];
        }
    }
}

// renderer.rs:2164-2174
extern "C" void oracle_bucket_max(const vvr_control_point* pts, uint32_t n, float* out64) {
    float lut[LUT_SIZE * 4];
    oracle_tf_lut(pts, n, lut);
    int per = LUT_SIZE / 64;
    for (int b = 0; b < 64; b++) {
        int end = std::min((b + 1) * per + 1, LUT_SIZE);
        float m = 0.0f;
        for (int i = b * per; i < end; i++) m = std::max(m, lut[4 * i + 3]);  // f32::max
        out64[b] = m;
    }
}

// ---- camera.rs ----------------------------------------------------------------

static V cam_forward(const vvr_camera& c) {  // camera.rs:35-39
    float sy = std::sin(c.yaw), cy = std::cos(c.yaw);
    float sp = std::sin(c.pitch), cp = std::cos(c.pitch);
    return V(cp * sy, cp * cy, -sp);
}

static V glam_normalize(V v) {  // glam: v * (1 / length)
    float len = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    float r = 1.0f / len;
    return V(v.x * r, v.y * r, v.z * r);
}

static void cam_basis(const vvr_camera& c, V& right, V& up) {  // camera.rs:45-50
    V f = cam_forward(c);
    right = glam_normalize(cross(f, V(0, 0, 1)));
    up = cross(right, f);
}

extern "C" void oracle_camera_framing(const float center[3], float diagonal, vvr_camera* out) {  // camera.rs:24-33
    float fov_y = 30.0f * (3.14159265358979323846f / 180.0f);  // f32::to_radians
    out->target[0] = center[0];
    out->target[1] = center[1];
    out->target[2] = center[2];
    out->distance = 0.55f * diagonal / std::tan(fov_y / 2.0f);
    out->yaw = 0.0f;
    out->pitch = 0.0f;
    out->fov_y = fov_y;
}

// ---- volume.rs: index <-> world (f64) -------------------------------------------

// General 4x4 inverse (cofactor expansion), column-major, f64.
static void inverse4(const double m[16], double inv[16]) {
    double a[4][4];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) a[r][c] = m[c * 4 + r];
    double cof[4][4];
    for (int r = 0; r < 4; r++) {
        for (int c = 0; c < 4; c++) {
            double sub[3][3];
            int si = 0;
            for (int i = 0; i < 4; i++) {
                if (i == r) continue;
                int sj = 0;
                for (int j = 0; j < 4; j++) {
                    if (j == c) continue;
                    sub[si][sj++] = a[i][j];
                }
                si++;
            }
            double det3 = sub[0][0] * (sub[1][1] * sub[2][2] - sub[1][2] * sub[2][1]) -
                          sub[0][1] * (sub[1][0] * sub[2][2] - sub[1][2] * sub[2][0]) +
                          sub[0][2] * (sub[1][0] * sub[2][1] - sub[1][1] * sub[2][0]);
            cof[r][c] = ((r + c) & 1) ? -det3 : det3;
        }
    }
    double det = 0;
    for (int c = 0; c < 4; c++) det += a[0][c] * cof[0][c];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) inv[r * 4 + c] = cof[r][c] / det;  // inv[c][r] = cof[r][c]/det, column-major
}

// volume.rs:23-35: index_to_world = [direction * diag(spacing) | origin].
static void index_to_world(const vvr_volume& v, double m[16]) {
    for (int c = 0; c < 3; c++) {
        for (int r = 0; r < 3; r++) m[c * 4 + r] = v.direction[c * 3 + r] * v.spacing[c];
        m[c * 4 + 3] = 0.0;
    }
    m[12] = v.origin[0];
    m[13] = v.origin[1];
    m[14] = v.origin[2];
    m[15] = 1.0;
}

extern "C" void oracle_geometry(const vvr_volume* vol, float world_to_index[16], float index_to_world_f[16],
                                float center[3], float* diagonal, float* min_spacing, float* max_spacing) {
    double i2w[16], w2i[16];
    index_to_world(*vol, i2w);
    inverse4(i2w, w2i);
    for (int i = 0; i < 16; i++) {
        world_to_index[i] = (float)w2i[i];
        index_to_world_f[i] = (float)i2w[i];
    }
    // renderer.rs:880-883: centre = index_to_world((dims - 1) / 2)
    double p[3];
    for (int a = 0; a < 3; a++) p[a] = ((double)vol->dims[a] - 1.0) * 0.5;
    for (int r = 0; r < 3; r++)
        center[r] = (float)(i2w[0 * 4 + r] * p[0] + i2w[1 * 4 + r] * p[1] + i2w[2 * 4 + r] * p[2] + i2w[12 + r]);
    double e2 = 0;
    for (int a = 0; a < 3; a++) {
        double e = std::max((double)vol->dims[a] - 1.0, 0.0) * vol->spacing[a];
        e2 += e * e;
    }
    *diagonal = (float)std::sqrt(e2);
    *min_spacing = (float)std::min(vol->spacing[0], std::min(vol->spacing[1], vol->spacing[2]));
    *max_spacing = (float)std::max(vol->spacing[0], std::max(vol->spacing[1], vol->spacing[2]));
}

// ---- renderer.rs: uniform packing --------------------------------------------------

// renderer.rs:80-87
static float tissue_majorant(const vvr_label_params& l, float anisotropy, float limit) {
    float g = rclamp(anisotropy, 0.0f, 0.95f);
    float largest = 0.0f;
    for (int i = 0; i < 256; i++) {
        if (l.colors[i][3] > 0.5f && !l.ghost[i] && l.optics[i][3] > 0.0f) {
            float m = 0.0f;
            for (int c = 0; c < 3; c++) m = std::max(m, l.optics[i][c]);
            largest = std::max(largest, m);
        }
    }
    return std::min(largest / (1.0f - g), limit);
}

// Lighting::light_direction (renderer.rs:299-307).
static V light_direction(const vvr_lighting& l, const vvr_camera& cam) {
    V right, up;
    cam_basis(cam, right, up);
    float se = std::sin(l.elevation), ce = std::cos(l.elevation);
    float sa = std::sin(l.azimuth), ca = std::cos(l.azimuth);
    V f = cam_forward(cam);
    V a = (-f) * ce * ca;
    V b = right * ce * sa;
    V c = up * se;
    return glam_normalize(a + b + c);
}

// Renderer::record's uniform fill (renderer.rs:1516-1660) for a renderer with
// volume `vol`, `accumulated` paths so far, labels uploaded or not, and a light
// map whose brightest column is env_peak_u when has_map.
extern "C" void oracle_pack_uniforms_env(const vvr_volume* vol, const vvr_frame_params* p, const float* bucket_max64,
                                         int has_labels, uint32_t accumulated, uint32_t width, uint32_t height,
                                         uint64_t frame, int has_map, float env_peak_u, vvr_uniforms* u) {
    std::memset(u, 0, sizeof(*u));
    float w2i[16], i2w[16], center[3], diagonal, min_sp, max_sp;
    oracle_geometry(vol, w2i, i2w, center, &diagonal, &min_sp, &max_sp);
    const vvr_camera& cam = p->camera;
    V right, up;
    cam_basis(cam, right, up);
    float tan_half = std::tan(cam.fov_y / 2.0f);
    float aspect = (float)width / (float)height;
    const vvr_lighting& light = p->lighting;
    V ld = light_direction(light, cam);
    const vvr_label_params& labels = p->labels;
    // Map without a light map set falls back to the studio (renderer.rs:1524-1527).
    uint32_t environment =
        light.environment == VVR_ENV_MAP && !has_map ? (uint32_t)VVR_ENV_STUDIO : light.environment;
    // The map's brightest column turned to the key light's azimuth (renderer.rs:1531-1534).
    float env_rotation = 0.0f;
    if (has_map) {
        float key_phi = std::atan2(std::sin(light.azimuth), -std::cos(light.azimuth));
        env_rotation = (env_peak_u - 0.5f) * 6.28318530717958647692f - key_phi + light.env_rotation;
    }

    std::memcpy(u->world_to_index, w2i, sizeof(w2i));
    V f = cam_forward(cam);
    V eye = V(cam.target[0], cam.target[1], cam.target[2]) - f * cam.distance;
    u->eye[0] = eye.x, u->eye[1] = eye.y, u->eye[2] = eye.z, u->eye[3] = 1.0f;
    u->forward[0] = f.x, u->forward[1] = f.y, u->forward[2] = f.z;
    V r = right * tan_half * aspect;
    V uu = up * tan_half;
    u->right[0] = r.x, u->right[1] = r.y, u->right[2] = r.z;
    u->up[0] = uu.x, u->up[1] = uu.y, u->up[2] = uu.z;
    for (int a = 0; a < 3; a++) u->dims[a] = (float)vol->dims[a];
    u->dims[3] = p->step_mm;
    u->flags[0] = (uint32_t)frame;
    u->flags[1] = p->srgb_encode ? 1u : 0u;
    u->flags[2] = width;
    u->flags[3] = height;
    u->labels[0] = has_labels ? (float)labels.mode : 0.0f;
    u->labels[1] = labels.tint;
    u->labels[2] = labels.extinction;
    u->labels[3] = labels.ghost_opacity;
    u->light_dir[0] = ld.x, u->light_dir[1] = ld.y, u->light_dir[2] = ld.z, u->light_dir[3] = light.intensity;
    float amb = environment == VVR_ENV_STUDIO ? 0.5f * light.ambient : light.ambient;
    u->ambient[0] = u->ambient[1] = u->ambient[2] = amb;
    u->ambient[3] = std::exp2(light.exposure);
    u->material[0] = light.roughness;
    u->material[1] = light.metalness;
    u->material[2] = (float)light.tone_mapper;
    u->material[3] = light.light_size;
    u->lighting[0] = light.shadows ? 1u : 0u;
    u->lighting[1] = light.ambient_occlusion ? 1u : 0u;
    u->lighting[2] = p->preintegrate ? 1u : 0u;
    u->lighting[3] = environment;
    u->light_cells[0] = 2.0f * max_sp;  // SHADOW_DOWNSAMPLE
    u->light_cells[1] = 4.0f * max_sp;  // AO_DOWNSAMPLE

    // clip_uniforms (renderer.rs:1449-1466)
    uint32_t n = 0;
    for (int i = 0; i < 3; i++) {
        const vvr_clip_plane& cp = p->clipping.planes[i];
        if (!cp.enabled) continue;
        float sign = cp.cut_positive ? 1.0f : -1.0f;
        V normal(0, 0, 0);
        normal[(int)cp.plane] = sign;
        // a = mat3(index_to_world) (columns), b = translation
        V a0(i2w[0], i2w[1], i2w[2]), a1(i2w[4], i2w[5], i2w[6]), a2(i2w[8], i2w[9], i2w[10]);
        V b(i2w[12], i2w[13], i2w[14]);
        V at_n(dot(a0, normal), dot(a1, normal), dot(a2, normal));
        u->clip_planes[n][0] = at_n.x;
        u->clip_planes[n][1] = at_n.y;
        u->clip_planes[n][2] = at_n.z;
        u->clip_planes[n][3] = sign * cp.position - dot(normal, b);
        u->clip_normals[n][0] = normal.x;
        u->clip_normals[n][1] = normal.y;
        u->clip_normals[n][2] = normal.z;
        n++;
    }
    u->clip[0] = n;
    u->clip[1] = p->clipping.all ? 1u : 0u;

    uint32_t skip = p->skip_empty_space ? 1u : 0u;
    if (p->path_tracing_enabled) {
        u->path[0] = accumulated;
        u->path[1] = p->path_tracing.bounces;
        u->path[2] = std::max(p->path_tracing.samples_per_frame, 1u);
    }
    u->path[3] = skip;
    u->env[0] = env_rotation;
    u->env[1] = p->smooth_sampling ? 1.0f : 0.0f;
    u->env[2] = light.anisotropy;
    u->env[3] = light.chromatic;
    if (p->path_tracing_enabled) {
        float focus = cam.distance;
        if (p->path_tracing.has_focus) {
            V fp(p->path_tracing.focus[0], p->path_tracing.focus[1], p->path_tracing.focus[2]);
            focus = dot(fp - eye, f);
        }
        u->lens[0] = p->path_tracing.aperture;
        u->lens[1] = std::max(focus, 1.0f);
    }
    u->finish[0] = light.bloom;
    u->finish[1] = light.vignette;
    u->finish[2] = light.grain;
    u->finish[3] = light.auto_exposure ? 1.0f : 0.0f;
    for (int i = 0; i < 16; i++)
        for (int c = 0; c < 4; c++) u->tf_bucket_max[i][c] = bucket_max64[4 * i + c];
    u->subsurface[0] = light.subsurface * light.subsurface_strength;
    u->subsurface[1] = 10.0f / std::max(light.scattering_distance, 0.1f);
    u->subsurface[2] = 0.7f;  // SUBSURFACE_SPREAD
    u->subsurface[3] = 2.0f * max_sp;
    for (int i = 0; i < 256; i++) {
        for (int c = 0; c < 4; c++) u->label_colors[i][c] = labels.colors[i][c];
        if (labels.colors[i][3] > 0.5f && labels.ghost[i]) u->label_colors[i][3] = 2.0f;
        for (int c = 0; c < 4; c++) u->label_density[i][c] = labels.density[i][c];
        u->label_materials[i][0] = labels.materials[i][0];
        u->label_materials[i][1] = labels.materials[i][1];
        u->label_materials[i][2] = labels.materials[i][3];
        u->label_materials[i][3] = labels.materials[i][2] * light.subsurface_strength;
        u->label_surface[i][0] = labels.surface[i][0];
        u->label_surface[i][1] = labels.surface[i][1] * light.surface_detail;
        u->label_surface[i][2] = labels.surface[i][2] * light.surface_detail;
        u->label_surface[i][3] = labels.surface[i][3];
        for (int c = 0; c < 4; c++) u->label_optics[i][c] = labels.optics[i][c];
    }
    float limit = 1.0f / (2.0f * max_sp);  // TISSUE_MIN_FREE_PATH_VOXELS
    u->tissue[0] = tissue_majorant(labels, light.anisotropy, limit);
    u->tissue[1] = limit;
}

extern "C" void oracle_pack_uniforms(const vvr_volume* vol, const vvr_frame_params* p, const float* bucket_max64,
                                     int has_labels, uint32_t accumulated, uint32_t width, uint32_t height,
                                     uint64_t frame, vvr_uniforms* u) {
    oracle_pack_uniforms_env(vol, p, bucket_max64, has_labels, accumulated, width, height, frame, 0, 0.0f, u);
}

// ---- EnvironmentMap::prepare (environment.rs:56-180) and its upload (renderer.rs:2224-2290)
namespace {
const float kPiF = 3.14159265358979323846f;
float env_lum(const float* c) { return (0.2126f * c[0] + 0.7152f * c[1]) + 0.0722f * c[2]; }
void env_dir(float uu, float vv, float d[3]) {  // environment.rs:44-54
    float phi = (uu - 0.5f) * 2.0f * kPiF;
    float theta = vv * kPiF;
    d[0] = std::sin(theta) * std::sin(phi);
    d[1] = std::cos(theta);
    d[2] = std::sin(theta) * std::cos(phi);
}
}  // namespace

extern "C" float oracle_env_prepare(uint32_t width, uint32_t height, const float* rgb, uint16_t* radiance_out,
                                    float* table, uint16_t* diffuse_out) {
    const size_t w = width, h = height;
    auto sin_theta = [&](size_t y) { return std::sin(kPiF * ((float)y + 0.5f) / (float)h); };
    double sum[3] = {0, 0, 0}, weight = 0;
    for (size_t y = 0; y < h; y++) {
        double row[3] = {0, 0, 0};
        for (size_t x = 0; x < w; x++)
            for (int i = 0; i < 3; i++) row[i] += (double)rgb[3 * (y * w + x) + i];
        float st = sin_theta(y);
        for (int i = 0; i < 3; i++) sum[i] += row[i] * (double)st;
        weight += (double)st * (double)w;
    }
    float mean_c[3] = {(float)(sum[0] / weight), (float)(sum[1] / weight), (float)(sum[2] / weight)};
    float mean = std::max(env_lum(mean_c), 1e-6f);
    std::vector<float> radiance(4 * w * h);
    for (size_t i = 0; i < w * h; i++) {
        for (int k = 0; k < 3; k++) radiance[4 * i + k] = rgb[3 * i + k] / mean;
        radiance[4 * i + 3] = 1.0f;
    }
    std::vector<float> f(w * h);
    for (size_t i = 0; i < w * h; i++) f[i] = env_lum(rgb + 3 * i) * sin_theta(i / w);
    std::vector<double> row_sums(h, 0.0);
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) row_sums[y] += (double)f[y * w + x];
    double total = 0;
    for (double r : row_sums) total += r;
    total = std::max(total, 1e-12);
    double marginal = 0;
    for (size_t y = 0; y < h; y++) {
        float* row = table + 2 * y * (w + 1);
        double acc = 0;
        for (size_t x = 0; x < w; x++) {
            acc += (double)f[y * w + x];
            row[2 * x] = row_sums[y] > 0.0 ? (float)(acc / row_sums[y]) : (float)(x + 1) / (float)w;
            row[2 * x + 1] = (float)((double)f[y * w + x] / total);
        }
        row[2 * (w - 1)] = 1.0f;
        marginal += row_sums[y];
        row[2 * w] = (float)(marginal / total);
        row[2 * w + 1] = (float)(row_sums[y] / total);
    }
    table[2 * ((h - 1) * (w + 1) + w)] = 1.0f;
    double px = 0, pz = 0;
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            double l = (double)f[y * w + x];
            float d[3];
            env_dir(((float)x + 0.5f) / (float)w, ((float)y + 0.5f) / (float)h, d);
            px += l * l * (double)d[0];
            pz += l * l * (double)d[2];
        }
    float peak_u = 0.5f + (float)(std::atan2(px, pz) / (2.0 * 3.14159265358979323846));
    // Diffuse convolution (environment.rs:135-178).
    const size_t sw = std::min<size_t>(128, w), sh = std::max<size_t>(sw / 2, 1);
    std::vector<float> small_l(3 * sw * sh), small_d(3 * sw * sh);
    for (size_t i = 0; i < sw * sh; i++) {
        size_t sx = i % sw, sy = i / sw;
        size_t x0 = sx * w / sw, x1 = std::max((sx + 1) * w / sw, x0 + 1);
        size_t y0 = sy * h / sh, y1 = std::max((sy + 1) * h / sh, y0 + 1);
        float c[3] = {0, 0, 0};
        for (size_t y = y0; y < y1; y++)
            for (size_t x = x0; x < x1; x++)
                for (int k = 0; k < 3; k++) c[k] = c[k] + radiance[4 * (y * w + x) + k];
        float n = (float)((y1 - y0) * (x1 - x0));
        float uu = ((float)sx + 0.5f) / (float)sw, vv = ((float)sy + 0.5f) / (float)sh;
        float solid = (2.0f * kPiF / (float)sw) * (kPiF / (float)sh) * std::sin(kPiF * vv);
        for (int k = 0; k < 3; k++) small_l[3 * i + k] = c[k] / n * solid;
        env_dir(uu, vv, &small_d[3 * i]);
    }
    const size_t dw = 32, dh = 16;
    for (size_t i = 0; i < dw * dh; i++) {
        float nrm[3];
        env_dir(((float)(i % dw) + 0.5f) / (float)dw, ((float)(i / dw) + 0.5f) / (float)dh, nrm);
        float c[3] = {0, 0, 0};
        for (size_t j = 0; j < sw * sh; j++) {
            const float* d = &small_d[3 * j];
            float cs = (nrm[0] * d[0] + nrm[1] * d[1]) + nrm[2] * d[2];
            if (cs > 0.0f)
                for (int k = 0; k < 3; k++) c[k] = c[k] + small_l[3 * j + k] * cs;
        }
        for (int k = 0; k < 3; k++) diffuse_out[4 * i + k] = f32_to_f16(std::min(c[k] / kPiF, 60000.0f));
        diffuse_out[4 * i + 3] = f32_to_f16(1.0f);
    }
    for (size_t i = 0; i < 4 * w * h; i++) radiance_out[i] = f32_to_f16(std::min(radiance[i], 60000.0f));
    return peak_u;
}
