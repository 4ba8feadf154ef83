// Per-frame host logic of the renderer: parameter defaults, the orbit camera,
// volume geometry, the opacity cache key and the uniform block fill.
// Follows crates/render/src/renderer.rs:179-307, 1449-1660,
// crates/render/src/camera.rs and crates/volio/src/volume.rs.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "host.hpp"

namespace vvr {

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr uint32_t kShadowDownsample = 2;         // renderer.rs:35
constexpr uint32_t kAoDownsample = 4;             // renderer.rs:39
constexpr float kSubsurfaceSpread = 0.7f;         // renderer.rs:37
constexpr float kTissueMinFreePathVoxels = 2.0f;  // renderer.rs:92

struct F3 {
    float x, y, z;
};
F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 scale(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
F3 cross(F3 a, F3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
// glam's Vec3::normalize: self * (1 / length).
F3 normalize(F3 a) {
    const float r = 1.0f / std::sqrt(dot(a, a));
    return scale(a, r);
}

// OrbitCamera::forward / basis (camera.rs:35-50).
F3 forward(const vvr_camera& c) {
    const float sy = std::sin(c.yaw), cy = std::cos(c.yaw);
    const float sp = std::sin(c.pitch), cp = std::cos(c.pitch);
    return {cp * sy, cp * cy, -sp};
}
void basis(const vvr_camera& c, F3& right, F3& up) {
    const F3 f = forward(c);
    right = normalize(cross(f, F3{0.0f, 0.0f, 1.0f}));
    up = cross(right, f);
}

// Lighting::light_direction (renderer.rs:299-307).
F3 light_direction(const vvr_lighting& l, const vvr_camera& cam) {
    F3 right, up;
    basis(cam, right, up);
    const float se = std::sin(l.elevation), ce = std::cos(l.elevation);
    const float sa = std::sin(l.azimuth), ca = std::cos(l.azimuth);
    const F3 f = forward(cam);
    const F3 a = scale(scale(F3{-f.x, -f.y, -f.z}, ce), ca);
    const F3 b = scale(scale(right, ce), sa);
    const F3 c = scale(up, se);
    return normalize(add(add(a, b), c));
}

// tissue_majorant (renderer.rs:80-87).
float tissue_majorant(const vvr_label_params& l, float anisotropy, float limit) {
    const float g = std::min(std::max(anisotropy, 0.0f), 0.95f);
    float largest = 0.0f;
    for (int i = 0; i < 256; i++) {
        if (!(l.colors[i][3] > 0.5f) || l.ghost[i] || !(l.optics[i][3] > 0.0f)) continue;
        float m = 0.0f;
        for (int c = 0; c < 3; c++) m = std::max(m, l.optics[i][c]);
        largest = std::max(largest, m);
    }
    return std::min(largest / (1.0f - g), limit);
}

// 4x4 inverse of a column-major f64 matrix by cofactors.
void invert4(const double m[16], double out[16]) {
    auto at = [&](int r, int c) { return m[c * 4 + r]; };
    double cof[16];
    for (int r = 0; r < 4; r++) {
        for (int c = 0; c < 4; c++) {
            double s[9];
            int k = 0;
            for (int i = 0; i < 4; i++) {
                if (i == r) continue;
                for (int j = 0; j < 4; j++) {
                    if (j == c) continue;
                    s[k++] = at(i, j);
                }
            }
            const double d = s[0] * (s[4] * s[8] - s[5] * s[7]) - s[1] * (s[3] * s[8] - s[5] * s[6]) +
                             s[2] * (s[3] * s[7] - s[4] * s[6]);
            cof[r * 4 + c] = ((r + c) & 1) ? -d : d;
        }
    }
    double det = 0.0;
    for (int c = 0; c < 4; c++) det += at(0, c) * cof[c];
    if (det == 0.0) throw Error{VVR_EINVAL, "volume geometry is singular (zero spacing or degenerate direction)"};
    // inverse = adjugate / det; adjugate = transpose(cofactors); column-major store.
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out[r * 4 + c] = cof[r * 4 + c] / det;
}

}  // namespace

bool OpacityKey::operator==(const OpacityKey& o) const {
    return generation == o.generation && mode == o.mode && extinction_bits == o.extinction_bits &&
           std::memcmp(visible, o.visible, sizeof(visible)) == 0 && std::memcmp(ghost, o.ghost, sizeof(ghost)) == 0 &&
           ghost_opacity_bits == o.ghost_opacity_bits && margin == o.margin && smoothing == o.smoothing;
}

// Renderer::opacity_key (renderer.rs:1468-1495).
OpacityKey opacity_key(const vvr_label_params& l, uint64_t generation, bool has_labels, float min_spacing) {
    OpacityKey k{};
    for (int i = 0; i < 256; i++) {
        if (l.colors[i][3] > 0.5f) k.visible[i / 64] |= uint64_t(1) << (i % 64);
        if (l.ghost[i]) k.ghost[i / 64] |= uint64_t(1) << (i % 64);
    }
    k.generation = generation;
    k.mode = has_labels ? l.mode : uint32_t(VVR_LABELS_OFF);
    std::memcpy(&k.extinction_bits, &l.extinction, 4);
    std::memcpy(&k.ghost_opacity_bits, &l.ghost_opacity, 4);
    k.margin = int32_t(std::min(std::max(std::round(l.cut_margin_mm / min_spacing), 0.0f), 3.0f));
    k.smoothing = std::min(l.smoothing, 8u);
    return k;
}

Geometry make_geometry(const vvr_volume& v) {
    Geometry g{};
    for (int a = 0; a < 3; a++) g.dims[a] = v.dims[a];
    // Volume::index_to_world (volume.rs:23-31): [direction * diag(spacing) | origin].
    double i2w[16] = {0};
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) i2w[c * 4 + r] = v.direction[c * 3 + r] * v.spacing[c];
    i2w[12] = v.origin[0];
    i2w[13] = v.origin[1];
    i2w[14] = v.origin[2];
    i2w[15] = 1.0;
    double w2i[16];
    invert4(i2w, w2i);
    for (int i = 0; i < 16; i++) {
        g.world_to_index[i] = float(w2i[i]);
        g.index_to_world[i] = float(i2w[i]);
    }
    auto xform = [&](const double p[3], double out[3]) {
        for (int r = 0; r < 3; r++) out[r] = i2w[r] * p[0] + i2w[4 + r] * p[1] + i2w[8 + r] * p[2] + i2w[12 + r];
    };
    // renderer.rs:876-895: centre of the voxel centres; bounds of the voxel faces.
    double c[3], p[3];
    for (int a = 0; a < 3; a++) p[a] = (double(v.dims[a]) - 1.0) * 0.5;
    xform(p, c);
    for (int a = 0; a < 3; a++) {
        g.center[a] = float(c[a]);
        g.bounds_lo[a] = 3.4028235e38f;
        g.bounds_hi[a] = -3.4028235e38f;
    }
    for (int i = 0; i < 8; i++) {
        double q[3] = {(i & 1) ? double(v.dims[0]) - 0.5 : -0.5, (i & 2) ? double(v.dims[1]) - 0.5 : -0.5,
                       (i & 4) ? double(v.dims[2]) - 0.5 : -0.5};
        double w[3];
        xform(q, w);
        for (int a = 0; a < 3; a++) {
            g.bounds_lo[a] = std::min(g.bounds_lo[a], float(w[a]));
            g.bounds_hi[a] = std::max(g.bounds_hi[a], float(w[a]));
        }
    }
    double e2 = 0.0;
    for (int a = 0; a < 3; a++) {
        const double e = std::max(double(v.dims[a]) - 1.0, 0.0) * v.spacing[a];
        e2 += e * e;
    }
    g.diagonal = float(std::sqrt(e2));
    g.min_spacing = float(std::min(v.spacing[0], std::min(v.spacing[1], v.spacing[2])));
    g.max_spacing = float(std::max(v.spacing[0], std::max(v.spacing[1], v.spacing[2])));
    return g;
}

void pack_uniforms(const Geometry& g, const vvr_frame_params& p, const float bucket_max[64], bool has_labels,
                   uint32_t accumulated, uint32_t width, uint32_t height, uint64_t frame, vvr_uniforms* u,
                   const float* env_peak_u) {
    std::memset(u, 0, sizeof(*u));
    const vvr_camera& cam = p.camera;
    F3 right, up;
    basis(cam, right, up);
    const float tan_half = std::tan(cam.fov_y / 2.0f);
    const float aspect = float(width) / float(height);
    const vvr_lighting& light = p.lighting;
    const F3 ld = light_direction(light, cam);
    const vvr_label_params& labels = p.labels;
    // Map without a light map set falls back to the studio (renderer.rs:1524-1527).
    const uint32_t environment =
        light.environment == VVR_ENV_MAP && !env_peak_u ? uint32_t(VVR_ENV_STUDIO) : light.environment;

    std::memcpy(u->world_to_index, g.world_to_index, sizeof(u->world_to_index));
    const F3 f = forward(cam);
    const F3 eye = sub(F3{cam.target[0], cam.target[1], cam.target[2]}, scale(f, cam.distance));
    const F3 r = scale(scale(right, tan_half), aspect);
    const F3 uu = scale(up, tan_half);
    const float v4[5][4] = {{eye.x, eye.y, eye.z, 1.0f},
                            {f.x, f.y, f.z, 0.0f},
                            {r.x, r.y, r.z, 0.0f},
                            {uu.x, uu.y, uu.z, 0.0f},
                            {float(g.dims[0]), float(g.dims[1]), float(g.dims[2]), p.step_mm}};
    std::memcpy(u->eye, v4[0], 16);
    std::memcpy(u->forward, v4[1], 16);
    std::memcpy(u->right, v4[2], 16);
    std::memcpy(u->up, v4[3], 16);
    std::memcpy(u->dims, v4[4], 16);
    u->flags[0] = uint32_t(frame);
    u->flags[1] = p.srgb_encode ? 1u : 0u;
    u->flags[2] = width;
    u->flags[3] = height;
    u->labels[0] = has_labels ? float(labels.mode) : 0.0f;
    u->labels[1] = labels.tint;
    u->labels[2] = labels.extinction;
    u->labels[3] = labels.ghost_opacity;
    u->light_dir[0] = ld.x;
    u->light_dir[1] = ld.y;
    u->light_dir[2] = ld.z;
    u->light_dir[3] = light.intensity;
    // The studio's mean is half the dome's (renderer.rs:1574-1581).
    const float amb = environment == VVR_ENV_STUDIO ? 0.5f * light.ambient : light.ambient;
    u->ambient[0] = u->ambient[1] = u->ambient[2] = amb;
    u->ambient[3] = std::exp2(light.exposure);
    u->material[0] = light.roughness;
    u->material[1] = light.metalness;
    u->material[2] = float(light.tone_mapper);
    u->material[3] = light.light_size;
    u->lighting[0] = light.shadows ? 1u : 0u;
    u->lighting[1] = light.ambient_occlusion ? 1u : 0u;
    u->lighting[2] = p.preintegrate ? 1u : 0u;
    u->lighting[3] = environment;
    u->light_cells[0] = float(kShadowDownsample) * g.max_spacing;
    u->light_cells[1] = float(kAoDownsample) * g.max_spacing;

    // clip_uniforms (renderer.rs:1449-1466): planes in index space, distances in mm.
    uint32_t n = 0;
    for (int i = 0; i < 3; i++) {
        const vvr_clip_plane& cp = p.clipping.planes[i];
        if (!cp.enabled) continue;
        const float sign = cp.cut_positive ? 1.0f : -1.0f;
        F3 normal{0.0f, 0.0f, 0.0f};
        (&normal.x)[std::min<uint32_t>(cp.plane, 2u)] = sign;
        const float* m = g.index_to_world;
        const F3 a0{m[0], m[1], m[2]}, a1{m[4], m[5], m[6]}, a2{m[8], m[9], m[10]};
        const F3 b{m[12], m[13], m[14]};
        u->clip_planes[n][0] = dot(a0, normal);
        u->clip_planes[n][1] = dot(a1, normal);
        u->clip_planes[n][2] = dot(a2, normal);
        u->clip_planes[n][3] = sign * cp.position - dot(normal, b);
        u->clip_normals[n][0] = normal.x;
        u->clip_normals[n][1] = normal.y;
        u->clip_normals[n][2] = normal.z;
        n++;
    }
    u->clip[0] = n;
    u->clip[1] = p.clipping.all ? 1u : 0u;

    if (p.path_tracing_enabled) {
        u->path[0] = accumulated;
        u->path[1] = p.path_tracing.bounces;
        u->path[2] = std::max(p.path_tracing.samples_per_frame, 1u);
    }
    u->path[3] = p.skip_empty_space ? 1u : 0u;
    // The light map's brightest column turned to the key light's azimuth (renderer.rs:1531-1534).
    u->env[0] = 0.0f;
    if (env_peak_u) {
        const float key_phi = std::atan2(std::sin(light.azimuth), -std::cos(light.azimuth));
        u->env[0] = (*env_peak_u - 0.5f) * 6.28318530717958647692f - key_phi + light.env_rotation;
    }
    u->env[1] = p.smooth_sampling ? 1.0f : 0.0f;
    u->env[2] = light.anisotropy;
    u->env[3] = light.chromatic;
    if (p.path_tracing_enabled) {
        float focus = cam.distance;
        if (p.path_tracing.has_focus) {
            const F3 fp{p.path_tracing.focus[0], p.path_tracing.focus[1], p.path_tracing.focus[2]};
            focus = dot(sub(fp, eye), f);
        }
        u->lens[0] = p.path_tracing.aperture;
        u->lens[1] = std::max(focus, 1.0f);
    }
    u->finish[0] = light.bloom;
    u->finish[1] = light.vignette;
    u->finish[2] = light.grain;
    u->finish[3] = light.auto_exposure ? 1.0f : 0.0f;
    for (int i = 0; i < 64; i++) u->tf_bucket_max[i / 4][i % 4] = bucket_max[i];
    u->subsurface[0] = light.subsurface * light.subsurface_strength;
    u->subsurface[1] = 10.0f / std::max(light.scattering_distance, 0.1f);
    u->subsurface[2] = kSubsurfaceSpread;
    u->subsurface[3] = float(kShadowDownsample) * g.max_spacing;
    for (int i = 0; i < 256; i++) {
        std::memcpy(u->label_colors[i], labels.colors[i], 16);
        if (labels.colors[i][3] > 0.5f && labels.ghost[i]) u->label_colors[i][3] = 2.0f;
        std::memcpy(u->label_density[i], labels.density[i], 16);
        u->label_materials[i][0] = labels.materials[i][0];
        u->label_materials[i][1] = labels.materials[i][1];
        u->label_materials[i][2] = labels.materials[i][3];
        u->label_materials[i][3] = labels.materials[i][2] * light.subsurface_strength;
        u->label_surface[i][0] = labels.surface[i][0];
        u->label_surface[i][1] = labels.surface[i][1] * light.surface_detail;
        u->label_surface[i][2] = labels.surface[i][2] * light.surface_detail;
        u->label_surface[i][3] = labels.surface[i][3];
        std::memcpy(u->label_optics[i], labels.optics[i], 16);
    }
    const float limit = 1.0f / (kTissueMinFreePathVoxels * g.max_spacing);
    u->tissue[0] = tissue_majorant(labels, light.anisotropy, limit);
    u->tissue[1] = limit;
}

}  // namespace vvr

using namespace vvr;

extern "C" void vvr_label_params_default(vvr_label_params* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->mode = VVR_LABELS_OFF;
    o->tint = 0.7f;
    o->extinction = 0.5f;
    for (int i = 0; i < 256; i++) {
        for (int c = 0; c < 4; c++) o->colors[i][c] = 1.0f;
        o->materials[i][0] = 0.5f;
        o->materials[i][1] = 0.0f;
        o->materials[i][2] = 0.3f;
        o->materials[i][3] = 0.0f;
        o->density[i][0] = 0.0f;
        o->density[i][1] = 1e-3f;
        o->density[i][2] = 0.5f;
        o->density[i][3] = 1.0f;
    }
    o->ghost_opacity = 0.08f;
    o->cut_margin_mm = 2.0f;
    o->smoothing = 1;
}

extern "C" void vvr_lighting_default(vvr_lighting* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    const float deg = kPi / 180.0f;
    o->azimuth = -60.0f * deg;
    o->elevation = 40.0f * deg;
    o->intensity = 3.0f;
    o->ambient = 0.3f;
    o->environment = VVR_ENV_STUDIO;
    o->light_size = 12.0f * deg;
    o->env_rotation = 0.0f;
    o->anisotropy = 0.0f;
    o->chromatic = 1.0f;
    o->shadows = 1;
    o->ambient_occlusion = 1;
    o->exposure = 0.0f;
    o->auto_exposure = 1;
    o->bloom = 0.04f;
    o->vignette = 0.25f;
    o->grain = 0.012f;
    o->tone_mapper = VVR_TONE_ACES;
    o->roughness = 0.5f;
    o->metalness = 0.0f;
    o->subsurface = 0.3f;
    o->scattering_distance = 4.0f;
    o->subsurface_strength = 1.0f;
    o->surface_detail = 1.0f;
}

extern "C" void vvr_path_tracing_default(vvr_path_tracing* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->bounces = 4;
    o->samples_per_frame = 1;
    o->denoiser = VVR_DENOISE_ATROUS;
    o->aperture = 0.0f;
    o->has_focus = 0;
}

extern "C" void vvr_camera_framing(const float center[3], float diagonal, vvr_camera* out) {
    if (!out || !center) return;
    const float fov_y = 30.0f * (kPi / 180.0f);
    out->target[0] = center[0];
    out->target[1] = center[1];
    out->target[2] = center[2];
    out->distance = 0.55f * diagonal / std::tan(fov_y / 2.0f);
    out->yaw = 0.0f;
    out->pitch = 0.0f;
    out->fov_y = fov_y;
}

extern "C" void vvr_frame_params_default(vvr_frame_params* o, const vvr_camera* camera, float min_spacing) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    if (camera) o->camera = *camera;
    o->step_mm = 0.5f * min_spacing;
    o->srgb_encode = 1;
    vvr_label_params_default(&o->labels);
    vvr_lighting_default(&o->lighting);
    // The CLI's defaults differ from Lighting::default in the phase anisotropy (main.rs:175-176).
    o->lighting.anisotropy = 0.6f;
    o->preintegrate = 0;
    o->skip_empty_space = 1;
    for (int i = 0; i < 3; i++) o->clipping.planes[i].plane = uint32_t(i);
    o->path_tracing_enabled = 0;
    vvr_path_tracing_default(&o->path_tracing);
    o->smooth_sampling = 1;
}

extern "C" int vvr_frame_uniforms(const vvr_volume* volume, const vvr_tf* tf, const vvr_frame_params* params,
                                  int has_labels, uint32_t accumulated, uint32_t width, uint32_t height, uint64_t frame,
                                  void* out, size_t bytes) {
    if (!volume || !params || !out || bytes != sizeof(vvr_uniforms) || !width || !height) return VVR_EINVAL;
    if (!tf || !tf->n_points || !tf->points) return VVR_EINVAL;
    try {
        const Geometry g = make_geometry(*volume);
        const TfTables t = build_tf_tables(*tf);
        pack_uniforms(g, *params, t.bucket_max, has_labels != 0, accumulated, width, height, frame,
                      static_cast<vvr_uniforms*>(out));
    } catch (const Error& e) {
        return e.code;
    }
    return VVR_OK;
}

extern "C" size_t vvr_abi_sizeof(int which) {
    switch (which) {
        case VVR_ABI_VOLUME: return sizeof(vvr_volume);
        case VVR_ABI_CONTROL_POINT: return sizeof(vvr_control_point);
        case VVR_ABI_TF: return sizeof(vvr_tf);
        case VVR_ABI_SOFT_GRID: return sizeof(vvr_soft_grid);
        case VVR_ABI_CAMERA: return sizeof(vvr_camera);
        case VVR_ABI_LABEL_PARAMS: return sizeof(vvr_label_params);
        case VVR_ABI_LIGHTING: return sizeof(vvr_lighting);
        case VVR_ABI_CLIPPING: return sizeof(vvr_clipping);
        case VVR_ABI_PATH_TRACING: return sizeof(vvr_path_tracing);
        case VVR_ABI_FRAME_PARAMS: return sizeof(vvr_frame_params);
        case VVR_ABI_UNIFORMS: return sizeof(vvr_uniforms);
        case VVR_ABI_SERVER_CONFIG: return sizeof(vvr_server_config);
        case VVR_ABI_REQUEST: return sizeof(vvr_request);
        case VVR_ABI_COMPLETION: return sizeof(vvr_completion);
        case VVR_ABI_SERVER_STATS: return sizeof(vvr_server_stats);
        case VVR_ABI_PT_SHAPE: return sizeof(vvr_pt_shape);
        case VVR_ABI_ENV_MAP: return sizeof(vvr_env_map);
        case VVR_ABI_UNET_CONFIG: return sizeof(vvr_unet_config);
        default: return 0;
    }
}
