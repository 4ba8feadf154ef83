// The emission-absorption ray marcher (dvr.comp), for one frame and for a
// batch of frames. One 64-lane wavefront renders an 8 x 8 pixel tile, so the
// rays of a wave stay coherent and share cache lines of the bricked volumes.
// The path tracer is the wavefront engine of kernels_wavefront.hip.
#include <algorithm>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "device/lighting.hpp"
#include "device/pt.hpp"
#include "launch.hpp"

namespace vvr {

namespace {

// dvr.comp:15-21
__device__ __forceinline__ Sample labelled(const DevScene& s, Sample smp, V3 idx) {
    smp = apply_labels(s, smp, idx);
    if (measured(s.u, smp)) smp.extinction = dot(tissue_sigma(s.u, smp), V3(1.0f / 3.0f));
    return smp;
}

__device__ __forceinline__ void store_half4(uint16_t* out, size_t i, float a, float b, float c, float d) {
    const __half h[4] = {__float2half_rn(a), __float2half_rn(b), __float2half_rn(c), __float2half_rn(d)};
    uint2 w;
    w.x = uint32_t(__half_as_ushort(h[0])) | (uint32_t(__half_as_ushort(h[1])) << 16);
    w.y = uint32_t(__half_as_ushort(h[2])) | (uint32_t(__half_as_ushort(h[3])) << 16);
    reinterpret_cast<uint2*>(out)[i] = w;
}

// Pixel of lane `threadIdx.x` in an 8 x 8 tile.
__device__ __forceinline__ void tile_pixel(uint32_t& px, uint32_t& py) {
    px = blockIdx.x * 8u + (threadIdx.x & 7u);
    py = blockIdx.y * 8u + (threadIdx.x >> 3);
}

// dvr.comp:23-147. The lighting volumes (s.shadow, s.ao) are built before the
// frame when its uniforms turn them on.
__device__ __forceinline__ void dvr_pixel(const DevScene& sc, uint32_t px, uint32_t py, uint16_t* __restrict__ hdr,
                                          uint32_t width) {
    const vvr_uniforms* u = sc.u;
    const V3 dir = camera_ray(u, float(px) + 0.5f, float(py) + 0.5f);
    const V3 o_idx = world_to_index_point(u, V3(u->eye[0], u->eye[1], u->eye[2]));
    const M3 ti = to_index(u);
    const V3 d_idx = ti.mul(dir);
    float t0, t1;
    intersect_volume(u, o_idx, d_idx, t0, t1);
    const float step_mm = u->dims[3];
    const bool preintegrate = u->lighting[2] != 0u;
    const V3 l(u->light_dir[0], u->light_dir[1], u->light_dir[2]);
    const V3 v = -dir;
    const V3 light(u->light_dir[3]);
    const V3 ambient_light(u->ambient[0], u->ambient[1], u->ambient[2]);
    V3 radiance(0.0f);
    float transmittance = 1.0f;
    bool refined = false;
    if (t0 < t1) {
        float tt = t0 + ign(float(px), float(py)) * step_mm;
        float prev_density = density_view(sc, o_idx + d_idx * (tt - step_mm));
        for (; tt < t1 && transmittance > 0.01f; tt += step_mm) {
            V3 idx = o_idx + d_idx * tt;
            // Empty-space skipping (dvr.comp:118-125).
            if (brick_majorant(sc, idx) <= 0.0f) {
                tt += brick_exit(idx, d_idx);
                prev_density = -1.0f;
                continue;
            }
            int cap_plane;
            const float cut = clip_depth(u, idx, cap_plane);
            if (cut > 0.0f) {
                prev_density = -1.0f;
                continue;
            }
            const float density = density_view(sc, idx);
            const float front_density = prev_density;
            const bool segment = preintegrate && front_density >= 0.0f;
            Sample s = segment ? tf_segment(sc, front_density, density) : tf_point(sc, density);
            prev_density = density;
            s = labelled(sc, s, idx);
            if (s.extinction <= 0.0f) continue;
            float alpha = 1.0f - vexp(-s.extinction * step_mm);
            if (segment && alpha > 0.2f && transmittance > 0.5f && s.label == 0u &&
                fabsf(density - front_density) > 1e-4f) {
                idx = idx - d_idx * (step_mm * (1.0f - preint_half_depth(sc, front_density, density)));
            } else if (!refined && alpha > 0.2f && transmittance > 0.9f) {
                // First-surface bisection (dvr.comp:150-168).
                refined = true;
                float inside = 0.0f, outside = -step_mm;
                const float target = 0.5f * s.extinction;
                for (int i = 0; i < 5; i++) {
                    const float mid = 0.5f * (inside + outside);
                    const V3 at = idx + d_idx * mid;
                    const bool opaque = labelled(sc, tf_point(sc, density_view(sc, at)), at).extinction >= target;
                    inside = opaque ? mid : inside;
                    outside = opaque ? outside : mid;
                }
                idx = idx + d_idx * inside;
            }
            V3 n;
            float surfaceness;
            const float extinction_scale = surface_frame(sc, idx, s, v, cap_plane, cut, 1.5f * step_mm, n, surfaceness);
            apply_surface_texture(u, idx, s);
            if (s.ghost) alpha = 1.0f - vexp(-s.extinction * extinction_scale * step_mm);
            // dvr.comp:107-119: the coarse lighting volumes are read just outside the
            // surface, the diffused light a little below it.
            float shadow = 1.0f, ao = 1.0f;
            V3 diffused(1.0f);
            if (u->lighting[0] != 0u) {
                shadow = shadow_lookup(sc, index_to_tex(u, idx + ti.mul(n * (1.5f * u->light_cells[0] * surfaceness)))).w;
                const float4 d = shadow_lookup(sc, index_to_tex(u, idx - ti.mul(n * (u->subsurface[3] * surfaceness))));
                diffused = V3(d.x, d.y, d.z);
            }
            if (u->lighting[1] != 0u)
                ao = ao_lookup(sc, index_to_tex(u, idx + ti.mul(n * (u->light_cells[1] * surfaceness))));
            n = bump_normal(u, idx, s, n);
            n = dot(n, v) < 0.0f ? -n : n;
            const float nl = gmax(dot(n, l), 0.0f);
            V3 diffuse;
            const V3 specular = brdf_coated(n, v, l, s, diffuse);
            const V3 surface = (specular + diffuse) * (1.0f - s.subsurface) * nl * shadow + diffuse * s.subsurface * diffused;
            const V3 volume = s.albedo / (4.0f * kPi) * vmix(V3(shadow), diffused, s.subsurface);
            const V3 direct = vmix(volume, surface, surfaceness) * light;
            // dvr.comp:135-139: a light map's diffuse response where surfaces face it.
            const V3 ambient_env = u->lighting[3] == 2u
                                       ? vmix(ambient_light, env_map_diffuse(sc, n), surfaceness * (1.0f - s.subsurface))
                                       : ambient_light;
            const V3 ambient = s.albedo * ambient_env * ao;
            radiance = radiance + transmittance * alpha * (direct + ambient);
            transmittance *= 1.0f - alpha;
        }
    }
    store_half4(hdr, size_t(py) * width + px, radiance.x, radiance.y, radiance.z, transmittance);
}

__global__ void __launch_bounds__(64) k_dvr(DevScene sc, uint16_t* __restrict__ hdr, uint32_t width, uint32_t height) {
    uint32_t px, py;
    tile_pixel(px, py);
    if (px >= width || py >= height) return;
    dvr_pixel(sc, px, py, hdr, width);
}

// Several ray-marched frames in one launch: one workgroup per 8 x 8 tile of any
// job, jobs sorted by their first tile.
__global__ void __launch_bounds__(64) k_dvr_batched(const DvrJob* __restrict__ jobs, uint32_t n_jobs, uint32_t tile0) {
    const uint32_t t = tile0 + blockIdx.x;
    uint32_t lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_base <= t) lo = mid;
        else hi = mid - 1;
    }
    const DvrJob& job = jobs[lo];
    const uint32_t tile = t - job.tile_base;
    const uint32_t px = (tile % job.tiles_x) * 8u + (threadIdx.x & 7u);
    const uint32_t py = (tile / job.tiles_x) * 8u + (threadIdx.x >> 3);
    if (px >= job.width || py >= job.height) return;
    dvr_pixel(job.scene, px, py, job.hdr, job.width);
}

}  // namespace

void launch_dvr(const DevScene& s, uint16_t* hdr, uint32_t width, uint32_t height, hipStream_t st) {
    const dim3 grid((width + 7) / 8, (height + 7) / 8);
    hipLaunchKernelGGL(k_dvr, grid, dim3(64), 0, st, s, hdr, width, height);
}

void launch_dvr_batched(const DvrJob* jobs, uint32_t n_jobs, uint32_t total_tiles, hipStream_t st) {
    if (n_jobs == 0 || total_tiles == 0) return;
    // Launches of at most 2^24 tiles keep every grid far below 2^32 threads.
    constexpr uint32_t kMaxTiles = 1u << 24;
    for (uint32_t t0 = 0; t0 < total_tiles; t0 += kMaxTiles)
        hipLaunchKernelGGL(k_dvr_batched, dim3(std::min(kMaxTiles, total_tiles - t0)), dim3(64), 0, st, jobs, n_jobs,
                           t0);
}

}  // namespace vvr
