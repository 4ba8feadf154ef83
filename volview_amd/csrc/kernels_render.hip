// The per-frame kernels: the emission-absorption ray marcher (dvr.comp) and
// the path tracer (pt.comp). One 64-lane wavefront renders an 8 x 8 pixel tile,
// so the rays of a wave stay coherent and share cache lines of the bricked
// volumes.
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
            float alpha = 1.0f - expf(-s.extinction * step_mm);
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
            if (s.ghost) alpha = 1.0f - expf(-s.extinction * extinction_scale * step_mm);
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
__global__ void __launch_bounds__(64) k_dvr_batched(const DvrJob* __restrict__ jobs, uint32_t n_jobs) {
    uint32_t lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_base <= blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const DvrJob& job = jobs[lo];
    const uint32_t tile = blockIdx.x - job.tile_base;
    const uint32_t px = (tile % job.tiles_x) * 8u + (threadIdx.x & 7u);
    const uint32_t py = (tile / job.tiles_x) * 8u + (threadIdx.x >> 3);
    if (px >= job.width || py >= job.height) return;
    dvr_pixel(job.scene, px, py, job.hdr, job.width);
}

// Adds a pixel's path sums to the accumulation and guide images (restarting
// them for the first frame) and writes the running mean (pt.comp:554-560).
__device__ __forceinline__ void store_pixel(const PixelSums& ps, size_t i, float4* __restrict__ accum,
                                            float4* __restrict__ albedo, float4* __restrict__ normal,
                                            uint16_t* __restrict__ hdr, uint32_t first_path, uint32_t samples) {
    float4 a = make_float4(ps.radiance.x, ps.radiance.y, ps.radiance.z, ps.background);
    float4 b = make_float4(ps.albedo.x, ps.albedo.y, ps.albedo.z, ps.lum2);
    float4 c = make_float4(ps.normal.x, ps.normal.y, ps.normal.z, ps.depth);
    if (first_path != 0u) {
        const float4 pa = accum[i], pb = albedo[i], pc = normal[i];
        a = make_float4(pa.x + a.x, pa.y + a.y, pa.z + a.z, pa.w + a.w);
        b = make_float4(pb.x + b.x, pb.y + b.y, pb.z + b.z, pb.w + b.w);
        c = make_float4(pc.x + c.x, pc.y + c.y, pc.z + c.z, pc.w + c.w);
    }
    accum[i] = a;
    albedo[i] = b;
    normal[i] = c;
    const float n = float(first_path + samples);
    store_half4(hdr, i, a.x / n, a.y / n, a.z / n, a.w / n);
}

// pt.comp:373-561, one lane per pixel: `samples` paths at path indices
// first_path + k, each from its own PCG stream (the reference's seed of a frame
// of one path), summed in path order.
__global__ void __launch_bounds__(64) k_pt(DevScene sc, float4* __restrict__ accum, float4* __restrict__ albedo,
                                           float4* __restrict__ normal, uint16_t* __restrict__ hdr, uint32_t width,
                                           uint32_t height, uint32_t first_path, uint32_t samples) {
    uint32_t px, py;
    tile_pixel(px, py);
    if (px >= width || py >= height) return;
    const FrameConsts fc = frame_consts(sc.u);
    PixelSums ps;
    ps.clear();
    PathState st;
    for (uint32_t k = 0; k < samples; k++) {
        path_begin(sc, fc, px, py, first_path + k, st);
        while (path_bounce(sc, fc, st)) {
        }
        ps.add(st);
    }
    store_pixel(ps, size_t(py) * width + px, accum, albedo, normal, hdr, first_path, samples);
}

// Pixel of work item `item`: 8 x 8 tiles in row-major order, so the first paths
// of a wave are coherent.
__device__ __forceinline__ void item_pixel(uint32_t item, uint32_t tiles_x, uint32_t& px, uint32_t& py) {
    const uint32_t tile = item >> 6, w = item & 63u;
    px = (tile % tiles_x) * 8u + (w & 7u);
    py = (tile / tiles_x) * 8u + (w >> 3);
}

// The same paths as k_pt, traced by persistent waves with path regeneration:
// a lane whose path ends starts its pixel's next path at once, and a lane whose
// pixel is done takes the next pixel from a global queue. Lanes needing a pixel
// are counted with a wave ballot; one atomic per wave reserves their items and
// each lane takes its slot by its rank (mbcnt) among them. A lane traces its
// pixel's paths in index order and writes the pixel alone, so the images equal
// k_pt's.
// Waves per SIMD the register allocation targets: the kernel's peak register
// demand exceeds the 256 VGPRs of two waves, so this trades spills for occupancy.
#ifndef VVR_PT_WAVES_PER_EU
#define VVR_PT_WAVES_PER_EU 1
#endif

__global__ void __launch_bounds__(64, VVR_PT_WAVES_PER_EU) k_pt_persistent(DevScene sc, float4* __restrict__ accum,
                                                      float4* __restrict__ albedo, float4* __restrict__ normal,
                                                      uint16_t* __restrict__ hdr, uint32_t width, uint32_t height,
                                                      uint32_t first_path, uint32_t samples,
                                                      uint32_t* __restrict__ queue) {
    const uint32_t tiles_x = (width + 7u) / 8u;
    const uint32_t n_items = tiles_x * ((height + 7u) / 8u) * 64u;
    const FrameConsts fc = frame_consts(sc.u);
    const uint32_t lane = threadIdx.x;
    constexpr uint32_t kNone = 0xffffffffu;
    uint32_t item = kNone, k = 0, px = 0, py = 0;
    bool active = false, done = false;
    PixelSums ps;
    ps.clear();
    PathState st;
    for (;;) {
        // A lane whose pixel has all its paths writes it out.
        if (!active && item != kNone && k == samples) {
            if (px < width && py < height)
                store_pixel(ps, size_t(py) * width + px, accum, albedo, normal, hdr, first_path, samples);
            item = kNone;
        }
        // Lanes without a pixel take the next ones from the queue.
        const bool need = item == kNone && !done;
        const uint64_t mask = __ballot(need);
        if (mask) {
            const int leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t base = 0;
            if (int(lane) == leader) base = atomicAdd(queue, uint32_t(__popcll(mask)));
            base = __shfl(base, leader);
            if (need) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
                const uint32_t it = base + rank;
                if (it < n_items) {
                    item = it;
                    k = 0;
                    ps.clear();
                    item_pixel(it, tiles_x, px, py);
                } else {
                    done = true;
                }
            }
        }
        // Lanes with a pixel and no path in flight start its next path (off-image
        // pixels of partial tiles are skipped).
        if (!active && item != kNone && k < samples) {
            if (px < width && py < height) {
                path_begin(sc, fc, px, py, first_path + k, st);
                active = true;
            } else {
                k = samples;
            }
        }
        if (__ballot(active || (item != kNone)) == 0) break;
        if (active && !path_bounce(sc, fc, st)) {
            ps.add(st);
            k++;
            active = false;
        }
    }
}

}  // namespace

void launch_dvr(const DevScene& s, uint16_t* hdr, uint32_t width, uint32_t height, hipStream_t st) {
    const dim3 grid((width + 7) / 8, (height + 7) / 8);
    hipLaunchKernelGGL(k_dvr, grid, dim3(64), 0, st, s, hdr, width, height);
}

void launch_dvr_batched(const DvrJob* jobs, uint32_t n_jobs, uint32_t total_tiles, hipStream_t st) {
    if (n_jobs == 0 || total_tiles == 0) return;
    hipLaunchKernelGGL(k_dvr_batched, dim3(total_tiles), dim3(64), 0, st, jobs, n_jobs);
}

void launch_pt(const DevScene& s, float4* accum, float4* albedo, float4* normal, uint16_t* hdr, uint32_t width,
               uint32_t height, uint32_t first_path, uint32_t samples, hipStream_t st) {
    const dim3 grid((width + 7) / 8, (height + 7) / 8);
    hipLaunchKernelGGL(k_pt, grid, dim3(64), 0, st, s, accum, albedo, normal, hdr, width, height, first_path, samples);
}

int pt_persistent_waves_per_cu() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_pt_persistent, 64, 0) != hipSuccess) return 8;
    return blocks > 0 ? blocks : 1;
}

void launch_pt_persistent(const DevScene& s, float4* accum, float4* albedo, float4* normal, uint16_t* hdr,
                          uint32_t width, uint32_t height, uint32_t first_path, uint32_t samples, uint32_t* queue,
                          uint32_t waves, hipStream_t st) {
    hipLaunchKernelGGL(k_pt_persistent, dim3(waves), dim3(64), 0, st, s, accum, albedo, normal, hdr, width, height,
                       first_path, samples, queue);
}

}  // namespace vvr
