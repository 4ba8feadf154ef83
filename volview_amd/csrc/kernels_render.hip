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

// A camera ray of dvr.comp in index space (dvr.comp:23-40).
struct DvrRay {
    V3 dir, o_idx, d_idx;
    V3 inv_d;  // brick_exit_inverse(d_idx)
    float t0, t1;
};
__device__ __forceinline__ DvrRay dvr_ray(const vvr_uniforms* u, uint32_t px, uint32_t py) {
    DvrRay r;
    r.dir = camera_ray(u, float(px) + 0.5f, float(py) + 0.5f);
    r.o_idx = world_to_index_point(u, V3(u->eye[0], u->eye[1], u->eye[2]));
    r.d_idx = to_index(u).mul(r.dir);
    r.inv_d = brick_exit_inverse(r.d_idx);
    intersect_volume(u, r.o_idx, r.d_idx, r.t0, r.t1);
    return r;
}

// The march's loop state between two steps.
struct MarchState {
    float tt, transmittance, prev_density;
    bool refined;
};
// The first step's distance, jittered by the pixel's noise (dvr.comp:41-44), or
// the far end for a ray that misses the volume.
__device__ __forceinline__ float march_start(const DvrRay& r, float step_mm, uint32_t px, uint32_t py) {
    return r.t0 < r.t1 ? r.t0 + ign(float(px), float(py)) * step_mm : r.t1;
}
__device__ __forceinline__ MarchState march_begin(const DevScene& sc, const DvrRay& r, uint32_t px, uint32_t py) {
    const float step_mm = sc.u->dims[3];
    MarchState m;
    m.transmittance = 1.0f;
    m.refined = false;
    m.tt = march_start(r, step_mm, px, py);
    m.prev_density = -1.0f;
    if (r.t0 < r.t1) m.prev_density = density_view<true>(sc, r.o_idx + r.d_idx * (m.tt - step_mm));
    return m;
}

// The first surface (dvr.comp:150-168): five bisection steps back over the
// last step for where the extinction reaches half the sample's.
__device__ __forceinline__ V3 first_surface(const DevScene& sc, V3 idx, V3 d_idx, float extinction) {
    const float step_mm = sc.u->dims[3];
    float inside = 0.0f, outside = -step_mm;
    const float target = 0.5f * extinction;
    for (int i = 0; i < 5; i++) {
        const float mid = 0.5f * (inside + outside);
        const V3 at = idx + d_idx * mid;
        const bool opaque = labelled(sc, tf_point(sc, density_view<true>(sc, at)), at).extinction >= target;
        inside = opaque ? mid : inside;
        outside = opaque ? outside : mid;
    }
    return idx + d_idx * inside;
}

// dvr.comp:41-147 from state `m`: empty-space skipping (dvr.comp:118-125),
// clipping, the point or pre-integrated transfer function, labels, and the
// first-surface bisection (dvr.comp:150-168). Each visible sample goes to
// visit(idx, s, density, front, cap_plane, cut, alpha, transmittance) (front:
// the segment's front density, -1 for a point sample), which returns its
// opacity (a ghost's, from its surface frame); `stop()`, asked before each
// step, ends the march there (`m` then resumes it). True once the ray is done.
template <class Visit, class Stop>
__device__ __forceinline__ bool dvr_march(const DevScene& sc, const DvrRay& r, MarchState& m, Visit&& visit,
                                          Stop&& stop) {
    const vvr_uniforms* u = sc.u;
    const float step_mm = u->dims[3];
    const bool preintegrate = u->lighting[2] != 0u;
    const V3 o_idx = r.o_idx, d_idx = r.d_idx;
    for (; m.tt < r.t1 && m.transmittance > 0.01f; m.tt += step_mm) {
        if (stop()) return false;
        V3 idx = o_idx + d_idx * m.tt;
        if (brick_majorant(sc, idx) <= 0.0f) {
            m.tt += brick_exit_fast(idx, r.inv_d);
            m.prev_density = -1.0f;
            continue;
        }
        int cap_plane;
        const float cut = clip_depth(u, idx, cap_plane);
        if (cut > 0.0f) {
            m.prev_density = -1.0f;
            continue;
        }
        const float density = density_view<true>(sc, idx);
        const float front_density = m.prev_density;
        const bool segment = preintegrate && front_density >= 0.0f;
        Sample s = segment ? tf_segment(sc, front_density, density) : tf_point(sc, density);
        m.prev_density = density;
        s = labelled(sc, s, idx);
        if (s.extinction <= 0.0f) continue;
        const float alpha = 1.0f - vexp(-s.extinction * step_mm);
        if (segment && alpha > 0.2f && m.transmittance > 0.5f && s.label == 0u &&
            fabsf(density - front_density) > 1e-4f) {
            idx = idx - d_idx * (step_mm * (1.0f - preint_half_depth(sc, front_density, density)));
        } else if (!m.refined && alpha > 0.2f && m.transmittance > 0.9f) {
            m.refined = true;
            idx = first_surface(sc, idx, d_idx, s.extinction);
        }
        m.transmittance *=
            1.0f - visit(idx, s, density, segment ? front_density : -1.0f, cap_plane, cut, alpha, m.transmittance);
    }
    return true;
}

// dvr.comp:95-145: a visible sample's shading frame, surface texture and
// lighting; returns direct + ambient, and `alpha` becomes a ghost's opacity.
// The lighting volumes (s.shadow, s.ao) are built before the frame when its
// uniforms turn them on.
__device__ inline V3 dvr_shade(const DevScene& sc, V3 dir, V3 idx, Sample s, int cap_plane, float cut, float& alpha) {
    const vvr_uniforms* u = sc.u;
    const float step_mm = u->dims[3];
    const M3 ti = to_index(u);
    const V3 l(u->light_dir[0], u->light_dir[1], u->light_dir[2]);
    const V3 v = -dir;
    const V3 light(u->light_dir[3]);
    const V3 ambient_light(u->ambient[0], u->ambient[1], u->ambient[2]);
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
    if (u->lighting[1] != 0u) ao = ao_lookup(sc, index_to_tex(u, idx + ti.mul(n * (u->light_cells[1] * surfaceness))));
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
    return direct + ambient;
}

// The march from `m` with each sample shaded at once (dvr.comp as one loop),
// adding to `radiance`; the pixel's RGBA16F.
__device__ __forceinline__ void dvr_march_shaded(const DevScene& sc, const DvrRay& r, MarchState& m, V3 radiance,
                                                 uint16_t* __restrict__ hdr, size_t pixel) {
    dvr_march(
        sc, r, m,
        [&](V3 idx, const Sample& s, float, float, int cap_plane, float cut, float alpha, float transmittance) {
            const V3 e = dvr_shade(sc, r.dir, idx, s, cap_plane, cut, alpha);
            radiance = radiance + transmittance * alpha * e;
            return alpha;
        },
        [] { return false; });
    store_half4(hdr, pixel, radiance.x, radiance.y, radiance.z, m.transmittance);
}

// dvr.comp:23-147 for one pixel.
__device__ __forceinline__ void dvr_pixel(const DevScene& sc, uint32_t px, uint32_t py, uint16_t* __restrict__ hdr,
                                          uint32_t width) {
    const DvrRay r = dvr_ray(sc.u, px, py);
    MarchState m = march_begin(sc, r, px, py);
    dvr_march_shaded(sc, r, m, V3(0.0f), hdr, size_t(py) * width + px);
}

__global__ void __launch_bounds__(64) k_dvr(DevScene sc, uint16_t* __restrict__ hdr, uint32_t width, uint32_t height) {
    uint32_t px, py;
    tile_pixel(px, py);
    if (px >= width || py >= height) return;
    dvr_pixel(sc, px, py, hdr, width);
}

// ---- the frame in passes ---------------------------------------------------------
//
// Shading a sample costs many times a march step, and in the one-loop kernel a
// wave shades whenever any of its lanes reached a visible sample, mostly with
// few lanes on. Here each round marches every pixel that is not done to its
// next kDvrSamples visible samples and records them (the shading position,
// its weight transmittance x alpha, and what rebuilds the sample: density,
// front density of a pre-integrated segment, label) in a list; the shading runs
// one sample per lane over the list, and each pixel adds its samples' light in
// march order, as the one loop does. After the last round, the pixels with more
// to march go on with the one loop. The march needs each sample's opacity, so a
// frame that may draw a ghost (whose opacity comes from its shading frame) uses
// the one loop.

__device__ __forceinline__ uint32_t lane_rank64(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Appends `value` for the lanes with `push` set: one atomic per wave.
__device__ __forceinline__ void list_push(bool push, uint32_t value, uint32_t* __restrict__ list,
                                          uint32_t* __restrict__ count) {
    const uint64_t mask = __ballot(push);
    if (!mask) return;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (int(__lane_id()) == leader) base = atomicAdd(count, uint32_t(__popcll(mask)));
    base = __shfl(base, leader);
    if (push) list[base + lane_rank64(mask)] = value;
}

// Clears the sample list's count and the segmented march's item counter.
__global__ void k_dvr_reset(uint32_t* __restrict__ count) {
    if (threadIdx.x < 4u) count[threadIdx.x] = 0u;
}

// Pixel state after a round: (tt, transmittance, previous density, flags).
constexpr uint32_t kMarchRefined = 1u, kMarchDone = 2u;  // flags; samples of the round << 8, round << 16
__device__ __forceinline__ MarchState unpack_march(float4 st) {
    MarchState m;
    m.tt = st.x, m.transmittance = st.y, m.prev_density = st.z;
    m.refined = (__float_as_uint(st.w) & kMarchRefined) != 0u;
    return m;
}

// One round of the march (round 0 from the camera, later rounds from the
// state, pixels that are done skipped). Record of sample k of pixel p at
// rec[2 (p kDvrSamples + k)]: (shading position, weight), (density, front
// density of a pre-integrated segment or -1, cut, label | (cap plane + 1) << 8).
__global__ void __launch_bounds__(64) k_dvr_march(DevScene sc, DvrPasses P, uint32_t width, uint32_t height,
                                                  uint32_t round) {
    uint32_t px, py;
    tile_pixel(px, py);
    if (px >= width || py >= height) return;
    const uint32_t pixel = py * width + px;
    const DvrRay r = dvr_ray(sc.u, px, py);
    MarchState m;
    if (round == 0u) {
        m = march_begin(sc, r, px, py);
    } else {
        const float4 st = P.state[pixel];
        if (__float_as_uint(st.w) & kMarchDone) return;
        m = unpack_march(st);
    }
    uint32_t n = 0;
    float4* __restrict__ rec = P.rec + size_t(pixel) * kDvrSamples * 2u;
    const bool done = dvr_march(
        sc, r, m,
        [&](V3 idx, const Sample& s, float density, float front, int cap_plane, float cut, float alpha,
            float transmittance) {
            float4* q = rec + n * 2u;
            q[0] = make_float4(idx.x, idx.y, idx.z, transmittance * alpha);
            q[1] = make_float4(density, front, cut, __uint_as_float(s.label | (uint32_t(cap_plane + 1) << 8)));
            list_push(true, pixel * kDvrSamples + n, P.list, P.count);
            n++;
            return alpha;
        },
        [&] { return n == kDvrSamples; });
    const uint32_t flags = (m.refined ? kMarchRefined : 0u) | (done ? kMarchDone : 0u) | (n << 8) | (round << 16);
    P.state[pixel] = make_float4(m.tt, m.transmittance, m.prev_density, __uint_as_float(flags));
}

// ---- the first round's march as parallel segments -------------------------------------
//
// A ray's march is a chain of dependent steps, and the frame waits for its longest
// rays: on DVR-1 a few hundred waves march ~200 cubic-fetch steps each while most
// finish at once. The step positions depend only on the ray and the majorant
// grid (a step into an empty brick jumps to its exit), not on what the samples
// hold, so the loop is split: k_dvr_walk runs each ray's loop over the majorant
// alone and keeps the distance at every kDvrSegIters-th iteration;
// k_dvr_segments runs the full loop body for every (ray, segment) at once and
// records the visible samples; k_dvr_merge takes each ray's records in march
// order and does what the loop does with them (transmittance, the stop after
// kDvrSamples samples, the first-surface bisection), so the records, states
// and list equal k_dvr_march's bit for bit. Rays past kDvrSegMax segments, or
// whose segment found the pool full, go on from that segment's state in
// k_dvr_continue.

constexpr uint32_t kSegFull = 15u;  // seg_out count of a segment that found the pool full

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) v = max(v, uint32_t(__shfl_xor(int(v), d)));
    return v;
}

// Each ray's loop over the majorant alone: the distance at the start of every
// segment; the tile's (tile, segment) items for k_dvr_segments.
__global__ void __launch_bounds__(64) k_dvr_walk(DevScene sc, DvrPasses P, uint32_t width, uint32_t height) {
    uint32_t px, py;
    tile_pixel(px, py);
    const uint32_t lane = threadIdx.x, npx = width * height;
    uint32_t nseg = 0;
    if (px < width && py < height) {
        const uint32_t pixel = py * width + px;
        const DvrRay r = dvr_ray(sc.u, px, py);
        const float step_mm = sc.u->dims[3];
        float tt = march_start(r, step_mm, px, py);
        int cached = -2;  // the brick whose majorant `maj` holds (brick_index gives -1 and up)
        float maj = 0.0f;
        while (tt < r.t1 && nseg < P.seg_max) {
            P.seg_tt[size_t(nseg) * npx + pixel] = tt;
            nseg++;
            for (uint32_t k = 0; k < kDvrSegIters && tt < r.t1; k++, tt += step_mm) {
                const V3 idx = r.o_idx + r.d_idx * tt;
                const int b = brick_index(sc, idx);
                if (b != cached) maj = brick_majorant_at(sc, b), cached = b;
                if (maj <= 0.0f) tt += brick_exit_fast(idx, r.inv_d);
            }
        }
        P.walk[pixel] = make_float2(tt, __uint_as_float(nseg));
    }
    const uint32_t items = wave_max(nseg);
    uint32_t base = 0;
    if (lane == 0u && items) base = atomicAdd(P.seg_n, items);
    base = uint32_t(__shfl(int(base), 0));
    if (lane < items) P.seg_items[base + lane] = (blockIdx.y * gridDim.x + blockIdx.x) | (lane << 22);
}

__device__ __forceinline__ uint32_t wave_exclusive_sum(uint32_t v, uint32_t& total) {
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = uint32_t(__shfl_up(int(inc), d));
        if (int(__lane_id()) >= d) inc += o;
    }
    total = uint32_t(__shfl(int(inc), 63));
    return inc - v;
}

// The listed (tile, segment) items, one per wave at a time (wave w takes items
// w, w + waves, ...): each ray of the tile marches that segment's loop
// iterations and its visible samples go to the wave's own part of the pool,
// contiguous per ray and segment. (Items and pool space taken from shared
// counters made the waves wait on contended atomics: 3x slower.)
__global__ void __launch_bounds__(64) k_dvr_segments(DevScene sc, DvrPasses P, uint32_t width, uint32_t height,
                                                     uint32_t tiles_x) {
    __shared__ float4 s_rec[kDvrSegIters][64];
    const uint32_t lane = threadIdx.x, npx = width * height;
    const vvr_uniforms* u = sc.u;
    const float step_mm = u->dims[3];
    const uint32_t items = P.seg_n[0];
    const uint32_t chunk = P.pool_cap / gridDim.x;  // this wave's part of the pool
    float4* __restrict__ part_pool = P.pool + size_t(blockIdx.x) * chunk;
    uint32_t used = 0;
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
        const uint32_t e = P.seg_items[item], tile = e & 0x3fffffu, seg = e >> 22;
        const uint32_t px = (tile % tiles_x) * 8u + (lane & 7u), py = (tile / tiles_x) * 8u + (lane >> 3);
        const uint32_t pixel = py * width + px;
        bool active = px < width && py < height;
        if (active) active = seg < __float_as_uint(P.walk[pixel].y);
        uint32_t n = 0;
        if (active) {
            const DvrRay r = dvr_ray(u, px, py);
            float tt = P.seg_tt[size_t(seg) * npx + pixel];
            // The majorant of the next step's brick is loaded before this step's
            // work (a step into an empty brick loads it again from its exit).
            float maj = brick_majorant(sc, r.o_idx + r.d_idx * tt);
            for (uint32_t k = 0; k < kDvrSegIters && tt < r.t1; k++) {
                const V3 idx = r.o_idx + r.d_idx * tt;
                if (maj <= 0.0f) {
                    tt += brick_exit_fast(idx, r.inv_d);
                    tt += step_mm;
                    maj = brick_majorant(sc, r.o_idx + r.d_idx * tt);
                    continue;
                }
                const float step_at = tt;
                tt += step_mm;
                maj = brick_majorant(sc, r.o_idx + r.d_idx * tt);
                int cap_plane;
                if (clip_depth(u, idx, cap_plane) > 0.0f) continue;
                const float density = density_view<true>(sc, idx);
                const Sample smp = labelled(sc, tf_point(sc, density), idx);
                if (smp.extinction <= 0.0f) continue;
                s_rec[n][lane] = make_float4(step_at, smp.extinction, density, __uint_as_float(smp.label));
                n++;
            }
        }
        uint32_t total;
        const uint32_t first = wave_exclusive_sum(n, total);
        const uint32_t base = used;
        const bool fits = base + total <= chunk;
        if (fits) used += total;
        if (active) {
            if (fits) {
                for (uint32_t q = 0; q < n; q++) part_pool[base + first + q] = s_rec[q][lane];
                P.seg_out[size_t(seg) * npx + pixel] = ((blockIdx.x * chunk + base + first) << 4) | n;
            } else {
                P.seg_out[size_t(seg) * npx + pixel] = kSegFull;
            }
        }
    }
}

// Each ray's records in march order, as dvr_march's loop takes its samples:
// the same records, list entries and state as k_dvr_march (round 0). The
// first-surface bisection only moves its sample's shading position, so it runs
// after the loop, with the wave's lanes together.
__global__ void __launch_bounds__(64) k_dvr_merge(DevScene sc, DvrPasses P, uint32_t width, uint32_t height) {
    __shared__ uint32_t s_so[kDvrSegMax][64];  // every segment's pool entry, loaded before any is used
    uint32_t px, py;
    tile_pixel(px, py);
    const bool valid = px < width && py < height;
    const vvr_uniforms* u = sc.u;
    const uint32_t lane = threadIdx.x, pixel = py * width + px, npx = width * height;
    const float step_mm = u->dims[3];
    uint32_t n = 0, refine_n = kDvrSamples;
    float refine_ext = 0.0f;
    DvrRay r;
    if (valid) {
        const float2 w = P.walk[pixel];
        const uint32_t nseg = __float_as_uint(w.y);
        uint32_t so[kDvrSegMax];
#pragma unroll
        for (uint32_t k = 0; k < kDvrSegMax; k++) so[k] = k < nseg ? P.seg_out[size_t(k) * npx + pixel] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kDvrSegMax; k++) s_so[k][lane] = so[k];
        r = dvr_ray(u, px, py);
        float4* __restrict__ rec = P.rec + size_t(pixel) * kDvrSamples * 2u;
        float transmittance = 1.0f, tt = w.x;
        bool refined = false, ended = false, done = false;
        for (uint32_t seg = 0; seg < nseg && !ended; seg++) {
            const uint32_t out = s_so[seg][lane];
            if ((out & 15u) == kSegFull) {  // not marched: the one loop goes on from here
                tt = P.seg_tt[size_t(seg) * npx + pixel];
                ended = true;
                break;
            }
            for (uint32_t q = 0; q < (out & 15u); q++) {
                const float4 v = P.pool[(out >> 4) + q];
                const float at = v.x, extinction = v.y, density = v.z;
                const uint32_t label = __float_as_uint(v.w);
                const V3 idx = r.o_idx + r.d_idx * at;
                int cap_plane;
                const float cut = clip_depth(u, idx, cap_plane);
                const float alpha = 1.0f - vexp(-extinction * step_mm);
                if (!refined && alpha > 0.2f && transmittance > 0.9f) {
                    refined = true;
                    refine_n = n, refine_ext = extinction;
                }
                float4* qr = rec + n * 2u;
                qr[0] = make_float4(idx.x, idx.y, idx.z, transmittance * alpha);
                qr[1] = make_float4(density, -1.0f, cut, __uint_as_float(label | (uint32_t(cap_plane + 1) << 8)));
                n++;
                transmittance *= 1.0f - alpha;
                // The loop's checks before its next step: the ray ends when it is
                // opaque (or past its end); after kDvrSamples samples it stops.
                tt = at + step_mm;
                if (!(transmittance > 0.01f) || !(tt < r.t1)) {
                    done = ended = true;
                    break;
                }
                if (n == kDvrSamples) {
                    ended = true;
                    break;
                }
            }
        }
        if (!ended) {  // every segment taken: done if the walk reached the ray's end
            tt = w.x;
            done = !(tt < r.t1);
        }
        const uint32_t flags = (refined ? kMarchRefined : 0u) | (done ? kMarchDone : 0u) | (n << 8);
        P.state[pixel] = make_float4(tt, transmittance, -1.0f, __uint_as_float(flags));
    }
    // The pixel's samples on the shading list (one atomic per wave).
    uint32_t total;
    const uint32_t first = wave_exclusive_sum(n, total);
    uint32_t base = 0;
    if (lane == 0u && total) base = atomicAdd(P.count, total);
    base = uint32_t(__shfl(int(base), 0));
    for (uint32_t k = 0; k < n; k++) P.list[base + first + k] = pixel * kDvrSamples + k;
    if (refine_n < kDvrSamples) {
        float4* qr = P.rec + (size_t(pixel) * kDvrSamples + refine_n) * 2u;
        const float4 a = qr[0];
        const V3 idx = first_surface(sc, V3(a.x, a.y, a.z), r.d_idx, refine_ext);
        qr[0] = make_float4(idx.x, idx.y, idx.z, a.w);
    }
}

// The listed samples' light (direct + ambient), one per lane, over the
// record's shading position (its weight kept). The sample is rebuilt as the
// march made it: the point or segment transfer function, then its label
// (with_label; the shading reads no extinction, the one field labelled()
// changes after it).
__global__ void __launch_bounds__(64) k_dvr_shade(DevScene sc, DvrPasses P, uint32_t width) {
    const uint32_t n = *P.count;
    const uint32_t mode = uint32_t(sc.u->labels[0]);
    for (uint32_t q = blockIdx.x * 64u + __lane_id(); q < n; q += gridDim.x * 64u) {
        const uint32_t slot = P.list[q];
        const uint32_t pixel = slot / kDvrSamples;
        const V3 dir = camera_ray(sc.u, float(pixel % width) + 0.5f, float(pixel / width) + 0.5f);
        float4* __restrict__ rec = P.rec + size_t(slot) * 2u;
        const float4 a = rec[0], b = rec[1];
        const uint32_t bits = __float_as_uint(b.w);
        const Sample s0 = b.y >= 0.0f ? tf_segment(sc, b.y, b.x) : tf_point(sc, b.x);
        const Sample s = with_label(sc, s0, bits & 0xffu, mode, false);
        float alpha = 0.0f;
        const V3 e = dvr_shade(sc, dir, V3(a.x, a.y, a.z), s, int(bits >> 8) - 1, b.z, alpha);
        rec[0] = make_float4(e.x, e.y, e.z, a.w);
    }
}

// The round's light of each pixel that marched in it, added in march order
// to its running radiance; the RGBA16F of the pixels that are done.
__global__ void __launch_bounds__(64) k_dvr_resolve(DvrPasses P, uint16_t* __restrict__ hdr, uint32_t width,
                                                    uint32_t height, uint32_t round) {
    uint32_t px, py;
    tile_pixel(px, py);
    if (px >= width || py >= height) return;
    const uint32_t pixel = py * width + px;
    const float4 st = P.state[pixel];
    const uint32_t flags = __float_as_uint(st.w);
    if ((flags >> 16) != round) return;  // done in an earlier round
    const float4* __restrict__ rec = P.rec + size_t(pixel) * kDvrSamples * 2u;
    V3 radiance(0.0f);
    if (round != 0u) {
        const float4 q = P.radiance[pixel];
        radiance = V3(q.x, q.y, q.z);
    }
    for (uint32_t k = 0; k < ((flags >> 8) & 0xffu); k++) {
        const float4 e = rec[k * 2u];  // light, weight
        radiance = radiance + e.w * V3(e.x, e.y, e.z);
    }
    if (flags & kMarchDone) store_half4(hdr, pixel, radiance.x, radiance.y, radiance.z, st.y);
    else P.radiance[pixel] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
}

// After the last round: the pixels with more to march go on with the one loop.
__global__ void __launch_bounds__(64) k_dvr_continue(DevScene sc, DvrPasses P, uint16_t* __restrict__ hdr,
                                                     uint32_t width, uint32_t height) {
    uint32_t px, py;
    tile_pixel(px, py);
    if (px >= width || py >= height) return;
    const uint32_t pixel = py * width + px;
    const float4 st = P.state[pixel];
    if (__float_as_uint(st.w) & kMarchDone) return;
    const DvrRay r = dvr_ray(sc.u, px, py);
    MarchState m = unpack_march(st);
    const float4 rad = P.radiance[pixel];
    dvr_march_shaded(sc, r, m, V3(rad.x, rad.y, rad.z), hdr, pixel);
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

void launch_dvr_passes(const DevScene& s, const DvrPasses& P, uint16_t* hdr, uint32_t width, uint32_t height,
                       int cus, bool segmented, hipStream_t st) {
    const dim3 tiles((width + 7) / 8, (height + 7) / 8);
    const uint32_t shade_waves = std::max(std::min((width * height * kDvrSamples + 63u) / 64u, uint32_t(cus) * 16u), 1u);
    for (uint32_t round = 0; round < kDvrRounds; round++) {
        hipLaunchKernelGGL(k_dvr_reset, dim3(1), dim3(64), 0, st, P.count);
        if (round == 0 && segmented) {
            hipLaunchKernelGGL(k_dvr_walk, tiles, dim3(64), 0, st, s, P, width, height);
            // 32 waves per CU, or one per item a frame this small can have: each wave's
            // part of the pool is the pool over the waves.
            const uint32_t seg_waves = uint32_t(std::min<size_t>(size_t(std::max(cus, 1)) * 32u,
                                                                 size_t(tiles.x) * tiles.y * kDvrSegMax));
            hipLaunchKernelGGL(k_dvr_segments, dim3(seg_waves), dim3(64), 0, st, s, P, width, height, tiles.x);
            hipLaunchKernelGGL(k_dvr_merge, tiles, dim3(64), 0, st, s, P, width, height);
        } else {
            hipLaunchKernelGGL(k_dvr_march, tiles, dim3(64), 0, st, s, P, width, height, round);
        }
        hipLaunchKernelGGL(k_dvr_shade, dim3(shade_waves), dim3(64), 0, st, s, P, width);
        hipLaunchKernelGGL(k_dvr_resolve, tiles, dim3(64), 0, st, P, hdr, width, height, round);
    }
    hipLaunchKernelGGL(k_dvr_continue, tiles, dim3(64), 0, st, s, P, hdr, width, height);
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
