// Internals of libvvr shared by the renderer and the server translation units:
// error handling, device allocation, the wavefront engine's device state and
// the renderer object behind the C ABI. Not part of the C ABI.
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <vector>

#include "host.hpp"
#include "launch.hpp"

namespace vvr {

inline thread_local std::string g_last_error;

inline void check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? VVR_ENOMEM : VVR_EDEVICE;
    throw Error{code, std::string(what) + ": " + hipGetErrorString(e)};
}

template <class T> T* dalloc(size_t count, const char* what) {
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{VVR_ENOMEM, std::string("cannot allocate ") + what + " (" + std::to_string(count * sizeof(T)) +
                                    " bytes): " + hipGetErrorString(e)};
    }
    return static_cast<T*>(p);
}

inline void dfree(void* p) {
    if (p) (void)hipFree(p);
}

inline uint32_t env_u32(const char* name, uint32_t fallback) {
    const char* v = std::getenv(name);
    if (!v || !*v) return fallback;
    const unsigned long x = std::strtoul(v, nullptr, 10);
    return x ? uint32_t(std::min<unsigned long>(x, 1ul << 30)) : fallback;
}

// Whether a label of a frame with uniforms `u` may be drawn as a ghost (a
// label colour's alpha above 1.5 in a label mode).
inline bool labels_may_draw_ghosts(const vvr_uniforms& u) {
    if (uint32_t(u.labels[0]) != VVR_LABELS_OFF)
        for (int l = 1; l < 256; l++)
            if (u.label_colors[l][3] > 1.5f) return true;
    return false;
}
// Whether a frame with uniforms `u` reads the density (all but isolate mode,
// whose samples take the label's optics alone: density_unused on the device).
inline bool frame_reads_density(const vvr_uniforms& u) { return uint32_t(u.labels[0]) != VVR_LABELS_ISOLATE; }
// Such a frame can be ray marched in passes unless a label may be drawn as a
// ghost, whose opacity comes from its shading frame, which the march cannot know.
inline bool dvr_in_passes(const vvr_uniforms& u) { return !labels_may_draw_ghosts(u); }
// Whether its march runs as parallel segments: not when pre-integrated, whose
// samples need the previous step's density.
inline bool dvr_segmented(const vvr_uniforms& u) { return u.lighting[2] == 0u; }

// The buffers of launch_dvr_passes for frames of up to `px` pixels in `tiles`
// 8 x 8 tiles, carved from `mem` (dvr_passes_bytes of such a frame).
inline DvrPasses carve_dvr_passes(uint8_t* mem, size_t px, size_t tiles) {
    DvrPasses P{};
    uint8_t* p = mem;
    auto take = [&](size_t bytes) {
        uint8_t* q = p;
        p += (bytes + 255) / 256 * 256;
        return q;
    };
    P.rec = reinterpret_cast<float4*>(take(px * kDvrSamples * 32));
    P.state = reinterpret_cast<float4*>(take(px * 16));
    P.radiance = reinterpret_cast<float4*>(take(px * 16));
    P.list = reinterpret_cast<uint32_t*>(take(px * kDvrSamples * 4));
    P.count = reinterpret_cast<uint32_t*>(take(16));  // and the segment counters
    P.seg_n = P.count + 2;
    P.seg_tt = reinterpret_cast<float*>(take(px * kDvrSegMax * 4));
    P.seg_out = reinterpret_cast<uint32_t*>(take(px * kDvrSegMax * 4));
    P.walk = reinterpret_cast<float2*>(take(px * 8));
    P.pool = reinterpret_cast<float4*>(take(px * kDvrPoolPerPixel * 16));
    P.seg_items = reinterpret_cast<uint32_t*>(take(tiles * kDvrSegMax * 4));
    // VVR_DVR_POOL (records per pixel, 1 .. kDvrPoolPerPixel) shrinks the pool, so
    // that segments find it full and their rays go on in the one loop.
    const uint32_t per_px = std::max(1u, std::min(env_u32("VVR_DVR_POOL", kDvrPoolPerPixel), kDvrPoolPerPixel));
    P.pool_cap = uint32_t(std::min<size_t>(px * per_px, size_t(1) << 28));
    // VVR_DVR_SEGMENTS (1 .. kDvrSegMax) shortens the walk: longer rays go on in the one loop.
    P.seg_max = std::max(1u, std::min(env_u32("VVR_DVR_SEGMENTS", kDvrSegMax), kDvrSegMax));
    return P;
}

// Device state of the wavefront path tracer: the path slots, queues, the chunk
// sums of the requests in flight and the request table.
struct WfEngine {
    int device = 0;
    uint32_t slots = 0;
    uint8_t* slot_mem = nullptr;
    WfSlots S{};
    WfQueues Q{};
    float4* partial = nullptr;
    size_t partial_cap = 0;  // float4s
    PtRequest* d_reqs = nullptr;
    PtRequest* h_reqs = nullptr;  // pinned staging
    uint32_t req_cap = 0;
    uint32_t* h_live = nullptr;   // pinned, two read-back slots
    hipEvent_t live_ev[2] = {nullptr, nullptr};
    WfGrid grid{};
    uint32_t max_slots = 1u << 23;

    explicit WfEngine(int dev) : device(dev) {
        int cus = 0;
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "device attribute");
        grid = wf_grid(std::max(cus, 1));
        grid.flight_steps = int(std::min(env_u32("VVR_PT_FLIGHT_STEPS", uint32_t(grid.flight_steps)), 4096u));
        grid.shadow_steps = int(std::min(env_u32("VVR_PT_SHADOW_STEPS", uint32_t(grid.shadow_steps)), 4096u));
        if (const char* v = std::getenv("VVR_PT_TAIL")) grid.tail_paths = uint32_t(std::strtoul(v, nullptr, 10));  // 0: never
        // 8M slots (about 3 GiB of the 288): enough paths in flight that the long
        // random walks of a frame's last iterations still fill the device (PT-1
        // 201.8 against 207.1 ms per frame at 4M, with 24 / 32 tracking runs).
        max_slots = env_u32("VVR_PT_SLOTS", 1u << 23);
        check(hipHostMalloc((void**)&h_live, 2 * sizeof(uint32_t), hipHostMallocDefault), "hipHostMalloc");
        for (auto& e : live_ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    }
    ~WfEngine() {
        dfree(slot_mem), dfree(partial), dfree(d_reqs);
        if (h_reqs) (void)hipHostFree(h_reqs);
        if (h_live) (void)hipHostFree(h_live);
        for (auto& e : live_ev)
            if (e) (void)hipEventDestroy(e);
    }

    void reserve_slots(uint32_t n) {
        if (n <= slots) return;
        dfree(slot_mem);
        slot_mem = nullptr, slots = 0;
        constexpr int kQueues = 10;
        const size_t path_bytes = size_t(n) * kWfPathRecord * sizeof(float4);
        const size_t sums_bytes = size_t(n) * kWfSumsRecord * sizeof(float4);
        const size_t event_bytes = size_t(n) * kWfEventRecord * sizeof(float4);
        const size_t q_bytes = (size_t(n) * 4 + 255) / 256 * 256;
        const size_t sorted_bytes = (size_t(wf_sorted_capacity(n)) * 4 + 255) / 256 * 256;
        const size_t req_bytes = (size_t(n) * 2 + 255) / 256 * 256, bin_bytes = 4 * kWfBins * 4;
        slot_mem = dalloc<uint8_t>(path_bytes + sums_bytes + event_bytes + kQueues * q_bytes + sorted_bytes + req_bytes + 256 +
                                       bin_bytes,
                                   "path slots");
        uint8_t* p = slot_mem;
        float4* paths = reinterpret_cast<float4*>(p);
        p += path_bytes;
        float4* events = reinterpret_cast<float4*>(p);
        p += event_bytes;
        float4* sums = reinterpret_cast<float4*>(p);
        p += sums_bytes;
        auto sums_field = [&](int k) { return SlotField<float4>{sums + k, kWfSumsRecord}; };
        auto path_field = [&](int k) { return SlotField<float4>{paths + k, kWfPathRecord}; };
        auto event_field = [&](int k) { return SlotField<float4>{events + k, kWfEventRecord}; };
        S.p_origin = path_field(0), S.p_dir = path_field(1), S.p_thr = path_field(2), S.p_spec = path_field(3);
        S.p_rad = path_field(4);
        S.p_flight = SlotField<uint2>{reinterpret_cast<uint2*>(paths + 5), 2 * kWfPathRecord};
        S.c0 = sums_field(0), S.c1 = sums_field(1), S.c2 = sums_field(2), S.e_map = sums_field(3);
        S.e_maplt = sums_field(4);
        S.e_idx = event_field(0), S.e_s0 = event_field(1), S.e_s1 = event_field(2), S.e_s2 = event_field(3);
        S.e_n = event_field(4), S.e_light = event_field(5), S.e_lt = event_field(6);
        uint32_t** qs[kQueues] = {&Q.ff[0],   &Q.ff[1],   &Q.ev,    &Q.end[0], &Q.end[1],
                                  &Q.idle[0], &Q.idle[1], &Q.sh[0], &Q.sh[1],  &Q.sc};
        for (auto q : qs) *q = reinterpret_cast<uint32_t*>(p), p += q_bytes;
        Q.sorted = reinterpret_cast<uint32_t*>(p), p += sorted_bytes;
        S.req = reinterpret_cast<uint16_t*>(p), p += req_bytes;
        Q.counts = reinterpret_cast<uint32_t*>(p), p += 256;
        Q.bins = reinterpret_cast<uint32_t*>(p);
        slots = n;
    }

    void reserve(uint32_t n_reqs, size_t partial_f4) {
        if (partial_f4 > partial_cap) {
            dfree(partial);
            partial = nullptr, partial_cap = 0;
            partial = dalloc<float4>(partial_f4, "path-tracing chunk sums");
            partial_cap = partial_f4;
        }
        if (n_reqs > req_cap) {
            dfree(d_reqs);
            if (h_reqs) (void)hipHostFree(h_reqs);
            d_reqs = nullptr, h_reqs = nullptr, req_cap = 0;
            d_reqs = dalloc<PtRequest>(n_reqs, "request table");
            check(hipHostMalloc((void**)&h_reqs, n_reqs * sizeof(PtRequest), hipHostMallocDefault), "hipHostMalloc");
            req_cap = n_reqs;
        }
    }

    // The shape the next run uses (vvr_pt_shape; zero fields keep theirs).
    void set_shape(const vvr_pt_shape& s) {
        if (s.flight_steps > 4096 || s.shadow_steps > 4096) throw Error{VVR_EINVAL, "steps must be 1..4096"};
        if (s.slots && (s.slots < 64 || s.slots > (1u << 26))) throw Error{VVR_EINVAL, "slots must be 64..2^26"};
        if (s.slots) max_slots = (s.slots + 63u) / 64u * 64u;
        if (s.flight_steps) grid.flight_steps = int(s.flight_steps);
        if (s.shadow_steps) grid.shadow_steps = int(s.shadow_steps);
        if (s.tail_paths) grid.tail_paths = s.tail_paths == UINT32_MAX ? 0u : s.tail_paths;
    }
    vvr_pt_shape shape() const {
        return vvr_pt_shape{max_slots, uint32_t(grid.flight_steps), uint32_t(grid.shadow_steps),
                            grid.tail_paths ? grid.tail_paths : UINT32_MAX};
    }

    // Items of `samples` paths per pixel at `chunk` paths per item. Work items are
    // numbered in 32 bits: a request of 2^31 items or more is refused here, before
    // anything is allocated or queued.
    static void shape(PtRequest& q, uint32_t width, uint32_t height, uint32_t first, uint32_t samples,
                      uint32_t chunk) {
        q.width = width, q.height = height;
        q.tiles_x = (width + 7) / 8;
        q.n_px = q.tiles_x * ((height + 7) / 8) * 64u;
        q.first_path = first, q.samples = samples;
        q.chunk = std::min(std::max(chunk, 1u), samples);
        q.n_chunks = (samples + q.chunk - 1) / q.chunk;
        const uint64_t items = uint64_t(q.n_chunks) * q.n_px;
        if (items >= (1ull << 31))
            throw Error{VVR_ELIMIT, "too many paths in one frame: " + std::to_string(items) +
                                        " work items (the limit is 2^31; lower samples_per_frame)"};
        q.n_items = uint32_t(items);
    }

    // Renders the requests (scene, images and shape set) on `st`. Returns once
    // every path is traced or queued to be; the resolves are queued last.
    // env_map: some request renders with a light map; ghosts: a request's labels
    // may be drawn as ghosts; density: a request reads the density.
    void run(const PtRequest* reqs, uint32_t n, hipStream_t st, bool env_map = false, bool ghosts = true,
             bool density = true) {
        if (n == 0) return;
        size_t partial_f4 = 0;
        uint64_t items = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (uint64_t(reqs[i].n_chunks) * reqs[i].n_px != reqs[i].n_items)
                throw Error{VVR_ELIMIT, "request shape out of range"};
            partial_f4 += size_t(reqs[i].n_chunks) * 3u * reqs[i].n_px;
            items += reqs[i].n_items;
        }
        if (items >= (1ull << 31)) throw Error{VVR_ELIMIT, "too many paths in one frame (2^31 work items)"};
        if (n > kWfBins) throw Error{VVR_ELIMIT, "too many requests in one run"};
        reserve(n, partial_f4);
        const uint32_t want = uint32_t(std::min<uint64_t>(max_slots, (items + 63) / 64 * 64));
        reserve_slots(want);
        size_t off = 0;
        uint32_t base = 0;
        for (uint32_t i = 0; i < n; i++) {
            h_reqs[i] = reqs[i];
            h_reqs[i].partial = partial + off;
            h_reqs[i].item_base = base;
            off += size_t(reqs[i].n_chunks) * 3u * reqs[i].n_px;
            base += reqs[i].n_items;
        }
        check(hipMemcpyAsync(d_reqs, h_reqs, n * sizeof(PtRequest), hipMemcpyHostToDevice, st), "request upload");
        wf_init(S, Q, want, st);
        check(hipGetLastError(), "wavefront init");
        // Iterations in batches; the live-slot count after each batch is read
        // back while the next two batches run, so the device never waits on the host.
        constexpr uint32_t kBatch = 4;
        uint32_t iter = 0;
        auto launch_batch = [&](int k) {
            for (uint32_t j = 0; j < kBatch; j++, iter++)
                wf_iteration(d_reqs, n, base, S, Q, iter & 1u, grid, env_map, ghosts, density, st);
            check(hipGetLastError(), "wavefront iteration");
            check(hipMemcpyAsync(&h_live[k], Q.counts + kWfLIVE, 4, hipMemcpyDeviceToHost, st), "live count");
            check(hipEventRecord(live_ev[k], st), "hipEventRecord");
        };
        launch_batch(0);
        launch_batch(1);
        for (int k = 0;; k ^= 1) {
            check(hipEventSynchronize(live_ev[k]), "wavefront progress");
            if (h_live[k] == 0) break;
            // Fewer live slots than were filled: every item is handed out (regen
            // leaves a slot idle only then). Few paths left: finish them in one launch.
            if (n == 1 && h_live[k] < std::min(want, grid.tail_paths)) {
                wf_finish(d_reqs, S, Q, iter & 1u, h_live[k], ghosts, st);
                check(hipGetLastError(), "wavefront tail");
                break;
            }
            launch_batch(k);
        }
        for (uint32_t i = 0; i < n; i++) wf_resolve(h_reqs[i], st);
        check(hipGetLastError(), "wavefront resolve");
        // The staging above is read by copies that have completed (the loop
        // waited on events recorded after them).
    }
};

// The frame parameters record() and the server accept.
// What a lighting volume depends on beyond the opacity key (renderer.rs:1663-1667,
// 1707-1749): the clip planes as uploaded and the label colours and densities
// sample_point reads; the shadow volume also on the light direction and the
// subsurface terms. Words of the uniform block, compared exactly.
struct LightKey {
    OpacityKey opacity;
    std::vector<uint32_t> words;
    bool operator==(const LightKey& o) const { return opacity == o.opacity && words == o.words; }
    bool operator!=(const LightKey& o) const { return !(*this == o); }
};

inline LightKey light_key(const OpacityKey& opacity, const vvr_uniforms& u, bool shadow) {
    LightKey k{opacity, {}};
    auto add = [&](const void* p, size_t bytes) {
        const uint32_t* w = static_cast<const uint32_t*>(p);
        k.words.insert(k.words.end(), w, w + bytes / 4);
    };
    add(u.clip, sizeof(u.clip));
    add(u.clip_planes, sizeof(u.clip_planes));
    add(u.label_colors, sizeof(u.label_colors));
    add(u.label_density, sizeof(u.label_density));
    if (shadow) {
        add(u.light_dir, 3 * sizeof(float));
        add(&u.subsurface[1], 2 * sizeof(float));
    }
    return k;
}

inline void validate_frame(const vvr_frame_params& p) {
    const bool pt = p.path_tracing_enabled != 0;
    if (pt && p.path_tracing.denoiser > VVR_DENOISE_LEARNED) throw Error{VVR_EINVAL, "unknown denoiser"};
    if (!(p.step_mm > 0.0f)) throw Error{VVR_EINVAL, "step_mm must be positive"};
    if (p.labels.mode > VVR_LABELS_HYBRID) throw Error{VVR_EINVAL, "unknown label mode"};
}

}  // namespace vvr

// The ABI's opaque types are global (include/vvr.h); their members use vvr's names.
using namespace vvr;

struct vvr_ctx {
    int device;
};

// A server request's hold on its renderer; let_go returns true when the caller
// must delete the renderer (destroyed meanwhile and this was the last hold).
inline void hold_renderer(vvr_renderer* r);
inline bool let_go_renderer(vvr_renderer* r);

// Runs the learned denoiser on a path-traced frame's images into its HDR image (unet.cpp).
void unet_denoise_frame(vvr_unet* n, const float4* accum, const float4* albedo, const float4* normal, float samples,
                        uint32_t width, uint32_t height, uint16_t* hdr, hipStream_t st);

struct vvr_renderer {
    int device = 0;
    hipStream_t stream = nullptr;
    // Server requests that read this renderer hold it: a renderer destroyed
    // while held is deleted when its last request lets go (vvr_renderer_destroy).
    std::mutex hold_mu;
    int holds = 0;
    bool orphaned = false;
    Geometry geo{};
    BrickGrid grid{};
    int mdims[3] = {0, 0, 0};

    uint16_t* density = nullptr;
    float* density_f = nullptr;  // its values as floats (launch_upload_density)
    uint32_t* normals = nullptr;
    float* majorant = nullptr;
    uint32_t* maj_cells = nullptr;  // majorant build (modes off, tint): per micro-brick, its 2^3 cells' density
    uint8_t* maj_nz = nullptr;      // min | (65535 - max) << 16, and (tint) a bit per cell: some mask value is
                                    // non-zero
    uint16_t* maj_ranges = nullptr;  // per brick, its window's transfer-function bucket range
    // The density never changes after upload, so its cell records and brick ranges
    // are built once; the mask bits of maj_nz follow the mask (mask_builds counts
    // its rebuilds).
    bool maj_cells_ready = false;
    uint64_t mask_builds = 0, maj_nz_build = 0;
    uint8_t* labels = nullptr;
    uint32_t* mask = nullptr;    // image A: what rendering samples
    uint32_t* mask_b = nullptr;  // image B: blur partner
    uint32_t* mask_cells = nullptr;  // one bit per mask cell (launch_mask_cells)
    uint8_t* cell_labels = nullptr;  // and its uniform drawn label, or 0
    uint32_t* soft = nullptr;
    SoftGrid soft_grid{};
    float4* lut = nullptr;
    float4* preint = nullptr;
    vvr_uniforms* d_uniforms[2] = {nullptr, nullptr};
    vvr_uniforms* h_uniforms = nullptr;  // pinned, 2 slots (FRAMES_IN_FLIGHT, renderer.rs:14)
    hipEvent_t slot_done[2] = {nullptr, nullptr};
    bool slot_used[2] = {false, false};
    vvr_uniforms last_uniforms{};

    TfTables tf;
    bool has_labels = false;
    uint64_t generation = 0;
    std::optional<OpacityKey> mask_key, majorant_key;
    // Lighting volumes of the ray marcher, allocated on first use.
    uint2* shadow = nullptr;   // RGBA16F at ceil(dims / 2)
    uint16_t* ao = nullptr;    // R16F at ceil(dims / 4)
    std::optional<LightKey> shadow_key, ao_key;
    // Light map of Environment::Map (set_environment_map), absent by default.
    uint2* env_rad = nullptr;
    float2* env_table = nullptr;
    uint2* env_dif = nullptr;
    int env_w = 0, env_h = 0;
    bool has_env = false;
    float env_peak_u = 0.0f;
    const float* env_peak() const { return has_env ? &env_peak_u : nullptr; }

    uint32_t width = 0, height = 0;
    uint16_t* hdr = nullptr;
    float4* accum = nullptr;
    float4* albedo = nullptr;
    float4* normal = nullptr;
    float4* dn_scratch = nullptr;  // a-trous denoiser: 4 float4 images of the frame, allocated on first use
    uint32_t* output = nullptr;    // RGBA8 display image (tone-mapped)
    uint16_t* bloom = nullptr;     // glow images A, B (quarter), C, D (eighth resolution), RGBA16F
    float* exposure = nullptr;     // adapted auto exposure, persists across frames
    bool denoised = false;         // the HDR image holds the denoiser's output
    vvr_unet* unet = nullptr;      // the learned denoiser (VVR_DENOISE_LEARNED), owned by the caller
    std::unique_ptr<WfEngine> wf;  // the wavefront path tracer, created on the first path-traced frame
    uint32_t pt_chunk = 1;         // paths per wavefront work item (VVR_PT_CHUNK)
    uint32_t accumulated = 0;
    bool acc_valid = false;
    vvr_frame_params acc_params{};

    bool timing = false;
    struct Pending {
        int which;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    double total_ms[VVR_KERNEL_COUNT] = {};
    uint64_t launches[VVR_KERNEL_COUNT] = {};

    void use_device() const { check(hipSetDevice(device), "hipSetDevice"); }

    int shadow_dims(int a) const { return (int(geo.dims[a]) + 1) / 2; }  // SHADOW_DOWNSAMPLE (renderer.rs:35-36)
    int ao_dims(int a) const { return (int(geo.dims[a]) + 3) / 4; }      // AO_DOWNSAMPLE (renderer.rs:39-40)

    // Whether `p` needs the label masks or the majorant rebuilt.
    bool needs_opacity_rebuild(const vvr_frame_params& p) const {
        const OpacityKey key = opacity_key(p.labels, generation, has_labels, geo.min_spacing);
        if (!majorant_key || *majorant_key != key) return true;
        return key.mode != VVR_LABELS_OFF && (!mask_key || *mask_key != key);
    }
    // Whether `p`, a ray-marched frame with uniforms `hu` packed from it, needs a
    // lighting volume rebuilt.
    bool needs_lighting_rebuild(const vvr_frame_params& p, const vvr_uniforms& hu) const {
        if (p.path_tracing_enabled) return false;
        const OpacityKey key = opacity_key(p.labels, generation, has_labels, geo.min_spacing);
        if (p.lighting.ambient_occlusion && (!ao_key || *ao_key != light_key(key, hu, false))) return true;
        return p.lighting.shadows && (!shadow_key || *shadow_key != light_key(key, hu, true));
    }
    bool needs_prepare(const vvr_frame_params& p, const vvr_uniforms& hu) const {
        return needs_opacity_rebuild(p) || needs_lighting_rebuild(p, hu);
    }

    // The label masks and the majorant grid for the opacity key of `p`
    // (renderer.rs:1468-1495, 1671-1706), then for a ray-marched frame the
    // ambient-occlusion and shadow volumes (renderer.rs:1707-1749), each
    // rebuilt on `st` when its key changed. `du` is the device copy of the
    // uniforms `hu` packed from `p`. Timed when `st` is the renderer's stream.
    OpacityKey prepare(const vvr_frame_params& p, const vvr_uniforms& hu, const vvr_uniforms* du, hipStream_t st) {
        const OpacityKey opacity = opacity_key(p.labels, generation, has_labels, geo.min_spacing);
        const BrickGrid& g = grid;
        auto run = [&](int which, auto f) {
            if (st == stream) timed(which, f);
            else {
                f();
                check(hipGetLastError(), "kernel launch");
            }
        };
        if (opacity.mode != VVR_LABELS_OFF && (!mask_key || *mask_key != opacity)) {
            // Binary masks into B, then blur x (B -> A), y (A -> B), z (B -> A) (renderer.rs:1671-1694).
            // A, rewritten by the first blur, holds the soft grid's node densities meanwhile.
            const int radius = int(opacity.smoothing);
            run(VVR_KERNEL_MASK, [&] {
                launch_mask_binary(density, labels, soft_grid, opacity.margin, mask_b, mask,
                                   g.padded() * sizeof(uint32_t), g, du, st);
            });
            run(VVR_KERNEL_MASK, [&] { launch_mask_blur(mask_b, mask, 0, radius, g, st); });
            run(VVR_KERNEL_MASK, [&] { launch_mask_blur(mask, mask_b, 1, radius, g, st); });
            run(VVR_KERNEL_MASK, [&] { launch_mask_blur(mask_b, mask, 2, radius, g, st); });
            run(VVR_KERNEL_MASK,
                [&] { launch_mask_cells(mask, labels, g, int(opacity.mode), du, mask_cells, cell_labels, st); });
            mask_key = opacity;
            mask_builds++;
        }
        if (!majorant_key || *majorant_key != opacity) {
            // Cell records of the two-pass build, allocated on the first build that needs them.
            // A transfer-function change only reduces the kept records into bricks.
            const size_t nmb = size_t(g.bx) * g.by * g.bz;
            if (opacity.mode <= VVR_LABELS_TINT && !maj_cells) {
                maj_cells = dalloc<uint32_t>(8 * nmb, "majorant cells");
                maj_ranges = dalloc<uint16_t>(size_t(mdims[0]) * mdims[1] * mdims[2], "majorant brick ranges");
            }
            if (opacity.mode == VVR_LABELS_TINT && !maj_nz) maj_nz = dalloc<uint8_t>(nmb, "majorant mask cells");
            const MajorantCells cells{maj_cells, maj_nz, maj_ranges, !maj_cells_ready, maj_nz_build != mask_builds};
            run(VVR_KERNEL_MAJORANT, [&] {
                launch_majorant(density, opacity.mode != VVR_LABELS_OFF ? mask : nullptr, majorant, g, int(opacity.mode),
                                du, cells, st);
            });
            if (opacity.mode <= VVR_LABELS_TINT) maj_cells_ready = true;
            if (opacity.mode == VVR_LABELS_TINT) maj_nz_build = mask_builds;
            majorant_key = opacity;
        }
        if (p.path_tracing_enabled) return opacity;  // the path tracer computes its own lighting
        const size_t s_texels = size_t(shadow_dims(0)) * shadow_dims(1) * shadow_dims(2);
        const size_t a_texels = size_t(ao_dims(0)) * ao_dims(1) * ao_dims(2);
        if (p.lighting.ambient_occlusion) {
            const LightKey key = light_key(opacity, hu, false);
            if (!ao_key || *ao_key != key) {
                if (!ao) ao = dalloc<uint16_t>(a_texels, "ambient-occlusion volume");
                run(VVR_KERNEL_LIGHTING, [&] { launch_ao(scene(du), ao, st); });
                ao_key = key;
            }
        }
        if (p.lighting.shadows) {
            const LightKey key = light_key(opacity, hu, true);
            if (!shadow_key || *shadow_key != key) {
                if (!shadow) shadow = dalloc<uint2>(s_texels, "shadow volume");
                run(VVR_KERNEL_LIGHTING, [&] { launch_shadow_sweep(scene(du), shadow, hu, st); });
                shadow_key = key;
            }
        }
        return opacity;
    }

    DevScene scene(const vvr_uniforms* u) const {
        DevScene s;
        s.g = grid;
        s.density = density;
        s.density_f = density_f;
        s.labels = labels;
        s.mask = mask;
        s.mask_cells = mask_cells;
        s.cell_labels = cell_labels;
        s.mg = mask_cell_grid(grid);
        s.normals = normals;
        s.majorant = majorant;
        s.mx = mdims[0], s.my = mdims[1], s.mz = mdims[2];
        s.lut = lut;
        s.preint = preint;
        s.u = u;
        s.shadow = shadow;
        s.ao = ao;
        for (int a = 0; a < 3; a++) s.sdims[a] = shadow_dims(a), s.adims[a] = ao_dims(a);
        s.env_rad = env_rad, s.env_table = env_table, s.env_dif = env_dif;
        s.env_w = env_w, s.env_h = env_h;
        return s;
    }

    // Launch `f` bracketed by timing events when timing is on.
    template <class F> void timed(int which, F f) {
        hipEvent_t a = nullptr, b = nullptr;
        if (timing) {
            check(hipEventCreate(&a), "hipEventCreate");
            check(hipEventCreate(&b), "hipEventCreate");
            check(hipEventRecord(a, stream), "hipEventRecord");
        }
        f();
        check(hipGetLastError(), "kernel launch");
        if (timing) {
            check(hipEventRecord(b, stream), "hipEventRecord");
            pending.push_back({which, a, b});
        }
    }

    void collect_timing() {
        for (auto& p : pending) {
            check(hipEventSynchronize(p.b), "hipEventSynchronize");
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
            total_ms[p.which] += ms;
            launches[p.which] += 1;
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        pending.clear();
    }

    // Writes `u` into a uniform slot and returns its device copy.
    const vvr_uniforms* upload_uniforms(const vvr_uniforms& u, int slot) {
        if (slot_used[slot]) check(hipEventSynchronize(slot_done[slot]), "hipEventSynchronize");
        h_uniforms[slot] = u;
        check(hipMemcpyAsync(d_uniforms[slot], &h_uniforms[slot], sizeof(vvr_uniforms), hipMemcpyHostToDevice, stream),
              "uniform upload");
        check(hipEventRecord(slot_done[slot], stream), "hipEventRecord");
        slot_used[slot] = true;
        last_uniforms = u;
        return d_uniforms[slot];
    }

    void upload_tf(const vvr_tf& t) {
        tf = build_tf_tables(t);
        std::vector<float4> l(kLutSize), p(kLutSize);
        for (int i = 0; i < kLutSize; i++) {
            l[i] = make_float4(tf.lut_f16[4 * i], tf.lut_f16[4 * i + 1], tf.lut_f16[4 * i + 2], tf.lut_f16[4 * i + 3]);
            p[i] = make_float4(tf.preint[4 * i], tf.preint[4 * i + 1], tf.preint[4 * i + 2], tf.preint[4 * i + 3]);
        }
        check(hipMemcpy(lut, l.data(), kLutSize * sizeof(float4), hipMemcpyHostToDevice), "LUT upload");
        check(hipMemcpy(preint, p.data(), kLutSize * sizeof(float4), hipMemcpyHostToDevice), "pre-integration upload");
    }

    // Buffers of the ray marcher's passes at the frame size (launch_dvr_passes),
    // allocated by the first frame that uses them.
    uint8_t* dvr_mem = nullptr;
    DvrPasses dvr_passes{};
    int cus = 1;  // compute units of the device
    // Whether a frame with uniforms `u` is ray marched in passes (dvr_in_passes;
    // nor with VVR_DVR_PASSES=0, or buffers beyond 4 GiB).
    bool marches_in_passes(const vvr_uniforms& u) const {
        const char* env = std::getenv("VVR_DVR_PASSES");
        if ((env && env[0] == '0') || dvr_passes_bytes(width, height) > (size_t(4) << 30)) return false;
        return dvr_in_passes(u);
    }
    // Whether the march of such a frame runs as parallel segments
    // (dvr_segmented; VVR_DVR_PASSES=2: the march as one loop per ray).
    bool segments_march(const vvr_uniforms& u) const {
        const char* env = std::getenv("VVR_DVR_PASSES");
        return dvr_segmented(u) && !(env && env[0] == '2');
    }
    const DvrPasses& passes() {
        if (!dvr_mem) {
            dvr_mem = dalloc<uint8_t>(dvr_passes_bytes(width, height), "ray-marching passes");
            dvr_passes = carve_dvr_passes(dvr_mem, size_t(width) * height, dvr_tiles(width, height));
        }
        return dvr_passes;
    }

    void free_frame() {
        dfree(hdr), dfree(accum), dfree(albedo), dfree(normal), dfree(dn_scratch), dfree(output), dfree(bloom);
        dfree(dvr_mem);
        dvr_mem = nullptr;
        hdr = nullptr, accum = albedo = normal = dn_scratch = nullptr;
        output = nullptr, bloom = nullptr;
        denoised = false;
    }

    ~vvr_renderer() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        free_frame();
        wf.reset();
        dfree(density), dfree(density_f), dfree(normals), dfree(majorant), dfree(labels), dfree(mask), dfree(mask_b), dfree(mask_cells), dfree(cell_labels),
            dfree(soft);
        dfree(maj_cells), dfree(maj_nz), dfree(maj_ranges);
        dfree(shadow), dfree(ao), dfree(env_rad), dfree(env_table), dfree(env_dif);
        dfree(lut), dfree(preint), dfree(d_uniforms[0]), dfree(d_uniforms[1]), dfree(exposure);
        if (h_uniforms) (void)hipHostFree(h_uniforms);
        for (auto& e : slot_done)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace vvr {

// Runs an ABI call: exceptions become status codes and the thread's last error.
template <class F> int guarded(F f) {
    try {
        f();
        return VVR_OK;
    } catch (const Error& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return VVR_ENOMEM;
    } catch (...) {
        g_last_error = "unexpected internal error";
        return VVR_EDEVICE;
    }
}

}  // namespace vvr

inline void hold_renderer(vvr_renderer* r) {
    std::lock_guard<std::mutex> lk(r->hold_mu);
    r->holds++;
}

inline bool let_go_renderer(vvr_renderer* r) {
    std::lock_guard<std::mutex> lk(r->hold_mu);
    return --r->holds == 0 && r->orphaned;
}
