// The renderer object behind the C ABI: owns the device volumes and frame
// images, orders the passes of a frame and keeps the caches of the reference's
// Renderer (crates/render/src/renderer.rs:452-2126): the opacity-keyed mask and
// majorant volumes and the path-tracing accumulation.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "host.hpp"
#include "launch.hpp"

namespace vvr {
namespace {

thread_local std::string g_last_error;

void check(hipError_t e, const char* what) {
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

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

uint32_t env_u32(const char* name, uint32_t fallback) {
    const char* v = std::getenv(name);
    if (!v || !*v) return fallback;
    const unsigned long x = std::strtoul(v, nullptr, 10);
    return x ? uint32_t(std::min<unsigned long>(x, 1ul << 30)) : fallback;
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
    uint32_t max_slots = 1u << 22;

    explicit WfEngine(int dev) : device(dev) {
        int cus = 0;
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "device attribute");
        grid = wf_grid(std::max(cus, 1));
        // 4M slots (about 1 GiB): enough paths in flight that the long random
        // walks of a frame's last iterations still fill the device.
        max_slots = env_u32("VVR_PT_SLOTS", 1u << 22);
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
        constexpr int kF4 = 15, kQueues = 5;
        const size_t f4_bytes = size_t(n) * sizeof(float4), q_bytes = (size_t(n) * 4 + 255) / 256 * 256;
        slot_mem = dalloc<uint8_t>(kF4 * f4_bytes + kQueues * q_bytes + 256, "path slots");
        uint8_t* p = slot_mem;
        float4** f4s[kF4] = {&S.p_origin, &S.p_dir, &S.p_thr, &S.p_spec, &S.p_rad, &S.e_idx, &S.e_s0, &S.e_s1,
                             &S.e_s2,     &S.e_n,   &S.e_light, &S.e_lt, &S.c0,   &S.c1,   &S.c2};
        for (auto f : f4s) *f = reinterpret_cast<float4*>(p), p += f4_bytes;
        uint32_t** qs[kQueues] = {&Q.ff[0], &Q.ff[1], &Q.ev, &Q.end[0], &Q.end[1]};
        for (auto q : qs) *q = reinterpret_cast<uint32_t*>(p), p += q_bytes;
        Q.counts = reinterpret_cast<uint32_t*>(p);
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

    // Items of `samples` paths per pixel at `chunk` paths per item.
    static void shape(PtRequest& q, uint32_t width, uint32_t height, uint32_t first, uint32_t samples,
                      uint32_t chunk) {
        q.width = width, q.height = height;
        q.tiles_x = (width + 7) / 8;
        q.n_px = q.tiles_x * ((height + 7) / 8) * 64u;
        q.first_path = first, q.samples = samples;
        q.chunk = std::min(std::max(chunk, 1u), samples);
        q.n_chunks = (samples + q.chunk - 1) / q.chunk;
        q.n_items = q.n_chunks * q.n_px;
    }

    // Renders the requests (scene, images and shape set) on `st`. Returns once
    // every path is traced or queued to be; the resolves are queued last.
    void run(const PtRequest* reqs, uint32_t n, hipStream_t st) {
        if (n == 0) return;
        size_t partial_f4 = 0;
        uint64_t items = 0;
        for (uint32_t i = 0; i < n; i++) {
            partial_f4 += size_t(reqs[i].n_chunks) * 3u * reqs[i].n_px;
            items += reqs[i].n_items;
        }
        if (items >= (1ull << 31)) throw Error{VVR_ELIMIT, "too many paths in one frame (2^31 work items)"};
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
            for (uint32_t j = 0; j < kBatch; j++, iter++) wf_iteration(d_reqs, n, base, S, Q, iter & 1u, grid, st);
            check(hipGetLastError(), "wavefront iteration");
            check(hipMemcpyAsync(&h_live[k], Q.counts + kWfLIVE, 4, hipMemcpyDeviceToHost, st), "live count");
            check(hipEventRecord(live_ev[k], st), "hipEventRecord");
        };
        launch_batch(0);
        launch_batch(1);
        for (int k = 0;; k ^= 1) {
            check(hipEventSynchronize(live_ev[k]), "wavefront progress");
            if (h_live[k] == 0) break;
            launch_batch(k);
        }
        for (uint32_t i = 0; i < n; i++) wf_resolve(h_reqs[i], st);
        check(hipGetLastError(), "wavefront resolve");
        // The staging above is read by copies that have completed (the loop
        // waited on events recorded after them).
    }
};

}  // namespace
}  // namespace vvr

using namespace vvr;

struct vvr_ctx {
    int device;
};

struct vvr_renderer {
    int device = 0;
    hipStream_t stream = nullptr;
    Geometry geo{};
    BrickGrid grid{};
    int mdims[3] = {0, 0, 0};

    uint16_t* density = nullptr;
    uint32_t* normals = nullptr;
    float* majorant = nullptr;
    uint8_t* labels = nullptr;
    uint32_t* mask = nullptr;    // image A: what rendering samples
    uint32_t* mask_b = nullptr;  // image B: blur partner
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

    uint32_t width = 0, height = 0;
    uint16_t* hdr = nullptr;
    float4* accum = nullptr;
    float4* albedo = nullptr;
    float4* normal = nullptr;
    // Path-tracing engine: the wavefront engine (default), or with
    // VVR_PT_KERNEL=persistent / simple the persistent megakernel or one lane per pixel.
    enum { kWavefront, kPersistent, kSimple } pt_kernel = kWavefront;
    std::unique_ptr<WfEngine> wf;
    uint32_t pt_chunk = 1;       // paths per wavefront work item (VVR_PT_CHUNK)
    uint32_t* queue = nullptr;   // work queue of the persistent path tracer
    uint32_t pt_waves = 0;       // its grid: waves resident on the device
    uint32_t accumulated = 0;
    bool acc_valid = false;
    vvr_frame_params acc_params{};

    bool timing = false;
    struct Pending {
        int which;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    double total_ms[VVR_KERNEL_COUNT] = {0, 0, 0, 0, 0};
    uint64_t launches[VVR_KERNEL_COUNT] = {0, 0, 0, 0, 0};

    void use_device() const { check(hipSetDevice(device), "hipSetDevice"); }

    DevScene scene(const vvr_uniforms* u) const {
        DevScene s;
        s.g = grid;
        s.density = density;
        s.labels = labels;
        s.mask = mask;
        s.normals = normals;
        s.majorant = majorant;
        s.mx = mdims[0], s.my = mdims[1], s.mz = mdims[2];
        s.lut = lut;
        s.preint = preint;
        s.u = u;
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

    void free_frame() {
        dfree(hdr), dfree(accum), dfree(albedo), dfree(normal);
        hdr = nullptr, accum = albedo = normal = nullptr;
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
        dfree(density), dfree(normals), dfree(majorant), dfree(labels), dfree(mask), dfree(mask_b), dfree(soft);
        dfree(lut), dfree(preint), dfree(d_uniforms[0]), dfree(d_uniforms[1]), dfree(queue);
        if (h_uniforms) (void)hipHostFree(h_uniforms);
        for (auto& e : slot_done)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

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

// Largest volume the 32-bit brick addressing covers, and a per-axis limit in the
// spirit of maxImageDimension3D (renderer.rs:561-567).
constexpr uint32_t kMaxDim = 8192;

void new_renderer(vvr_ctx* ctx, const vvr_volume* vol, const vvr_tf* tf, vvr_renderer** out) {
    if (!ctx || !vol || !out) throw Error{VVR_EINVAL, "null argument"};
    if (!vol->hu) throw Error{VVR_EINVAL, "volume has no voxel data"};
    for (int a = 0; a < 3; a++)
        if (vol->dims[a] == 0) throw Error{VVR_EINVAL, "volume has a zero dimension"};
    validate_tf(tf);
    for (int a = 0; a < 3; a++)
        if (vol->dims[a] > kMaxDim)
            throw Error{VVR_ELIMIT, "volume [" + std::to_string(vol->dims[0]) + ", " + std::to_string(vol->dims[1]) +
                                        ", " + std::to_string(vol->dims[2]) + "] exceeds the 3D volume limit of " +
                                        std::to_string(kMaxDim)};
    BrickGrid g = make_brick_grid(int(vol->dims[0]), int(vol->dims[1]), int(vol->dims[2]));
    if (g.padded() * 4 >= (size_t(1) << 32))
        throw Error{VVR_ELIMIT, "volume exceeds the 2^30 bricked voxels of 32-bit addressing"};

    auto r = std::make_unique<vvr_renderer>();
    r->device = ctx->device;
    r->use_device();
    r->geo = make_geometry(*vol);
    r->grid = g;
    for (int a = 0; a < 3; a++) r->mdims[a] = int((vol->dims[a] + kBrick - 1) / kBrick);
    check(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : r->slot_done) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    check(hipHostMalloc((void**)&r->h_uniforms, 2 * sizeof(vvr_uniforms), hipHostMallocDefault), "hipHostMalloc");
    for (auto& d : r->d_uniforms) d = dalloc<vvr_uniforms>(1, "uniforms");
    r->lut = dalloc<float4>(kLutSize, "transfer function");
    r->queue = dalloc<uint32_t>(1, "work queue");
    {
        int cus = 0;
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, r->device), "device attribute");
        r->pt_waves = uint32_t(std::max(cus, 1) * pt_persistent_waves_per_cu());
        const char* k = std::getenv("VVR_PT_KERNEL");
        if (k && std::strcmp(k, "simple") == 0) r->pt_kernel = vvr_renderer::kSimple;
        if (k && std::strcmp(k, "persistent") == 0) r->pt_kernel = vvr_renderer::kPersistent;
        r->pt_chunk = env_u32("VVR_PT_CHUNK", 1);
    }
    r->preint = dalloc<float4>(kLutSize, "pre-integration table");
    r->upload_tf(*tf);

    const size_t nvox = size_t(vol->dims[0]) * vol->dims[1] * vol->dims[2];
    r->density = dalloc<uint16_t>(g.padded(), "density");
    r->normals = dalloc<uint32_t>(g.padded(), "normals");
    r->majorant = dalloc<float>(size_t(r->mdims[0]) * r->mdims[1] * r->mdims[2], "extinction bounds");
    int16_t* hu = dalloc<int16_t>(nvox, "HU staging");
    hipError_t e = hipMemcpy(hu, vol->hu, nvox * sizeof(int16_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_upload_density(hu, r->density, g, r->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    dfree(hu);
    check(e, "density upload");
    // Normals (renderer.rs:910-942) with Uniforms::zeroed_with(world_to_index, dims).
    vvr_uniforms u{};
    std::memcpy(u.world_to_index, r->geo.world_to_index, sizeof(u.world_to_index));
    for (int a = 0; a < 3; a++) u.dims[a] = float(vol->dims[a]);
    const vvr_uniforms* du = r->upload_uniforms(u, 0);
    r->timed(VVR_KERNEL_NORMALS, [&] { launch_normals(r->density, r->normals, g, du, r->stream); });
    check(hipStreamSynchronize(r->stream), "normals");
    *out = r.release();
}

void set_labels(vvr_renderer* r, const uint8_t* labels, size_t n, const vvr_soft_grid* soft) {
    if (!r || !labels) throw Error{VVR_EINVAL, "null argument"};
    const uint32_t* d = r->geo.dims;
    const size_t nvox = size_t(d[0]) * d[1] * d[2];
    if (n != nvox)
        throw Error{VVR_EINVAL, "label volume has " + std::to_string(n) + " voxels, expected [" + std::to_string(d[0]) +
                                    ", " + std::to_string(d[1]) + ", " + std::to_string(d[2]) + "]"};
    if (soft) {
        if (!soft->data || !soft->dims[0] || !soft->dims[1] || !soft->dims[2])
            throw Error{VVR_EINVAL, "soft label grid size mismatch"};
        for (int a = 0; a < 3; a++)
            if (soft->src_axis[a] > 2) throw Error{VVR_EINVAL, "soft label grid axis out of range"};
    }
    r->use_device();
    check(hipStreamSynchronize(r->stream), "device wait");  // device_wait_idle (renderer.rs:1256)
    const BrickGrid& g = r->grid;
    if (!r->labels) r->labels = dalloc<uint8_t>(g.padded(), "labels");
    if (!r->mask) r->mask = dalloc<uint32_t>(g.padded(), "label masks");
    if (!r->mask_b) r->mask_b = dalloc<uint32_t>(g.padded(), "label masks (blur)");
    uint8_t* staging = dalloc<uint8_t>(nvox, "label staging");
    hipError_t e = hipMemcpy(staging, labels, nvox, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_upload_labels(staging, r->labels, g, r->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    dfree(staging);
    check(e, "label upload");
    dfree(r->soft);
    r->soft = nullptr;
    r->soft_grid = SoftGrid{};
    if (soft) {
        const size_t nodes = size_t(soft->dims[0]) * soft->dims[1] * soft->dims[2];
        r->soft = dalloc<uint32_t>(nodes, "soft labels");
        check(hipMemcpy(r->soft, soft->data, nodes * 4, hipMemcpyHostToDevice), "soft label upload");
        r->soft_grid.nodes = r->soft;
        for (int a = 0; a < 3; a++) {
            r->soft_grid.dims[a] = int(soft->dims[a]);
            r->soft_grid.push[a] = int(soft->src_axis[a]);
        }
        int flips = 0;
        for (int a = 0; a < 3; a++) flips |= (soft->flip[a] ? 1 : 0) << a;
        r->soft_grid.push[3] = flips | 8;
    } else {
        r->soft_grid.dims[0] = r->soft_grid.dims[1] = r->soft_grid.dims[2] = 1;
    }
    r->has_labels = true;
    r->generation++;
}

void record(vvr_renderer* r, uint64_t frame, const vvr_frame_params* p) {
    if (!r || !p) throw Error{VVR_EINVAL, "null argument"};
    if (!r->hdr) throw Error{VVR_ESTATE, "resize() must be called before record()"};
    const bool pt = p->path_tracing_enabled != 0;
    if (!pt && (p->lighting.shadows || p->lighting.ambient_occlusion))
        throw Error{VVR_EINVAL, "shadow and ambient-occlusion volumes are not available: set lighting.shadows and "
                                "lighting.ambient_occlusion to 0 for ray marching"};
    if (pt && p->path_tracing.denoiser != VVR_DENOISE_OFF)
        throw Error{VVR_EINVAL, "no denoiser is available: set path_tracing.denoiser to VVR_DENOISE_OFF"};
    if (!(p->step_mm > 0.0f)) throw Error{VVR_EINVAL, "step_mm must be positive"};
    if (p->labels.mode > VVR_LABELS_HYBRID) throw Error{VVR_EINVAL, "unknown label mode"};
    r->use_device();

    // Accumulation restarts when anything but the denoiser changes (renderer.rs:1535-1544).
    vvr_frame_params key = *p;
    key.path_tracing.denoiser = VVR_DENOISE_OFF;
    if (pt && (!r->acc_valid || std::memcmp(&key, &r->acc_params, sizeof(key)) != 0)) {
        r->accumulated = 0;
        r->acc_params = key;
        r->acc_valid = true;
    }
    const int slot = int(frame % 2);
    vvr_uniforms u;
    pack_uniforms(r->geo, *p, r->tf.bucket_max, r->has_labels, r->accumulated, r->width, r->height, frame, &u);
    const vvr_uniforms* du = r->upload_uniforms(u, slot);

    const OpacityKey opacity = opacity_key(p->labels, r->generation, r->has_labels, r->geo.min_spacing);
    const BrickGrid& g = r->grid;
    if (opacity.mode != VVR_LABELS_OFF && (!r->mask_key || *r->mask_key != opacity)) {
        // Binary masks into B, then blur x (B -> A), y (A -> B), z (B -> A) (renderer.rs:1671-1694).
        const int radius = int(opacity.smoothing);
        r->timed(VVR_KERNEL_MASK, [&] {
            launch_mask_binary(r->density, r->labels, r->soft_grid, opacity.margin, r->mask_b, g, du, r->stream);
        });
        r->timed(VVR_KERNEL_MASK, [&] { launch_mask_blur(r->mask_b, r->mask, 0, radius, g, r->stream); });
        r->timed(VVR_KERNEL_MASK, [&] { launch_mask_blur(r->mask, r->mask_b, 1, radius, g, r->stream); });
        r->timed(VVR_KERNEL_MASK, [&] { launch_mask_blur(r->mask_b, r->mask, 2, radius, g, r->stream); });
        r->mask_key = opacity;
    }
    if (!r->majorant_key || *r->majorant_key != opacity) {
        r->timed(VVR_KERNEL_MAJORANT, [&] {
            launch_majorant(r->density, opacity.mode != VVR_LABELS_OFF ? r->mask : nullptr, r->majorant, g,
                            int(opacity.mode), du, r->stream);
        });
        r->majorant_key = opacity;
    }
    const DevScene sc = r->scene(du);
    if (pt) {
        const uint32_t spf = std::max(p->path_tracing.samples_per_frame, 1u);
        const uint32_t first = r->accumulated;
        if (r->pt_kernel == vvr_renderer::kWavefront) {
            if (!r->wf) r->wf = std::make_unique<WfEngine>(r->device);
            PtRequest q{};
            q.scene = sc;
            q.accum = r->accum, q.albedo = r->albedo, q.normal = r->normal, q.hdr = r->hdr;
            WfEngine::shape(q, r->width, r->height, first, spf, r->pt_chunk);
            r->timed(VVR_KERNEL_PT, [&] { r->wf->run(&q, 1, r->stream); });
        } else if (r->pt_kernel == vvr_renderer::kSimple) {
            r->timed(VVR_KERNEL_PT, [&] {
                launch_pt(sc, r->accum, r->albedo, r->normal, r->hdr, r->width, r->height, first, spf, r->stream);
            });
        } else {
            check(hipMemsetAsync(r->queue, 0, sizeof(uint32_t), r->stream), "queue reset");
            r->timed(VVR_KERNEL_PT, [&] {
                launch_pt_persistent(sc, r->accum, r->albedo, r->normal, r->hdr, r->width, r->height, first, spf,
                                     r->queue, r->pt_waves, r->stream);
            });
        }
        r->accumulated += spf;
    } else {
        r->timed(VVR_KERNEL_DVR, [&] { launch_dvr(sc, r->hdr, r->width, r->height, r->stream); });
    }
}

}  // namespace

extern "C" {

const char* vvr_last_error(void) { return g_last_error.c_str(); }
const char* vvr_version(void) { return "vvr 0.1.0 (gfx950)"; }

int vvr_ctx_create(int device, vvr_ctx** out) {
    return guarded([&] {
        if (!out) throw Error{VVR_EINVAL, "null argument"};
        int n = 0;
        check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n)
            throw Error{VVR_EINVAL, "HIP device " + std::to_string(device) + " not present (" + std::to_string(n) +
                                        " devices)"};
        check(hipSetDevice(device), "hipSetDevice");
        *out = new vvr_ctx{device};
    });
}

void vvr_ctx_destroy(vvr_ctx* ctx) { delete ctx; }

int vvr_renderer_new(vvr_ctx* ctx, const vvr_volume* volume, const vvr_tf* tf, vvr_renderer** out) {
    return guarded([&] { new_renderer(ctx, volume, tf, out); });
}

int vvr_renderer_set_labels(vvr_renderer* r, const uint8_t* labels, size_t n, const vvr_soft_grid* soft) {
    return guarded([&] { set_labels(r, labels, n, soft); });
}

int vvr_renderer_set_transfer_function(vvr_renderer* r, const vvr_tf* tf) {
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        validate_tf(tf);
        r->use_device();
        check(hipStreamSynchronize(r->stream), "device wait");
        r->upload_tf(*tf);
        r->generation++;
    });
}

int vvr_renderer_resize(vvr_renderer* r, uint32_t width, uint32_t height) {
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        if (!width || !height || width > 16384 || height > 16384)
            throw Error{VVR_EINVAL, "image size must be 1..16384 per axis"};
        r->use_device();
        check(hipStreamSynchronize(r->stream), "device wait");
        r->free_frame();
        const size_t n = size_t(width) * height;
        r->hdr = dalloc<uint16_t>(4 * n, "hdr");
        r->accum = dalloc<float4>(n, "accumulation");
        r->albedo = dalloc<float4>(n, "albedo guide");
        r->normal = dalloc<float4>(n, "normal guide");
        check(hipMemsetAsync(r->hdr, 0, 8 * n, r->stream), "hipMemset");
        r->width = width;
        r->height = height;
        r->acc_valid = false;
        r->accumulated = 0;
    });
}

int vvr_renderer_record(vvr_renderer* r, uint64_t frame, const vvr_frame_params* params) {
    return guarded([&] { record(r, frame, params); });
}

int vvr_renderer_synchronize(vvr_renderer* r) {
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        r->use_device();
        check(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
    });
}

uint32_t vvr_renderer_accumulated_samples(const vvr_renderer* r) { return r ? r->accumulated : 0u; }

int vvr_renderer_default_camera(const vvr_renderer* r, vvr_camera* out) {
    return guarded([&] {
        if (!r || !out) throw Error{VVR_EINVAL, "null argument"};
        vvr_camera_framing(r->geo.center, r->geo.diagonal, out);
    });
}

int vvr_renderer_geometry(const vvr_renderer* r, float lo[3], float hi[3], float* min_spacing, float center[3],
                          float* diagonal) {
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        for (int a = 0; a < 3; a++) {
            if (lo) lo[a] = r->geo.bounds_lo[a];
            if (hi) hi[a] = r->geo.bounds_hi[a];
            if (center) center[a] = r->geo.center[a];
        }
        if (min_spacing) *min_spacing = r->geo.min_spacing;
        if (diagonal) *diagonal = r->geo.diagonal;
    });
}

void vvr_renderer_destroy(vvr_renderer* r) { delete r; }

int vvr_renderer_read_hdr(vvr_renderer* r, float* rgba) {
    return guarded([&] {
        if (!r || !rgba) throw Error{VVR_EINVAL, "null argument"};
        if (!r->hdr) throw Error{VVR_ESTATE, "resize() must be called first"};
        r->use_device();
        const size_t n = size_t(r->width) * r->height * 4;
        std::vector<uint16_t> h(n);
        check(hipMemcpyAsync(h.data(), r->hdr, n * 2, hipMemcpyDeviceToHost, r->stream), "hdr read-back");
        check(hipStreamSynchronize(r->stream), "hdr read-back");
        for (size_t i = 0; i < n; i++) rgba[i] = f16_bits_to_f32(h[i]);
    });
}

int vvr_renderer_read_accum(vvr_renderer* r, float* accum, float* albedo, float* normal) {
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        if (!r->accum) throw Error{VVR_ESTATE, "resize() must be called first"};
        r->use_device();
        const size_t bytes = size_t(r->width) * r->height * sizeof(float4);
        if (accum) check(hipMemcpyAsync(accum, r->accum, bytes, hipMemcpyDeviceToHost, r->stream), "read-back");
        if (albedo) check(hipMemcpyAsync(albedo, r->albedo, bytes, hipMemcpyDeviceToHost, r->stream), "read-back");
        if (normal) check(hipMemcpyAsync(normal, r->normal, bytes, hipMemcpyDeviceToHost, r->stream), "read-back");
        check(hipStreamSynchronize(r->stream), "read-back");
    });
}

int vvr_renderer_read_denoised(vvr_renderer* r, float* rgba) {
    (void)rgba;
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        throw Error{VVR_ESTATE, "no denoiser is available"};
    });
}

int vvr_renderer_read_volume(vvr_renderer* r, int which, void* out, size_t bytes) {
    return guarded([&] {
        if (!r || !out) throw Error{VVR_EINVAL, "null argument"};
        r->use_device();
        const size_t nvox = size_t(r->geo.dims[0]) * r->geo.dims[1] * r->geo.dims[2];
        const void* src = nullptr;
        int elem = 0;
        switch (which) {
            case VVR_VOL_DENSITY: src = r->density, elem = 2; break;
            case VVR_VOL_LABELS: src = r->labels, elem = 1; break;
            case VVR_VOL_NORMALS: src = r->normals, elem = 4; break;
            case VVR_VOL_MASK: src = r->mask, elem = 4; break;
            case VVR_VOL_MAJORANT: {
                const size_t nb = size_t(r->mdims[0]) * r->mdims[1] * r->mdims[2];
                if (bytes != nb * 4) throw Error{VVR_EINVAL, "majorant read-back size mismatch"};
                check(hipMemcpyAsync(out, r->majorant, bytes, hipMemcpyDeviceToHost, r->stream), "read-back");
                check(hipStreamSynchronize(r->stream), "read-back");
                return;
            }
            default: throw Error{VVR_EINVAL, "unknown volume"};
        }
        if (!src) throw Error{VVR_ESTATE, "that volume does not exist yet (set_labels first)"};
        if (bytes != nvox * size_t(elem)) throw Error{VVR_EINVAL, "read-back size mismatch"};
        void* tmp = dalloc<uint8_t>(nvox * elem, "read-back staging");
        launch_debrick(src, tmp, r->grid, elem, r->stream);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, bytes, hipMemcpyDeviceToHost, r->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        dfree(tmp);
        check(e, "read-back");
    });
}

int vvr_renderer_read_majorant(vvr_renderer* r, float* out, size_t n) {
    return vvr_renderer_read_volume(r, VVR_VOL_MAJORANT, out, n * sizeof(float));
}

int vvr_renderer_read_uniforms(vvr_renderer* r, void* out, size_t bytes) {
    return guarded([&] {
        if (!r || !out) throw Error{VVR_EINVAL, "null argument"};
        if (bytes != sizeof(vvr_uniforms)) throw Error{VVR_EINVAL, "uniform block is 21200 bytes"};
        std::memcpy(out, &r->last_uniforms, sizeof(vvr_uniforms));
    });
}

int vvr_renderer_set_timing(vvr_renderer* r, int enabled) {
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        r->use_device();
        r->collect_timing();
        r->timing = enabled != 0;
        for (int i = 0; i < VVR_KERNEL_COUNT; i++) r->total_ms[i] = 0.0, r->launches[i] = 0;
    });
}

int vvr_renderer_kernel_time(vvr_renderer* r, int which, double* total_ms, uint64_t* launches) {
    return guarded([&] {
        if (!r || which < 0 || which >= VVR_KERNEL_COUNT) throw Error{VVR_EINVAL, "invalid argument"};
        r->use_device();
        r->collect_timing();
        if (total_ms) *total_ms = r->total_ms[which];
        if (launches) *launches = r->launches[which];
    });
}

}  // extern "C"
