// The renderer object behind the C ABI: owns the device volumes and frame
// images, orders the passes of a frame and keeps the caches of the reference's
// Renderer (crates/render/src/renderer.rs:452-2126): the opacity-keyed mask and
// majorant volumes and the path-tracing accumulation.
#include "renderer_impl.hpp"

using namespace vvr;

namespace {

// Largest volume the 32-bit brick addressing covers, and a per-axis limit in the
// spirit of maxImageDimension3D (renderer.rs:561-567).
constexpr uint32_t kMaxDim = 8192;

// Server requests read a held renderer's volumes, tables and light map from the
// server's stream at any time: calls that change them are refused meanwhile.
void refuse_if_held(vvr_renderer* r, const char* what) {
    std::lock_guard<std::mutex> lk(r->hold_mu);
    if (r->holds > 0)
        throw Error{VVR_ESTATE, std::string(what) + " on a renderer that server requests still read (" +
                                    std::to_string(r->holds) + " queued or running)"};
}

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
    check(hipDeviceGetAttribute(&r->cus, hipDeviceAttributeMultiprocessorCount, r->device), "device attribute");
    r->cus = std::max(r->cus, 1);
    r->geo = make_geometry(*vol);
    r->grid = g;
    for (int a = 0; a < 3; a++) r->mdims[a] = int((vol->dims[a] + kBrick - 1) / kBrick);
    check(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : r->slot_done) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    check(hipHostMalloc((void**)&r->h_uniforms, 2 * sizeof(vvr_uniforms), hipHostMallocDefault), "hipHostMalloc");
    for (auto& d : r->d_uniforms) d = dalloc<vvr_uniforms>(1, "uniforms");
    r->lut = dalloc<float4>(kLutSize, "transfer function");
    // Auto exposure, adapted over frames: 0 = not yet metered (renderer.rs:791-818).
    r->exposure = dalloc<float>(1, "exposure");
    check(hipMemsetAsync(r->exposure, 0, sizeof(float), r->stream), "hipMemset");
    r->pt_chunk = env_u32("VVR_PT_CHUNK", 1);
    r->preint = dalloc<float4>(kLutSize, "pre-integration table");
    r->upload_tf(*tf);

    const size_t nvox = size_t(vol->dims[0]) * vol->dims[1] * vol->dims[2];
    r->density = dalloc<uint16_t>(g.padded(), "density");
    r->density_f = dalloc<float>(row_volume_texels(g.nx, g.ny, g.nz), "density values");
    r->normals = dalloc<uint32_t>(g.padded(), "normals");
    r->majorant = dalloc<float>(size_t(r->mdims[0]) * r->mdims[1] * r->mdims[2], "extinction bounds");
    int16_t* hu = dalloc<int16_t>(nvox, "HU staging");
    hipError_t e = hipMemcpy(hu, vol->hu, nvox * sizeof(int16_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_upload_density(hu, r->density, r->density_f, g, r->stream);
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
    refuse_if_held(r, "set_labels");
    r->use_device();
    check(hipStreamSynchronize(r->stream), "device wait");  // device_wait_idle (renderer.rs:1256)
    const BrickGrid& g = r->grid;
    if (!r->labels) r->labels = dalloc<uint8_t>(g.padded(), "labels");
    if (!r->mask) r->mask = dalloc<uint32_t>(g.padded(), "label masks");
    if (!r->mask_b) r->mask_b = dalloc<uint32_t>(g.padded(), "label masks (blur)");
    if (!r->mask_cells) r->mask_cells = dalloc<uint32_t>(mask_cell_grid(g).padded() / 32, "label mask cells");
    if (!r->cell_labels) r->cell_labels = dalloc<uint8_t>(mask_cell_grid(g).padded(), "label mask cell labels");
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

WfEngine& wavefront(vvr_renderer* r) {
    if (!r->wf) r->wf = std::make_unique<WfEngine>(r->device);
    return *r->wf;
}

void record(vvr_renderer* r, uint64_t frame, const vvr_frame_params* p);

// The autotuner: coordinate search over the step bounds, then the slots, each
// candidate timed by HIP events around whole frames on the renderer's stream.
void autotune(vvr_renderer* r, const vvr_frame_params* p, uint32_t trials, vvr_pt_shape* chosen) {
    if (!r || !p || !chosen) throw Error{VVR_EINVAL, "null argument"};
    if (!p->path_tracing_enabled) throw Error{VVR_EINVAL, "the autotuner times path-traced frames"};
    r->use_device();
    trials = std::max(trials, 1u);
    WfEngine& e = wavefront(r);
    hipEvent_t a = nullptr, b = nullptr;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    uint64_t frame = 1;
    auto frame_ms = [&]() {
        check(hipEventRecord(a, r->stream), "hipEventRecord");
        record(r, frame++, p);
        check(hipEventRecord(b, r->stream), "hipEventRecord");
        check(hipEventSynchronize(b), "autotune frame");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        return ms;
    };
    auto time_shape = [&](const vvr_pt_shape& s) {
        e.set_shape(s);
        float best = 1e30f;
        for (uint32_t t = 0; t < trials; t++) best = std::min(best, frame_ms());
        return best;
    };
    try {
        frame_ms();  // derived volumes, slot arrays
        // The current shape is the incumbent; a candidate replaces the best so far
        // only when it is 2 % faster, so that timing noise does not move the choice.
        vvr_pt_shape best = e.shape();
        float best_ms = time_shape(best);
        auto consider = [&](const vvr_pt_shape& s) {
            if (s.slots == best.slots && s.flight_steps == best.flight_steps && s.shadow_steps == best.shadow_steps &&
                s.tail_paths == best.tail_paths)
                return;
            const float ms = time_shape(s);
            if (ms < 0.98f * best_ms) best_ms = ms, best = s;
        };
        const uint32_t steps[][2] = {{8, 16}, {12, 24}, {16, 32}, {24, 32}, {24, 48}, {32, 64}};
        const vvr_pt_shape s0 = best;
        for (const auto& st : steps) consider(vvr_pt_shape{s0.slots, st[0], st[1], s0.tail_paths});
        const vvr_pt_shape s1 = best;
        for (uint32_t slots : {1u << 20, 1u << 21, 1u << 22, 1u << 23})
            consider(vvr_pt_shape{slots, s1.flight_steps, s1.shadow_steps, s1.tail_paths});
        const vvr_pt_shape s2 = best;
        for (uint32_t tail : {1u << 15, 1u << 16, 1u << 17, 1u << 18, UINT32_MAX})
            consider(vvr_pt_shape{s2.slots, s2.flight_steps, s2.shadow_steps, tail});
        e.set_shape(best);
        *chosen = best;
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    check(hipEventDestroy(a), "hipEventDestroy");
    check(hipEventDestroy(b), "hipEventDestroy");
}

void record(vvr_renderer* r, uint64_t frame, const vvr_frame_params* p) {
    if (!r || !p) throw Error{VVR_EINVAL, "null argument"};
    if (!r->hdr) throw Error{VVR_ESTATE, "resize() must be called before record()"};
    validate_frame(*p);
    refuse_if_held(r, "record");
    const bool pt = p->path_tracing_enabled != 0;
    if (pt && p->path_tracing.denoiser == VVR_DENOISE_LEARNED && !r->unet)
        throw Error{VVR_EINVAL, "no learned denoiser is attached (vvr_renderer_set_denoiser)"};
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
    pack_uniforms(r->geo, *p, r->tf.bucket_max, r->has_labels, r->accumulated, r->width, r->height, frame, &u,
                  r->env_peak());
    const vvr_uniforms* du = r->upload_uniforms(u, slot);
    r->prepare(*p, u, du, r->stream);
    const DevScene sc = r->scene(du);
    if (pt) {
        const uint32_t spf = std::max(p->path_tracing.samples_per_frame, 1u);
        const uint32_t first = r->accumulated;
        PtRequest q{};
        q.scene = sc;
        q.accum = r->accum, q.albedo = r->albedo, q.normal = r->normal, q.hdr = r->hdr;
        WfEngine::shape(q, r->width, r->height, first, spf, r->pt_chunk);  // throws before anything is queued
        wavefront(r);
        r->timed(VVR_KERNEL_PT, [&] {
            r->wf->run(&q, 1, r->stream, u.lighting[3] == 2u, labels_may_draw_ghosts(u), frame_reads_density(u));
        });
        r->accumulated += spf;
        // renderer.rs:1795-1797: the a-trous denoiser replaces the HDR image with
        // its filtered version of everything accumulated so far.
        r->denoised = p->path_tracing.denoiser != VVR_DENOISE_OFF;
        if (p->path_tracing.denoiser == VVR_DENOISE_ATROUS) {
            if (!r->dn_scratch) r->dn_scratch = dalloc<float4>(4 * size_t(r->width) * r->height, "denoiser images");
            r->timed(VVR_KERNEL_DENOISE, [&] {
                launch_denoise(r->accum, r->albedo, r->normal, r->dn_scratch, r->hdr, r->width, r->height,
                               float(r->accumulated), r->stream);
            });
        } else if (p->path_tracing.denoiser == VVR_DENOISE_LEARNED) {
            r->timed(VVR_KERNEL_DENOISE, [&] {
                unet_denoise_frame(r->unet, r->accum, r->albedo, r->normal, float(r->accumulated), r->width, r->height,
                                   r->hdr, r->stream);
            });
        }
    } else {
        r->denoised = false;
        if (r->marches_in_passes(u)) {
            const DvrPasses& passes = r->passes();
            r->timed(VVR_KERNEL_DVR,
                     [&] {
                         launch_dvr_passes(sc, passes, r->hdr, r->width, r->height, r->cus, r->segments_march(u),
                                           r->stream);
                     });
        } else {
            r->timed(VVR_KERNEL_DVR, [&] { launch_dvr(sc, r->hdr, r->width, r->height, r->stream); });
        }
    }
    // Camera finishing and tone mapping into the RGBA8 output (renderer.rs:1801-1811).
    r->timed(VVR_KERNEL_POST, [&] {
        launch_post(du, r->hdr, r->width, r->height, r->bloom, r->exposure, r->output, r->stream);
    });
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
        refuse_if_held(r, "set_transfer_function");
        r->use_device();
        check(hipStreamSynchronize(r->stream), "device wait");
        r->upload_tf(*tf);
        r->generation++;
    });
}

int vvr_renderer_set_environment_map(vvr_renderer* r, const vvr_env_map* map) {
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        if (map && (!map->rgb || map->width < 2 || map->height < 1 || map->width > 16384 || map->height > 8192))
            throw Error{VVR_EINVAL, "light map must be 2..16384 x 1..8192 RGB texels"};
        refuse_if_held(r, "set_environment_map");
        r->use_device();
        check(hipStreamSynchronize(r->stream), "device wait");
        dfree(r->env_rad), dfree(r->env_table), dfree(r->env_dif);
        r->env_rad = nullptr, r->env_table = nullptr, r->env_dif = nullptr;
        r->has_env = false;
        r->env_w = r->env_h = 0;
        r->acc_valid = false;  // renderer.rs:1299: accumulation restarts
        if (!map) return;
        const EnvTables t = prepare_environment(map->width, map->height, map->rgb);
        r->env_rad = dalloc<uint2>(size_t(t.width) * t.height, "light map");
        r->env_table = dalloc<float2>(size_t(t.width + 1) * t.height, "light map sampling");
        r->env_dif = dalloc<uint2>(32 * 16, "light map diffuse");
        check(hipMemcpy(r->env_rad, t.radiance.data(), t.radiance.size() * 2, hipMemcpyHostToDevice), "upload");
        check(hipMemcpy(r->env_table, t.table.data(), t.table.size() * 4, hipMemcpyHostToDevice), "upload");
        check(hipMemcpy(r->env_dif, t.diffuse.data(), t.diffuse.size() * 2, hipMemcpyHostToDevice), "upload");
        r->env_w = int(t.width), r->env_h = int(t.height);
        r->env_peak_u = t.peak_u;
        r->has_env = true;
    });
}

int vvr_renderer_set_denoiser(vvr_renderer* r, vvr_unet* net) {
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        refuse_if_held(r, "set_denoiser");
        r->use_device();
        check(hipStreamSynchronize(r->stream), "device wait");
        r->unet = net;
    });
}

int vvr_renderer_read_environment(vvr_renderer* r, float* peak_u, float* table, size_t table_floats) {
    return guarded([&] {
        if (!r || !peak_u) throw Error{VVR_EINVAL, "null argument"};
        if (!r->has_env) throw Error{VVR_ESTATE, "no light map is set"};
        *peak_u = r->env_peak_u;
        if (table) {
            const size_t n = 2 * size_t(r->env_w + 1) * r->env_h;
            if (table_floats != n) throw Error{VVR_EINVAL, "sampling table size mismatch"};
            r->use_device();
            check(hipMemcpy(table, r->env_table, n * 4, hipMemcpyDeviceToHost), "read-back");
        }
    });
}

int vvr_renderer_resize(vvr_renderer* r, uint32_t width, uint32_t height) {
    return guarded([&] {
        if (!r) throw Error{VVR_EINVAL, "null argument"};
        if (!width || !height || width > 16384 || height > 16384)
            throw Error{VVR_EINVAL, "image size must be 1..16384 per axis"};
        refuse_if_held(r, "resize");
        r->use_device();
        check(hipStreamSynchronize(r->stream), "device wait");
        r->free_frame();
        const size_t n = size_t(width) * height;
        r->hdr = dalloc<uint16_t>(4 * n, "hdr");
        r->accum = dalloc<float4>(n, "accumulation");
        r->albedo = dalloc<float4>(n, "albedo guide");
        r->normal = dalloc<float4>(n, "normal guide");
        r->output = dalloc<uint32_t>(n, "output");
        r->bloom = dalloc<uint16_t>(bloom_bytes(width, height) / 2, "bloom images");
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

int vvr_renderer_get_pt_shape(vvr_renderer* r, vvr_pt_shape* out) {
    return guarded([&] {
        if (!r || !out) throw Error{VVR_EINVAL, "null argument"};
        r->use_device();
        *out = wavefront(r).shape();
    });
}

int vvr_renderer_set_pt_shape(vvr_renderer* r, const vvr_pt_shape* shape) {
    return guarded([&] {
        if (!r || !shape) throw Error{VVR_EINVAL, "null argument"};
        r->use_device();
        wavefront(r).set_shape(*shape);
    });
}

int vvr_renderer_autotune(vvr_renderer* r, const vvr_frame_params* params, uint32_t trials, vvr_pt_shape* chosen) {
    return guarded([&] { autotune(r, params, trials, chosen); });
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

void vvr_renderer_destroy(vvr_renderer* r) {
    if (!r) return;
    {
        std::lock_guard<std::mutex> lk(r->hold_mu);
        if (r->holds > 0) {  // server requests still read it: the last one deletes it
            r->orphaned = true;
            return;
        }
    }
    delete r;
}

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

int vvr_renderer_read_rgba(vvr_renderer* r, uint8_t* rgba) {
    return guarded([&] {
        if (!r || !rgba) throw Error{VVR_EINVAL, "null argument"};
        if (!r->output) throw Error{VVR_ESTATE, "resize() must be called first"};
        r->use_device();
        check(hipMemcpyAsync(rgba, r->output, size_t(r->width) * r->height * 4, hipMemcpyDeviceToHost, r->stream),
              "output read-back");
        check(hipStreamSynchronize(r->stream), "output read-back");
    });
}

int vvr_renderer_read_exposure(vvr_renderer* r, float* exposure) {
    return guarded([&] {
        if (!r || !exposure) throw Error{VVR_EINVAL, "null argument"};
        r->use_device();
        check(hipMemcpyAsync(exposure, r->exposure, sizeof(float), hipMemcpyDeviceToHost, r->stream), "read-back");
        check(hipStreamSynchronize(r->stream), "read-back");
    });
}

int vvr_render_to_rgba(vvr_renderer* r, const vvr_frame_params* params, uint32_t width, uint32_t height,
                       uint32_t frames, uint8_t* rgba) {
    int st = vvr_renderer_resize(r, width, height);
    for (uint32_t f = 1; st == VVR_OK && f <= std::max(frames, 1u); f++) st = vvr_renderer_record(r, f, params);
    return st == VVR_OK ? vvr_renderer_read_rgba(r, rgba) : st;
}

int vvr_renderer_read_denoised(vvr_renderer* r, float* rgba) {
    {
        const int st = guarded([&] {
            if (!r || !rgba) throw Error{VVR_EINVAL, "null argument"};
            if (!r->denoised) throw Error{VVR_ESTATE, "the last frame was not denoised"};
        });
        if (st != VVR_OK) return st;
    }
    return vvr_renderer_read_hdr(r, rgba);
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
            case VVR_VOL_SHADOW:
            case VVR_VOL_AO: {  // linear already
                const bool sh = which == VVR_VOL_SHADOW;
                const void* lv = sh ? static_cast<const void*>(r->shadow) : static_cast<const void*>(r->ao);
                if (!lv) throw Error{VVR_ESTATE, "that lighting volume has not been built (record a ray-marched frame "
                                                 "with it on first)"};
                size_t n = 1;
                for (int a = 0; a < 3; a++) n *= size_t(sh ? r->shadow_dims(a) : r->ao_dims(a));
                if (bytes != n * (sh ? 8 : 2)) throw Error{VVR_EINVAL, "lighting read-back size mismatch"};
                check(hipMemcpyAsync(out, lv, bytes, hipMemcpyDeviceToHost, r->stream), "read-back");
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
