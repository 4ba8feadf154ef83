// Host-side launchers of the HIP kernels (kernels_volume.hip, kernels_render.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include "device/types.hpp"

namespace vvr {

// HU (linear, int16) -> unorm16 micro-bricks (transfer.rs:121-124, renderer.rs:577).
// HU -> the unorm16 density and its float values (c / 65535), bricked.
void launch_upload_density(const int16_t* hu, uint16_t* out, float* out_f, const BrickGrid& g, hipStream_t st);
// uint8 labels (linear) -> micro-bricks.
void launch_upload_labels(const uint8_t* labels, uint8_t* out, const BrickGrid& g, hipStream_t st);
// Micro-bricks -> linear, element size 1, 2 or 4 bytes.
void launch_debrick(const void* in, void* out, const BrickGrid& g, int elem_bytes, hipStream_t st);

// normals.comp
void launch_normals(const uint16_t* density, uint32_t* normals, const BrickGrid& g, const vvr_uniforms* u,
                    hipStream_t st);

// mask.comp, pass 0 (binary / soft masks) into `dst`.
struct SoftGrid {
    const uint32_t* nodes;  // RGBA8 per node (best, runner-up, margin x 16, overridden), grid axis 0 fastest
    int dims[3];
    int push[4];            // source axes, flip bits | 8 when present
};
// `scratch` (scratch_bytes, not `dst`) may be overwritten: it holds the soft
// grid's node densities when they fit (4 bytes per node).
void launch_mask_binary(const uint16_t* density, const uint8_t* labels, SoftGrid soft, int margin, uint32_t* dst,
                        void* scratch, size_t scratch_bytes, const BrickGrid& g, const vvr_uniforms* u,
                        hipStream_t st);
// mask.comp, passes 1-3: binomial blur of radius r along `axis`.
void launch_mask_blur(const uint32_t* src, uint32_t* dst, int axis, int radius, const BrickGrid& g, hipStream_t st);
// The cell grid of DevScene::mask_cells (n + 1 cells per axis), its bits
// (mask_cell_grid(g).padded() / 32 words) and label bytes (padded() bytes) for
// label mode `mode` and the label flags of `u`.
inline BrickGrid mask_cell_grid(const BrickGrid& g) { return make_brick_grid(g.nx + 1, g.ny + 1, g.nz + 1); }
void launch_mask_cells(const uint32_t* mask, const uint8_t* labels, const BrickGrid& g, int mode,
                       const vvr_uniforms* u, uint32_t* bits, uint8_t* cell_labels, hipStream_t st);

// What the majorant build of modes off and tint keeps per volume.
struct MajorantCells {
    uint32_t* cells;   // per micro-brick, its 2^3 cells' density min | (65535 - max) << 16
    uint8_t* nz;       // tint: per micro-brick, a bit per cell set when some mask value there is non-zero
    uint16_t* ranges;  // per brick, the transfer-function buckets first | last << 8 of its window's density
    bool density;      // rebuild cells and ranges (else they hold the density's: it never changes)
    bool mask;         // tint: rebuild nz (else it holds the current mask's bits)
};

// majorant.comp: `mask` may be null when mode is 0. Mode off with the density's
// ranges kept only maps each brick's range through the transfer function.
void launch_majorant(const uint16_t* density, const uint32_t* mask, float* out, const BrickGrid& g, int mode,
                     const vvr_uniforms* u, const MajorantCells& c, hipStream_t st);

// dvr.comp: hdr = RGBA16F (width x height x 4 halves).
void launch_dvr(const DevScene& s, uint16_t* hdr, uint32_t width, uint32_t height, hipStream_t st);
// The same frame in passes: rounds of march (each pixel to its next
// kDvrSamples visible samples), shading of the listed samples, per-pixel sums;
// then the rest of the long rays in one loop. For frames without ghosted
// labels; buffers of dvr_passes_bytes(width, height); `cus` sizes the shading.
// With `segmented` (frames without pre-integration) the march of the first
// round runs as segments of kDvrSegIters loop iterations, up to kDvrSegMax per
// ray, marched in parallel: a walk over the majorant grid finds each segment's
// start, the segments record their visible samples in a pool (kDvrPoolPerPixel
// records per pixel on average), and a merge composites each ray's records in
// march order.
constexpr uint32_t kDvrSamples = 32;  // samples a pixel records per round
constexpr uint32_t kDvrRounds = 1;
constexpr uint32_t kDvrSegIters = 8;
constexpr uint32_t kDvrSegMax = 64;
constexpr uint32_t kDvrPoolPerPixel = 32;
static_assert(kDvrSegMax <= 64, "a tile's wave lists its segment items one per lane");
static_assert(kDvrSegIters < 15, "a segment's record count and the full-pool mark share 4 bits");
struct DvrPasses {
    float4* rec;        // kDvrSamples records of 2 vectors per pixel
    float4* state;      // per pixel: the march state after a round
    float4* radiance;   // per pixel: the light of its samples so far
    uint32_t* list;     // the round's samples (pixel x kDvrSamples + k)
    uint32_t* count;    // their number
    // Segmented march: [segment][pixel] start distance and pool entry
    // (first record << 4 | records, 15 = the pool was full); per pixel the
    // walk's end (distance, segments); the pool of (distance, extinction,
    // density, label) records.
    float* seg_tt;
    uint32_t* seg_out;
    float2* walk;
    float4* pool;
    uint32_t pool_cap;  // records; each wave of k_dvr_segments fills its own equal part
    // The (tile, segment) items that hold a marching ray (tile | segment << 22)
    // and two counters: items listed, items taken.
    uint32_t* seg_items;
    uint32_t* seg_n;
    uint32_t seg_max;  // segments walked per ray, 1 .. kDvrSegMax
};
inline size_t dvr_tiles(uint32_t width, uint32_t height) { return size_t((width + 7) / 8) * ((height + 7) / 8); }
inline size_t dvr_passes_bytes_for(size_t pixels, size_t tiles) {
    return pixels * (kDvrSamples * 32 + 16 + 16 + kDvrSamples * 4 + kDvrSegMax * 8 + 8 + kDvrPoolPerPixel * 16) +
           tiles * kDvrSegMax * 4 + 12 * 256;  // each buffer rounded up to 256 bytes
}
inline size_t dvr_passes_bytes(uint32_t width, uint32_t height) {
    return dvr_passes_bytes_for(size_t(width) * height, dvr_tiles(width, height));
}
void launch_dvr_passes(const DevScene& s, const DvrPasses& P, uint16_t* hdr, uint32_t width, uint32_t height,
                       int cus, bool segmented, hipStream_t st);
// Lighting volumes of the ray marcher (shadow.comp, ao.comp) into s.sdims /
// s.adims texel arrays; `hu` is the host copy of the uniforms s.u points to.
void shadow_sweep_axis(const vvr_uniforms& hu, int& axis, int& sign);
void launch_shadow_sweep(const DevScene& s, uint2* out, const vvr_uniforms& hu, hipStream_t st);
void launch_ao(const DevScene& s, uint16_t* out, hipStream_t st);
// The a-trous denoiser over a path-traced frame of `samples` accumulated paths:
// reads the accumulation and guide images, writes the HDR image. `scratch`
// holds 4 float4 images of the frame.
// Camera finishing (bloom.comp) and tone mapping (tonemap.comp) of the HDR
// image into RGBA8 `out`; `bloom` holds bloom_bytes() of glow images and
// `exposure` the adapted exposure (one float, 0 = not yet metered).
void launch_post(const vvr_uniforms* du, const uint16_t* hdr, uint32_t width, uint32_t height, uint16_t* bloom,
                 float* exposure, uint32_t* out, hipStream_t st);
size_t bloom_bytes(uint32_t width, uint32_t height);
// ---- the learned denoiser's U-Net (kernels_unet.hip) ----------------------------
// Tensors are NHWC bf16 given as (pointer, channel stride, channel offset).
struct UnetConv {  // out = act(conv3x3(in) + bias), optionally masked by mask > 0
    const uint16_t* in;
    int in_cs, in_co, cin;  // cin: multiple of 16
    const uint16_t* wts;    // [co rounded up to 32][9][cin]
    const float* bias;      // [cout] or null
    uint16_t* out;          // bf16 output or null
    int out_cs, out_co, cout;
    float* out_f32;  // [pixels][cout] or null
    const uint16_t* mask;
    int mask_cs, mask_co;
    int n, h, w, relu, co_blocks;
};
struct UnetWgrad {  // dw[co][tap][ci] += ..., db[co] += ... (FP32)
    const uint16_t* x;
    int x_cs, x_co, cin;
    const uint16_t* dy;
    int dy_cs, cout;
    float* dw;
    float* db;
    int n, h, w, chunk, co_blocks;
};
struct UnetEw {  // pooling, upsampling, slice copies, gradient sums
    const uint16_t* a;
    int a_cs, a_co;
    const uint16_t* b;
    int b_cs, b_co;
    const uint16_t* c_in;
    int c_cs, c_co;
    const uint16_t* mask;
    int mask_cs, mask_co;
    uint16_t* out;
    int out_cs, out_co;
    int n, h, w, c;
};
enum { kUnetPool = 0, kUnetUp = 1, kUnetCopy = 2, kUnetCombine = 3 };
void launch_conv3x3(const UnetConv& c, hipStream_t st);
void launch_conv_wgrad(const UnetWgrad& g, hipStream_t st);
void launch_wtrans(const uint16_t* w, int cout, int cin, uint16_t* wt, hipStream_t st);
void launch_unet_ew(int op, const UnetEw& e, hipStream_t st);
void launch_l1_loss(const float* pred, const float* target, int npix, uint16_t* grad, float* loss, hipStream_t st);
void launch_adam(float* w, float* m, float* v, const float* g, uint16_t* w_bf, int n, float lr, int step,
                 hipStream_t st);
void launch_f32_to_bf16(const float* a, uint16_t* out, int n, hipStream_t st);
void launch_unet_input(const float4* accum, const float4* albedo, const float4* normal, int w, int h, float n,
                       uint16_t* x, int wp, int hp, float* alb_out, hipStream_t st);
void launch_unet_target(const float4* t, const float* alb, int npix, float* out, hipStream_t st);
// The allocation-free denoise's scratch on `device` (which must be current): a
// static device array of *bytes bytes, its address looked up once per device.
hipError_t unet_scratch(int device, void** ptr, size_t* bytes);
void launch_unet_output(const float* pred, const float* alb, int wp, const float4* accum, float n, int w, int h,
                        uint16_t* hdr, hipStream_t st);

void launch_denoise(const float4* accum, const float4* albedo, const float4* normal, float4* scratch, uint16_t* hdr,
                    uint32_t width, uint32_t height, float samples, hipStream_t st);

// dvr.comp for n_jobs frames in one launch (`jobs` in device memory, sorted by
// tile_base; total_tiles = the last job's tile_base + its tile count).
void launch_dvr_batched(const DvrJob* jobs, uint32_t n_jobs, uint32_t total_tiles, hipStream_t st);

// Wavefront path tracer (kernels_wavefront.hip). One iteration runs the five
// stages over the queues; `reqs` is a device array of n_reqs requests with
// item ranges [item_base, item_base + n_items) covering [0, total_items).
struct WfStageGrid {
    int regen, extend, shade, shadow, scatter;  // single-wave workgroups per stage
};
struct WfGrid {
    WfStageGrid single, batched;  // one request / several (their kernels' occupancies differ)
    // Free-flight grids of the ghost-free kernels: a single request's, reading the
    // density or not (isolate mode), and the several requests' in isolate mode.
    int extend_no_ghosts, extend_isolate, batched_extend_isolate;
    int sort;                     // 256-lane workgroups of the queue sort
    // Tracking-loop iterations a free flight / shadow ray runs per launch before
    // it is suspended and re-queued, so that a wave's lanes do not idle behind
    // its longest ray (4096: never suspended).
    int flight_steps, shadow_steps;
    // Single-request engines: the live paths under which the frame's tail runs
    // in one launch (wf_finish); 0: never.
    uint32_t tail_paths;
};
WfGrid wf_grid(int cus);
// All `slots` slots empty on the end queue of iteration 0, item queue at 0.
void wf_init(const WfSlots& S, const WfQueues& Q, uint32_t slots, hipStream_t st);
// env_map: a request in flight renders with a light map (adds its stage).
// ghosts: the (single) request's labels may be drawn as ghosts (else its free
// flights run without the ghost branch); density: it reads the density (else,
// in isolate mode, its free flights run without the fetches).
void wf_iteration(const PtRequest* reqs, uint32_t n_reqs, uint32_t total_items, const WfSlots& S, const WfQueues& Q,
                  uint32_t cur, const WfGrid& g, bool env_map, bool ghosts, bool density, hipStream_t st);
// Served engine: all slots idle; one iteration drawing work from the active
// requests of the table `reqs` (see k_wf_regen_served).
void wf_init_served(const WfSlots& S, const WfQueues& Q, uint32_t slots, hipStream_t st);
// isolate: every request with paths in flight is in isolate mode and draws no ghost.
void wf_iteration_served(const PtRequest* reqs, const WfActive& A, const WfServed& V, const WfSlots& S,
                         const WfQueues& Q, uint32_t cur, const WfGrid& g, bool env_map, bool isolate,
                         hipStream_t st);
// Chunk sums -> accumulation, guides and HDR images of one request.
// Single-request engine, every item handed out: runs the paths still in flight
// (at most `paths`) to their ends in one launch, after the iterations queued so
// far; `cur` is the next iteration's; `ghosts` as for wf_iteration.
void wf_finish(const PtRequest* reqs, const WfSlots& S, const WfQueues& Q, uint32_t cur, uint32_t paths, bool ghosts,
               hipStream_t st);
// Served engine, after the iterations queued so far on `st` (`cur` the next
// one's): the paths of request-table entry `entry` leave the queues. With
// `drop` they are abandoned (slots idle, items counted in done); otherwise up to
// `cap` are listed in `list` / *count (zeroed by the caller) for wf_tail.
void wf_tail_gather(const WfSlots& S, const WfQueues& Q, uint32_t cur, uint32_t entry, bool drop, uint32_t* list,
                    uint32_t* count, uint32_t cap, uint32_t* done, uint32_t slots, hipStream_t st);
// Runs the listed paths to their ends (the same results as iterations), on a
// stream of its own beside the iterations; finished items counted in done.
void wf_tail(const PtRequest* reqs, const WfSlots& S, uint32_t entry, const uint32_t* list, const uint32_t* count,
             uint32_t cap, uint32_t* done, hipStream_t st);
// After wf_tail has ended: its slots back on the end queue of iteration `cur`.
void wf_tail_return(const WfQueues& Q, uint32_t cur, const uint32_t* list, const uint32_t* count, uint32_t cap,
                    hipStream_t st);
void wf_resolve(const PtRequest& r, hipStream_t st);

}  // namespace vvr
