/*
 * vvr.h — C ABI of the MI355X volume renderer (majorant build, ray marcher,
 * path tracer). Every entry point mirrors one call of volview's
 * `render::Renderer` (crates/render/src/renderer.rs) or its offscreen loop
 * (crates/render/src/offscreen.rs); the comment on each cites what it replaces.
 *
 * Conventions
 *   - Status: 0 on success, a negative VVR_E* code otherwise; the message of the
 *     last failure on the calling thread is vvr_last_error(). Nothing throws
 *     across this boundary and nothing panics.
 *   - Ownership: the caller owns every host buffer; inputs are copied during the
 *     call. The library owns all device memory.
 *   - Threading: one renderer is used by one thread at a time (the reference's
 *     `&mut self`). record() is asynchronous on the renderer's HIP stream;
 *     read_*() and synchronize() wait for it.
 *   - Volume layout on input: voxel (i, j, k) at i + dims[0] * (j + dims[1] * k),
 *     as volio::Volume (crates/volio/src/volume.rs:5-7).
 */
#ifndef VVR_H
#define VVR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VVR_OK 0
#define VVR_EINVAL -1  /* size mismatch or invalid argument (reference: bail!) */
#define VVR_ELIMIT -2  /* device limit exceeded (renderer.rs:561-575)        */
#define VVR_ENOMEM -3  /* device allocation failed                            */
#define VVR_EDEVICE -4 /* HIP runtime or kernel launch failure                */
#define VVR_ESTATE -5  /* API misuse, e.g. record() before resize()           */
#define VVR_ECANCELED -6 /* a served request cancelled before it rendered     */

typedef struct vvr_ctx vvr_ctx;
typedef struct vvr_renderer vvr_renderer;

/* volio::Volume (crates/volio/src/volume.rs:9-19). `direction` is column-major:
 * column c is the unit patient-space direction of index axis c. */
typedef struct vvr_volume {
    uint32_t dims[3];
    double spacing[3];
    double origin[3];
    double direction[9];
    const int16_t* hu;
} vvr_volume;

/* transfer::ControlPoint (crates/render/src/transfer.rs:8-15). */
typedef struct vvr_control_point {
    float hu;
    float color[3];
    float extinction;
} vvr_control_point;

/* transfer::TransferFunction (transfer.rs:26-30): piecewise linear over HU. */
typedef struct vvr_tf {
    uint32_t n_points;
    const vvr_control_point* points;
} vvr_tf;

/* renderer::SoftLabelGrid (renderer.rs:104-114). `data` holds dims[0]*dims[1]*dims[2]
 * nodes of 4 bytes (best, runner-up, margin x 16, overridden), grid axis 0 fastest. */
typedef struct vvr_soft_grid {
    uint32_t dims[3];
    uint32_t src_axis[3];
    uint8_t flip[3];
    const uint8_t* data;
} vvr_soft_grid;

/* camera::OrbitCamera (crates/render/src/camera.rs:10-17). */
typedef struct vvr_camera {
    float target[3];
    float distance;
    float yaw;
    float pitch;
    float fov_y;
} vvr_camera;

enum { VVR_LABELS_OFF = 0, VVR_LABELS_TINT = 1, VVR_LABELS_ISOLATE = 2, VVR_LABELS_HYBRID = 3 };
enum { VVR_ENV_DOME = 0, VVR_ENV_STUDIO = 1, VVR_ENV_MAP = 2 };
enum { VVR_TONE_NONE = 0, VVR_TONE_ACES = 1, VVR_TONE_AGX = 2 };
enum { VVR_DENOISE_OFF = 0, VVR_DENOISE_ATROUS = 1, VVR_DENOISE_LEARNED = 2 };

/* renderer::LabelParams (renderer.rs:132-177). */
typedef struct vvr_label_params {
    uint32_t mode;
    float tint;
    float extinction;
    float colors[256][4];    /* rgb albedo, a: 0 hidden / 1 shown        */
    float materials[256][4]; /* roughness, metalness, subsurface, coat   */
    uint8_t ghost[256];
    float ghost_opacity;
    float cut_margin_mm;
    uint32_t smoothing;
    float density[256][4]; /* hybrid ramp start/end, median, IQR       */
    float surface[256][4]; /* pattern, colour variation, relief, size  */
    float optics[256][4];  /* measured mu_s' rgb (1/mm), w: 1 measured */
} vvr_label_params;

/* renderer::Lighting (renderer.rs:223-297). */
typedef struct vvr_lighting {
    float azimuth;
    float elevation;
    float intensity;
    float ambient;
    uint32_t environment;
    float light_size;
    float env_rotation;
    float anisotropy;
    float chromatic;
    uint32_t shadows;
    uint32_t ambient_occlusion;
    float exposure;
    uint32_t auto_exposure;
    float bloom;
    float vignette;
    float grain;
    uint32_t tone_mapper;
    float roughness;
    float metalness;
    float subsurface;
    float scattering_distance;
    float subsurface_strength;
    float surface_detail;
} vvr_lighting;

/* renderer::ClipPlane / Clipping (renderer.rs:396-428). plane: 0 sagittal, 1 coronal, 2 axial. */
typedef struct vvr_clip_plane {
    uint32_t enabled;
    uint32_t plane;
    uint32_t cut_positive;
    float position;
} vvr_clip_plane;

typedef struct vvr_clipping {
    vvr_clip_plane planes[3];
    uint32_t all;
} vvr_clipping;

/* renderer::PathTracing (renderer.rs:358-372). */
typedef struct vvr_path_tracing {
    uint32_t bounces;
    uint32_t samples_per_frame;
    uint32_t denoiser;
    float aperture;
    uint32_t has_focus;
    float focus[3];
} vvr_path_tracing;

/* renderer::FrameParams (renderer.rs:309-331). path_tracing is Some when
 * path_tracing_enabled != 0. */
typedef struct vvr_frame_params {
    vvr_camera camera;
    float step_mm;
    uint32_t srgb_encode;
    vvr_label_params labels;
    vvr_lighting lighting;
    uint32_t preintegrate;
    uint32_t skip_empty_space;
    vvr_clipping clipping;
    uint32_t path_tracing_enabled;
    vvr_path_tracing path_tracing;
    uint32_t smooth_sampling;
} vvr_frame_params;

/* ---- defaults (the reference's `impl Default`) ---------------------------- */

/* LabelParams::default (renderer.rs:179-196). */
void vvr_label_params_default(vvr_label_params* out);
/* Lighting::default (renderer.rs:255-283). */
void vvr_lighting_default(vvr_lighting* out);
/* PathTracing::default (renderer.rs:374-389), with the a-trous denoiser. */
void vvr_path_tracing_default(vvr_path_tracing* out);
/* FrameParams as Settings::frame_params builds it (crates/app/src/viewer.rs:374-398)
 * with the CLI defaults (crates/app/src/main.rs:84-176): step 0.5 x min spacing,
 * skip on, cubic sampling on, no labels, ray marching. */
void vvr_frame_params_default(vvr_frame_params* out, const vvr_camera* camera, float min_spacing);
/* OrbitCamera::framing (camera.rs:24-33). */
void vvr_camera_framing(const float center[3], float diagonal, vvr_camera* out);

/* ---- host-side arithmetic the reference does on the CPU ------------------- */

/* transfer::hu_to_unorm16 (transfer.rs:121-124). */
uint16_t vvr_hu_to_unorm16(int16_t hu);
/* TransferFunction::evaluate (transfer.rs:72-86): rgba = albedo, extinction. */
void vvr_tf_evaluate(const vvr_tf* tf, float hu, float rgba[4]);
/* TransferFunction::to_lut (transfer.rs:90-97): 4096 x rgba f32. */
int vvr_tf_lut(const vvr_tf* tf, float* lut4096x4);
/* The fp16 bits uploaded for the LUT (renderer.rs:2177-2195, half::f16::from_f32). */
int vvr_tf_lut_f16(const vvr_tf* tf, uint16_t* lut4096x4);
/* renderer::bucket_max (renderer.rs:2164-2174): 64 f32. */
int vvr_tf_bucket_max(const vvr_tf* tf, float* out64);
/* TransferFunction::cumulative_integrals (transfer.rs:103-117): 4096 x rgba f32. */
int vvr_tf_cumulative_integrals(const vvr_tf* tf, float* out4096x4);
/* TransferFunction::presets (transfer.rs:32-69): 0 bone, 1 angio, 2 soft tissue.
 * Returns the number of points written (<= cap), or VVR_EINVAL. */
int vvr_tf_preset(uint32_t index, vvr_control_point* points, uint32_t cap);

/* The 21,200-byte uniform block record() writes for `params` on a renderer of
 * `volume` (renderer.rs:1516-1660), computed on the host without a device:
 * has_labels as after set_labels, `accumulated` paths so far, a width x height
 * frame. For FFI binders and parity checks of the packing. */
int vvr_frame_uniforms(const vvr_volume* volume, const vvr_tf* tf, const vvr_frame_params* params, int has_labels,
                       uint32_t accumulated, uint32_t width, uint32_t height, uint64_t frame, void* out, size_t bytes);

enum { VVR_ABI_VOLUME = 0, VVR_ABI_CONTROL_POINT = 1, VVR_ABI_TF = 2, VVR_ABI_SOFT_GRID = 3, VVR_ABI_CAMERA = 4,
       VVR_ABI_LABEL_PARAMS = 5, VVR_ABI_LIGHTING = 6, VVR_ABI_CLIPPING = 7, VVR_ABI_PATH_TRACING = 8,
       VVR_ABI_FRAME_PARAMS = 9, VVR_ABI_UNIFORMS = 10, VVR_ABI_SERVER_CONFIG = 11, VVR_ABI_REQUEST = 12,
       VVR_ABI_COMPLETION = 13, VVR_ABI_SERVER_STATS = 14, VVR_ABI_PT_SHAPE = 15, VVR_ABI_ENV_MAP = 16,
       VVR_ABI_UNET_CONFIG = 17 };
/* sizeof of an ABI struct, for binding checks (0 for an unknown id). */
size_t vvr_abi_sizeof(int which);

/* ---- context and renderer -------------------------------------------------- */

/* Context::new (crates/render/src/context.rs:37): selects HIP device `device`. */
int vvr_ctx_create(int device, vvr_ctx** out);
void vvr_ctx_destroy(vvr_ctx* ctx);

/* Renderer::new (renderer.rs:555-906): HU -> unorm16 and upload, derived volumes,
 * normals pass. Fails on the device limits like the reference. */
int vvr_renderer_new(vvr_ctx* ctx, const vvr_volume* volume, const vvr_tf* tf, vvr_renderer** out);
/* Renderer::set_labels (renderer.rs:1241-1279). n must equal the voxel count;
 * soft may be NULL. Invalidates the mask and majorant volumes. */
int vvr_renderer_set_labels(vvr_renderer* r, const uint8_t* labels, size_t n, const vvr_soft_grid* soft);
/* Renderer::set_transfer_function (renderer.rs:1226-1235). */
int vvr_renderer_set_transfer_function(vvr_renderer* r, const vvr_tf* tf);
/* Renderer::resize (renderer.rs:1313-1408): (re)allocates the frame images and
 * restarts accumulation. */
int vvr_renderer_resize(vvr_renderer* r, uint32_t width, uint32_t height);
/* Renderer::record (renderer.rs:1499-1812): mask and majorant passes when the
 * opacity key changed, then the ray marcher or one path-tracing frame (plus the
 * a-trous filter when selected). Asynchronous. VVR_ESTATE before resize().
 * A frame of k paths per pixel traces sample indices accumulated ..
 * accumulated + k - 1, each seeded as the reference seeds a frame of one path
 * at that index (pt.comp:378). */
int vvr_renderer_record(vvr_renderer* r, uint64_t frame, const vvr_frame_params* params);
/* Waits for all recorded work (the fence wait of Context::one_shot, context.rs:207-230). */
int vvr_renderer_synchronize(vvr_renderer* r);
/* Renderer::accumulated_samples (renderer.rs:1217-1219). */
uint32_t vvr_renderer_accumulated_samples(const vvr_renderer* r);
/* Renderer::default_camera (renderer.rs:944-946). */
int vvr_renderer_default_camera(const vvr_renderer* r, vvr_camera* out);
/* Renderer::bounds / min_spacing / center / diagonal fields (renderer.rs:546-551). */
int vvr_renderer_geometry(const vvr_renderer* r, float bounds_lo[3], float bounds_hi[3], float* min_spacing,
                          float center[3], float* diagonal);
/* Renderer::destroy (renderer.rs:2072-2125). A renderer that server requests
 * still read (queued or running) is deleted when the last of them ends; the
 * handle must not be used after this call either way. */
void vvr_renderer_destroy(vvr_renderer* r);

/* ---- read-back (synchronising) --------------------------------------------- */

/* The linear HDR image (RGBA16F in the reference, binding 8) widened to f32:
 * width*height*4 floats. Ray marching: radiance rgb, transmittance. Path
 * tracing: accumulated mean, background share. */
int vvr_renderer_read_hdr(vvr_renderer* r, float* rgba);
/* Path-tracing images (bindings 17-19): accumulated radiance+background, albedo
 * guide+luminance^2, normal guide+depth; width*height*4 floats each. Any may be NULL. */
int vvr_renderer_read_accum(vvr_renderer* r, float* accum, float* albedo, float* normal);
/* The RGBA8 display image of the last record(): camera finishing and tone
 * mapping of the HDR image (Renderer::output after record_post,
 * renderer.rs:1801-1866; offscreen.rs:52-76 reads it back). 4 bytes a pixel. */
int vvr_renderer_read_rgba(vvr_renderer* r, uint8_t* rgba);
/* The adapted auto-exposure factor (bloom.comp pass 1; 0 = not yet metered). */
int vvr_renderer_read_exposure(vvr_renderer* r, float* exposure);
/* render_to_rgba (offscreen.rs:11-78): resize, record `frames` frames (path
 * tracing accumulates them), then read the RGBA8 image back. */
int vvr_render_to_rgba(vvr_renderer* r, const vvr_frame_params* params, uint32_t width, uint32_t height,
                       uint32_t frames, uint8_t* rgba);
/* The denoised image of the last frame (renderer.rs:1795-1797: with
 * VVR_DENOISE_ATROUS the HDR image holds denoise.comp's output); -5 when the
 * last frame was not denoised. */
int vvr_renderer_read_denoised(vvr_renderer* r, float* rgba);

enum {
    VVR_VOL_DENSITY = 0,  /* uint16 unorm, 1 per voxel           */
    VVR_VOL_LABELS = 1,   /* uint8, 1 per voxel                  */
    VVR_VOL_NORMALS = 2,  /* int8 snorm, 4 per voxel             */
    VVR_VOL_MASK = 3,     /* uint16 unorm, 2 per voxel (r, g)    */
    VVR_VOL_MAJORANT = 4, /* float, 1 per 8^3 brick              */
    VVR_VOL_SHADOW = 5,   /* RGBA16F bits, ceil(dims / 2) texels  */
    VVR_VOL_AO = 6        /* R16F bits, ceil(dims / 4) texels     */
};
/* A derived volume in the reference's layout (x fastest), for parity checks.
 * `bytes` must equal the volume's size. */
int vvr_renderer_read_volume(vvr_renderer* r, int which, void* out, size_t bytes);
/* Renderer's extinction bounds (majorant.comp output), f32 per brick. */
int vvr_renderer_read_majorant(vvr_renderer* r, float* out, size_t n);
/* The 21,200-byte uniform block written for the last record() (renderer.rs:42-74,
 * 1554-1660), for parity checks of the host-side packing. */
int vvr_renderer_read_uniforms(vvr_renderer* r, void* out, size_t bytes);

/* ---- light map ---------------------------------------------------------------- */

/* environment::EnvironmentMap (environment.rs:12-17): linear-light RGB radiance,
 * width x height texels, rows from the top (zenith) down. */
typedef struct vvr_env_map {
    uint32_t width;
    uint32_t height;
    const float* rgb;
} vvr_env_map;
/* Renderer::set_environment_map (renderer.rs:1283-1302): sets (or with NULL
 * removes) the HDR light map of VVR_ENV_MAP, importance-sampled by the path
 * tracer and convolved for the ray marcher; restarts accumulation. Copied at the
 * call; waits for the renderer's stream. */
int vvr_renderer_set_environment_map(vvr_renderer* r, const vvr_env_map* map);
/* The set light map's brightest column (environment.rs:113-125) and, for parity
 * checks, its (width + 1) x height x 2 sampling table (table may be NULL). */
int vvr_renderer_read_environment(vvr_renderer* r, float* peak_u, float* table, size_t table_floats);

/* ---- the learned denoiser --------------------------------------------------------
 * No counterpart in the reference (which calls Intel OIDN there, renderer.rs:351-357):
 * a convolutional U-Net over the path tracer's accumulation and guide images,
 * trained on the GPU by this library's bf16 MFMA kernels; the a-trous filter is
 * its fallback and comparison point. */
typedef struct vvr_unet vvr_unet;
typedef struct vvr_unet_config {
    uint32_t levels;       /* 2..5 resolution levels */
    uint32_t channels[5];  /* per level, multiples of 16 */
    uint32_t seed;         /* of the He-normal initial weights */
} vvr_unet_config;
int vvr_unet_create(vvr_ctx* ctx, const vvr_unet_config* config, vvr_unet** out);
void vvr_unet_destroy(vvr_unet* net);
/* Parameters, FP32, per layer W[cout][3x3][cin] then b[cout], layers in the order
 * encoder (two per level), decoder (two per level, deepest first), output. */
size_t vvr_unet_num_params(const vvr_unet* net);
int vvr_unet_get_params(vvr_unet* net, float* out, size_t count);
int vvr_unet_set_params(vvr_unet* net, const float* params, size_t count);
/* The gradients of the last train_step's loss, in the parameter order. */
int vvr_unet_get_grads(vvr_unet* net, float* out, size_t count);
/* One step on `batch` frames of width x height (multiples of 2^(levels-1)):
 * host float4 images of `samples` paths (accumulation, albedo and normal guides,
 * as vvr_renderer_read_accum returns them) and the reference radiance (float4,
 * rgb used). Computes the L1 loss of the log irradiance and the gradients; with
 * lr > 0 takes an Adam step. */
int vvr_unet_train_step(vvr_unet* net, uint32_t batch, uint32_t width, uint32_t height, const float* accum,
                        const float* albedo, const float* normal, float samples, const float* target, float lr,
                        float* loss);
/* The denoised frame (rgba floats: radiance, background share) of host images. */
int vvr_unet_denoise(vvr_unet* net, uint32_t width, uint32_t height, const float* accum, const float* albedo,
                     const float* normal, float samples, float* rgba);
/* No counterpart in the reference. For callers that cannot allocate device or
 * host memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid it): vvr_unet_denoise, same inputs and bit-identical output, with every
 * buffer of the call in a static device array of 64 MiB per GPU that the library
 * holds while loaded. Enough for the 5-level 32-128-channel network on frames of
 * about 96 Ki pixels (256 x 384); larger frames fail with VVR_ELIMIT. Calls on one
 * device share the array, so calls that overlap on one device race. */
int vvr_unet_denoise_noalloc(vvr_unet* net, uint32_t width, uint32_t height, const float* accum,
                             const float* albedo, const float* normal, float samples, float* rgba);
/* Frames recorded with VVR_DENOISE_LEARNED run `net` (owned by the caller, kept
 * alive while attached; NULL detaches). */
int vvr_renderer_set_denoiser(vvr_renderer* r, vvr_unet* net);

/* ---- measurement ------------------------------------------------------------ */

enum { VVR_KERNEL_MAJORANT = 0, VVR_KERNEL_NORMALS = 1, VVR_KERNEL_MASK = 2, VVR_KERNEL_DVR = 3, VVR_KERNEL_PT = 4,
       VVR_KERNEL_LIGHTING = 5, VVR_KERNEL_DENOISE = 6, VVR_KERNEL_POST = 7, VVR_KERNEL_COUNT = 8 };
/* Enables HIP-event timing of each kernel on the stream it is launched on. */
int vvr_renderer_set_timing(vvr_renderer* r, int enabled);
/* Total milliseconds and launches of kernel `which` since timing was enabled. */
int vvr_renderer_kernel_time(vvr_renderer* r, int which, double* total_ms, uint64_t* launches);

/* ---- shape of the wavefront path tracer, and its autotuner -------------------
 * No counterpart in the reference, whose megakernel runs one pixel per lane. */
typedef struct vvr_pt_shape {
    uint32_t slots;        /* paths in flight, 64 .. 2^26, rounded up to 64 (0: keep) */
    uint32_t flight_steps; /* tracking iterations a free flight runs per launch before it is
                              suspended and resumed in the next iteration, 1 .. 4096 (0: keep) */
    uint32_t shadow_steps; /* the same for a shadow ray (0: keep) */
    uint32_t tail_paths;   /* once every work item of a frame is handed out and fewer paths than
                              this are in flight, they run to their ends in one launch
                              (UINT32_MAX: never; 0: keep) */
} vvr_pt_shape;
int vvr_renderer_get_pt_shape(vvr_renderer* r, vvr_pt_shape* out);
/* Frames are the same with any shape; only their speed changes. */
int vvr_renderer_set_pt_shape(vvr_renderer* r, const vvr_pt_shape* shape);
/* Times the path-tracing frames of `params` (path_tracing_enabled, the renderer
 * resized) with candidate shapes and keeps the fastest, written to `chosen`: the
 * current shape, step pairs (8,16) (12,24) (16,32) (24,32) (24,48) (32,64) at the
 * current slots, then 1, 2, 4 and 8 Mi slots with the best pair, then tail thresholds of 32,
 * 64, 128 and 256 Ki paths and none with the best so far. Each candidate's time is the
 * least of `trials` frames (HIP events); a candidate replaces the best so far only
 * when it is 2 % faster. The trial frames are recorded as frames 1, 2, ... and
 * accumulate like any others. -1 for a renderer not on the wavefront engine. */
int vvr_renderer_autotune(vvr_renderer* r, const vvr_frame_params* params, uint32_t trials, vvr_pt_shape* chosen);

/* ---- serving: many requests on one device ------------------------------------
 * No counterpart in the reference, which renders one frame per renderer at a
 * time (the plain loop of offscreen.rs:32-36). A server keeps the volumes of
 * its callers' renderers resident and renders requests of any of them side by
 * side: ray-marched frames in batched launches, path-traced requests as work
 * items of one shared wavefront engine that admits new requests while others
 * are in flight (continuous batching) up to an HBM budget. Each request's result
 * is the same as rendering it alone on its renderer with samples_per_frame =
 * pt_increment, frame after frame. A request holds its renderer from
 * submission to its final frame; meanwhile record, resize, set_labels,
 * set_transfer_function, set_environment_map and set_denoiser on that renderer
 * return VVR_ESTATE, and destroy defers the deletion to the last request.
 * Work recorded on the renderer before the submission completes before the
 * server touches the renderer's volumes. */
typedef struct vvr_server vvr_server;

typedef struct vvr_server_config {
    uint64_t hbm_budget;   /* bytes of request images and chunk sums admitted at once (0: 64 GiB) */
    uint32_t slots;        /* path slots of the wavefront engine (0: 4,194,304) */
    uint32_t pt_increment; /* paths per pixel per delivered frame of a path-traced request (0: all at once) */
    uint32_t max_requests; /* requests admitted at once, at most 256 (0: 256) */
    uint32_t flight_steps; /* vvr_pt_shape's, for the server's engine (0: 24) */
    uint32_t shadow_steps; /* (0: 48) */
    uint32_t iterations;   /* wavefront iterations queued per pass between the host's looks at the
                              counters, 1..64 (0: 2) */
} vvr_server_config;

typedef struct vvr_request {
    vvr_renderer* volume;            /* its volume, labels and transfer function */
    const vvr_frame_params* params;  /* path tracing when path_tracing_enabled: samples_per_frame paths per pixel */
    uint32_t width, height;
    uint64_t tag;                    /* the caller's, returned in completions */
} vvr_request;

typedef struct vvr_completion {
    uint64_t id, tag;
    int32_t status;        /* VVR_OK, or the error that ended the request */
    uint32_t spp;          /* paths per pixel in this frame (path tracing), 0 for ray marching */
    uint32_t final_frame;  /* 1 for the request's last frame */
    uint32_t pad;
    uint64_t submit_ns;    /* steady-clock times: submission, admission, frame on the host */
    uint64_t admit_ns;
    uint64_t frame_ns;
} vvr_completion;

typedef struct vvr_server_stats {
    uint64_t submitted, completed, frames;
    uint64_t pt_paths;       /* paths traced into delivered frames */
    uint64_t iterations;     /* wavefront iterations */
    uint64_t hbm_in_use;     /* bytes charged to admitted requests */
    uint64_t hbm_peak;
    uint64_t pt_paths_abandoned; /* paths begun for increments of cancelled requests that never resolved
                                    (traced to their ends, or dropped in flight at the cancellation) */
} vvr_server_stats;

int vvr_server_create(vvr_ctx* ctx, const vvr_server_config* config, vvr_server** out);
/* Chooses the engine shape of a server for a request mix: renders `requests`
 * (a sample of the mix, e.g. a batch it will serve; their frames are discarded)
 * on trial servers over path slots (1, 2, 4, 8 Mi) and wavefront iterations per
 * pass (1, 2, 4, 8), and writes the fastest into config->slots and
 * config->iterations; the other fields are kept and used by every trial. The
 * requests' renderers must stay alive for the call. No counterpart in the
 * reference: it serves one request at a time. */
int vvr_server_autotune(vvr_ctx* ctx, const vvr_request* requests, uint32_t n, vvr_server_config* config);
/* Queues a request; returns at once with its id. Thread-safe. */
int vvr_server_submit(vvr_server* s, const vvr_request* request, uint64_t* id);
/* Up to `cap` frame completions (in completion order), waiting up to wait_ms
 * for the first. *n is the number returned. */
int vvr_server_poll(vvr_server* s, vvr_completion* out, uint32_t cap, uint32_t* n, uint32_t wait_ms);
/* The request's latest delivered frame as float RGBA (linear radiance +
 * transmittance or background): n_floats must equal width * height * 4 of the
 * request. Frames land in one of two pinned buffers in turn, so a read never
 * sees a frame that is still being copied, and a read racing the next frame's
 * delivery returns either frame whole. */
int vvr_server_read_hdr(vvr_server* s, uint64_t id, float* rgba, size_t n_floats);
/* Stops a request: one not admitted yet ends with status VVR_ECANCELED. A
 * path-traced one in flight stops handing out paths at once, abandons the
 * increment in flight and ends with a final completion whose frame is the last
 * delivered one (status VVR_ECANCELED if it delivered none). A ray-marched one
 * in flight finishes. */
int vvr_server_cancel(vvr_server* s, uint64_t id);
/* Frees a request after its final frame. */
int vvr_server_release(vvr_server* s, uint64_t id);
int vvr_server_get_stats(vvr_server* s, vvr_server_stats* out);
/* Waits for admitted requests to finish, then stops. */
void vvr_server_destroy(vvr_server* s);

/* ---- synthetic phantoms (workload generator) -------------------------------- */

/* Seeded CT phantom: air, body ellipsoid, lungs, skull and rib shells, a
 * contrast vessel (int16 HU, 1 mm voxels, x fastest). */
int vvr_phantom_ct(const uint32_t dims[3], uint64_t seed, int16_t* hu);
/* Up to n_organs seeded ellipsoidal organs (labels 1..n): label volume, the soft
 * grid on the volume grid (4 bytes per voxel) and the fp16 margin (signed
 * distance / 2 mm). Returns the number of organs placed. Any output may be NULL. */
int vvr_phantom_organs(const uint32_t dims[3], uint64_t seed, uint32_t n_organs, uint8_t* labels, uint8_t* soft,
                       uint16_t* margin_f16);

/* ---- errors ----------------------------------------------------------------- */

const char* vvr_last_error(void);
const char* vvr_version(void);

#ifdef __cplusplus
}
#endif
#endif /* VVR_H */
