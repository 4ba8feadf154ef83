/* Test infrastructure only: the CPU oracle of the volume-rendering hot path.
 * Nothing in the shipped library includes, links or calls this.
 *
 * C entry points of liboracle.so, called from tests/ through ctypes. Volumes are
 * in the reference's layout: voxel (i, j, k) at i + dims[0] * (j + dims[1] * k). */
#ifndef VVR_ORACLE_API_H
#define VVR_ORACLE_API_H

#include <stddef.h>
#include <stdint.h>

#include "../include/vvr.h"
#include "../include/vvr_uniforms.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The textures the shaders sample (binding numbers of shaders/common.glsl). */
typedef struct oracle_scene {
    uint32_t dims[3];
    const uint16_t* density;  /* binding 1, R16_UNORM                              */
    const uint16_t* lut;      /* binding 2, 4096 x RGBA16F bits                    */
    const float* preint;      /* binding 12, 4096 x RGBA32F                        */
    const uint8_t* labels;    /* binding 4, R8_UINT (NULL: 1-voxel placeholder 0)  */
    const int8_t* normals;    /* binding 5, RGBA8_SNORM                            */
    const uint16_t* mask;     /* binding 13, RG16_UNORM (NULL outside label modes) */
    const float* majorant;    /* binding 22, R32F, ceil(dims / 8)                  */
    const uint16_t* shadow;   /* binding 6, RGBA16F bits, ceil(dims / 2) (NULL: off) */
    const uint16_t* ao;       /* binding 7, R16F bits, ceil(dims / 4) (NULL: off)    */
    /* Light map (bindings 25-27; NULL without one): RGBA16F radiance env_w x env_h,
     * RG32F sampling table (env_w + 1) x env_h, RGBA16F diffuse 32 x 16. */
    const uint16_t* env_radiance;
    const float* env_table;
    const uint16_t* env_diffuse;
    uint32_t env_w, env_h;
} oracle_scene;

uint16_t oracle_hu_to_unorm16(int16_t hu);
void oracle_tf_evaluate(const vvr_control_point* pts, uint32_t n, float hu, float out[4]);
void oracle_tf_lut(const vvr_control_point* pts, uint32_t n, float* lut4096x4);
void oracle_tf_lut_f16(const vvr_control_point* pts, uint32_t n, uint16_t* out);
void oracle_tf_cumulative(const vvr_control_point* pts, uint32_t n, float* out);
void oracle_bucket_max(const vvr_control_point* pts, uint32_t n, float* out64);
void oracle_camera_framing(const float center[3], float diagonal, vvr_camera* out);
void oracle_geometry(const vvr_volume* vol, float world_to_index[16], float index_to_world[16], float center[3],
                     float* diagonal, float* min_spacing, float* max_spacing);
void oracle_pack_uniforms(const vvr_volume* vol, const vvr_frame_params* p, const float* bucket_max64, int has_labels,
                          uint32_t accumulated, uint32_t width, uint32_t height, uint64_t frame, vvr_uniforms* u);
/* As oracle_pack_uniforms with a light map whose brightest column is env_peak_u
 * (has_map 0: none). */
void oracle_pack_uniforms_env(const vvr_volume* vol, const vvr_frame_params* p, const float* bucket_max64,
                              int has_labels, uint32_t accumulated, uint32_t width, uint32_t height, uint64_t frame,
                              int has_map, float env_peak_u, vvr_uniforms* u);
/* EnvironmentMap::prepare and upload_environment (environment.rs:56-180,
 * renderer.rs:2224-2290) of a width x height linear RGB map, rows from the
 * zenith: radiance RGBA16F bits (width x height x 4), the RG32F sampling table
 * ((width + 1) x height x 2), the diffuse map RGBA16F bits (32 x 16 x 4); returns
 * the brightest column peak_u. */
float oracle_env_prepare(uint32_t width, uint32_t height, const float* rgb, uint16_t* radiance, float* table,
                         uint16_t* diffuse);
/* sample_env (pt.comp:203-226) with a PCG stream seeded `seed`: direction, pdf;
 * and env_pdf (pt.comp:193-199) of a direction. */
void oracle_sample_env(const oracle_scene* s, const vvr_uniforms* u, uint32_t seed, uint32_t n, float* dirs,
                       float* pdfs);
float oracle_env_pdf(const oracle_scene* s, const vvr_uniforms* u, const float d[3]);

uint16_t oracle_f32_to_f16(float f);
float oracle_f16_to_f32(uint16_t h);

/* majorant.comp over all bricks. */
void oracle_majorant(const oracle_scene* s, const vvr_uniforms* u, float* out);
/* normals.comp: 4 snorm8 per voxel. */
void oracle_normals(const oracle_scene* s, const vvr_uniforms* u, int8_t* out);
/* mask.comp, all four passes: 2 unorm16 per voxel (the final image A).
 * soft: NULL or grid of soft_dims nodes x 4 bytes; soft_push = pc.soft
 * (xyz source axes, w flip bits | 8 present); margin = pc.params.x, radius = pc.params.z. */
void oracle_mask(const oracle_scene* s, const vvr_uniforms* u, const uint8_t* soft, const uint32_t soft_dims[3],
                 const int32_t soft_push[4], int32_t margin, int32_t radius, uint16_t* out);
/* The binary pass (pass 0) alone, float before quantisation: 2 per voxel. */
void oracle_mask_binary(const oracle_scene* s, const vvr_uniforms* u, const uint8_t* soft, const uint32_t soft_dims[3],
                        const int32_t soft_push[4], int32_t margin, float* out);

/* dvr.comp for `n` pixels (x, y pairs): rgba (radiance, transmittance) before the
 * RGBA16F store. precision 32 or 64. */
void oracle_dvr(const oracle_scene* s, const vvr_uniforms* u, const uint32_t* pixels, size_t n, int precision,
                float* rgba);
/* shadow.comp swept over every slice as Renderer::record_shadow_sweep does
 * (renderer.rs:1907-1940): RGBA16F bits, ceil(dims / 2) texels x fastest. */
void oracle_shadow(const oracle_scene* s, const vvr_uniforms* u, int precision, uint16_t* out);
/* ao.comp: R16F bits, ceil(dims / 4) texels x fastest. */
void oracle_ao(const oracle_scene* s, const vvr_uniforms* u, int precision, uint16_t* out);
/* lighting_lookup (common.glsl:154-156) of the shadow (which = 0: rgba) or
 * ambient-occlusion (which = 1: r) volume at texture coordinate tc. */
void oracle_lighting_lookup(const oracle_scene* s, const vvr_uniforms* u, int which, const float tc[3], int precision,
                            float out[4]);
/* denoise.comp over a whole image, as Renderer::record_denoise runs it
 * (renderer.rs:1868-1904): prepare, `levels` a-trous levels (the reference runs
 * 5), finish. Images are width x height float4 row-major: the accumulation and
 * guide images and their sample count. hdr: the finished image before its
 * RGBA16F store; last (may be NULL): the last filtered irradiance + variance. */
void oracle_denoise(const float* accum, const float* albedo, const float* normal, uint32_t width, uint32_t height,
                    float samples, int levels, int precision, float* hdr, float* last);
/* Camera finishing: bloom.comp's 7 passes (renderer.rs:1815-1866) and
 * tonemap.comp over an HDR image (width x height RGBA floats, the RGBA16F
 * values). exposure: the adapted auto exposure in and out (0 = not yet
 * metered). out: RGBA8 (UNORM stores round to nearest). bloom_a / bloom_c (may
 * be NULL): the final quarter- and eighth-resolution glow images. */
void oracle_post(const oracle_scene* s, const vvr_uniforms* u, const float* hdr, uint32_t width, uint32_t height,
                 float* exposure, uint8_t* out, float* bloom_a, float* bloom_c);
float oracle_linear_to_srgb(float c);
float oracle_aces(float x);
void oracle_agx(const float in[3], float out[3]);
uint32_t oracle_hash01_bits(uint32_t v);
/* pt.comp for `n` pixels: the values the dispatch adds to the accumulation,
 * albedo-guide and normal-guide images (4 floats each per pixel). */
void oracle_pt(const oracle_scene* s, const vvr_uniforms* u, const uint32_t* pixels, size_t n, int precision,
               float* accum, float* albedo, float* normal);

/* Building blocks, exposed for closed-form checks. */
float oracle_density_cubic(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], int precision);
void oracle_density_gradient_cubic(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], int precision,
                                   float out[3]);
float oracle_density_linear(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], int precision);
uint32_t oracle_pcg_sequence(uint32_t seed, uint32_t n, uint32_t* out);
void oracle_rand_sequence(uint32_t seed, uint32_t n, float* out);
float oracle_phase_hg(float g, float cos_t);
void oracle_sample_hg(float g, uint32_t seed, uint32_t n, float* cos_out);
float oracle_ggx_distribution(float nh, float a2);
void oracle_single_scattering_albedo(const float a[3], float out[3]);
/* free_flight in a homogeneous scene: 1 if the ray collided, distance to the collision. */
int oracle_free_flight(const oracle_scene* s, const vvr_uniforms* u, const float origin_idx[3], const float dir[3],
                       uint32_t seed, float* t_out, float spectral[3]);
void oracle_transmittance(const oracle_scene* s, const vvr_uniforms* u, const float origin_idx[3], const float dir[3],
                          uint32_t seed, float out[3]);

/* Shading library (shading.glsl): brdf_coated's specular and diffuse parts for a
 * surface of these materials (brdf alone when coat = 0); smith; the path
 * tracer's lobe pdf (bsdf_pdf) and lobe sampling (pt.comp:491-502), n samples from
 * a PCG stream seeded `seed`; sample_light in the frame of `dir`; mis. */
void oracle_brdf(const float n[3], const float v[3], const float l[3], const float albedo[3], float roughness,
                 float metalness, float coat, int precision, float spec[3], float diffuse[3]);
double oracle_smith(double nv, double nl, double a);
float oracle_bsdf_pdf(const float n[3], const float v[3], const float l[3], float roughness, float metalness,
                      float coat);
void oracle_bsdf_sample(const float n[3], const float v[3], float roughness, float metalness, float coat,
                        uint32_t seed, uint32_t count, float* dirs);
void oracle_sample_light(const vvr_uniforms* u, const float dir[3], uint32_t seed, uint32_t count, float* dirs);
float oracle_mis(float a, float b);
/* Procedural texture (shading.glsl:175-342): pcg3d hash, gradient noise, fbm,
 * cellular noise, relief height; the texture's albedo / roughness / coat
 * change and the bump-mapped normal of label `label` at index idx. */
void oracle_hash3(const int32_t c[3], float out[3]);
float oracle_gradient_noise(const float p[3], int precision);
float oracle_fbm(const float p[3], int precision);
void oracle_cellular(const float p[3], int precision, float* nearest, float* border, int32_t cell[3]);
float oracle_surface_height(const float p[3], const float params[4], int precision);
void oracle_index_to_world(const vvr_uniforms* u, const float idx[3], float out[3]);
void oracle_surface_texture(const vvr_uniforms* u, uint32_t label, const float idx[3], const float albedo[3],
                            float roughness, float coat, float out_albedo[3], float* out_roughness, float* out_coat);
void oracle_bump_normal(const vvr_uniforms* u, uint32_t label, const float idx[3], const float n[3], float out[3]);
/* Labels (common.glsl:232-351): the 21-candidate search (1 and the label when
 * found), with_label of a transfer-function sample of `density` (albedo out,
 * extinction returned), apply_labels at idx (label returned). */
int oracle_drawn_label_near(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], int ghost,
                            uint32_t* label);
float oracle_with_label(const oracle_scene* s, const vvr_uniforms* u, float density, uint32_t label, int ghost,
                        float albedo[3]);
int oracle_apply_labels(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], float* extinction,
                        int* ghost);
/* clip_depth (common.glsl:357-373): mm inside the cut, the bounding plane. */
float oracle_clip_depth(const vvr_uniforms* u, const float idx[3], int* plane);
/* Pre-integration (common.glsl:195-230): tf_segment's albedo and extinction,
 * preint_half_depth. */
void oracle_tf_segment(const oracle_scene* s, const vvr_uniforms* u, float front, float back, int precision,
                       float out[4]);
float oracle_preint_half_depth(const oracle_scene* s, const vvr_uniforms* u, float front, float back, int precision);
/* surface_frame (shading.glsl:73-144) of a sample of `label` (ghost or not) seen
 * from v: normal, surfaceness, and the returned extinction scale. */
float oracle_surface_frame(const oracle_scene* s, const vvr_uniforms* u, const float idx[3], uint32_t label,
                           int ghost, const float v[3], int cap_plane, float cut, float cap_depth, int precision,
                           float n[3], float* surfaceness);
/* The FP32 oracle's log and exp of n values (the render path's model of GLSL's). */
void oracle_vlog_vexp(const float* x, size_t n, float* log_out, float* exp_out);
/* The render path's pow (vec.hpp vpow) of each pair. */
void oracle_vpow(const float* x, const float* y, size_t n, float* out);
/* The kernels' x / 6 without the division (scene.hpp div6_normal / div6: x / 2
 * times RN(1/3) plus one FMA-exact correction; a true division for |x| < 2^-125),
 * checked against x / 6.0f. mode 0: over the float bit patterns first .. last
 * (stride apart) as dividends, through div6; mode 1: those patterns as the
 * fraction f of the cubic weights w0..w3 (common.glsl:88-95), each numerator
 * through the kernels' form; mode 2: the uncorrected product x RN(1/6) in place of
 * the kernels' form (a check that the comparison can fail). Returns the number of
 * mismatching bit patterns. */
uint64_t oracle_div6_mismatches(uint32_t first, uint32_t last, uint32_t stride, int mode);
/* tissue_sigma (shading.glsl:161-168) of a sample of `label` with `extinction`. */
void oracle_tissue_sigma(const vvr_uniforms* u, uint32_t label, float extinction, float out[3]);
/* The PCG stream of pixel (px, py) after pt.comp's seeding and warm-up draw. */
void oracle_pixel_stream(const vvr_uniforms* u, uint32_t px, uint32_t py, uint32_t n, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif
