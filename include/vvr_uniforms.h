/*
 * vvr_uniforms.h — the per-frame parameter block every kernel reads: a byte-for-byte
 * mirror of the reference's std140 `Uniforms` block (crates/render/src/renderer.rs:42-74
 * and shaders/common.glsl:4-33; all members are 16-byte aligned, 21,200 bytes).
 * Matrices are column-major like GLSL and glam.
 */
#ifndef VVR_UNIFORMS_H
#define VVR_UNIFORMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vvr_uniforms {
    float world_to_index[16]; /* patient mm -> continuous voxel index           */
    float eye[4];             /* xyz camera position (mm)                       */
    float forward[4];         /* xyz view direction                             */
    float right[4];           /* image-plane right x tan(fov/2) x aspect        */
    float up[4];              /* image-plane up x tan(fov/2)                    */
    float dims[4];            /* xyz volume dims, w ray step (mm)               */
    uint32_t flags[4];        /* x frame, y sRGB, zw output size                */
    float labels[4];          /* x mode, y tint, z label extinction, w ghost op.*/
    float light_dir[4];       /* xyz towards the key light, w irradiance        */
    float ambient[4];         /* rgb ambient radiance, w exposure multiplier    */
    float material[4];        /* x rough, y metal, z tone mapper, w light radius*/
    uint32_t lighting[4];     /* x shadows, y AO, z preintegrate, w environment */
    float light_cells[4];     /* x shadow, y AO cell size (mm)                  */
    uint32_t clip[4];         /* x plane count, y all                           */
    float clip_planes[3][4];  /* index space: dot(xyz, idx) - w                 */
    float clip_normals[3][4]; /* world normal into the cut side                 */
    float subsurface[4];      /* x amount, y indirect ext. scale, z blur, w depth*/
    uint32_t path[4];         /* x paths so far, y bounces, z paths/frame, w skip*/
    float env[4];             /* x map rotation, y cubic, z HG g, w chromatic   */
    float lens[4];            /* x lens radius, y focus distance                */
    float finish[4];          /* bloom, vignette, grain, auto exposure          */
    float tissue[4];          /* x measured-tissue bound, y its limit (1/mm)    */
    float tf_bucket_max[16][4];
    float label_colors[256][4];
    float label_materials[256][4]; /* x rough, y metal, z coat, w subsurface */
    float label_density[256][4];
    float label_surface[256][4];
    float label_optics[256][4];
} vvr_uniforms;

#ifdef __cplusplus
static_assert(sizeof(vvr_uniforms) == 21200, "must match the reference's Uniforms block");
}
#endif
#endif
