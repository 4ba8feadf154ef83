// Plain data types shared by the host code and the kernels (no device code,
// so host translation units compile them with a plain C++ compiler).
#pragma once
#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>
#include <stddef.h>
#include <stdint.h>

#include "../../../include/vvr_uniforms.h"

#if defined(__HIPCC__)
#define VVR_HD __host__ __device__ __forceinline__
#else
#define VVR_HD inline
#endif

namespace vvr {

// 4^3 micro-brick layout of a volume of nx x ny x nz voxels.
struct BrickGrid {
    int nx, ny, nz;
    int bx, by, bz;          // bricks per axis, ceil(n / 4)
    uint32_t ystride;        // bx * 64
    uint32_t zstride;        // bx * by * 64
    VVR_HD uint32_t xoff(int x) const { return (uint32_t(x >> 2) << 6) | uint32_t(x & 3); }
    VVR_HD uint32_t yoff(int y) const { return uint32_t(y >> 2) * ystride + (uint32_t(y & 3) << 2); }
    VVR_HD uint32_t zoff(int z) const { return uint32_t(z >> 2) * zstride + (uint32_t(z & 3) << 4); }
    VVR_HD uint32_t at(int x, int y, int z) const { return xoff(x) + yoff(y) + zoff(z); }
    VVR_HD size_t padded() const { return size_t(zstride) * size_t(bz); }
};

// The ray marcher's float copy of the density (DevScene::density_f): linear, x
// fastest, with a border of kRowPad replicated texels on each side of every
// axis, so that the 4 texels of any cubic-window row are consecutive.
constexpr int kRowPad = 2;
VVR_HD size_t row_volume_texels(int nx, int ny, int nz) {
    return size_t(nx + 2 * kRowPad) * size_t(ny + 2 * kRowPad) * size_t(nz + 2 * kRowPad);
}

inline BrickGrid make_brick_grid(int nx, int ny, int nz) {
    BrickGrid g;
    g.nx = nx, g.ny = ny, g.nz = nz;
    g.bx = (nx + 3) / 4, g.by = (ny + 3) / 4, g.bz = (nz + 3) / 4;
    g.ystride = uint32_t(g.bx) * 64u;
    g.zstride = uint32_t(g.bx) * uint32_t(g.by) * 64u;
    return g;
}

// Everything the render kernels read, by value in the kernel arguments.
struct DevScene {
    BrickGrid g;
    const uint16_t* __restrict__ density;  // unorm16
    const float* __restrict__ density_f;   // the same values as floats, c / 65535, in rows (row_volume_texels)
    const uint8_t* __restrict__ labels;    // label per voxel
    const uint32_t* __restrict__ mask;     // RG16: r low half, g high half
    // One bit per unit cell of the mask (the 8 clamped corners a sample there
    // interpolates; cells 0..n per axis, in 4^3-cell micro-bricks of grid mg): 0
    // when no corner reaches the frame's solid or ghost threshold, so that no
    // label can be drawn there (null: not built).
    const uint32_t* __restrict__ mask_cells;
    BrickGrid mg;
    // Per mask cell, in the same layout: L when all 8 corners are solid (at least
    // the solid threshold) and carry the same label L, drawn in the frame's mode
    // (then the label search returns L: its first candidate is a corner); else 0.
    const uint8_t* __restrict__ cell_labels;
    const uint32_t* __restrict__ normals;  // RGBA8 snorm, x in the low byte
    const float* __restrict__ majorant;    // ceil(dims / 8), x fastest
    int mx, my, mz;
    const float4* __restrict__ lut;     // 4096 transfer-function entries, fp16-rounded
    const float4* __restrict__ preint;  // 4096 cumulative integrals
    const vvr_uniforms* __restrict__ u;
    // Lighting volumes of the ray marcher (null when the frame does not use them):
    // RGBA16F diffused light + direct transmittance at ceil(dims / 2), R16F
    // ambient visibility at ceil(dims / 4), both x fastest.
    const uint2* __restrict__ shadow;
    const uint16_t* __restrict__ ao;
    int sdims[3], adims[3];
    // Light map (Environment::Map; null without one): RGBA16F radiance env_w x
    // env_h, RG32F sampling table (env_w + 1) x env_h, RGBA16F diffuse 32 x 16.
    const uint2* __restrict__ env_rad;
    const float2* __restrict__ env_table;
    const uint2* __restrict__ env_dif;
    int env_w, env_h;
};

// One ray-marched frame of a batched launch: its tiles are workgroups
// tile_base .. tile_base + tiles_x * ceil(height / 8) - 1.
struct DvrJob {
    DevScene scene;
    uint16_t* hdr;
    uint32_t width, height, tiles_x, tile_base;
};

// One path-tracing request of the wavefront engine: a frame of `samples` paths
// per pixel at path indices first_path .. on one scene, cut into work items of
// `chunk` consecutive paths of one pixel. Items are numbered chunk-major over
// 8 x 8 pixel tiles: item = c * n_px + tile * 64 + pixel in tile.
struct PtRequest {
    DevScene scene;
    float4* accum;
    float4* albedo;
    float4* normal;
    uint16_t* hdr;
    float4* partial;  // [n_chunks][3][n_px] chunk sums
    uint32_t width, height, tiles_x;
    uint32_t first_path, samples, chunk, n_chunks, n_px;
    uint32_t item_base, n_items;  // global item range of this request
};

// Path slots of the wavefront engine, structure of float4 arrays (one float4
// per slot per field; the w lanes carry packed integers).
// One field of the path slots, read and written as field[slot]. The fields a
// stage touches together are interleaved slot by slot: each slot's fields form
// a record of `stride` elements, so one lane's loads of a path fall in one or
// two cache lines (queues visit slots in no order that would coalesce lanes).
template <class T> struct SlotField {
    T* base;
    uint32_t stride;
    __device__ __forceinline__ T& operator[](uint32_t i) const { return base[size_t(i) * stride]; }
};

// Records of 8 float4s, one 128-byte line each: a path's state and suspended
// loop (what extend reads), its chunk sums and light-map sample, its event.
constexpr uint32_t kWfPathRecord = 8, kWfSumsRecord = 8, kWfEventRecord = 8;

struct WfSlots {
    SlotField<float4> p_origin;  // origin (index space), last_pdf
    SlotField<float4> p_dir;     // direction (world), rng state
    SlotField<float4> p_thr;     // throughput, hero | surfaces << 2 | volumes << 8 | bounce << 18
    SlotField<float4> p_spec;    // spectral ratios, global item index
    SlotField<float4> p_rad;     // radiance, k | request << 16
    SlotField<uint2> p_flight;   // a suspended tracking loop: distance along the ray (f32 bits), iteration
                                 // (free flight: iteration kWfNewFlight for a flight not begun; the shadow
                                 // ray's, whose running transmittance is in e_lt)
    SlotField<float4> c0;        // chunk sums: radiance, background
    SlotField<float4> c1;        //             albedo, luminance^2
    SlotField<float4> c2;        //             normal, depth
    SlotField<float4> e_map;     // light-map sample: direction, pdf
    SlotField<float4> e_maplt;   // its radiance / pdf x transmittance
    SlotField<float4> e_idx;     // event position, cut
    SlotField<float4> e_s0;      // sample albedo, extinction
    SlotField<float4> e_s1;      // roughness, metalness, subsurface, coat
    SlotField<float4> e_s2;      // density, label | ghost << 8 | (cap_plane + 1) << 9 | surface event << 12
    SlotField<float4> e_n;       // shading normal
    SlotField<float4> e_light;   // key-light direction
    SlotField<float4> e_lt;      // its transmittance
    uint16_t* req;               // request of the slot's path (engines over several requests)
};

// Counters in WfQueues::counts: queue lengths, the item queue head and the live
// slots of a single run, and the idle queues of the served engine.
enum {
    kWfFF0 = 0,
    kWfFF1 = 1,
    kWfEV = 2,
    kWfEND0 = 3,
    kWfEND1 = 4,
    kWfHEAD = 5,
    kWfLIVE = 6,
    kWfIDLE0 = 7,
    kWfIDLE1 = 8,
    kWfSORTFF = 9,   // entries of the sorted flight queue, padding included
    kWfSORTEV = 10,  // entries of the sorted event queue, padding included
    kWfSH0 = 11,
    kWfSH1 = 12,
    kWfSC = 13,
    kWfCounts = 16
};

struct WfQueues {
    uint32_t* ff[2];    // slots to extend
    uint32_t* ev;       // slots at an event (shade, shadow, scatter)
    uint32_t* end[2];   // slots whose path ended (or that have no item)
    uint32_t* idle[2];  // served engine: slots waiting for work
    uint32_t* sh[2];    // slots whose shadow ray was suspended (resumed by the next shadow stage)
    uint32_t* sc;       // slots whose shadow ray is done (light-map stage, scatter)
    uint32_t* counts;   // kWfCounts counters
    // Engines over several requests: ff or ev binned by request, each request's
    // run padded with kWfNone to a multiple of 64 entries so that every wave
    // holds one request, and the bins' counts and cursors (kWfBins each for the
    // ff sort, then for the ev sort).
    uint32_t* sorted;
    uint32_t* bins;
};

constexpr uint32_t kWfBins = 256;  // request indices the queue sort bins
constexpr uint32_t kWfNone = 0xffffffffu;
constexpr uint32_t kWfNewFlight = 0xffffffffu;  // p_flight.y of a free flight not begun
// Entries of the sorted queue: a queue of `slots` entries plus the padding.
constexpr uint32_t wf_sorted_capacity(uint32_t slots) { return slots + kWfBins * 64u; }

// Requests the served engine draws work items from this iteration: entries of
// its request table, passed by value.
constexpr int kWfMaxActive = 64;
struct WfActive {
    uint32_t n;
    uint16_t r[kWfMaxActive];
};

// The served engine's per-request counters (indexed by request-table entry)
// and which idle queue holds the idle slots; with `revive` set the iteration
// offers them work.
constexpr uint32_t kWfCounterStride = 32;  // uint32s between two requests' counters (one 128-byte line each)
struct WfServed {
    uint32_t* head;  // items taken, at r * kWfCounterStride
    uint32_t* done;  // items finished, at r * kWfCounterStride
    uint32_t idle_in;
    uint32_t revive;
};

}  // namespace vvr
