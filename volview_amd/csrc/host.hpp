// Internal host-side declarations of libvvr (not part of the C ABI).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vvr.h"
#include "../../include/vvr_uniforms.h"

namespace vvr {

constexpr float kHuMin = -1024.0f;  // transfer.rs:2
constexpr float kHuMax = 3071.0f;   // transfer.rs:3
constexpr int kLutSize = 4096;      // transfer.rs:4
constexpr int kBrick = 8;           // majorant brick edge (renderer.rs:30, common.glsl:57)

struct Error {
    int code;
    std::string msg;
};

// Transfer-function tables (transfer.rs:72-117, renderer.rs:2164-2195).
struct TfTables {
    std::vector<float> lut;      // 4096 x rgba, f32 (to_lut)
    std::vector<float> lut_f16;  // the same after rounding through binary16, widened back
    std::vector<float> preint;   // 4096 x rgba (cumulative_integrals)
    float bucket_max[64];
};
void validate_tf(const vvr_tf* tf);
void tf_evaluate(const vvr_tf& tf, float hu, float out[4]);
TfTables build_tf_tables(const vvr_tf& tf);
uint16_t f32_to_f16_bits(float f);
float f16_bits_to_f32(uint16_t h);
uint16_t hu_to_unorm16(int16_t hu);

// Volume geometry (crates/volio/src/volume.rs:21-40, renderer.rs:876-895).
struct Geometry {
    uint32_t dims[3];
    float world_to_index[16];
    float index_to_world[16];
    float bounds_lo[3], bounds_hi[3];
    float center[3];
    float diagonal;
    float min_spacing, max_spacing;
};
Geometry make_geometry(const vvr_volume& v);

// Renderer::opacity_key (renderer.rs:1468-1495).
struct OpacityKey {
    uint64_t generation;
    uint32_t mode;
    uint32_t extinction_bits;
    uint64_t visible[4];
    uint64_t ghost[4];
    uint32_t ghost_opacity_bits;
    int32_t margin;
    uint32_t smoothing;
    bool operator==(const OpacityKey& o) const;
    bool operator!=(const OpacityKey& o) const { return !(*this == o); }
};
OpacityKey opacity_key(const vvr_label_params& l, uint64_t generation, bool has_labels, float min_spacing);

// The uniform fill of Renderer::record (renderer.rs:1516-1660).
// env_peak_u: the light map's brightest column (environment.rs:113-125), null without a map.
void pack_uniforms(const Geometry& g, const vvr_frame_params& p, const float bucket_max[64], bool has_labels,
                   uint32_t accumulated, uint32_t width, uint32_t height, uint64_t frame, vvr_uniforms* u,
                   const float* env_peak_u = nullptr);

// EnvironmentMap::prepare and upload_environment (environment.rs:56-180,
// renderer.rs:2224-2290): a width x height linear RGB light map, rows from the
// zenith, as the shaders read it.
struct EnvTables {
    uint32_t width = 0, height = 0;
    std::vector<uint16_t> radiance;  // RGBA16F bits, mean radiance 1 over the sphere
    std::vector<float> table;        // (width + 1) x height RG: conditional CDF + texel probability; marginal
    std::vector<uint16_t> diffuse;   // 32 x 16 RGBA16F bits: cosine-weighted mean radiance
    float peak_u = 0.0f;             // image column of the brightest region
};
EnvTables prepare_environment(uint32_t width, uint32_t height, const float* rgb);

// Seeded synthetic phantoms.
void phantom_ct(const uint32_t dims[3], uint64_t seed, int16_t* hu);
int phantom_organs(const uint32_t dims[3], uint64_t seed, uint32_t n_organs, uint8_t* labels, uint8_t* soft,
                   uint16_t* margin_f16);

}  // namespace vvr
