// Transfer functions and density normalisation: the host-side tables the
// kernels read. Follows crates/render/src/transfer.rs and the table uploads
// of crates/render/src/renderer.rs:2164-2195.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "host.hpp"

namespace vvr {

namespace {
// Rust's f32::clamp (without its panic on inverted bounds).
float rust_clamp(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

struct Preset {
    const char* name;
    std::vector<vvr_control_point> pts;
};
// transfer.rs:32-69
const std::vector<Preset>& presets() {
    static const std::vector<Preset> p = {
        {"bone",
         {{-1024.0f, {0.0f, 0.0f, 0.0f}, 0.0f},
          {180.0f, {0.9f, 0.6f, 0.45f}, 0.0f},
          {300.0f, {0.95f, 0.85f, 0.7f}, 0.6f},
          {1200.0f, {1.0f, 0.97f, 0.9f}, 3.0f},
          {3071.0f, {1.0f, 1.0f, 1.0f}, 4.0f}}},
        {"angio",
         {{-1024.0f, {0.0f, 0.0f, 0.0f}, 0.0f},
          {120.0f, {0.7f, 0.1f, 0.08f}, 0.0f},
          {200.0f, {0.85f, 0.15f, 0.1f}, 0.5f},
          {450.0f, {0.95f, 0.3f, 0.2f}, 1.5f},
          {600.0f, {0.95f, 0.88f, 0.75f}, 2.5f},
          {3071.0f, {1.0f, 1.0f, 1.0f}, 4.0f}}},
        {"soft tissue",
         {{-1024.0f, {0.0f, 0.0f, 0.0f}, 0.0f},
          {-200.0f, {0.9f, 0.75f, 0.5f}, 0.0f},
          {-80.0f, {0.95f, 0.8f, 0.5f}, 0.004f},
          {-20.0f, {0.8f, 0.5f, 0.4f}, 0.004f},
          {40.0f, {0.75f, 0.3f, 0.25f}, 0.03f},
          {150.0f, {0.85f, 0.2f, 0.15f}, 0.15f},
          {300.0f, {0.95f, 0.3f, 0.2f}, 0.8f},
          {500.0f, {0.95f, 0.88f, 0.75f}, 2.0f},
          {3071.0f, {1.0f, 1.0f, 1.0f}, 4.0f}}},
    };
    return p;
}
}  // namespace

void validate_tf(const vvr_tf* tf) {
    if (!tf || tf->n_points == 0 || !tf->points) throw Error{VVR_EINVAL, "transfer function has no control points"};
}

// TransferFunction::evaluate (transfer.rs:72-86).
void tf_evaluate(const vvr_tf& tf, float hu, float out[4]) {
    const vvr_control_point* p = tf.points;
    uint32_t n = tf.n_points;
    auto put = [&](const vvr_control_point& c) {
        out[0] = c.color[0];
        out[1] = c.color[1];
        out[2] = c.color[2];
        out[3] = c.extinction;
    };
    if (hu <= p[0].hu) return put(p[0]);
    for (uint32_t i = 1; i < n; i++) {
        if (hu <= p[i].hu) {
            float t = (hu - p[i - 1].hu) / std::max(p[i].hu - p[i - 1].hu, 1e-6f);
            const float a[4] = {p[i - 1].color[0], p[i - 1].color[1], p[i - 1].color[2], p[i - 1].extinction};
            const float b[4] = {p[i].color[0], p[i].color[1], p[i].color[2], p[i].extinction};
            for (int c = 0; c < 4; c++) out[c] = a[c] + (b[c] - a[c]) * t;
            return;
        }
    }
    put(p[n - 1]);
}

uint16_t f32_to_f16_bits(float f) {
    // Round to nearest even, like half::f16::from_f32 (renderer.rs:2181).
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const int32_t exp = int32_t((x >> 23) & 0xffu);
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xff) return uint16_t(sign | 0x7c00u | (man ? (0x200u | (man >> 13)) : 0u));
    const int32_t e = exp - 112;  // rebias 127 -> 15
    if (e >= 31) return uint16_t(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return uint16_t(sign);
        man |= 0x800000u;
        const int shift = 14 - e;
        uint32_t q = man >> shift;
        const uint32_t r = man & ((1u << shift) - 1u);
        const uint32_t h = 1u << (shift - 1);
        q += (r > h || (r == h && (q & 1u))) ? 1u : 0u;
        return uint16_t(sign | q);
    }
    uint32_t q = (uint32_t(e) << 10) | (man >> 13);
    const uint32_t r = man & 0x1fffu;
    q += (r > 0x1000u || (r == 0x1000u && (q & 1u))) ? 1u : 0u;
    return uint16_t(sign | q);
}

float f16_bits_to_f32(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1fu;
    const uint32_t man = h & 0x3ffu;
    float f;
    if (exp == 0) {
        f = std::ldexp(float(man), -24);
        uint32_t b;
        std::memcpy(&b, &f, 4);
        b |= sign;
        std::memcpy(&f, &b, 4);
        return f;
    }
    uint32_t b = exp == 31 ? (sign | 0x7f800000u | (man << 13)) : (sign | ((exp + 112u) << 23) | (man << 13));
    std::memcpy(&f, &b, 4);
    return f;
}

// transfer::hu_to_unorm16 (transfer.rs:121-124).
uint16_t hu_to_unorm16(int16_t hu) {
    const float t = rust_clamp((float(hu) - kHuMin) / (kHuMax - kHuMin), 0.0f, 1.0f);
    return uint16_t(std::round(t * 65535.0f));
}

TfTables build_tf_tables(const vvr_tf& tf) {
    TfTables t;
    t.lut.resize(kLutSize * 4);
    // to_lut (transfer.rs:90-97): entry i at HU_MIN + i / 4095 * (HU_MAX - HU_MIN).
    for (int i = 0; i < kLutSize; i++) {
        const float s = float(i) / float(kLutSize - 1);
        tf_evaluate(tf, kHuMin + s * (kHuMax - kHuMin), &t.lut[4 * i]);
    }
    t.lut_f16.resize(kLutSize * 4);
    for (int i = 0; i < kLutSize * 4; i++) t.lut_f16[i] = f16_bits_to_f32(f32_to_f16_bits(t.lut[i]));
    // cumulative_integrals (transfer.rs:103-117): trapezoid rule in f64.
    t.preint.assign(kLutSize * 4, 0.0f);
    const double dx = 1.0 / double(kLutSize - 1);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 1; i < kLutSize; i++) {
        const float* a = &t.lut[4 * (i - 1)];
        const float* b = &t.lut[4 * i];
        const double wa[4] = {double(a[0] * a[3]), double(a[1] * a[3]), double(a[2] * a[3]), double(a[3])};
        const double wb[4] = {double(b[0] * b[3]), double(b[1] * b[3]), double(b[2] * b[3]), double(b[3])};
        for (int c = 0; c < 4; c++) {
            acc[c] += 0.5 * (wa[c] + wb[c]) * dx;
            t.preint[4 * i + c] = float(acc[c]);
        }
    }
    // bucket_max (renderer.rs:2164-2174): each 1/64 of the LUT plus the next bucket's first entry.
    const int per = kLutSize / 64;
    for (int b = 0; b < 64; b++) {
        const int end = std::min((b + 1) * per + 1, kLutSize);
        float m = 0.0f;
        for (int i = b * per; i < end; i++) m = std::max(m, t.lut[4 * i + 3]);
        t.bucket_max[b] = m;
    }
    return t;
}

}  // namespace vvr

using namespace vvr;

extern "C" uint16_t vvr_hu_to_unorm16(int16_t hu) { return hu_to_unorm16(hu); }

extern "C" void vvr_tf_evaluate(const vvr_tf* tf, float hu, float rgba[4]) {
    if (!tf || !tf->n_points || !tf->points || !rgba) return;
    tf_evaluate(*tf, hu, rgba);
}

extern "C" int vvr_tf_lut(const vvr_tf* tf, float* out) {
    if (!tf || !tf->n_points || !tf->points || !out) return VVR_EINVAL;
    TfTables t = build_tf_tables(*tf);
    std::memcpy(out, t.lut.data(), t.lut.size() * sizeof(float));
    return VVR_OK;
}

extern "C" int vvr_tf_lut_f16(const vvr_tf* tf, uint16_t* out) {
    if (!tf || !tf->n_points || !tf->points || !out) return VVR_EINVAL;
    TfTables t = build_tf_tables(*tf);
    for (size_t i = 0; i < t.lut.size(); i++) out[i] = f32_to_f16_bits(t.lut[i]);
    return VVR_OK;
}

extern "C" int vvr_tf_bucket_max(const vvr_tf* tf, float* out) {
    if (!tf || !tf->n_points || !tf->points || !out) return VVR_EINVAL;
    TfTables t = build_tf_tables(*tf);
    std::memcpy(out, t.bucket_max, sizeof(t.bucket_max));
    return VVR_OK;
}

extern "C" int vvr_tf_cumulative_integrals(const vvr_tf* tf, float* out) {
    if (!tf || !tf->n_points || !tf->points || !out) return VVR_EINVAL;
    TfTables t = build_tf_tables(*tf);
    std::memcpy(out, t.preint.data(), t.preint.size() * sizeof(float));
    return VVR_OK;
}

extern "C" int vvr_tf_preset(uint32_t index, vvr_control_point* points, uint32_t cap) {
    const auto& p = presets();
    if (index >= p.size()) return VVR_EINVAL;
    const auto& pts = p[index].pts;
    uint32_t n = std::min<uint32_t>(cap, uint32_t(pts.size()));
    if (points)
        for (uint32_t i = 0; i < n; i++) points[i] = pts[i];
    return int(pts.size() <= cap ? pts.size() : n);
}
