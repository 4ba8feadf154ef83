// Test infrastructure only: the CPU oracle of the volume-rendering hot path.
// Nothing in the shipped library includes, links or calls this.
//
// Scalar and vector helpers with GLSL semantics (GLSL 4.60 spec, section 8),
// templated on the real type R so one restatement serves as the FP32 oracle
// (R = float, strict IEEE: built without FMA contraction) and the FP64 oracle.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>

namespace oracle {

template <class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a) : x(a), y(a), z(a) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> operator+(V3<R> a, R s) { return {a.x + s, a.y + s, a.z + s}; }
template <class R> inline V3<R> operator-(V3<R> a, R s) { return {a.x - s, a.y - s, a.z - s}; }
template <class R> inline V3<R> operator-(R s, V3<R> a) { return {s - a.x, s - a.y, s - a.z}; }
template <class R> inline V3<R> operator+(R s, V3<R> a) { return {s + a.x, s + a.y, s + a.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }

// GLSL min/max: min(x, y) = y < x ? y : x.
template <class R> inline R gmin(R a, R b) { return b < a ? b : a; }
template <class R> inline R gmax(R a, R b) { return a < b ? b : a; }
template <class R> inline R clampr(R x, R lo, R hi) { return gmin(gmax(x, lo), hi); }
// GLSL's transcendental built-ins on the render path. GLSL leaves their precision
// to the implementation (log 3 ulp, exp 3 + 2|x| ulp). The FP32 oracle fixes one
// model for each, which the kernels repeat operation for operation:
//  - log and exp: float algorithms. Cody and Waite's reduction to a mantissa m in
//    [sqrt(1/2), sqrt(2)), log(m) = 2 atanh(s) with s = (m - 1) / (m + 1) by its
//    series to s^9 (<= 2 ulp); exp by reduction to r = x - k ln 2 and the Taylor
//    polynomial of e^r to r^7 (<= 1 ulp), scaled by two exact powers of two;
//  - pow(x, y) = exp(y log x) with those two (GLSL defines it through exp2 and log2);
//  - sin, cos, acos, atan2: evaluated in double and rounded once (the correctly
//    rounded value, barring ties closer than a double ulp).
// The FP64 oracle uses the double functions throughout.
inline float f32_from_bits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline uint32_t f32_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
constexpr float kLn2Hi = 0.693145751953125f;  // ln 2 with 11 trailing zero bits
constexpr float kLn2Lo = 1.42860677e-06f;     // ln 2 - kLn2Hi

inline float vlog_f32(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    int e = 0;
    if (x < 1.17549435e-38f) x *= 16777216.0f, e = -24;
    const uint32_t b = f32_bits(x);
    e += int(b >> 23) - 127;
    float m = f32_from_bits((b & 0x7fffffu) | 0x3f800000u);
    if (m > 1.41421356f) m = m * 0.5f, e += 1;
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    const float poly = z * (0.666666667f + z * (0.4f + z * (0.285714286f + z * 0.222222222f)));
    const float lm = (s + s) + s * poly;
    const float fe = float(e);
    return fe * kLn2Hi + (fe * kLn2Lo + lm);
}

inline float vexp_f32(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return INFINITY;
    if (x < -103.972084f) return 0.0f;
    const float kf = std::rint(x * 1.44269504f);
    const float r = (x - kf * kLn2Hi) - kf * kLn2Lo;
    const float p =
        1.0f + r * (1.0f + r * (0.5f + r * (0.166666667f + r * (0.0416666667f + r * (0.00833333333f +
                                                                                      r * (0.00138888889f +
                                                                                           r * 0.000198412698f))))));
    const int k = int(kf), k1 = k / 2, k2 = k - k1;
    return p * f32_from_bits(uint32_t(k1 + 127) << 23) * f32_from_bits(uint32_t(k2 + 127) << 23);
}

template <class R> inline R vlog(R x) {
    if constexpr (std::is_same_v<R, float>) return vlog_f32(x);
    else return std::log(x);
}
template <class R> inline R vexp(R x) {
    if constexpr (std::is_same_v<R, float>) return vexp_f32(x);
    else return std::exp(x);
}
// GLSL's pow (exp2(y log2 x), undefined for x < 0) as exp(y log x) of the base
// clamped at 0: a base that rounding left just below 0 (1 - dot(v, h) of unit
// vectors) gives 0, not NaN.
template <class R> inline R vpow(R x, R y) {
    const R b = x > R(0) ? x : R(0);
    if constexpr (std::is_same_v<R, float>) return vexp_f32(y * vlog_f32(b));
    else return std::pow(b, y);
}
template <class R> inline R cr_sin(R x) { return R(std::sin(double(x))); }
template <class R> inline R cr_cos(R x) { return R(std::cos(double(x))); }
template <class R> inline R cr_acos(R x) { return R(std::acos(double(x))); }
template <class R> inline R cr_atan2(R y, R x) { return R(std::atan2(double(y), double(x))); }

template <class R> inline R mixr(R a, R b, R t) { return a * (R(1) - t) + b * t; }
template <class R> inline R fract(R x) { return x - std::floor(x); }
template <class R> inline R smoothstep(R e0, R e1, R x) {
    R t = clampr((x - e0) / (e1 - e0), R(0), R(1));
    return t * t * (R(3) - R(2) * t);
}
template <class R> inline V3<R> vmin(V3<R> a, V3<R> b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
template <class R> inline V3<R> vmax(V3<R> a, V3<R> b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
template <class R> inline V3<R> vclamp(V3<R> a, R lo, R hi) {
    return {clampr(a.x, lo, hi), clampr(a.y, lo, hi), clampr(a.z, lo, hi)};
}
template <class R> inline V3<R> vmix(V3<R> a, V3<R> b, R t) {
    return {mixr(a.x, b.x, t), mixr(a.y, b.y, t), mixr(a.z, b.z, t)};
}
template <class R> inline V3<R> vmix(V3<R> a, V3<R> b, V3<R> t) {
    return {mixr(a.x, b.x, t.x), mixr(a.y, b.y, t.y), mixr(a.z, b.z, t.z)};
}
template <class R> inline V3<R> vfloor(V3<R> a) { return {std::floor(a.x), std::floor(a.y), std::floor(a.z)}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return a / length(a); }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class R> inline R max3(V3<R> a) { return gmax(a.x, gmax(a.y, a.z)); }
template <class R> inline V3<R> reflect(V3<R> i, V3<R> n) { return i - R(2) * dot(n, i) * n; }

// Column-major 3x3 (the upper-left block of a GLSL mat4).
template <class R> struct M3 {
    V3<R> c[3];
    V3<R> mul(V3<R> v) const {  // M * v
        return V3<R>(c[0].x * v.x + c[1].x * v.y + c[2].x * v.z, c[0].y * v.x + c[1].y * v.y + c[2].y * v.z,
                     c[0].z * v.x + c[1].z * v.y + c[2].z * v.z);
    }
    V3<R> mul_t(V3<R> v) const {  // transpose(M) * v
        return V3<R>(dot(c[0], v), dot(c[1], v), dot(c[2], v));
    }
};

// IEEE binary16 <-> binary32, round to nearest even (the `half` crate's from_f32).
inline uint16_t f32_to_f16(float f) {
    uint32_t x;
    __builtin_memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) return uint16_t(sign | 0x7c00u | (man ? 0x200u | (man >> 13) : 0u));
    int e = int(exp) - 127 + 15;
    if (e >= 31) return uint16_t(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return uint16_t(sign);
        man |= 0x800000u;
        int shift = 14 - e;
        uint32_t half_man = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (half_man & 1u))) half_man++;
        return uint16_t(sign | half_man);
    }
    uint32_t h = sign | (uint32_t(e) << 10) | (man >> 13);
    uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return uint16_t(h);
}

inline float f16_to_f32(uint16_t h) {
    uint32_t sign = uint32_t(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    uint32_t x;
    if (exp == 0) {
        if (man == 0) {
            x = sign;
        } else {
            int e = -1;
            do {
                e++;
                man <<= 1;
            } while (!(man & 0x400u));
            x = sign | uint32_t(127 - 15 - e) << 23 | ((man & 0x3ffu) << 13);
        }
    } else if (exp == 31) {
        x = sign | 0x7f800000u | (man << 13);
    } else {
        x = sign | ((exp + 112u) << 23) | (man << 13);
    }
    float f;
    __builtin_memcpy(&f, &x, 4);
    return f;
}

}  // namespace oracle
