// Small float vector type for the kernels, with GLSL semantics for min/max/mix.
#pragma once
#include <hip/hip_runtime.h>

namespace vvr {

struct V3 {
    float x, y, z;
    __device__ __forceinline__ V3() : x(0.f), y(0.f), z(0.f) {}
    __device__ __forceinline__ explicit V3(float a) : x(a), y(a), z(a) {}
    __device__ __forceinline__ V3(float a, float b, float c) : x(a), y(b), z(c) {}
    __device__ __forceinline__ float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 operator/(V3 a, V3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ V3 operator+(V3 a, float s) { return V3(a.x + s, a.y + s, a.z + s); }
__device__ __forceinline__ V3 operator-(V3 a, float s) { return V3(a.x - s, a.y - s, a.z - s); }
__device__ __forceinline__ V3 operator-(float s, V3 a) { return V3(s - a.x, s - a.y, s - a.z); }
__device__ __forceinline__ V3 operator+(float s, V3 a) { return V3(s + a.x, s + a.y, s + a.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }

__device__ __forceinline__ float gmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float gmax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
// GLSL mix, a (1 - t) + b t, with its one contraction pinned: b t is fused into
// the sum. (Accumulating c0 w0 + c1 w1 + ... in this form, with w = 1 - t and t
// at two taps and 0 elsewhere, gives exactly mix of the two taps.)
// GLSL mix (x (1 - a) + y a), each operation rounded on its own: the library is
// built without floating-point contraction, so this is never fused.
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
// The transcendental built-ins of the render path. GLSL leaves their precision to
// the implementation (log 3 ulp, exp 3 + 2|x| ulp); here every one has a single
// definition that the CPU oracle repeats operation for operation, so both give
// the same bits and the path tracer's random walks take the same decisions.
//
// log and exp, the hot ones (every delta-tracking step, every ray-marching
// sample), are float algorithms: Cody and Waite's reduction to a mantissa in
// [sqrt(1/2), sqrt(2)) with log(m) = 2 atanh(s), s = (m - 1) / (m + 1), by its
// series to s^9 (<= 2 ulp); exp by reduction to r = x - k ln 2, |r| <= ln 2 / 2,
// and its Taylor polynomial to r^7 (<= 1 ulp), scaled by two exact powers of two.
// pow is exp(y log x) with those two. sin, cos, acos and atan2 are evaluated in
// double and rounded once to float: the correctly rounded value except within a
// double ulp of a float rounding boundary.
constexpr float kLn2Hi = 0.693145751953125f;  // ln 2 with 11 trailing zero bits: k ln2_hi is exact
constexpr float kLn2Lo = 1.42860677e-06f;     // ln 2 - kLn2Hi

__device__ __forceinline__ float vlog(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    int e = 0;
    if (x < 1.17549435e-38f) x *= 16777216.0f, e = -24;  // subnormal: scale by 2^24
    const uint32_t b = __float_as_uint(x);
    e += int(b >> 23) - 127;
    float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
    if (m > 1.41421356f) m = m * 0.5f, e += 1;
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    const float poly = z * (0.666666667f + z * (0.4f + z * (0.285714286f + z * 0.222222222f)));
    const float lm = (s + s) + s * poly;
    const float fe = float(e);
    return fe * kLn2Hi + (fe * kLn2Lo + lm);
}

__device__ __forceinline__ float vexp(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return __builtin_inff();
    if (x < -103.972084f) return 0.0f;
    const float kf = rintf(x * 1.44269504f);
    const float r = (x - kf * kLn2Hi) - kf * kLn2Lo;
    const float p =
        1.0f + r * (1.0f + r * (0.5f + r * (0.166666667f + r * (0.0416666667f + r * (0.00833333333f +
                                                                                      r * (0.00138888889f +
                                                                                           r * 0.000198412698f))))));
    const int k = int(kf), k1 = k / 2, k2 = k - k1;
    return p * __uint_as_float(uint32_t(k1 + 127) << 23) * __uint_as_float(uint32_t(k2 + 127) << 23);
}

// GLSL defines pow(x, y) as exp2(y log2(x)), undefined for x < 0; here exp(y
// log(x)) with the two above, of the base clamped at 0: a base that rounding
// left just below 0 (1 - dot(v, h) of unit vectors) gives 0, not NaN.
__device__ __forceinline__ float vpow(float x, float y) { return vexp(y * vlog(x > 0.0f ? x : 0.0f)); }
__device__ __forceinline__ float cr_sin(float x) { return float(sin(double(x))); }
__device__ __forceinline__ float cr_cos(float x) { return float(cos(double(x))); }
__device__ __forceinline__ float cr_acos(float x) { return float(acos(double(x))); }
__device__ __forceinline__ float cr_atan2(float y, float x) { return float(atan2(double(y), double(x))); }

__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return V3(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return V3(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }
__device__ __forceinline__ V3 vclamp(V3 a, float lo, float hi) {
    return V3(clampf(a.x, lo, hi), clampf(a.y, lo, hi), clampf(a.z, lo, hi));
}
__device__ __forceinline__ V3 vmix(V3 a, V3 b, float t) { return V3(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)); }
__device__ __forceinline__ V3 vfloor(V3 a) { return V3(floorf(a.x), floorf(a.y), floorf(a.z)); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a / length(a); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ V3 reflect(V3 i, V3 n) { return i - 2.0f * dot(n, i) * n; }
__device__ __forceinline__ float component(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// Column-major 3x3.
struct M3 {
    V3 c0, c1, c2;
    __device__ __forceinline__ V3 mul(V3 v) const {
        return V3(c0.x * v.x + c1.x * v.y + c2.x * v.z, c0.y * v.x + c1.y * v.y + c2.y * v.z,
                  c0.z * v.x + c1.z * v.y + c2.z * v.z);
    }
    __device__ __forceinline__ V3 mul_t(V3 v) const { return V3(dot(c0, v), dot(c1, v), dot(c2, v)); }
};

constexpr float kPi = 3.14159265358979f;

}  // namespace vvr
