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
__device__ __forceinline__ float mixf(float a, float b, float t) { return fmaf(b, t, a * (1.0f - t)); }
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
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
