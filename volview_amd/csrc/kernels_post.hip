// Camera finishing of the HDR image into the RGBA8 output (Renderer::record_post,
// renderer.rs:1815-1866): bloom.comp's seven passes and tonemap.comp.
//
//   k_bloom_down4   quarter-resolution Karis average of 4x4 HDR pixels, with
//                   the covered share (pass 0)
//   k_meter         auto exposure: the covered pixels' mean log luminance,
//                   adapted 15 % a frame in log space (pass 1). Lane t of one
//                   wave sums texels t, t + 64, ... in order, then a tree, as the
//                   reference's 64-invocation group does; the whole workgroup
//                   computes the terms into LDS ahead of the sums.
//   k_blur9         9-tap Gaussian along x or y (passes 2, 3, 5, 6)
//   k_bloom_down2   eighth-resolution 2x2 average (pass 4)
//   k_tonemap       exposure, bloom mix, vignette, ACES / AgX, background,
//                   sRGB, film grain; RGBA8 store rounding to nearest
//
// The glow images are RGBA16F (8 bytes a texel) as in the reference; the
// adapted exposure is one float that persists across frames.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "device/scene.hpp"
#include "launch.hpp"

namespace vvr {

namespace {

__device__ __forceinline__ float lumf(float r, float g, float b) { return (r * 0.2126f + g * 0.7152f) + b * 0.0722f; }
__device__ __forceinline__ float hbits(uint32_t b) { return __half2float(__ushort_as_half((unsigned short)b)); }

__device__ __forceinline__ float4 load_h4(const uint2* img, size_t i) {
    const uint2 w = img[i];
    return make_float4(hbits(w.x & 0xffffu), hbits(w.x >> 16), hbits(w.y & 0xffffu), hbits(w.y >> 16));
}
__device__ __forceinline__ void store_h4(uint2* img, size_t i, float a, float b, float c, float d) {
    uint2 w;
    w.x = uint32_t(__half_as_ushort(__float2half_rn(a))) | (uint32_t(__half_as_ushort(__float2half_rn(b))) << 16);
    w.y = uint32_t(__half_as_ushort(__float2half_rn(c))) | (uint32_t(__half_as_ushort(__float2half_rn(d))) << 16);
    img[i] = w;
}

// bloom.comp:104-122
__global__ void __launch_bounds__(64) k_bloom_down4(const uint2* __restrict__ hdr, int w, int h, uint2* __restrict__ a,
                                                    int qw, int qh) {
    const int x = blockIdx.x * 8 + int(threadIdx.x & 7u), y = blockIdx.y * 8 + int(threadIdx.x >> 3);
    if (x >= qw || y >= qh) return;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f, weights = 0.0f, coverage = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float4 p = load_h4(hdr, size_t(min(y * 4 + j, h - 1)) * w + min(x * 4 + i, w - 1));
            const float wt = 1.0f / (1.0f + lumf(p.x, p.y, p.z));
            sr += p.x * wt, sg += p.y * wt, sb += p.z * wt;
            weights += wt;
            coverage += 1.0f - p.w;
        }
    store_h4(a, size_t(y) * qw + x, sr / weights, sg / weights, sb / weights, coverage / 16.0f);
}

// bloom.comp:60-98. The sums keep the reference's order: lane t of one 64-lane
// group adds texels t, t + 64, ... in turn. The terms (coverage x log luminance)
// do not depend on that order, so the workgroup's 16 waves compute a chunk of
// them into LDS, and the first wave then adds its lanes' runs from LDS.
constexpr int kMeterThreads = 1024, kMeterPer = 8, kMeterChunk = kMeterThreads * kMeterPer;
__global__ void __launch_bounds__(kMeterThreads) k_meter(const uint2* __restrict__ a, int n,
                                                         float* __restrict__ exposure) {
    __shared__ float2 terms[kMeterChunk];  // 64 KiB
    const int t = int(threadIdx.x);
    float s = 0.0f, wsum = 0.0f;
    for (int c0 = 0; c0 < n; c0 += kMeterChunk) {
#pragma unroll
        for (int k = 0; k < kMeterPer; k++) {
            const int j = k * kMeterThreads + t;
            float2 v = make_float2(0.0f, 0.0f);
            if (c0 + j < n) {
                const float4 c = load_h4(a, size_t(c0 + j));
                const float term = c.w * logf(fmaxf(lumf(c.x, c.y, c.z), 1e-4f));
                v = make_float2(term, c.w);
            }
            terms[j] = v;
        }
        __syncthreads();
        if (t < 64) {
            const int m = min(kMeterChunk, n - c0);
            for (int j = t; j < m; j += 64) {
                const float2 v = terms[j];
                s += v.x;
                wsum += v.y;
            }
        }
        __syncthreads();
    }
    if (t >= 64) return;
    // The tree within the first wave: lane t adds lane t + stride's sums.
    for (int stride = 32; stride > 0; stride >>= 1) {
        const float o = __shfl_down(s, stride), ow = __shfl_down(wsum, stride);
        if (t < stride) s += o, wsum += ow;
    }
    if (t == 0) {
        const float target = wsum > 1.0f ? fminf(fmaxf(0.24f / expf(s / wsum), 0.25f), 4.0f) : 1.0f;
        const float prev = *exposure;
        *exposure = prev <= 0.0f ? target : expf(mixf(logf(prev), logf(target), 0.15f));
    }
}

// bloom.comp:46-54
__global__ void __launch_bounds__(64) k_blur9(const uint2* __restrict__ src, uint2* __restrict__ dst, int w, int h,
                                              int along_x) {
    const int x = blockIdx.x * 8 + int(threadIdx.x & 7u), y = blockIdx.y * 8 + int(threadIdx.x >> 3);
    if (x >= w || y >= h) return;
    const float k[5] = {0.2270f, 0.1945f, 0.1216f, 0.0540f, 0.0162f};
    const float4 c = load_h4(src, size_t(y) * w + x);
    float r = c.x * k[0], g = c.y * k[0], b = c.z * k[0], a = c.w * k[0];
#pragma unroll
    for (int i = 1; i < 5; i++) {
        const int px = along_x ? min(x + i, w - 1) : x, py = along_x ? y : min(y + i, h - 1);
        const int mx = along_x ? max(x - i, 0) : x, my = along_x ? y : max(y - i, 0);
        const float4 p = load_h4(src, size_t(py) * w + px), m = load_h4(src, size_t(my) * w + mx);
        r += k[i] * (p.x + m.x), g += k[i] * (p.y + m.y), b += k[i] * (p.z + m.z), a += k[i] * (p.w + m.w);
    }
    store_h4(dst, size_t(y) * w + x, r, g, b, a);
}

// bloom.comp:126-131
__global__ void __launch_bounds__(64) k_bloom_down2(const uint2* __restrict__ a, int qw, int qh, uint2* __restrict__ c,
                                                    int ew, int eh) {
    const int x = blockIdx.x * 8 + int(threadIdx.x & 7u), y = blockIdx.y * 8 + int(threadIdx.x >> 3);
    if (x >= ew || y >= eh) return;
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float4 v = load_h4(a, size_t(min(y * 2 + (i >> 1), qh - 1)) * qw + min(x * 2 + (i & 1), qw - 1));
        s[0] += v.x, s[1] += v.y, s[2] += v.z, s[3] += v.w;
    }
    store_h4(c, size_t(y) * ew + x, s[0] * 0.25f, s[1] * 0.25f, s[2] * 0.25f, s[3] * 0.25f);
}

// The renderer's linear clamp-to-edge sampler on a glow image at uv.
__device__ __forceinline__ float4 bilinear_h4(const uint2* img, int w, int h, float u, float v) {
    const float tx = u * float(w) - 0.5f, ty = v * float(h) - 0.5f;
    const float fx = floorf(tx), fy = floorf(ty);
    const float ax = tx - fx, ay = ty - fy;
    const int x0 = int(fx), y0 = int(fy);
    const int xa = min(max(x0, 0), w - 1), xb = min(max(x0 + 1, 0), w - 1);
    const int ya = min(max(y0, 0), h - 1), yb = min(max(y0 + 1, 0), h - 1);
    const float4 p00 = load_h4(img, size_t(ya) * w + xa), p10 = load_h4(img, size_t(ya) * w + xb);
    const float4 p01 = load_h4(img, size_t(yb) * w + xa), p11 = load_h4(img, size_t(yb) * w + xb);
    auto m = [](float4 p, float4 q, float t) {
        return make_float4(mixf(p.x, q.x, t), mixf(p.y, q.y, t), mixf(p.z, q.z, t), mixf(p.w, q.w, t));
    };
    return m(m(p00, p10, ax), m(p01, p11, ax), ay);
}

__device__ __forceinline__ float hash01(uint32_t v) {  // tonemap.comp:13-17
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return float((word >> 22u) ^ word) * (1.0f / 4294967296.0f);
}

__device__ __forceinline__ float aces(float x) {  // tonemap.comp:20-22
    return fminf(fmaxf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f), 1.0f);
}

__device__ __forceinline__ V3 agx(V3 c) {  // tonemap.comp:25-42
    const V3 i0(0.842479062253094f, 0.0423282422610123f, 0.0423756549057051f);
    const V3 i1(0.0784335999999992f, 0.878468636469772f, 0.0784336f);
    const V3 i2(0.0792237451477643f, 0.0791661274605434f, 0.879142973793104f);
    const V3 o0(1.19687900512017f, -0.0528968517574562f, -0.0529716355144438f);
    const V3 o1(-0.0980208811401368f, 1.15190312990417f, -0.0980434501171241f);
    const V3 o2(-0.0990297440797205f, -0.0989611768448433f, 1.15107367264116f);
    const float min_ev = -12.47393f, max_ev = 4.026069f;
    const V3 t = i0 * c.x + i1 * c.y + i2 * c.z;
    float v[3] = {t.x, t.y, t.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float l = fminf(fmaxf(log2f(fmaxf(v[k], 1e-10f)), min_ev), max_ev);
        l = (l - min_ev) / (max_ev - min_ev);
        const float l2 = l * l, l4 = l2 * l2;
        v[k] = 15.5f * l4 * l2 - 40.14f * l4 * l + 31.96f * l4 - 6.868f * l2 * l + 0.4298f * l2 + 0.1191f * l - 0.00232f;
    }
    const V3 o = o0 * v[0] + o1 * v[1] + o2 * v[2];
    return V3(powf(fminf(fmaxf(o.x, 0.0f), 1.0f), 2.2f), powf(fminf(fmaxf(o.y, 0.0f), 1.0f), 2.2f),
              powf(fminf(fmaxf(o.z, 0.0f), 1.0f), 2.2f));
}

__device__ __forceinline__ float srgb(float c) {  // common.glsl:444-447
    c = fminf(fmaxf(c, 0.0f), 1.0f);
    return c < 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
}

__device__ __forceinline__ uint32_t unorm8(float c) { return uint32_t(floorf(fminf(fmaxf(c, 0.0f), 1.0f) * 255.0f + 0.5f)); }

// tonemap.comp:48-85
__global__ void __launch_bounds__(64) k_tonemap(const vvr_uniforms* __restrict__ u, const uint2* __restrict__ hdr,
                                                const uint2* __restrict__ a, int qw, int qh,
                                                const uint2* __restrict__ c8, int ew, int eh,
                                                const float* __restrict__ exposure, uint32_t* __restrict__ out) {
    const int w = int(u->flags[2]), h = int(u->flags[3]);
    const int x = blockIdx.x * 8 + int(threadIdx.x & 7u), y = blockIdx.y * 8 + int(threadIdx.x >> 3);
    if (x >= w || y >= h) return;
    const size_t p = size_t(y) * w + x;
    const float4 hp = load_h4(hdr, p);
    float ex = u->ambient[3];
    if (u->finish[3] > 0.5f) ex *= *exposure;
    V3 c(hp.x, hp.y, hp.z);
    const float uu = (float(x) + 0.5f) / float(w), vv = (float(y) + 0.5f) / float(h);
    if (u->finish[0] > 0.0f) {
        const float4 na = bilinear_h4(a, qw, qh, uu, vv), wd = bilinear_h4(c8, ew, eh, uu, vv);
        const V3 glow(0.6f * na.x + 0.4f * wd.x, 0.6f * na.y + 0.4f * wd.y, 0.6f * na.z + 0.4f * wd.z);
        c = vmix(c, glow, u->finish[0]);
    }
    const float dx = uu * 2.0f - 1.0f, dy = vv * 2.0f - 1.0f;
    const float r = sqrtf(dx * dx + dy * dy) * 0.7071f;
    const float vignette = 1.0f - u->finish[1] * powf(r, 2.5f);
    c = c * (ex * vignette);
    const uint32_t op = uint32_t(u->material[2]);
    if (op == 1u) c = V3(aces(c.x), aces(c.y), aces(c.z));
    else if (op == 2u) c = agx(c);
    else c = V3(fminf(fmaxf(c.x, 0.0f), 1.0f), fminf(fmaxf(c.y, 0.0f), 1.0f), fminf(fmaxf(c.z, 0.0f), 1.0f));
    if (u->lighting[3] < 1u) {  // background (common.glsl:449-455): display-referred
        const V3 d = camera_ray(u, float(x) + 0.5f, float(y) + 0.5f);
        const float t = 0.5f + 0.5f * -d.z;
        c = c + vmix(V3(0.02f, 0.022f, 0.028f), V3(0.08f, 0.085f, 0.1f), t) * hp.w;
    }
    if (u->flags[1] != 0u) c = V3(srgb(c.x), srgb(c.y), srgb(c.z));
    if (u->finish[2] > 0.0f) {
        const uint32_t px = uint32_t(x), py = uint32_t(y);
        const float n = hash01(px + 4099u * py) + hash01(px * 7919u + py + 1u) - 1.0f;
        const float yl = fminf(fmaxf(lumf(c.x, c.y, c.z), 0.0f), 1.0f);
        const float g = n * u->finish[2] * (0.3f + 2.8f * yl * (1.0f - yl));
        c = V3(fminf(fmaxf(c.x + g, 0.0f), 1.0f), fminf(fmaxf(c.y + g, 0.0f), 1.0f), fminf(fmaxf(c.z + g, 0.0f), 1.0f));
    }
    out[p] = unorm8(c.x) | (unorm8(c.y) << 8) | (unorm8(c.z) << 16) | (255u << 24);
}

dim3 tiles8(int w, int h) { return dim3((w + 7) / 8, (h + 7) / 8); }

}  // namespace

void launch_post(const vvr_uniforms* du, const uint16_t* hdr, uint32_t width, uint32_t height, uint16_t* bloom,
                 float* exposure, uint32_t* out, hipStream_t st) {
    const int w = int(width), h = int(height);
    const int qw = (w + 3) / 4, qh = (h + 3) / 4, ew = (qw + 1) / 2, eh = (qh + 1) / 2;
    uint2* A = reinterpret_cast<uint2*>(bloom);
    uint2* B = A + size_t(qw) * qh;
    uint2* C = B + size_t(qw) * qh;
    uint2* D = C + size_t(ew) * eh;
    const uint2* H = reinterpret_cast<const uint2*>(hdr);
    hipLaunchKernelGGL(k_bloom_down4, tiles8(qw, qh), dim3(64), 0, st, H, w, h, A, qw, qh);
    hipLaunchKernelGGL(k_meter, dim3(1), dim3(kMeterThreads), 0, st, A, qw * qh, exposure);
    hipLaunchKernelGGL(k_blur9, tiles8(qw, qh), dim3(64), 0, st, A, B, qw, qh, 1);
    hipLaunchKernelGGL(k_blur9, tiles8(qw, qh), dim3(64), 0, st, B, A, qw, qh, 0);
    hipLaunchKernelGGL(k_bloom_down2, tiles8(ew, eh), dim3(64), 0, st, A, qw, qh, C, ew, eh);
    hipLaunchKernelGGL(k_blur9, tiles8(ew, eh), dim3(64), 0, st, C, D, ew, eh, 1);
    hipLaunchKernelGGL(k_blur9, tiles8(ew, eh), dim3(64), 0, st, D, C, ew, eh, 0);
    hipLaunchKernelGGL(k_tonemap, tiles8(w, h), dim3(64), 0, st, du, H, A, qw, qh, C, ew, eh, exposure, out);
}

size_t bloom_bytes(uint32_t width, uint32_t height) {
    const size_t qw = (width + 3) / 4, qh = (height + 3) / 4, ew = (qw + 1) / 2, eh = (qh + 1) / 2;
    return (2 * qw * qh + 2 * ew * eh) * 8;
}

}  // namespace vvr
