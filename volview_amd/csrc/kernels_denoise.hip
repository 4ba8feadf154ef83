// Edge-avoiding a-trous denoiser of the progressive path tracer (denoise.comp,
// run as Renderer::record_denoise does, renderer.rs:1868-1904):
//
//   k_dn_prepare  irradiance (mean radiance / mean albedo) and its variance
//                 (denoise.comp:63-80), plus each pixel's mean albedo and unit
//                 normal + depth, computed once here instead of at every tap
//                 of every level (the same values: a function of the pixel)
//   k_dn_filter   one level: 5x5 B3-spline taps `step` pixels apart, weighted
//                 by normal, depth, albedo and luminance similarity
//                 (denoise.comp:91-139); levels at steps 1, 2, 4, 8, 16
//   k_dn_finish   irradiance x mean albedo and the background share into the
//                 RGBA16F image (denoise.comp:84-89)
//
// Images are row-major float4 of the frame. Workgroups are 16 x 16 pixel tiles;
// the prepare pass stages the tile's luminances with a 2-pixel apron in LDS.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "launch.hpp"

namespace vvr {

namespace {

constexpr int kTile = 16;
constexpr int kLevels = 5;  // DENOISE_LEVELS (renderer.rs:32-33)

__device__ __forceinline__ float lum3(float r, float g, float b) { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }
__device__ __forceinline__ int clampi(int v, int hi) { return min(max(v, 0), hi); }

struct DnImages {
    const float4* accum;
    const float4* albedo;
    const float4* normal;
    float4* ping;
    float4* pong;
    float4* g_nd;   // unit normal (0 where undefined) and depth, per pixel
    float4* g_alb;  // mean albedo floored at 0.02, per pixel
    uint16_t* hdr;
    int w, h;
    float n;  // samples accumulated
};

__global__ void __launch_bounds__(256) k_dn_prepare(DnImages D) {
    __shared__ float L[kTile + 4][kTile + 4];  // luminance of the mean radiance, 2-pixel apron
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
    for (int i = threadIdx.x; i < (kTile + 4) * (kTile + 4); i += 256) {
        const int lx = i % (kTile + 4), ly = i / (kTile + 4);
        const int qx = clampi(x0 + lx - 2, D.w - 1), qy = clampi(y0 + ly - 2, D.h - 1);
        const float4 a = D.accum[size_t(qy) * D.w + qx];
        L[ly][lx] = lum3(a.x / D.n, a.y / D.n, a.z / D.n);
    }
    __syncthreads();
    const int x = x0 + tx, y = y0 + ty;
    if (x >= D.w || y >= D.h) return;
    const size_t p = size_t(y) * D.w + x;
    const float4 a = D.accum[p], g = D.albedo[p], nd = D.normal[p];
    const float ar = fmaxf(g.x / D.n, 0.02f), ag = fmaxf(g.y / D.n, 0.02f), ab = fmaxf(g.z / D.n, 0.02f);
    const float rr = a.x / D.n, rg = a.y / D.n, rb = a.z / D.n;
    const float l = lum3(rr, rg, rb);
    const float temporal = fmaxf(g.w / D.n - l * l, 0.0f) / D.n;
    float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 5; dy++)
#pragma unroll
        for (int dx = 0; dx < 5; dx++) {
            const float lq = L[ty + dy][tx + dx];
            s1 += lq;
            s2 += lq * lq;
        }
    const float spatial = fmaxf(s2 / 25.0f - (s1 / 25.0f) * (s1 / 25.0f), 0.0f);
    const float variance = fmaxf(temporal, spatial * fminf(1.0f, 4.0f / D.n));
    const float al = lum3(ar, ag, ab);
    D.ping[p] = make_float4(rr / ar, rg / ag, rb / ab, variance / (al * al));
    D.g_alb[p] = make_float4(ar, ag, ab, 0.0f);
    const float gx = nd.x / D.n, gy = nd.y / D.n, gz = nd.z / D.n;
    const float len = sqrtf(gx * gx + gy * gy + gz * gz);
    D.g_nd[p] = len > 1e-3f ? make_float4(gx / len, gy / len, gz / len, nd.w / D.n)
                            : make_float4(0.0f, 0.0f, 0.0f, nd.w / D.n);
}

__global__ void __launch_bounds__(256) k_dn_filter(DnImages D, const float4* __restrict__ in, float4* __restrict__ out,
                                                   int step) {
    const int x = blockIdx.x * kTile + int(threadIdx.x % kTile), y = blockIdx.y * kTile + int(threadIdx.x / kTile);
    if (x >= D.w || y >= D.h) return;
    const size_t p = size_t(y) * D.w + x;
    const float4 center = in[p], cg = D.g_nd[p], ca = D.g_alb[p];
    const float l_center = lum3(center.x, center.y, center.z);
    // The variance steering the luminance weight, blurred 3x3 (clamped to the image).
    float bv = 0.0f;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++)
            bv += in[size_t(clampi(y + dy, D.h - 1)) * D.w + clampi(x + dx, D.w - 1)].w *
                  float((2 - abs(dx)) * (2 - abs(dy))) / 16.0f;
    const float sigma_l = 4.0f * sqrtf(fmaxf(bv, 0.0f)) + 1e-4f;
    const float kernel[3] = {3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    const bool c_flat = cg.x * cg.x + cg.y * cg.y + cg.z * cg.z < 0.5f;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f, var = 0.0f, weights = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; dy++)
#pragma unroll
        for (int dx = -2; dx <= 2; dx++) {
            const int qx = x + dx * step, qy = y + dy * step;
            if (qx < 0 || qy < 0 || qx >= D.w || qy >= D.h) continue;
            const size_t q = size_t(qy) * D.w + qx;
            const float4 c = in[q], g = D.g_nd[q], qa = D.g_alb[q];
            const float dn = cg.x * g.x + cg.y * g.y + cg.z * g.z;
            float w_normal = powf(fmaxf(dn, 0.0f), 64.0f);
            const bool q_flat = g.x * g.x + g.y * g.y + g.z * g.z < 0.5f;
            if (c_flat || q_flat) w_normal = (c_flat && q_flat) ? 1.0f : 0.0f;
            const float w_depth = expf(-fabsf(cg.w - g.w) / (0.02f * cg.w + 1e-3f));
            const float da_r = ca.x - qa.x, da_g = ca.y - qa.y, da_b = ca.z - qa.z;
            const float w_albedo = expf(-(da_r * da_r + da_g * da_g + da_b * da_b) / 0.01f);
            const float w_lum = expf(-fabsf(l_center - lum3(c.x, c.y, c.z)) / sigma_l);
            const float wt = kernel[abs(dx)] * kernel[abs(dy)] * w_normal * w_depth * w_albedo * w_lum;
            sr += wt * c.x;
            sg += wt * c.y;
            sb += wt * c.z;
            var += wt * wt * c.w;
            weights += wt;
        }
    out[p] = weights > 0.0f ? make_float4(sr / weights, sg / weights, sb / weights, var / (weights * weights)) : center;
}

__global__ void __launch_bounds__(256) k_dn_finish(DnImages D, const float4* __restrict__ in) {
    const int x = blockIdx.x * kTile + int(threadIdx.x % kTile), y = blockIdx.y * kTile + int(threadIdx.x / kTile);
    if (x >= D.w || y >= D.h) return;
    const size_t p = size_t(y) * D.w + x;
    const float4 v = in[p], a = D.g_alb[p];
    const float bg = D.accum[p].w / D.n;
    const __half h[4] = {__float2half_rn(v.x * a.x), __float2half_rn(v.y * a.y), __float2half_rn(v.z * a.z),
                         __float2half_rn(bg)};
    uint2 word;
    word.x = uint32_t(__half_as_ushort(h[0])) | (uint32_t(__half_as_ushort(h[1])) << 16);
    word.y = uint32_t(__half_as_ushort(h[2])) | (uint32_t(__half_as_ushort(h[3])) << 16);
    reinterpret_cast<uint2*>(D.hdr)[p] = word;
}

}  // namespace

void launch_denoise(const float4* accum, const float4* albedo, const float4* normal, float4* scratch, uint16_t* hdr,
                    uint32_t width, uint32_t height, float samples, hipStream_t st) {
    const size_t np = size_t(width) * height;
    DnImages D{accum, albedo, normal, scratch, scratch + np, scratch + 2 * np, scratch + 3 * np, hdr,
               int(width), int(height), samples};
    const dim3 grid((width + kTile - 1) / kTile, (height + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_dn_prepare, grid, dim3(256), 0, st, D);
    float4* bufs[2] = {D.ping, D.pong};  // level k reads bufs[k % 2], writes the other
    for (int level = 0; level < kLevels; level++)
        hipLaunchKernelGGL(k_dn_filter, grid, dim3(256), 0, st, D, bufs[level % 2], bufs[(level + 1) % 2], 1 << level);
    hipLaunchKernelGGL(k_dn_finish, grid, dim3(256), 0, st, D, bufs[kLevels % 2]);
}

}  // namespace vvr
