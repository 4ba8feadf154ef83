// The learned denoiser's convolutional U-Net: bf16 MFMA forward and backward
// kernels with FP32 accumulation, and the elementwise kernels around them.
//
// Activations are NHWC bf16 (channels fastest, counts multiples of 16); a
// tensor argument is (pointer, channel stride, channel offset) so that
// concatenations are channel slices of one buffer. Weights of a 3x3 layer are
// bf16 [co_pad][9][ci] (co_pad = co rounded up to 32, zero rows), taps in
// row-major order of (dy, dx) in {-1, 0, 1}^2; zero padding at image borders.
//
//   k_conv3x3     out = act(conv(in) + bias): implicit GEMM, M = pixels, N = out
//                 channels, K = 9 x in channels, one v_mfma_f32_32x32x16_bf16 per
//                 16 K. A wave owns 32 consecutive pixels and up to 4 blocks of
//                 32 output channels; the A fragment (8 channels of one pixel per
//                 lane) is loaded once per K step and reused across the blocks.
//                 Also the input gradient, with flipped transposed weights and
//                 the mask of a ReLU'd input fused into the epilogue.
//   k_conv_wgrad  dW[co][tap][ci] += sum over pixels dY[p][co] x[p + tap][ci],
//                 db[co] += sum dY: M = out channels, N = in channels, K = pixels.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>

#include "launch.hpp"

namespace vvr {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

// Row of MFMA output register r for lane-half hi (0 or 1): D[8(r/4) + 4 hi + r%4][lane % 32].
__device__ __forceinline__ int drow(int r, int hi) { return 8 * (r >> 2) + 4 * hi + (r & 3); }

template <int NCOB>
__global__ void __launch_bounds__(256) k_conv3x3(UnetConv c) {
    const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
    const int col = lane & 31, hi = lane >> 5;
    const int npix = c.n * c.h * c.w;
    const int p = int(blockIdx.x) * 128 + wave * 32 + col;  // this lane's A row (pixel)
    const int cob0 = int(blockIdx.y) * NCOB;                 // first 32-channel output block
    int px = 0, py = 0, pn = 0;
    if (p < npix) {
        px = p % c.w;
        py = (p / c.w) % c.h;
        pn = p / (c.w * c.h);
    }
    f32x16 acc[NCOB];
#pragma unroll
    for (int b = 0; b < NCOB; b++) acc[b] = f32x16{};
    const int nb = min(NCOB, c.co_blocks - cob0);
    // K steps of 16 (tap-major, then input channels), each step's fragments
    // loaded one step ahead so that their latency overlaps the MFMAs.
    const int csteps = c.cin >> 4, total = 9 * csteps;
    int lt = 0, lc = 0;  // tap and channel step of the next load
    bf16x8 a_nx, w_nx[NCOB];
    auto load_step = [&]() {
        const int yy = py + lt / 3 - 1, xx = px + lt % 3 - 1;
        const bool inside = p < npix && yy >= 0 && yy < c.h && xx >= 0 && xx < c.w;
        const int ci = 16 * lc + 8 * hi;
        a_nx = inside ? load8(c.in + size_t((pn * c.h + yy) * c.w + xx) * c.in_cs + c.in_co + ci) : zero8();
#pragma unroll
        for (int b = 0; b < NCOB; b++)
            w_nx[b] = b < nb ? load8(c.wts + (size_t((cob0 + b) * 32 + col) * 9 + lt) * c.cin + ci) : zero8();
        if (++lc == csteps) lc = 0, lt++;
    };
    load_step();
    for (int s = 0; s < total; s++) {
        const bf16x8 a = a_nx;
        bf16x8 w[NCOB];
#pragma unroll
        for (int b = 0; b < NCOB; b++) w[b] = w_nx[b];
        if (s + 1 < total) load_step();
#pragma unroll
        for (int b = 0; b < NCOB; b++)
            if (b < nb) acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, w[b], acc[b], 0, 0, 0);
    }
    // Epilogue: lane holds output channel (cob0 + b) * 32 + col of 16 pixels.
    const int pbase = int(blockIdx.x) * 128 + wave * 32;
#pragma unroll
    for (int b = 0; b < NCOB; b++) {
        if (b >= nb) continue;
        const int co = (cob0 + b) * 32 + col;
        if (co >= c.cout) continue;
        const float bias = c.bias ? c.bias[co] : 0.0f;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int q = pbase + drow(r, hi);
            if (q >= npix) continue;
            float v = acc[b][r] + bias;
            if (c.relu) v = fmaxf(v, 0.0f);
            if (c.mask && bf2f(c.mask[size_t(q) * c.mask_cs + c.mask_co + co]) <= 0.0f) v = 0.0f;
            if (c.out) c.out[size_t(q) * c.out_cs + c.out_co + co] = f2bf(v);
            if (c.out_f32) c.out_f32[size_t(q) * c.cout + co] = v;
        }
    }
}

// k_conv3x3 over 2-D tiles of 8 x 32 output pixels (4 waves, each 2 rows of 32)
// for layers with many pixels. Per chunk of 32 input channels the workgroup
// stages in LDS the input tile with a one-pixel halo (10 x 34 pixels) and the
// chunk's weights for all 9 taps, then runs the 9 taps' MFMAs from LDS: each
// input pixel is fetched once per chunk instead of once per tap. LDS rows are
// padded to 80 bytes so that the 16-byte fragment reads of 8 lanes cover the 32
// banks once.
constexpr int kTileH = 8, kTileW = 32, kChunk = 32, kRow = kChunk + 8;  // kRow: bf16s per LDS row
// 16-byte loads per thread for a chunk's input tile and weights (4 per pixel or row).
constexpr int kInLoads = ((kTileH + 2) * (kTileW + 2) * (kChunk / 8) + 255) / 256;
template <int NCOB>
constexpr int kWtLoads = (9 * NCOB * 32 * (kChunk / 8) + 255) / 256;
constexpr int kTileIn = (kTileH + 2) * (kTileW + 2) * kRow;  // bf16s of the staged input tile

// The tile's input channels kc .. kc + 8 parts into registers: pixel
// (y0 - 1 + ty, x0 - 1 + tx), zero outside the image. All loads are issued
// before any is used, so their latencies overlap.
__device__ __forceinline__ void tile_input_load(const UnetConv& c, int n, int x0, int y0, int kc, int parts,
                                                uint4 (&v)[kInLoads]) {
    const int items = (kTileH + 2) * (kTileW + 2) * parts;
#pragma unroll
    for (int j = 0; j < kInLoads; j++) {
        const int k = int(threadIdx.x) + 256 * j;
        v[j] = make_uint4(0u, 0u, 0u, 0u);
        if (k < items) {
            const int pix = k / parts, part = k - pix * parts;
            const int ty = pix / (kTileW + 2), tx = pix - ty * (kTileW + 2);
            const int y = y0 - 1 + ty, x = x0 - 1 + tx;
            if (y >= 0 && y < c.h && x >= 0 && x < c.w)
                v[j] = *reinterpret_cast<const uint4*>(c.in + size_t((n * c.h + y) * c.w + x) * c.in_cs + c.in_co +
                                                       kc + part * 8);
        }
    }
}
__device__ __forceinline__ void tile_input_store(int parts, const uint4 (&v)[kInLoads], uint16_t* s_in) {
    const int items = (kTileH + 2) * (kTileW + 2) * parts;
#pragma unroll
    for (int j = 0; j < kInLoads; j++) {
        const int k = int(threadIdx.x) + 256 * j;
        if (k < items) {
            const int pix = k / parts, part = k - pix * parts;
            *reinterpret_cast<uint4*>(s_in + pix * kRow + part * 8) = v[j];
        }
    }
}

// The chunk's weights into LDS: row (tap, b * 32 + r) holds input channels kc .. kc + 8 parts.
template <int NCOB>
__device__ __forceinline__ void tile_weights_stage(const UnetConv& c, int cob0, int nb, int kc, int parts,
                                                   uint16_t* s_w) {
    const int items = 9 * NCOB * 32 * parts;
    uint4 v[kWtLoads<NCOB>];
#pragma unroll
    for (int j = 0; j < kWtLoads<NCOB>; j++) {
        const int k = int(threadIdx.x) + 256 * j;
        v[j] = make_uint4(0u, 0u, 0u, 0u);
        if (k < items) {
            const int row = k / parts, part = k - row * parts;
            const int tap = row / (NCOB * 32), r = row - tap * (NCOB * 32);
            if (r < nb * 32)
                v[j] = *reinterpret_cast<const uint4*>(c.wts + (size_t(cob0 * 32 + r) * 9 + tap) * c.cin + kc +
                                                       part * 8);
        }
    }
#pragma unroll
    for (int j = 0; j < kWtLoads<NCOB>; j++) {
        const int k = int(threadIdx.x) + 256 * j;
        if (k < items) {
            const int row = k / parts, part = k - row * parts;
            *reinterpret_cast<uint4*>(s_w + row * kRow + part * 8) = v[j];
        }
    }
}

// The 9 taps' MFMAs over cc (16 or 32) staged channels, unrolled (the second
// 16 channels behind a uniform test) so that the fragments' LDS reads can be
// issued ahead of the MFMAs using them. Wave w computes rows 2w, 2w + 1.
template <int NCOB>
__device__ __forceinline__ void tile_mfma(const uint16_t* s_in, const uint16_t* s_w, int nb, int cc,
                                          f32x16 (&acc)[2][NCOB]) {
    const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
    const int col = lane & 31, hi = lane >> 5;
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
        const int dy = tap / 3, dx = tap % 3;  // tile coordinates are shifted by the halo
#pragma unroll
        for (int ci = 0; ci < kChunk; ci += 16) {
            if (ci >= cc) break;
            bf16x8 a[2];
#pragma unroll
            for (int pb = 0; pb < 2; pb++) {
                const int ty = 2 * wave + pb + dy, tx = col + dx;
                a[pb] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(s_in + (ty * (kTileW + 2) + tx) * kRow + ci + 8 * hi));
            }
#pragma unroll
            for (int b = 0; b < NCOB; b++) {
                if (b < nb) {
                    const bf16x8 w = __builtin_bit_cast(
                        bf16x8,
                        *reinterpret_cast<const uint4*>(s_w + (tap * NCOB * 32 + b * 32 + col) * kRow + ci + 8 * hi));
#pragma unroll
                    for (int pb = 0; pb < 2; pb++)
                        acc[pb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[pb], w, acc[pb][b], 0, 0, 0);
                }
            }
        }
    }
}

// Epilogue: D row drow(r, hi) is pixel x0 + drow of row y0 + 2 wave + pb.
template <int NCOB>
__device__ __forceinline__ void tile_epilogue(const UnetConv& c, int n, int x0, int y0, int cob0, int nb,
                                              const f32x16 (&acc)[2][NCOB]) {
    const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
    const int col = lane & 31, hi = lane >> 5;
#pragma unroll
    for (int pb = 0; pb < 2; pb++) {
        const int y = y0 + 2 * wave + pb;
        if (y >= c.h) continue;
#pragma unroll
        for (int b = 0; b < NCOB; b++) {
            if (b >= nb) continue;
            const int co = (cob0 + b) * 32 + col;
            if (co >= c.cout) continue;
            const float bias = c.bias ? c.bias[co] : 0.0f;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int x = x0 + drow(r, hi);
                if (x >= c.w) continue;
                const size_t q = size_t(n * c.h + y) * c.w + x;
                float v = acc[pb][b][r] + bias;
                if (c.relu) v = fmaxf(v, 0.0f);
                if (c.mask && bf2f(c.mask[q * c.mask_cs + c.mask_co + co]) <= 0.0f) v = 0.0f;
                if (c.out) c.out[q * c.out_cs + c.out_co + co] = f2bf(v);
                if (c.out_f32) c.out_f32[q * c.cout + co] = v;
            }
        }
    }
}

template <int NCOB>
__global__ void __launch_bounds__(256) k_conv3x3_tile(UnetConv c, int co_groups) {
    __shared__ uint16_t s_in[kTileIn];
    __shared__ uint16_t s_w[9 * NCOB * 32 * kRow];
    const int x0 = int(blockIdx.x) * kTileW, y0 = int(blockIdx.y) * kTileH;
    const int n = int(blockIdx.z) / co_groups, cob0 = (int(blockIdx.z) % co_groups) * NCOB;
    const int nb = min(NCOB, c.co_blocks - cob0);
    f32x16 acc[2][NCOB];
#pragma unroll
    for (int pb = 0; pb < 2; pb++)
#pragma unroll
        for (int b = 0; b < NCOB; b++) acc[pb][b] = f32x16{};
    for (int kc = 0; kc < c.cin; kc += kChunk) {
        const int cc = min(kChunk, c.cin - kc);  // 16 or 32
        const int parts = cc >> 3;                // 16-byte parts per pixel or weight row
        uint4 vin[kInLoads];
        tile_input_load(c, n, x0, y0, kc, parts, vin);
        tile_weights_stage<NCOB>(c, cob0, nb, kc, parts, s_w);
        tile_input_store(parts, vin, s_in);
        __syncthreads();
        tile_mfma<NCOB>(s_in, s_w, nb, cc, acc);
        __syncthreads();  // before the next chunk overwrites the tiles
    }
    tile_epilogue<NCOB>(c, n, x0, y0, cob0, nb, acc);
}

// k_conv3x3_tile for layers of at most 32 input and 32 output channels (one
// chunk, one output block), with persistent workgroups: a workgroup stages the
// weights once, then walks tiles t = blockIdx.x, + gridDim.x, ... and loads the
// next tile's input into registers while the current tile's MFMAs and
// epilogue run. Tile t is column t % tiles_x, row (t / tiles_x) % tiles_y,
// image t / (tiles_x tiles_y).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_conv3x3_stream(UnetConv c, int tiles_x, int tiles_y, int n_tiles) {
    __shared__ uint16_t s_in[kTileIn];
    __shared__ uint16_t s_w[9 * 32 * kRow];
    int t = int(blockIdx.x);  // the grid has at most n_tiles workgroups
    const int cc = c.cin, parts = cc >> 3;
    tile_weights_stage<1>(c, 0, 1, 0, parts, s_w);
    auto origin = [&](int tile, int& n, int& x0, int& y0) {
        x0 = (tile % tiles_x) * kTileW;
        y0 = ((tile / tiles_x) % tiles_y) * kTileH;
        n = tile / (tiles_x * tiles_y);
    };
    int n, x0, y0;
    origin(t, n, x0, y0);
    uint4 vin[kInLoads];
    tile_input_load(c, n, x0, y0, 0, parts, vin);
    for (;;) {
        tile_input_store(parts, vin, s_in);  // every wave is past the last tile's MFMAs
        __syncthreads();
        const int next = t + int(gridDim.x);
        int nn = 0, nx0 = 0, ny0 = 0;
        if (next < n_tiles) {  // uniform
            origin(next, nn, nx0, ny0);
            tile_input_load(c, nn, nx0, ny0, 0, parts, vin);
        }
        f32x16 acc[2][1] = {{f32x16{}}, {f32x16{}}};
        tile_mfma<1>(s_in, s_w, 1, cc, acc);
        tile_epilogue<1>(c, n, x0, y0, 0, 1, acc);
        if (next >= n_tiles) break;
        __syncthreads();  // before the next tile overwrites s_in
        t = next, n = nn, x0 = nx0, y0 = ny0;
    }
}

// Weight gradient. One wave per (tap, 32 output channels, 32 input channels)
// and chunk of pixels; K = 16 pixels per MFMA, A = dY^T (lane: output channel,
// 8 consecutive pixels), B = shifted x (lane: input channel, 8 pixels).
__global__ void __launch_bounds__(64) k_conv_wgrad(UnetWgrad g) {
    const int lane = int(threadIdx.x), col = lane & 31, hi = lane >> 5;
    const int job = int(blockIdx.y);  // (tap, co block, ci block)
    const int tap = job % 9, cob = (job / 9) % g.co_blocks, cib = job / (9 * g.co_blocks);
    const int npix = g.n * g.h * g.w;
    const int p0 = int(blockIdx.x) * g.chunk, p1 = min(p0 + g.chunk, npix);
    const int co = cob * 32 + col, ci = cib * 32 + col;
    const bool co_ok = co < g.cout, ci_ok = ci < g.cin;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    f32x16 acc{};
    float bsum = 0.0f;
    for (int k = p0; k < p1; k += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int p = k + 8 * hi + j;
            uint16_t va = 0, vb = 0;
            if (p < p1) {
                if (co_ok) va = g.dy[size_t(p) * g.dy_cs + co];
                const int x = p % g.w, y = (p / g.w) % g.h, n = p / (g.w * g.h);
                const int xx = x + dx, yy = y + dy;
                if (ci_ok && xx >= 0 && xx < g.w && yy >= 0 && yy < g.h)
                    vb = g.x[size_t((n * g.h + yy) * g.w + xx) * g.x_cs + g.x_co + ci];
            }
            a[j] = __builtin_bit_cast(__bf16, va);
            b[j] = __builtin_bit_cast(__bf16, vb);
            if (tap == 0 && cib == 0) bsum += bf2f(va);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    // D[row = out channel][col = in channel].
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int o = cob * 32 + drow(r, hi);
        if (o < g.cout && ci_ok) atomicAdd(&g.dw[(size_t(o) * 9 + tap) * g.cin + ci], acc[r]);
    }
    if (tap == 0 && cib == 0 && co_ok) {
        bsum += __shfl_xor(bsum, 32);
        if (hi == 0) atomicAdd(&g.db[co], bsum);
    }
}

// Flipped transposed weights for the input gradient: wt[ci][8 - tap][co] = w[co][tap][ci],
// with ci padded to 32 rows and co to the stored count.
__global__ void k_wtrans(const uint16_t* __restrict__ w, int cout, int cin, uint16_t* __restrict__ wt, int cout_pad16,
                         int cin_pad32) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= cin_pad32 * 9 * cout_pad16) return;
    const int co = i % cout_pad16, tap = (i / cout_pad16) % 9, ci = i / (cout_pad16 * 9);
    wt[i] = (ci < cin && co < cout) ? w[(size_t(co) * 9 + (8 - tap)) * cin + ci] : uint16_t(0);
}

// The elementwise passes work on 8 channels (16 bytes) per thread: channel
// counts, strides and offsets are multiples of 16.
struct Vec8 {
    float v[8];
};
__device__ __forceinline__ Vec8 ld8(const uint16_t* p) {
    const uint4 u = *reinterpret_cast<const uint4*>(p);
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
    Vec8 r;
#pragma unroll
    for (int k = 0; k < 4; k++) r.v[2 * k] = bf2f(uint16_t(w[k] & 0xffffu)), r.v[2 * k + 1] = bf2f(uint16_t(w[k] >> 16));
    return r;
}
__device__ __forceinline__ void st8(uint16_t* p, const Vec8& r) {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; k++) w[k] = uint32_t(f2bf(r.v[2 * k])) | (uint32_t(f2bf(r.v[2 * k + 1])) << 16);
    *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
}

// Average 2x2 pooling: out (n, h/2, w/2) from in (n, h, w), channels c.
__global__ void k_pool2(UnetEw e) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int ow = e.w / 2, oh = e.h / 2, c8 = e.c >> 3;
    if (i >= e.n * oh * ow * c8) return;
    const int ch = (i % c8) * 8, q = i / c8, x = q % ow, y = (q / ow) % oh, n = q / (ow * oh);
    Vec8 s{};
    for (int k = 0; k < 4; k++) {
        const size_t src = size_t((n * e.h + 2 * y + (k >> 1)) * e.w + 2 * x + (k & 1));
        const Vec8 t = ld8(e.a + src * e.a_cs + e.a_co + ch);
#pragma unroll
        for (int j = 0; j < 8; j++) s.v[j] += t.v[j];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) s.v[j] *= 0.25f;
    st8(e.out + size_t(q) * e.out_cs + e.out_co + ch, s);
}

// Nearest 2x upsampling: out (n, 2h, 2w) from in (n, h, w) (h, w: the input's).
__global__ void k_up2(UnetEw e) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int ow = 2 * e.w, oh = 2 * e.h, c8 = e.c >> 3;
    if (i >= e.n * oh * ow * c8) return;
    const int ch = (i % c8) * 8, q = i / c8, x = q % ow, y = (q / ow) % oh, n = q / (ow * oh);
    const size_t src = size_t((n * e.h + y / 2) * e.w + x / 2);
    *reinterpret_cast<uint4*>(e.out + size_t(q) * e.out_cs + e.out_co + ch) =
        *reinterpret_cast<const uint4*>(e.a + src * e.a_cs + e.a_co + ch);
}

// Channel-slice copy at one resolution.
__global__ void k_copy(UnetEw e) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int c8 = e.c >> 3;
    if (i >= e.n * e.h * e.w * c8) return;
    const int ch = (i % c8) * 8, q = i / c8;
    *reinterpret_cast<uint4*>(e.out + size_t(q) * e.out_cs + e.out_co + ch) =
        *reinterpret_cast<const uint4*>(e.a + size_t(q) * e.a_cs + e.a_co + ch);
}

// Gradient at a ReLU'd tensor T (n, h, w): mask(T > 0) x (a at the same pixel
// + the 2x2 pooling's share of b at half resolution + the sum of c's 2x2
// children at double resolution), each term optional.
__global__ void k_grad_combine(UnetEw e) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int c8 = e.c >> 3;
    if (i >= e.n * e.h * e.w * c8) return;
    const int ch = (i % c8) * 8, q = i / c8, x = q % e.w, y = (q / e.w) % e.h, n = q / (e.w * e.h);
    Vec8 s{};
    if (e.a) s = ld8(e.a + size_t(q) * e.a_cs + e.a_co + ch);
    if (e.b) {
        const size_t src = size_t((n * (e.h / 2) + y / 2) * (e.w / 2) + x / 2);
        const Vec8 t = ld8(e.b + src * e.b_cs + e.b_co + ch);
#pragma unroll
        for (int j = 0; j < 8; j++) s.v[j] += 0.25f * t.v[j];
    }
    if (e.c_in) {
        for (int k = 0; k < 4; k++) {
            const size_t src = size_t((n * 2 * e.h + 2 * y + (k >> 1)) * (2 * e.w) + 2 * x + (k & 1));
            const Vec8 t = ld8(e.c_in + src * e.c_cs + e.c_co + ch);
#pragma unroll
            for (int j = 0; j < 8; j++) s.v[j] += t.v[j];
        }
    }
    if (e.mask) {
        const Vec8 m = ld8(e.mask + size_t(q) * e.mask_cs + e.mask_co + ch);
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (m.v[j] <= 0.0f) s.v[j] = 0.0f;
    }
    st8(e.out + size_t(q) * e.out_cs + e.out_co + ch, s);
}

// L1 loss in the log domain over 3 channels and its gradient into a 16-channel
// bf16 buffer (channels 3.. zero): loss += |pred - target| / count.
__global__ void k_l1_loss(const float* __restrict__ pred, const float* __restrict__ target, int npix,
                          uint16_t* __restrict__ grad, float* __restrict__ loss, float scale) {
    const int q = int(blockIdx.x * blockDim.x + threadIdx.x);
    float l = 0.0f;
    if (q < npix) {
        for (int ch = 0; ch < 16; ch++) {
            float g = 0.0f;
            if (ch < 3) {
                const float d = pred[size_t(q) * 3 + ch] - target[size_t(q) * 3 + ch];
                l += fabsf(d);
                g = d > 0.0f ? scale : (d < 0.0f ? -scale : 0.0f);
            }
            grad[size_t(q) * 16 + ch] = f2bf(g);
        }
    }
    // One atomic per wave.
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63u) == 0u) atomicAdd(loss, l * scale);
}

// Adam (Kingma and Ba 2015) on FP32 master weights, then their bf16 copy.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, uint16_t* __restrict__ w_bf, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi, v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w_bf) w_bf[i] = f2bf(wi);
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ out, int n) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) out[i] = f2bf(a[i]);
}

// Network input from a path-traced frame's accumulation and guide images (n
// samples): log1p of the demodulated irradiance (as the a-trous filter
// demodulates, denoise.comp:37-39), the mean albedo, the mean unit normal, the
// background share; 16 channels (10 used). Images are (h, w) float4 row-major;
// the input tensor is (hp, wp) >= (h, w), zero outside.
__global__ void k_unet_input(const float4* __restrict__ accum, const float4* __restrict__ albedo,
                             const float4* __restrict__ normal, int w, int h, float n, uint16_t* __restrict__ x,
                             int wp, int hp, float* __restrict__ alb_out) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= wp * hp) return;
    const int px = i % wp, py = i / wp;
    float f[16] = {};
    float ar = 1.0f, ag = 1.0f, ab = 1.0f;
    if (px < w && py < h) {
        const size_t p = size_t(py) * w + px;
        const float4 a = accum[p], g = albedo[p], nd = normal[p];
        ar = fmaxf(g.x / n, 0.02f), ag = fmaxf(g.y / n, 0.02f), ab = fmaxf(g.z / n, 0.02f);
        f[0] = log1pf(fmaxf(a.x / n, 0.0f) / ar);
        f[1] = log1pf(fmaxf(a.y / n, 0.0f) / ag);
        f[2] = log1pf(fmaxf(a.z / n, 0.0f) / ab);
        f[3] = ar, f[4] = ag, f[5] = ab;
        const float gx = nd.x / n, gy = nd.y / n, gz = nd.z / n;
        const float len = sqrtf(gx * gx + gy * gy + gz * gz);
        if (len > 1e-3f) f[6] = gx / len, f[7] = gy / len, f[8] = gz / len;
        f[9] = a.w / n;
    }
    uint32_t w16[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w16[k] = uint32_t(f2bf(f[2 * k])) | (uint32_t(f2bf(f[2 * k + 1])) << 16);
    uint4* xo = reinterpret_cast<uint4*>(x + size_t(i) * 16);  // the pixel's 16 channels: two 16-byte stores
    xo[0] = make_uint4(w16[0], w16[1], w16[2], w16[3]);
    xo[1] = make_uint4(w16[4], w16[5], w16[6], w16[7]);
    if (alb_out) {
        alb_out[size_t(i) * 3] = ar, alb_out[size_t(i) * 3 + 1] = ag, alb_out[size_t(i) * 3 + 2] = ab;
    }
}

// Training target: log1p of the reference radiance demodulated by the noisy
// frame's albedo (the network's output is modulated by the same albedo).
__global__ void k_unet_target(const float4* __restrict__ t, const float* __restrict__ alb, int npix,
                              float* __restrict__ out) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= npix) return;
    const float4 v = t[i];
    out[size_t(i) * 3] = log1pf(fmaxf(v.x, 0.0f) / alb[size_t(i) * 3]);
    out[size_t(i) * 3 + 1] = log1pf(fmaxf(v.y, 0.0f) / alb[size_t(i) * 3 + 1]);
    out[size_t(i) * 3 + 2] = log1pf(fmaxf(v.z, 0.0f) / alb[size_t(i) * 3 + 2]);
}

// The denoised frame: expm1 of the predicted log irradiance times the mean
// albedo, the background share as alpha, into RGBA16F.
__global__ void k_unet_output(const float* __restrict__ pred, const float* __restrict__ alb, int wp,
                              const float4* __restrict__ accum, float n, int w, int h, uint16_t* __restrict__ hdr) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= w * h) return;
    const int px = i % w, py = i / w;
    const size_t q = size_t(py) * wp + px;
    float c[4];
    for (int k = 0; k < 3; k++) c[k] = expm1f(fmaxf(pred[q * 3 + k], 0.0f)) * alb[q * 3 + k];
    c[3] = accum[i].w / n;
    uint2 word;
    word.x = uint32_t(__half_as_ushort(__float2half_rn(c[0]))) | (uint32_t(__half_as_ushort(__float2half_rn(c[1]))) << 16);
    word.y = uint32_t(__half_as_ushort(__float2half_rn(c[2]))) | (uint32_t(__half_as_ushort(__float2half_rn(c[3]))) << 16);
    reinterpret_cast<uint2*>(hdr)[i] = word;
}

int blocks(long n, int t = 256) { return int((n + t - 1) / t); }

}  // namespace

void launch_conv3x3(const UnetConv& c, hipStream_t st) {
    const int npix = c.n * c.h * c.w;
    UnetConv k = c;
    k.co_blocks = (c.cout + 31) / 32;
    // Output blocks per wave: up to 4 (A fragments reused across them), fewer
    // when that leaves under 2048 waves (deep levels: latency wants waves).
    // Many pixels: 2-D tiles with the input and weights in LDS, 1 or 2 output
    // blocks per workgroup, while that leaves 2048 waves.
    const int tiles_x = (c.w + kTileW - 1) / kTileW, tiles_y = (c.h + kTileH - 1) / kTileH;
    const int tiles = tiles_x * tiles_y * c.n;
    const char* stream_env = std::getenv("VVR_UNET_STREAM");  // "0": per-tile workgroups only
    if (k.co_blocks == 1 && c.cin <= kChunk && tiles * 4 >= 2048 && !(stream_env && stream_env[0] == '0')) {
        // One chunk and one output block: persistent workgroups, three per CU
        // (their LDS), each walking tiles.
        static const int cus = [] {
            int dev = 0, n = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
                n = 256;
            return n;
        }();
        const int grid = std::min(tiles, 3 * cus);
        hipLaunchKernelGGL(k_conv3x3_stream, dim3(grid), dim3(256), 0, st, k, tiles_x, tiles_y, tiles);
        return;
    }
    for (int nc : {2, 1}) {
        const int groups = (k.co_blocks + nc - 1) / nc;
        if (nc > k.co_blocks || tiles * 4 * groups < 2048) continue;
        const dim3 grid((c.w + kTileW - 1) / kTileW, (c.h + kTileH - 1) / kTileH, c.n * groups);
        if (nc == 2) hipLaunchKernelGGL(k_conv3x3_tile<2>, grid, dim3(256), 0, st, k, groups);
        else hipLaunchKernelGGL(k_conv3x3_tile<1>, grid, dim3(256), 0, st, k, groups);
        return;
    }
    const int pix_waves = (npix + 31) / 32;
    int ncob = std::min(4, k.co_blocks);
    while (ncob > 1 && pix_waves * ((k.co_blocks + ncob - 1) / ncob) < 2048) ncob--;
    const dim3 grid((npix + 127) / 128, (k.co_blocks + ncob - 1) / ncob);
    if (ncob == 1) hipLaunchKernelGGL(k_conv3x3<1>, grid, dim3(256), 0, st, k);
    else if (ncob == 2) hipLaunchKernelGGL(k_conv3x3<2>, grid, dim3(256), 0, st, k);
    else if (ncob == 3) hipLaunchKernelGGL(k_conv3x3<3>, grid, dim3(256), 0, st, k);
    else hipLaunchKernelGGL(k_conv3x3<4>, grid, dim3(256), 0, st, k);
}

void launch_conv_wgrad(const UnetWgrad& g, hipStream_t st) {
    UnetWgrad k = g;
    k.co_blocks = (g.cout + 31) / 32;
    const int npix = g.n * g.h * g.w;
    const int chunks = (npix + k.chunk - 1) / k.chunk;
    const dim3 grid(chunks, 9 * k.co_blocks * ((g.cin + 31) / 32));
    hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(64), 0, st, k);
}

void launch_wtrans(const uint16_t* w, int cout, int cin, uint16_t* wt, hipStream_t st) {
    const int cout16 = (cout + 15) / 16 * 16, cin32 = (cin + 31) / 32 * 32;
    hipLaunchKernelGGL(k_wtrans, dim3(blocks(long(cin32) * 9 * cout16)), dim3(256), 0, st, w, cout, cin, wt, cout16,
                       cin32);
}

void launch_unet_ew(int op, const UnetEw& e, hipStream_t st) {
    long n = long(e.n) * e.h * e.w * (e.c / 8);  // threads: 8 channels each
    if (op == kUnetPool) n = long(e.n) * (e.h / 2) * (e.w / 2) * (e.c / 8);
    if (op == kUnetUp) n = long(e.n) * (2 * e.h) * (2 * e.w) * (e.c / 8);
    auto k = op == kUnetPool ? k_pool2 : op == kUnetUp ? k_up2 : op == kUnetCopy ? k_copy : k_grad_combine;
    hipLaunchKernelGGL(k, dim3(blocks(n)), dim3(256), 0, st, e);
}

void launch_l1_loss(const float* pred, const float* target, int npix, uint16_t* grad, float* loss, hipStream_t st) {
    hipLaunchKernelGGL(k_l1_loss, dim3(blocks(npix)), dim3(256), 0, st, pred, target, npix, grad, loss,
                       1.0f / (3.0f * float(npix)));
}

void launch_adam(float* w, float* m, float* v, const float* g, uint16_t* w_bf, int n, float lr, int step,
                 hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, float(step)), c2 = 1.0f - powf(b2, float(step));
    hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, st, w, m, v, g, w_bf, n, lr, b1, b2, 1e-8f, c1, c2);
}

void launch_f32_to_bf16(const float* a, uint16_t* out, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(n)), dim3(256), 0, st, a, out, n);
}

void launch_unet_input(const float4* accum, const float4* albedo, const float4* normal, int w, int h, float n,
                       uint16_t* x, int wp, int hp, float* alb_out, hipStream_t st) {
    hipLaunchKernelGGL(k_unet_input, dim3(blocks(long(wp) * hp)), dim3(256), 0, st, accum, albedo, normal, w, h, n, x,
                       wp, hp, alb_out);
}

void launch_unet_target(const float4* t, const float* alb, int npix, float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_unet_target, dim3(blocks(npix)), dim3(256), 0, st, t, alb, npix, out);
}

void launch_unet_output(const float* pred, const float* alb, int wp, const float4* accum, float n, int w, int h,
                        uint16_t* hdr, hipStream_t st) {
    hipLaunchKernelGGL(k_unet_output, dim3(blocks(long(w) * h)), dim3(256), 0, st, pred, alb, wp, accum, n, w, h, hdr);
}

// ---- scratch of the allocation-free denoise (vvr_unet_denoise_noalloc) -------------
//
// A file-scope device array, one copy per GPU for as long as the library is
// loaded. Host code never passes its name to a kernel or a copy (on the host the
// name stands for nothing on the device): it uses the address hipGetSymbolAddress
// returns for the current device, looked up once per device and kept.
constexpr size_t kUnetScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kUnetScratchBytes];

hipError_t unet_scratch(int device, void** ptr, size_t* bytes) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* per_device[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!per_device[device]) {
        int current = -1;
        hipError_t e = hipGetDevice(&current);
        if (e != hipSuccess) return e;
        if (current != device) return hipErrorInvalidDevice;  // the caller makes `device` current
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        per_device[device] = p;
    }
    *ptr = per_device[device];
    *bytes = kUnetScratchBytes;
    return hipSuccess;
}

}  // namespace vvr
