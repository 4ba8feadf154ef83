// The learned denoiser (no counterpart in the reference; its comparison point is
// denoise.comp): a convolutional U-Net over the path tracer's accumulation and
// guide images, trained on the GPU by its own bf16 MFMA forward and backward
// kernels (kernels_unet.hip) with FP32 master weights and Adam.
//
// Levels l = 0 .. L-1 at resolution / 2^l with c_l channels:
//   encoder  e_l = relu(conv(relu(conv(l == 0 ? x : pool(e_{l-1})))))
//   decoder  d_l = relu(conv(relu(conv([up(d_{l+1}) | e_l])))), d_{L-1} = e_{L-1}
//   output   conv(d_0) -> 3 channels: log1p of the demodulated irradiance
// Input x: 16 channels (log1p irradiance, mean albedo, mean normal, background
// share, zeros). Loss: L1 against log1p of the target's irradiance
// (demodulated by the noisy frame's own albedo guide).
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>

#include "renderer_impl.hpp"

namespace vvr {
namespace {

struct View {  // an NHWC bf16 tensor slice
    uint16_t* p = nullptr;
    int cs = 0, co = 0, c = 0;
};

struct Layer {
    int cin = 0, cout = 0, co_pad = 0;  // co_pad: cout rounded up to 32
    size_t w_off = 0, b_off = 0;        // into the flat parameter arrays (padded layout)
    size_t w_n = 0;                     // co_pad * 9 * cin
    uint16_t* w_bf = nullptr;           // bf16 weights
    uint16_t* wt_bf = nullptr;          // flipped transposed, for the input gradient
    float* dw = nullptr;                // into the gradient array
    float* db = nullptr;
    const float* bias = nullptr;
};

int pad32(int c) { return (c + 31) / 32 * 32; }
int pad16(int c) { return (c + 15) / 16 * 16; }

}  // namespace
}  // namespace vvr

using namespace vvr;

struct vvr_unet {
    int device = 0;
    hipStream_t stream = nullptr;
    // Every use of the network's buffers is serialised: host calls under `mu`,
    // device work after `busy`, the event recorded after the previous use on
    // whichever stream it ran (the network's own or a renderer's).
    std::mutex mu;
    hipEvent_t busy = nullptr;
    int L = 0;
    int ch[5] = {};
    std::vector<Layer> layers;  // enc (2 per level), dec (2 per level, deepest first), out
    size_t n_params = 0;        // padded layout
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t* w_bf = nullptr;
    uint16_t* wt_bf = nullptr;
    int step = 0;
    // activations and gradients, for `cap_pix` pixels at level 0
    size_t cap_pix = 0, cap_elems = 0;
    int B = 0, H = 0, W = 0;
    uint16_t* act = nullptr;
    uint16_t* grad = nullptr;
    float *pred = nullptr, *target = nullptr, *alb = nullptr, *loss = nullptr;
    float4 *h_accum = nullptr, *h_albedo = nullptr, *h_normal = nullptr, *h_target = nullptr;  // staging (device)
    size_t stage_pix = 0;
    // views, per level
    View x0, ea[5], e[5], p[5], cat[5], da[5], d[5];
    View gx0, gea[5], ge[5], gp[5], gcat[5], gda[5], gd[5], gout;

    explicit vvr_unet(int dev) : device(dev) {}
    ~vvr_unet() {
        (void)hipSetDevice(device);
        if (busy) (void)hipEventSynchronize(busy);
        if (stream) (void)hipStreamSynchronize(stream);
        if (busy) (void)hipEventDestroy(busy);
        dfree(w), dfree(g), dfree(m), dfree(v), dfree(w_bf), dfree(wt_bf), dfree(act), dfree(grad);
        dfree(pred), dfree(target), dfree(alb), dfree(loss), dfree(h_accum), dfree(h_albedo), dfree(h_normal);
        dfree(h_target);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Device work on `s` starts after the previous use / the next use waits for `s`.
    void acquire(hipStream_t s) { check(hipStreamWaitEvent(s, busy, 0), "hipStreamWaitEvent"); }
    void release(hipStream_t s) { check(hipEventRecord(busy, s), "hipEventRecord"); }

    int enc(int l, int k) const { return 2 * l + k; }
    int dec(int l, int k) const { return 2 * L + 2 * (L - 2 - l) + k; }
    int out_layer() const { return 2 * L + 2 * (L - 1); }

    // Assigns the views for B images of H x W within `a` / `gr` (or, with null
    // bases, only counts); returns the elements of each block.
    size_t walk(uint16_t* a, uint16_t* gr) {
        const size_t pix0 = size_t(B) * H * W;
        auto px = [&](int l) { return pix0 >> (2 * l); };
        size_t used = 0;
        auto take = [&](View& va, View& vg, size_t pix, int c) {
            va = View{a ? a + used : nullptr, c, 0, c};
            vg = View{gr ? gr + used : nullptr, c, 0, c};
            used += pix * size_t(c);
        };
        View unused;
        take(x0, gx0, px(0), 16);
        for (int l = 0; l < L; l++) {
            const int c = ch[l];
            take(ea[l], gea[l], px(l), c);
            if (l > 0) take(p[l], gp[l], px(l), ch[l - 1]);
            if (l <= L - 2) {
                take(cat[l], gcat[l], px(l), ch[l + 1] + c);
                e[l] = View{cat[l].p, ch[l + 1] + c, ch[l + 1], c};  // the skip half of the concatenation
                take(unused, ge[l], px(l), c);                       // its gradient needs its own storage
                take(da[l], gda[l], px(l), c);
                take(d[l], gd[l], px(l), c);
            } else {
                take(e[l], ge[l], px(l), c);
            }
        }
        take(unused, gout, px(0), 16);
        return used;
    }

    // The forward pass's views only (no gradients) for one image of h x w within
    // `a` (or, with a null base, only counts); returns the elements used.
    size_t walk_forward(uint16_t* a, int h, int w) {
        const size_t pix0 = size_t(h) * w;
        auto px = [&](int l) { return pix0 >> (2 * l); };
        size_t used = 0;
        auto take = [&](View& v, size_t pix, int c) {
            v = View{a ? a + used : nullptr, c, 0, c};
            used += pix * size_t(c);
        };
        take(x0, px(0), 16);
        for (int l = 0; l < L; l++) {
            const int c = ch[l];
            take(ea[l], px(l), c);
            if (l > 0) take(p[l], px(l), ch[l - 1]);
            if (l <= L - 2) {
                take(cat[l], px(l), ch[l + 1] + c);
                e[l] = View{cat[l].p, ch[l + 1] + c, ch[l + 1], c};
                take(da[l], px(l), c);
                take(d[l], px(l), c);
            } else {
                take(e[l], px(l), c);
            }
        }
        return used;
    }

    // Buffers for b images of h x w (multiples of 2^(L-1)).
    void shape(int b, int h, int w) {
        if (h % (1 << (L - 1)) || w % (1 << (L - 1)))
            throw Error{VVR_EINVAL, "image size must be a multiple of 2^(levels - 1)"};
        B = b, H = h, W = w;
        const size_t pix0 = size_t(b) * h * w;
        const size_t total = walk(nullptr, nullptr);
        if (pix0 > cap_pix || total > cap_elems) {
            check(hipEventSynchronize(busy), "device wait");
            check(hipStreamSynchronize(stream), "device wait");
            dfree(act), dfree(grad), dfree(pred), dfree(target), dfree(alb);
            act = grad = nullptr, pred = target = alb = nullptr;
            act = dalloc<uint16_t>(total, "U-Net activations");
            grad = dalloc<uint16_t>(total, "U-Net gradients");
            pred = dalloc<float>(pix0 * 3, "U-Net output");
            target = dalloc<float>(pix0 * 3, "U-Net target");
            alb = dalloc<float>(pix0 * 3, "U-Net albedo");
            cap_pix = pix0, cap_elems = total;
        }
        walk(act, grad);
    }

    void conv(const Layer& ly, View in, View out, float* out_f32, bool relu, int level, View mask = View{}) {
        UnetConv c{};
        c.in = in.p, c.in_cs = in.cs, c.in_co = in.co, c.cin = ly.cin;
        c.wts = ly.w_bf, c.bias = ly.bias;
        c.out = out.p, c.out_cs = out.cs, c.out_co = out.co, c.cout = ly.cout;
        c.out_f32 = out_f32;
        c.mask = mask.p, c.mask_cs = mask.cs, c.mask_co = mask.co;
        c.n = B, c.h = H >> level, c.w = W >> level, c.relu = relu ? 1 : 0;
        launch_conv3x3(c, stream);
    }

    // dx = conv(dy, flipped transposed w), optionally masked by a ReLU'd input.
    void conv_dx(const Layer& ly, View dy, View dx, int level, View mask = View{}) {
        UnetConv c{};
        c.in = dy.p, c.in_cs = dy.cs, c.in_co = dy.co, c.cin = pad16(ly.cout);
        c.wts = ly.wt_bf;
        c.out = dx.p, c.out_cs = dx.cs, c.out_co = dx.co, c.cout = ly.cin;
        c.mask = mask.p, c.mask_cs = mask.cs, c.mask_co = mask.co;
        c.n = B, c.h = H >> level, c.w = W >> level;
        launch_conv3x3(c, stream);
    }

    void wgrad(const Layer& ly, View x, View dy, int level) {
        UnetWgrad k{};
        k.x = x.p, k.x_cs = x.cs, k.x_co = x.co, k.cin = ly.cin;
        k.dy = dy.p, k.dy_cs = dy.cs, k.cout = ly.cout;
        k.dw = ly.dw, k.db = ly.db;
        k.n = B, k.h = H >> level, k.w = W >> level, k.chunk = 1024;
        launch_conv_wgrad(k, stream);
    }

    void ew(int op, View a, View out, int level, int c, View b = View{}, View cin = View{}, View mask = View{}) {
        UnetEw k{};
        k.a = a.p, k.a_cs = a.cs, k.a_co = a.co;
        k.b = b.p, k.b_cs = b.cs, k.b_co = b.co;
        k.c_in = cin.p, k.c_cs = cin.cs, k.c_co = cin.co;
        k.mask = mask.p, k.mask_cs = mask.cs, k.mask_co = mask.co;
        k.out = out.p, k.out_cs = out.cs, k.out_co = out.co;
        k.n = B, k.h = H >> level, k.w = W >> level, k.c = c;
        launch_unet_ew(op, k, stream);
    }

    void forward() {
        for (int l = 0; l < L; l++) {
            if (l > 0) ew(kUnetPool, e[l - 1], p[l], l - 1, ch[l - 1]);
            conv(layers[enc(l, 0)], l == 0 ? x0 : p[l], ea[l], nullptr, true, l);
            conv(layers[enc(l, 1)], ea[l], e[l], nullptr, true, l);
        }
        for (int l = L - 2; l >= 0; l--) {
            const View below = l + 1 == L - 1 ? e[l + 1] : d[l + 1];
            ew(kUnetUp, below, View{cat[l].p, cat[l].cs, 0, ch[l + 1]}, l + 1, ch[l + 1]);
            conv(layers[dec(l, 0)], cat[l], da[l], nullptr, true, l);
            conv(layers[dec(l, 1)], da[l], d[l], nullptr, true, l);
        }
        conv(layers[out_layer()], d[0], View{}, pred, false, 0);
    }

    void backward() {
        check(hipMemsetAsync(g, 0, n_params * sizeof(float), stream), "gradient reset");
        for (auto& ly : layers) launch_wtrans(ly.w_bf, ly.cout, ly.cin, ly.wt_bf, stream);
        const Layer& lo = layers[out_layer()];
        wgrad(lo, d[0], gout, 0);
        conv_dx(lo, gout, gd[0], 0, d[0]);
        for (int l = 0; l <= L - 2; l++) {
            const Layer& b = layers[dec(l, 1)];
            const Layer& a = layers[dec(l, 0)];
            wgrad(b, da[l], gd[l], l);
            conv_dx(b, gd[l], gda[l], l, da[l]);
            wgrad(a, cat[l], gda[l], l);
            conv_dx(a, gda[l], gcat[l], l);
            if (l + 1 <= L - 2)  // the decoder output below: its gradient through the upsampling
                ew(kUnetCombine, View{}, gd[l + 1], l + 1, ch[l + 1], View{}, View{gcat[l].p, gcat[l].cs, 0, 0},
                   d[l + 1]);
        }
        for (int l = L - 1; l >= 0; l--) {
            View skip = l <= L - 2 ? View{gcat[l].p, gcat[l].cs, ch[l + 1], ch[l]} : View{};
            View from_pool = l < L - 1 ? gp[l + 1] : View{};
            View from_up = l == L - 1 ? View{gcat[L - 2].p, gcat[L - 2].cs, 0, 0} : View{};
            ew(kUnetCombine, skip, ge[l], l, ch[l], from_pool, from_up, e[l]);
            const Layer& b = layers[enc(l, 1)];
            const Layer& a = layers[enc(l, 0)];
            wgrad(b, ea[l], ge[l], l);
            conv_dx(b, ge[l], gea[l], l, ea[l]);
            wgrad(a, l == 0 ? x0 : p[l], gea[l], l);
            if (l > 0) conv_dx(a, gea[l], gp[l], l);
        }
    }

    void stage(size_t pix) {
        if (pix <= stage_pix) return;
        check(hipEventSynchronize(busy), "device wait");
        check(hipStreamSynchronize(stream), "device wait");
        dfree(h_accum), dfree(h_albedo), dfree(h_normal), dfree(h_target);
        h_accum = dalloc<float4>(pix, "U-Net staging");
        h_albedo = dalloc<float4>(pix, "U-Net staging");
        h_normal = dalloc<float4>(pix, "U-Net staging");
        h_target = dalloc<float4>(pix, "U-Net staging");
        stage_pix = pix;
    }
};

namespace {

void create_unet(vvr_ctx* ctx, const vvr_unet_config* cfg, vvr_unet** out) {
    if (!ctx || !cfg || !out) throw Error{VVR_EINVAL, "null argument"};
    if (cfg->levels < 2 || cfg->levels > 5) throw Error{VVR_EINVAL, "levels must be 2..5"};
    for (uint32_t l = 0; l < cfg->levels; l++)
        if (cfg->channels[l] == 0 || cfg->channels[l] % 16 || cfg->channels[l] > 256)
            throw Error{VVR_EINVAL, "channels must be positive multiples of 16 up to 256"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto n = std::make_unique<vvr_unet>(ctx->device);
    check(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreateWithFlags(&n->busy, hipEventDisableTiming), "hipEventCreate");
    n->L = int(cfg->levels);
    for (int l = 0; l < n->L; l++) n->ch[l] = int(cfg->channels[l]);
    const int L = n->L;
    auto add = [&](int cin, int cout) {
        Layer ly;
        ly.cin = cin, ly.cout = cout, ly.co_pad = pad32(cout);
        n->layers.push_back(ly);
    };
    for (int l = 0; l < L; l++) {
        add(l == 0 ? 16 : n->ch[l - 1], n->ch[l]);
        add(n->ch[l], n->ch[l]);
    }
    for (int l = L - 2; l >= 0; l--) {
        add(n->ch[l + 1] + n->ch[l], n->ch[l]);
        add(n->ch[l], n->ch[l]);
    }
    add(n->ch[0], 3);
    size_t off = 0, wt_total = 0, wbf_total = 0;
    for (auto& ly : n->layers) {
        ly.w_n = size_t(ly.co_pad) * 9 * ly.cin;
        ly.w_off = off;
        off += ly.w_n;
        ly.b_off = off;
        off += size_t(ly.co_pad);
        wbf_total += ly.w_n;
        wt_total += size_t(pad32(ly.cin)) * 9 * pad16(ly.cout);
    }
    n->n_params = off;
    n->w = dalloc<float>(off, "U-Net weights");
    n->g = dalloc<float>(off, "U-Net gradients");
    n->m = dalloc<float>(off, "Adam moments");
    n->v = dalloc<float>(off, "Adam moments");
    n->w_bf = dalloc<uint16_t>(off, "U-Net bf16 weights");  // same flat layout as w (biases unused)
    n->wt_bf = dalloc<uint16_t>(wt_total, "U-Net transposed weights");
    n->loss = dalloc<float>(1, "loss");
    size_t wt_off = 0;
    for (auto& ly : n->layers) {
        ly.w_bf = n->w_bf + ly.w_off;
        ly.wt_bf = n->wt_bf + wt_off;
        wt_off += size_t(pad32(ly.cin)) * 9 * pad16(ly.cout);
        ly.dw = n->g + ly.w_off;
        ly.db = n->g + ly.b_off;
        ly.bias = n->w + ly.b_off;
    }
    // He-normal initial weights (zero padding rows), zero biases, from the seed.
    std::vector<float> h(off, 0.0f);
    std::mt19937 rng(cfg->seed ? cfg->seed : 1u);
    for (auto& ly : n->layers) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / float(9 * ly.cin)));
        for (int co = 0; co < ly.cout; co++)
            for (size_t k = 0; k < size_t(9) * ly.cin; k++) h[ly.w_off + size_t(co) * 9 * ly.cin + k] = nd(rng);
    }
    check(hipMemcpy(n->w, h.data(), off * 4, hipMemcpyHostToDevice), "weights upload");
    check(hipMemset(n->m, 0, off * 4), "hipMemset");
    check(hipMemset(n->v, 0, off * 4), "hipMemset");
    launch_f32_to_bf16(n->w, n->w_bf, int(off), n->stream);
    check(hipStreamSynchronize(n->stream), "weights");
    *out = n.release();
}

// Flat parameter order of the ABI: per layer W[cout][9][cin] then b[cout]
// (no padding rows).
size_t abi_params(const vvr_unet* n) {
    size_t t = 0;
    for (const auto& ly : n->layers) t += size_t(ly.cout) * 9 * ly.cin + size_t(ly.cout);
    return t;
}

template <class F> void each_param(const vvr_unet* n, F f) {  // f(abi index, padded index, count)
    size_t a = 0;
    for (const auto& ly : n->layers) {
        const size_t wn = size_t(ly.cout) * 9 * ly.cin;
        f(a, ly.w_off, wn);
        a += wn;
        f(a, ly.b_off, size_t(ly.cout));
        a += size_t(ly.cout);
    }
}

void copy_out(vvr_unet* n, const float* dev, float* host, size_t count) {
    if (count != abi_params(n)) throw Error{VVR_EINVAL, "parameter count mismatch"};
    std::vector<float> h(n->n_params);
    check(hipMemcpy(h.data(), dev, n->n_params * 4, hipMemcpyDeviceToHost), "read-back");
    each_param(n, [&](size_t a, size_t p, size_t c) { std::memcpy(host + a, h.data() + p, c * 4); });
}

// Uploads a batch of frames and prepares the input (and the target's log irradiance).
void upload_batch(vvr_unet* n, uint32_t batch, uint32_t w, uint32_t h, const float* accum, const float* albedo,
                  const float* normal, float samples, const float* target) {
    const size_t pix = size_t(batch) * w * h;
    n->stage(pix);
    const size_t bytes = pix * sizeof(float4);
    check(hipMemcpyAsync(n->h_accum, accum, bytes, hipMemcpyHostToDevice, n->stream), "upload");
    check(hipMemcpyAsync(n->h_albedo, albedo, bytes, hipMemcpyHostToDevice, n->stream), "upload");
    check(hipMemcpyAsync(n->h_normal, normal, bytes, hipMemcpyHostToDevice, n->stream), "upload");
    // Stacked images are one tall image for the input pass (per-pixel).
    launch_unet_input(n->h_accum, n->h_albedo, n->h_normal, int(w), int(h * batch), samples, n->x0.p, int(w),
                      int(h * batch), n->alb, n->stream);
    if (target) {
        check(hipMemcpyAsync(n->h_target, target, bytes, hipMemcpyHostToDevice, n->stream), "upload");
        launch_unet_target(n->h_target, n->alb, int(pix), n->target, n->stream);
    }
}

// The network on a frame's accumulation and guide images into an RGBA16F image,
// padding the frame to the network's multiple; on stream `st`, the lock held.
void denoise_frame(vvr_unet* n, const float4* accum, const float4* albedo, const float4* normal, float samples,
                   uint32_t width, uint32_t height, uint16_t* hdr, hipStream_t st) {
    const int m = 1 << (n->L - 1);
    const int wp = int((width + m - 1) / m * m), hp = int((height + m - 1) / m * m);
    n->shape(1, hp, wp);
    // Run on the renderer's stream so the frame's order holds.
    hipStream_t saved = n->stream;
    n->stream = st;
    launch_unet_input(accum, albedo, normal, int(width), int(height), samples, n->x0.p, wp, hp, n->alb, st);
    n->forward();
    launch_unet_output(n->pred, n->alb, wp, accum, samples, int(width), int(height), hdr, st);
    n->stream = saved;
}

}  // namespace

extern "C" {

int vvr_unet_create(vvr_ctx* ctx, const vvr_unet_config* config, vvr_unet** out) {
    return guarded([&] { create_unet(ctx, config, out); });
}

void vvr_unet_destroy(vvr_unet* n) { delete n; }

size_t vvr_unet_num_params(const vvr_unet* n) { return n ? abi_params(n) : 0; }

int vvr_unet_get_params(vvr_unet* n, float* out, size_t count) {
    return guarded([&] {
        if (!n || !out) throw Error{VVR_EINVAL, "null argument"};
        std::lock_guard<std::mutex> lock(n->mu);
        check(hipSetDevice(n->device), "hipSetDevice");
        check(hipEventSynchronize(n->busy), "device wait");
        check(hipStreamSynchronize(n->stream), "device wait");
        copy_out(n, n->w, out, count);
    });
}

int vvr_unet_get_grads(vvr_unet* n, float* out, size_t count) {
    return guarded([&] {
        if (!n || !out) throw Error{VVR_EINVAL, "null argument"};
        std::lock_guard<std::mutex> lock(n->mu);
        check(hipSetDevice(n->device), "hipSetDevice");
        check(hipEventSynchronize(n->busy), "device wait");
        check(hipStreamSynchronize(n->stream), "device wait");
        copy_out(n, n->g, out, count);
    });
}

int vvr_unet_set_params(vvr_unet* n, const float* in, size_t count) {
    return guarded([&] {
        if (!n || !in) throw Error{VVR_EINVAL, "null argument"};
        if (count != abi_params(n)) throw Error{VVR_EINVAL, "parameter count mismatch"};
        std::lock_guard<std::mutex> lock(n->mu);
        check(hipSetDevice(n->device), "hipSetDevice");
        check(hipEventSynchronize(n->busy), "device wait");
        check(hipStreamSynchronize(n->stream), "device wait");
        std::vector<float> h(n->n_params, 0.0f);
        each_param(n, [&](size_t a, size_t p, size_t c) { std::memcpy(h.data() + p, in + a, c * 4); });
        check(hipMemcpy(n->w, h.data(), n->n_params * 4, hipMemcpyHostToDevice), "upload");
        check(hipMemset(n->m, 0, n->n_params * 4), "hipMemset");
        check(hipMemset(n->v, 0, n->n_params * 4), "hipMemset");
        n->step = 0;
        launch_f32_to_bf16(n->w, n->w_bf, int(n->n_params), n->stream);
        check(hipStreamSynchronize(n->stream), "upload");
    });
}

int vvr_unet_train_step(vvr_unet* n, uint32_t batch, uint32_t width, uint32_t height, const float* accum,
                        const float* albedo, const float* normal, float samples, const float* target, float lr,
                        float* loss) {
    return guarded([&] {
        if (!n || !accum || !albedo || !normal || !target) throw Error{VVR_EINVAL, "null argument"};
        if (!batch || !width || !height || !(samples > 0.0f)) throw Error{VVR_EINVAL, "empty batch"};
        std::lock_guard<std::mutex> lock(n->mu);
        check(hipSetDevice(n->device), "hipSetDevice");
        n->acquire(n->stream);
        n->shape(int(batch), int(height), int(width));
        upload_batch(n, batch, width, height, accum, albedo, normal, samples, target);
        n->forward();
        check(hipMemsetAsync(n->loss, 0, 4, n->stream), "loss reset");
        launch_l1_loss(n->pred, n->target, int(size_t(batch) * width * height), n->gout.p, n->loss, n->stream);
        n->backward();
        if (lr > 0.0f) {
            n->step++;
            launch_adam(n->w, n->m, n->v, n->g, n->w_bf, int(n->n_params), lr, n->step, n->stream);
        }
        float l = 0.0f;
        check(hipMemcpyAsync(&l, n->loss, 4, hipMemcpyDeviceToHost, n->stream), "loss read-back");
        n->release(n->stream);
        check(hipStreamSynchronize(n->stream), "train step");
        if (loss) *loss = l;
    });
}

int vvr_unet_denoise(vvr_unet* n, uint32_t width, uint32_t height, const float* accum, const float* albedo,
                     const float* normal, float samples, float* rgba) {
    return guarded([&] {
        if (!n || !accum || !albedo || !normal || !rgba) throw Error{VVR_EINVAL, "null argument"};
        std::lock_guard<std::mutex> lock(n->mu);
        check(hipSetDevice(n->device), "hipSetDevice");
        const size_t pix = size_t(width) * height;
        n->stage(pix);
        n->acquire(n->stream);
        const size_t bytes = pix * sizeof(float4);
        check(hipMemcpyAsync(n->h_accum, accum, bytes, hipMemcpyHostToDevice, n->stream), "upload");
        check(hipMemcpyAsync(n->h_albedo, albedo, bytes, hipMemcpyHostToDevice, n->stream), "upload");
        check(hipMemcpyAsync(n->h_normal, normal, bytes, hipMemcpyHostToDevice, n->stream), "upload");
        uint16_t* hdr = dalloc<uint16_t>(4 * pix, "U-Net frame");
        denoise_frame(n, n->h_accum, n->h_albedo, n->h_normal, samples, width, height, hdr, n->stream);
        n->release(n->stream);
        std::vector<uint16_t> h(4 * pix);
        hipError_t e = hipMemcpyAsync(h.data(), hdr, 8 * pix, hipMemcpyDeviceToHost, n->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
        dfree(hdr);
        check(e, "denoised read-back");
        for (size_t i = 0; i < 4 * pix; i++) rgba[i] = f16_bits_to_f32(h[i]);
    });
}

// vvr_unet_denoise without any allocation: every buffer of the call (staged
// inputs, activations, network output, the RGBA16F frame) lies in the library's
// static per-device scratch; the frame's read-back is widened to float in the
// caller's buffer. Calls on one device share that scratch, so calls that overlap
// on one device race; the network's own lock orders calls on one network.
int vvr_unet_denoise_noalloc(vvr_unet* n, uint32_t width, uint32_t height, const float* accum, const float* albedo,
                             const float* normal, float samples, float* rgba) {
    return guarded([&] {
        if (!n || !accum || !albedo || !normal || !rgba) throw Error{VVR_EINVAL, "null argument"};
        if (!width || !height || width > 16384 || height > 16384)
            throw Error{VVR_EINVAL, "frame size must be 1..16384 per axis"};
        std::lock_guard<std::mutex> lock(n->mu);
        check(hipSetDevice(n->device), "hipSetDevice");
        void* base = nullptr;
        size_t cap = 0;
        check(unet_scratch(n->device, &base, &cap), "denoiser scratch");
        const int m = 1 << (n->L - 1);
        const int wp = int((width + m - 1) / m * m), hp = int((height + m - 1) / m * m);
        const size_t pix = size_t(width) * height, pix0 = size_t(wp) * hp;
        auto align = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t stage_b = align(pix * sizeof(float4)), hdr_b = align(pix * 8), f3_b = align(pix0 * 3 * 4);
        const size_t act_b = align(n->walk_forward(nullptr, hp, wp) * 2);
        const size_t need = 3 * stage_b + hdr_b + 2 * f3_b + act_b;
        if (need > cap)
            throw Error{VVR_ELIMIT, "a " + std::to_string(width) + " x " + std::to_string(height) + " frame needs " +
                                        std::to_string(need) + " bytes of scratch; the allocation-free denoiser has " +
                                        std::to_string(cap) + " (use vvr_unet_denoise)"};
        uint8_t* q = static_cast<uint8_t*>(base);
        float4* in_accum = reinterpret_cast<float4*>(q);
        float4* in_albedo = reinterpret_cast<float4*>(q + stage_b);
        float4* in_normal = reinterpret_cast<float4*>(q + 2 * stage_b);
        uint16_t* hdr = reinterpret_cast<uint16_t*>(q + 3 * stage_b);
        float* pred = reinterpret_cast<float*>(q + 3 * stage_b + hdr_b);
        float* alb = reinterpret_cast<float*>(q + 3 * stage_b + hdr_b + f3_b);
        uint16_t* act = reinterpret_cast<uint16_t*>(q + 3 * stage_b + hdr_b + 2 * f3_b);
        n->acquire(n->stream);
        const size_t bytes = pix * sizeof(float4);
        check(hipMemcpyAsync(in_accum, accum, bytes, hipMemcpyHostToDevice, n->stream), "upload");
        check(hipMemcpyAsync(in_albedo, albedo, bytes, hipMemcpyHostToDevice, n->stream), "upload");
        check(hipMemcpyAsync(in_normal, normal, bytes, hipMemcpyHostToDevice, n->stream), "upload");
        // The network's views and output pointers point into the scratch for this
        // call; the next call of any other entry point lays them out again.
        float* const saved_pred = n->pred;
        float* const saved_alb = n->alb;
        n->B = 1, n->H = hp, n->W = wp;
        n->walk_forward(act, hp, wp);
        n->pred = pred, n->alb = alb;
        launch_unet_input(in_accum, in_albedo, in_normal, int(width), int(height), samples, n->x0.p, wp, hp, alb,
                          n->stream);
        n->forward();
        launch_unet_output(pred, alb, wp, in_accum, samples, int(width), int(height), hdr, n->stream);
        n->pred = saved_pred, n->alb = saved_alb;
        hipError_t e = hipGetLastError();
        n->release(n->stream);
        // RGBA16F into the first half of the caller's buffer, then widened in place
        // from the end: float i covers halves 2i and 2i + 1, read before it (byte
        // copies, so the compiler keeps that order).
        uint8_t* out = reinterpret_cast<uint8_t*>(rgba);
        if (e == hipSuccess) e = hipMemcpyAsync(out, hdr, 8 * pix, hipMemcpyDeviceToHost, n->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
        check(e, "denoised read-back");
        for (size_t i = 4 * pix; i-- > 0;) {
            uint16_t h;
            std::memcpy(&h, out + 2 * i, 2);
            const float f = f16_bits_to_f32(h);
            std::memcpy(out + 4 * i, &f, 4);
        }
    });
}

}  // extern "C"

// Runs the learned denoiser on a renderer's images into its HDR image
// (VVR_DENOISE_LEARNED), on the renderer's or the server's stream.
void unet_denoise_frame(vvr_unet* n, const float4* accum, const float4* albedo, const float4* normal, float samples,
                        uint32_t width, uint32_t height, uint16_t* hdr, hipStream_t st) {
    std::lock_guard<std::mutex> lock(n->mu);
    n->acquire(st);
    denoise_frame(n, accum, albedo, normal, samples, width, height, hdr, st);
    n->release(st);
}
