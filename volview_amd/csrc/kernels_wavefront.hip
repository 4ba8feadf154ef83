// Wavefront (stream) path tracer: the bounce loop of pt.comp:427-545 cut into
// stages, each its own kernel over a queue of the paths that are at that stage:
//
//   regen    paths that ended: add them to their pixel chunk, start the next
//            path of the chunk or take a new chunk from the global item queue
//   extend   free flight (delta tracking)               -> event or end
//   shade    shading frame, surface decision, light sample
//   shadow   ratio-tracked shadow ray
//   scatter  next-event estimate, next direction, roulette -> extend or end
//
// A free flight or shadow ray runs at most a set number of tracking iterations
// per launch; one that is not done is suspended (its loop state stored in its
// slot) and resumed by the same stage of the next iteration. Lanes then do not
// idle behind the longest ray of their wave, and each launch's waves are
// compacted anew. A path's random numbers are drawn in the same order either way.
//
// Each kernel keeps few registers live and runs coherent waves; queues are
// compacted with a wave ballot, one atomic per wave and the lanes' mbcnt ranks.
// A path draws its random numbers in the same order as the megakernel, so it
// traces the same path. Work items are chunks of consecutive paths of one
// pixel; a chunk's sums go to its own slot of a partial buffer and are added in
// chunk order, so frames are deterministic.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "device/pt.hpp"
#include "launch.hpp"

namespace vvr {

namespace {

constexpr uint32_t kNone = kWfNone;

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ V3 xyz(float4 v) { return V3(v.x, v.y, v.z); }
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Appends `value` for the lanes with `push` set: one atomic per wave.
__device__ __forceinline__ void wave_push(bool push, uint32_t value, uint32_t* __restrict__ queue,
                                          uint32_t* __restrict__ count) {
    const uint64_t mask = __ballot(push);
    if (!mask) return;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (int(__lane_id()) == leader) base = atomicAdd(count, uint32_t(__popcll(mask)));
    base = __shfl(base, leader);
    if (push) queue[base + lane_rank(mask)] = value;
}

// Reserves one value of *counter for each lane with `take` set; returns it.
__device__ __forceinline__ uint32_t wave_take(bool take, uint32_t* __restrict__ counter) {
    const uint64_t mask = __ballot(take);
    if (!mask) return kNone;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (int(__lane_id()) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
    base = __shfl(base, leader);
    return take ? base + lane_rank(mask) : kNone;
}

template <bool kBatched>
__device__ __forceinline__ const PtRequest& request(const PtRequest* __restrict__ reqs, uint32_t r) {
    return kBatched ? reqs[r] : reqs[0];
}

// Request of global item `item` (requests are sorted by item_base).
template <bool kBatched>
__device__ __forceinline__ uint32_t request_of_item(const PtRequest* __restrict__ reqs, uint32_t n_reqs,
                                                    uint32_t item) {
    if (!kBatched) return 0u;
    uint32_t lo = 0, hi = n_reqs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (reqs[mid].item_base <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Runs f(r) for the lanes with `on` set, once per distinct request r among
// them, with r wave-uniform: the request's scene is then read with scalar
// loads, as in a single-request launch. A wave of one request takes one pass.
template <bool kBatched, class F>
__device__ __forceinline__ void for_each_request(bool on, uint32_t req, F&& f) {
    if (!kBatched) {
        if (on) f(0u);
        return;
    }
    bool pending = on;
    for (;;) {
        const uint64_t m = __ballot(pending);
        if (!m) break;
        const uint32_t r = __builtin_amdgcn_readfirstlane(__shfl(req, __ffsll((unsigned long long)m) - 1));
        const bool mine = pending && req == r;
        if (mine) f(r);
        pending = pending && !mine;
    }
}

// Runs f(r) for the lanes with `on` set, r the request of the wave's first
// such lane, read into a scalar register. For waves of one request only: the
// sorted queues (each request's run padded to whole waves) and single-request
// engines. Unlike for_each_request, f is inlined once and not in a loop, which
// keeps the register-heavy stages at two waves per SIMD without spills.
template <bool kBatched, class F>
__device__ __forceinline__ void with_wave_request(bool on, uint32_t req, F&& f) {
    if (!kBatched) {
        if (on) f(0u);
        return;
    }
    const uint64_t m = __ballot(on);
    if (!m) return;
    const uint32_t r = __builtin_amdgcn_readfirstlane(__shfl(req, __ffsll((unsigned long long)m) - 1));
    if (on) f(r);
}

__device__ __forceinline__ uint32_t pack_counters(const PathState& st) {
    return uint32_t(st.hero) | (st.surfaces << 2) | (st.volumes << 8) | (st.bounce << 18);
}

struct SlotRef {
    uint32_t item, k, req;
};

__device__ __forceinline__ void load_path(const WfSlots& S, uint32_t i, PathState& st, SlotRef& ref) {
    const float4 a = S.p_origin[i], b = S.p_dir[i], c = S.p_thr[i], d = S.p_spec[i], e = S.p_rad[i];
    st.origin = xyz(a);
    st.last_pdf = a.w;
    st.dir = xyz(b);
    st.rng.state = fbits(b.w);
    st.throughput = xyz(c);
    const uint32_t cnt = fbits(c.w);
    st.hero = int(cnt & 3u);
    st.surfaces = (cnt >> 2) & 63u;
    st.volumes = (cnt >> 8) & 1023u;
    st.bounce = cnt >> 18;
    st.spectral = xyz(d);
    ref.item = fbits(d.w);
    st.radiance = xyz(e);
    ref.k = fbits(e.w) & 0xffffu;
    ref.req = fbits(e.w) >> 16;
}

__device__ __forceinline__ void store_path(const WfSlots& S, uint32_t i, const PathState& st, const SlotRef& ref) {
    S.p_origin[i] = f4(st.origin, st.last_pdf);
    S.p_dir[i] = f4(st.dir, bitsf(st.rng.state));
    S.p_thr[i] = f4(st.throughput, bitsf(pack_counters(st)));
    S.p_spec[i] = f4(st.spectral, bitsf(ref.item));
    S.p_rad[i] = f4(st.radiance, bitsf(ref.k | (ref.req << 16)));
}

__device__ __forceinline__ void load_event(const WfSlots& S, uint32_t i, EventState& ev) {
    const float4 a = S.e_idx[i], b = S.e_s0[i], c = S.e_s1[i], d = S.e_s2[i];
    ev.idx = xyz(a);
    ev.cut = a.w;
    ev.s.albedo = xyz(b);
    ev.s.extinction = b.w;
    ev.s.roughness = c.x;
    ev.s.metalness = c.y;
    ev.s.subsurface = c.z;
    ev.s.coat = c.w;
    ev.s.density = d.x;
    const uint32_t bits = fbits(d.y);
    ev.s.label = bits & 0xffu;
    ev.s.ghost = ((bits >> 8) & 1u) != 0u;
    ev.cap_plane = int((bits >> 9) & 3u) - 1;
    ev.surface_event = ((bits >> 12) & 1u) != 0u;
}

__device__ __forceinline__ void store_event(const WfSlots& S, uint32_t i, const EventState& ev) {
    S.e_idx[i] = f4(ev.idx, ev.cut);
    S.e_s0[i] = f4(ev.s.albedo, ev.s.extinction);
    S.e_s1[i] = make_float4(ev.s.roughness, ev.s.metalness, ev.s.subsurface, ev.s.coat);
    const uint32_t bits = ev.s.label | (ev.s.ghost ? 1u << 8 : 0u) | (uint32_t(ev.cap_plane + 1) << 9) |
                          (ev.surface_event ? 1u << 12 : 0u);
    S.e_s2[i] = make_float4(ev.s.density, bitsf(bits), 0.0f, 0.0f);
}

__device__ __forceinline__ void add4(float4* p, float4 v) {
    const float4 a = *p;
    *p = make_float4(a.x + v.x, a.y + v.y, a.z + v.z, a.w + v.w);
}

// Pixel (possibly outside the image) and path index of sample k of item `item`.
__device__ __forceinline__ void item_coords(const PtRequest& R, uint32_t item, uint32_t k, uint32_t& px, uint32_t& py,
                                            uint32_t& chunk_index, uint32_t& p_item, uint32_t& path_index,
                                            uint32_t& n_paths) {
    const uint32_t within = item - R.item_base;
    chunk_index = within / R.n_px;
    p_item = within - chunk_index * R.n_px;
    const uint32_t tile = p_item >> 6, w = p_item & 63u;
    px = (tile % R.tiles_x) * 8u + (w & 7u);
    py = (tile / R.tiles_x) * 8u + (w >> 3);
    const uint32_t first = chunk_index * R.chunk;
    n_paths = min(R.chunk, R.samples - first);
    path_index = R.first_path + first + k;
}

__device__ __forceinline__ void store_partial(const PtRequest& R, uint32_t chunk_index, uint32_t p_item, float4 c0,
                                              float4 c1, float4 c2) {
    float4* base = R.partial + size_t(chunk_index) * 3u * R.n_px;
    base[p_item] = c0;
    base[R.n_px + p_item] = c1;
    base[2u * R.n_px + p_item] = c2;
}

// The count and cursor arrays of both queue sorts to zero (one wave of the
// iteration's first kernel; the sorts run after it on the same stream).
__device__ __forceinline__ void clear_bins(const WfQueues& Q) {
    for (uint32_t i = threadIdx.x; i < 4u * kWfBins; i += blockDim.x) Q.bins[i] = 0u;
}

// Counting sort of a queue by the request of each slot's path, so that the
// stages' waves mostly hold paths of one request (each distinct request in a
// wave costs the stage one more pass). The order within a request is
// arbitrary: no result depends on queue order. kNone entries (slots a served
// tail took) are left out. Pass 1: requests' counts.
__global__ void __launch_bounds__(256) k_wf_bin_count(const uint32_t* __restrict__ in,
                                                      const uint32_t* __restrict__ count,
                                                      const uint16_t* __restrict__ req, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kWfBins];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t n = *count;
    for (uint32_t q = blockIdx.x * 256u + threadIdx.x; q < n; q += gridDim.x * 256u) {
        const uint32_t slot = in[q];
        if (slot != kNone) atomicAdd(&h[req[slot] & (kWfBins - 1u)], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// Pass 2: each workgroup ranks a chunk of entries within their requests in
// LDS, reserves each request's run with one atomic and writes the slots out.
// Runs start at multiples of 64 entries; workgroup 0 fills each run's tail up
// to the next multiple with kNone and stores the padded total in *out_count.
constexpr uint32_t kSortPer = 16;  // entries per lane and chunk
__global__ void __launch_bounds__(256) k_wf_bin_scatter(const uint32_t* __restrict__ in,
                                                        const uint32_t* __restrict__ count,
                                                        const uint16_t* __restrict__ req,
                                                        const uint32_t* __restrict__ hist, uint32_t* __restrict__ cursor,
                                                        uint32_t* __restrict__ out, uint32_t* __restrict__ out_count) {
    __shared__ uint32_t start[kWfBins], h[kWfBins], run[kWfBins];
    const uint32_t t = threadIdx.x;
    // Exclusive prefix of the padded counts: where each request's run starts.
    const uint32_t filled = hist[t];
    const uint32_t mine = (filled + 63u) & ~63u;
    start[t] = mine;
    __syncthreads();
    for (uint32_t off = 1; off < kWfBins; off <<= 1) {
        const uint32_t x = t >= off ? start[t - off] : 0u;
        __syncthreads();
        start[t] += x;
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        if (t == kWfBins - 1u) *out_count = start[t];
        for (uint32_t k = start[t] - mine + filled; k < start[t]; k++) out[k] = kNone;
    }
    start[t] -= mine;
    const uint32_t n = *count;
    constexpr uint32_t kChunk = 256u * kSortPer;
    for (uint32_t c = blockIdx.x * kChunk; c < n; c += gridDim.x * kChunk) {
        h[t] = 0u;
        __syncthreads();
        uint32_t slot[kSortPer], rank[kSortPer];
        uint16_t b[kSortPer];
#pragma unroll
        for (uint32_t j = 0; j < kSortPer; j++) {
            const uint32_t q = c + j * 256u + t;
            slot[j] = q < n ? in[q] : kNone;
            if (slot[j] != kNone) {
                b[j] = req[slot[j]] & (kWfBins - 1u);
                rank[j] = atomicAdd(&h[b[j]], 1u);
            }
        }
        __syncthreads();
        if (h[t]) run[t] = start[t] + atomicAdd(&cursor[t], h[t]);
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kSortPer; j++)
            if (slot[j] != kNone) out[run[b[j]] + rank[j]] = slot[j];
        __syncthreads();
    }
}

// Every slot empty and on the end queue; the item queue at 0.
__global__ void k_wf_init(WfSlots S, WfQueues Q, uint32_t slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kWfCounts && i != kWfEND0 && i != kWfLIVE) Q.counts[i] = 0u;
    if (i == kWfEND0 || i == kWfLIVE) Q.counts[i] = slots;
    if (i >= slots) return;
    Q.end[0][i] = i;
    S.p_spec[i] = make_float4(0.0f, 0.0f, 0.0f, bitsf(kNone));
    S.p_rad[i] = make_float4(0.0f, 0.0f, 0.0f, bitsf(0u));
}

// Stage 0: finish ended paths, start new ones.
template <bool kBatched>
__global__ void __launch_bounds__(64) k_wf_regen(const PtRequest* __restrict__ reqs, uint32_t n_reqs,
                                                 uint32_t total_items, WfSlots S, WfQueues Q, uint32_t cur) {
    const uint32_t nxt = cur ^ 1u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // queues this iteration's later stages fill
        Q.counts[kWfFF0 + nxt] = 0u;
        Q.counts[kWfEV] = 0u;
        Q.counts[kWfEND0 + nxt] = 0u;
        Q.counts[kWfSH0 + nxt] = 0u;
        Q.counts[kWfSC] = 0u;
    }
    if (kBatched && blockIdx.x == 0) clear_bins(Q);
    const uint32_t n = Q.counts[kWfEND0 + cur];
    const uint32_t lane = __lane_id();
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t q = base + lane;
        const bool active = q < n;
        const uint32_t slot = active ? Q.end[cur][q] : 0u;
        bool push = false, need = false;
        PathState st;
        SlotRef ref{kNone, 0u, 0u};
        float4 c0 = make_float4(0.f, 0.f, 0.f, 0.f), c1 = c0, c2 = c0;
        if (active) {
            load_path(S, slot, st, ref);
            need = true;
            if (ref.item != kNone) {
                const PtRequest& R = request<kBatched>(reqs, ref.req);
                c0 = S.c0[slot], c1 = S.c1[slot], c2 = S.c2[slot];
                // pt.comp:546-551: firefly clamp, sums of radiance and squared luminance.
                const V3 r = path_end(st);
                const float lum = dot(r, V3(0.2126f, 0.7152f, 0.0722f));
                c0 = make_float4(c0.x + r.x, c0.y + r.y, c0.z + r.z, c0.w);
                c1.w += lum * lum;
                ref.k++;
                uint32_t px, py, ci, pi, path_index, n_paths;
                item_coords(R, ref.item, ref.k, px, py, ci, pi, path_index, n_paths);
                if (ref.k < n_paths) {
                    const FrameConsts fc = frame_consts(R.scene.u);
                    path_begin(R.scene, fc, px, py, path_index, st);
                    push = true;
                    need = false;
                } else {
                    store_partial(R, ci, pi, c0, c1, c2);
                }
            }
        }
        // Lanes without a chunk take the next ones; chunks of pixels outside the
        // image are finished at once.
        while (__ballot(need)) {
            const uint32_t it = wave_take(need, &Q.counts[kWfHEAD]);
            if (need) {
                if (it >= total_items) {
                    ref.item = kNone;
                    need = false;
                } else {
                    ref.req = request_of_item<kBatched>(reqs, n_reqs, it);
                    const PtRequest& R = request<kBatched>(reqs, ref.req);
                    ref.item = it;
                    ref.k = 0u;
                    uint32_t px, py, ci, pi, path_index, n_paths;
                    item_coords(R, it, 0u, px, py, ci, pi, path_index, n_paths);
                    if (px < R.width && py < R.height) {
                        c0 = c1 = c2 = make_float4(0.f, 0.f, 0.f, 0.f);
                        const FrameConsts fc = frame_consts(R.scene.u);
                        path_begin(R.scene, fc, px, py, path_index, st);
                        push = true;
                        need = false;
                    }
                }
            }
        }
        if (active) {
            if (push) {
                store_path(S, slot, st, ref);
                S.p_flight[slot] = make_uint2(0u, kWfNewFlight);
                S.c0[slot] = c0, S.c1[slot] = c1, S.c2[slot] = c2;
                if (kBatched) S.req[slot] = uint16_t(ref.req);
            } else {
                S.p_spec[slot] = make_float4(0.f, 0.f, 0.f, bitsf(kNone));
            }
        }
        const bool idle = active && !push;
        const uint64_t gone = __ballot(idle);
        if (gone && int(lane) == __ffsll((unsigned long long)gone) - 1)
            atomicSub(&Q.counts[kWfLIVE], uint32_t(__popcll(gone)));
        wave_push(push, slot, Q.ff[cur], &Q.counts[kWfFF0 + cur]);
    }
}

// Served engine: every slot idle, nothing queued.
__global__ void k_wf_init_served(WfSlots S, WfQueues Q, uint32_t slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kWfCounts) Q.counts[i] = i == kWfIDLE0 ? slots : 0u;
    if (i >= slots) return;
    Q.idle[0][i] = i;
    S.p_spec[i] = make_float4(0.0f, 0.0f, 0.0f, bitsf(kNone));
    S.p_rad[i] = make_float4(0.0f, 0.0f, 0.0f, bitsf(0u));
}

// Stage 0 of the served engine: as k_wf_regen, but work items come from the
// active requests' own counters, a finished chunk counts towards its request's
// `done`, and slots without work wait on an idle queue (offered work again when
// the host revives them) instead of retiring.
__global__ void __launch_bounds__(64) k_wf_regen_served(const PtRequest* __restrict__ reqs, WfActive A, WfServed V,
                                                        WfSlots S, WfQueues Q, uint32_t cur) {
    const uint32_t nxt = cur ^ 1u;
    const uint32_t idle_out = V.revive ? (V.idle_in ^ 1u) : V.idle_in;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Q.counts[kWfFF0 + nxt] = 0u;
        Q.counts[kWfEV] = 0u;
        Q.counts[kWfEND0 + nxt] = 0u;
        Q.counts[kWfSH0 + nxt] = 0u;
        Q.counts[kWfSC] = 0u;
    }
    if (blockIdx.x == 0) clear_bins(Q);
    const uint32_t n_end = Q.counts[kWfEND0 + cur];
    const uint32_t n = n_end + (V.revive ? Q.counts[kWfIDLE0 + V.idle_in] : 0u);
    const uint32_t lane = __lane_id();
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t q = base + lane;
        const bool active = q < n;
        const uint32_t slot = active ? (q < n_end ? Q.end[cur][q] : Q.idle[V.idle_in][q - n_end]) : 0u;
        bool push = false, need = false;
        PathState st;
        SlotRef ref{kNone, 0u, 0u};
        float4 c0 = make_float4(0.f, 0.f, 0.f, 0.f), c1 = c0, c2 = c0;
        bool finishing = false;
        if (active) {
            need = true;
            // One round of loads: an idle slot holds no path (item kNone), and
            // its other fields are ignored.
            load_path(S, slot, st, ref);
            const float4 s0 = S.c0[slot], s1 = S.c1[slot], s2 = S.c2[slot];
            if (ref.item != kNone) {
                finishing = true;
                c0 = s0, c1 = s1, c2 = s2;
                const V3 r = path_end(st);  // pt.comp:546-551
                const float lum = dot(r, V3(0.2126f, 0.7152f, 0.0722f));
                c0 = make_float4(c0.x + r.x, c0.y + r.y, c0.z + r.z, c0.w);
                c1.w += lum * lum;
                ref.k++;
            }
        }
        for_each_request<true>(finishing, ref.req, [&](uint32_t rq) {
            const PtRequest& R = reqs[rq];
            uint32_t px, py, ci, pi, path_index, n_paths;
            item_coords(R, ref.item, ref.k, px, py, ci, pi, path_index, n_paths);
            const bool chunk_done = ref.k >= n_paths;
            if (!chunk_done) {
                path_begin(R.scene, frame_consts(R.scene.u), px, py, path_index, st);
                push = true;
                need = false;
            } else {
                store_partial(R, ci, pi, c0, c1, c2);
            }
            // One atomic per wave and request: per-lane atomics on one counter
            // serialise millions of updates an iteration.
            const uint64_t m = __ballot(chunk_done);
            if (m && int(__lane_id()) == __ffsll((unsigned long long)m) - 1)
                atomicAdd(&V.done[rq * kWfCounterStride], uint32_t(__popcll(m)));
        });
        // New items from the active requests in list order (the host keeps the
        // list in priority order): a request's work is handed out before the
        // next one's, and the next one fills the slots its tail leaves idle.
        // A request whose items are all handed out costs one atomic per wave
        // until the host drops it from the list after the batch (its head grows
        // by at most the slot count per iteration, far from wrapping).
        uint32_t a = 0;
        while (a < A.n && __ballot(need)) {
            const uint32_t r = A.r[a];
            const PtRequest& R = reqs[r];
            const uint32_t it = wave_take(need, &V.head[r * kWfCounterStride]);
            const bool got = need && it < R.n_items;
            const bool dry = need && !got;
            if (got) {
                ref.req = r;
                ref.item = it;
                ref.k = 0u;
                uint32_t px, py, ci, pi, path_index, n_paths;
                item_coords(R, it, 0u, px, py, ci, pi, path_index, n_paths);
                if (px < R.width && py < R.height) {
                    c0 = c1 = c2 = make_float4(0.f, 0.f, 0.f, 0.f);
                    path_begin(R.scene, frame_consts(R.scene.u), px, py, path_index, st);
                    push = true;
                    need = false;
                }
            }
            const uint64_t pad = __ballot(got && need);  // padding pixels: nothing to trace
            if (pad && int(lane) == __ffsll((unsigned long long)pad) - 1)
                atomicAdd(&V.done[r * kWfCounterStride], uint32_t(__popcll(pad)));
            if (__ballot(dry)) a++;
        }
        if (active) {
            if (push) {
                store_path(S, slot, st, ref);
                S.p_flight[slot] = make_uint2(0u, kWfNewFlight);
                S.c0[slot] = c0, S.c1[slot] = c1, S.c2[slot] = c2;
                S.req[slot] = uint16_t(ref.req);
            } else {
                S.p_spec[slot] = make_float4(0.f, 0.f, 0.f, bitsf(kNone));
            }
        }
        wave_push(push, slot, Q.ff[cur], &Q.counts[kWfFF0 + cur]);
        wave_push(active && !push, slot, Q.idle[idle_out], &Q.counts[kWfIDLE0 + idle_out]);
    }
}

// Stage 1: free flight over the `*n_in` entries of `in` (kNone: padding): new
// flights, and flights a previous launch suspended (p_flight); each runs at
// most `steps` tracking iterations, and one that is not done goes back on ff.
// Without the ghost branch and the density fetches (isolate mode) the kernel is
// held to 168 registers (three waves per SIMD, 14 spilled): PT-1 218.4 against
// 221.8 ms per frame at two waves; with the density read, three waves (24
// spilled, in the fetches) made Batched-32's path-traced requests 15 % slower.
template <bool kBatched, bool kGhosts = true, bool kDensity = true>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kGhosts || kDensity ? 2 : 3)))
k_wf_extend(const PtRequest* __restrict__ reqs, WfSlots S, WfQueues Q, const uint32_t* __restrict__ in,
            const uint32_t* __restrict__ n_in, uint32_t cur, int steps) {
    const uint32_t nxt = cur ^ 1u;
    const uint32_t n = *n_in;
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t q = base + __lane_id();
        const uint32_t slot = q < n ? in[q] : kNone;
        const bool active = slot != kNone;
        Flow flow = kFlowExit;
        PathState st;
        SlotRef ref{0u, 0u, 0u};
        EventState ev;
        uint2 fl = make_uint2(0u, kWfNewFlight);
        if (active) {
            load_path(S, slot, st, ref);
            fl = S.p_flight[slot];
        }
        with_wave_request<kBatched>(active, ref.req, [&](uint32_t r) {
            const PtRequest& R = reqs[r];
            Flight f = flight_begin(R.scene.u, st.origin, st.dir);
            if (fl.y != kWfNewFlight) f.tt = bitsf(fl.x), f.i = int(fl.y);
            flow = stage_extend_steps<kGhosts, kDensity>(R.scene, frame_consts(R.scene.u), st, ev, f, f.i + steps);
            fl = make_uint2(fbits(f.tt), uint32_t(f.i));
        });
        if (active) {
            if (flow == kFlowHit) {
                store_event(S, slot, ev);
            } else if (flow == kFlowSuspended) {
                S.p_flight[slot] = fl;
            } else {
                S.p_rad[slot] = f4(st.radiance, bitsf(ref.k | (ref.req << 16)));
                if (st.bounce == 0u) {  // background: pt.comp:434-438
                    add4(&S.c0[slot], make_float4(0.f, 0.f, 0.f, 1.0f));
                    add4(&S.c1[slot], make_float4(1.0f, 1.0f, 1.0f, 0.f));
                    add4(&S.c2[slot], make_float4(0.f, 0.f, 0.f, 1e4f));
                }
            }
            S.p_dir[slot].w = bitsf(st.rng.state);
            S.p_spec[slot] = f4(st.spectral, bitsf(ref.item));
        }
        wave_push(active && flow == kFlowHit, slot, Q.ev, &Q.counts[kWfEV]);
        wave_push(active && flow == kFlowSuspended, slot, Q.ff[nxt], &Q.counts[kWfFF0 + nxt]);
        wave_push(active && flow == kFlowExit, slot, Q.end[nxt], &Q.counts[kWfEND0 + nxt]);
    }
}

// Stage 2: shading frame and light sample, over the `*n_in` entries of `in`
// (kNone: padding).
template <bool kBatched>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_wf_shade(const PtRequest* __restrict__ reqs, WfSlots S, WfQueues Q, const uint32_t* __restrict__ in,
           const uint32_t* __restrict__ n_in) {
    const uint32_t n = *n_in;
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t q = base + __lane_id();
        const uint32_t slot = q < n ? in[q] : kNone;
        const bool active = slot != kNone;
        PathState st;  // the stage reads the direction, bounce and random stream
        EventState ev;
        uint32_t req = 0u;
        if (active) {
            const float4 d = S.p_dir[slot];
            st.dir = xyz(d);
            st.rng.state = fbits(d.w);
            st.bounce = fbits(S.p_thr[slot].w) >> 18;
            req = fbits(S.p_rad[slot].w) >> 16;
            load_event(S, slot, ev);
        }
        with_wave_request<kBatched>(active, req, [&](uint32_t r) {
            const PtRequest& R = reqs[r];
            stage_shade(R.scene, frame_consts(R.scene.u), st, ev);
        });
        if (active) {
            store_event(S, slot, ev);
            S.e_n[slot] = f4(ev.n, 0.0f);
            S.e_light[slot] = f4(ev.light_dir, 0.0f);
            S.p_dir[slot].w = bitsf(st.rng.state);
            if (st.bounce == 0u) {  // guides of the first event (pt.comp:459-462)
                add4(&S.c1[slot], f4(st.albedo, 0.0f));
                add4(&S.c2[slot], f4(st.normal, st.depth));
            }
        }
    }
}

// Stage 3: shadow ray: the events among the `*n_in` entries of `in` (new rays;
// kNone: padding), then the suspended rays of sh[cur]; each runs at most
// `steps` tracking iterations. Finished rays go on to the light-map stage and
// scatter (queue sc).
template <bool kBatched>
__global__ void __launch_bounds__(64) k_wf_shadow(const PtRequest* __restrict__ reqs, WfSlots S, WfQueues Q,
                                                  const uint32_t* __restrict__ in, const uint32_t* __restrict__ n_in,
                                                  uint32_t cur, int steps) {
    const uint32_t nxt = cur ^ 1u;
    const uint32_t n_new = *n_in;
    const uint32_t n = n_new + Q.counts[kWfSH0 + cur];
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t q = base + __lane_id();
        const bool resumed = q >= n_new;
        const uint32_t slot = q < n ? (resumed ? Q.sh[cur][q - n_new] : in[q]) : kNone;
        const bool active = slot != kNone;
        uint32_t req = 0u;
        Rng rng;
        rng.state = 0u;
        V3 origin(0.0f), dir(0.0f), result(1.0f);
        uint2 fl = make_uint2(0u, 0u);
        bool done = true;
        if (active) {
            req = fbits(S.p_rad[slot].w) >> 16;
            rng.state = fbits(S.p_dir[slot].w);
            origin = xyz(S.e_idx[slot]);
            dir = xyz(S.e_light[slot]);
            if (resumed) {
                result = xyz(S.e_lt[slot]);
                fl = S.p_flight[slot];
            }
        }
        for_each_request<kBatched>(active, req, [&](uint32_t r) {
            const DevScene& sc = reqs[r].scene;
            Flight f = flight_begin(sc.u, origin, dir);
            if (resumed) f.tt = bitsf(fl.x), f.i = int(fl.y);
            done = transmittance_steps(sc, f, result, rng, f.i + steps);
            fl = make_uint2(fbits(f.tt), uint32_t(f.i));
        });
        if (active) {
            S.e_lt[slot] = f4(result, 0.0f);
            if (!done) S.p_flight[slot] = fl;
            S.p_dir[slot].w = bitsf(rng.state);
        }
        wave_push(active && done, slot, Q.sc, &Q.counts[kWfSC]);
        wave_push(active && !done, slot, Q.sh[nxt], &Q.counts[kWfSH0 + nxt]);
    }
}

// Stage 3b, for frames with a light map: its sample and shadow ray (for the
// paths whose key-light shadow ray is done).
template <bool kBatched>
__global__ void __launch_bounds__(64) k_wf_shadow_map(const PtRequest* __restrict__ reqs, WfSlots S, WfQueues Q,
                                                      const uint32_t* __restrict__ in) {
    const uint32_t n = Q.counts[kWfSC];
    for (uint32_t q = blockIdx.x * 64u + __lane_id(); q < n; q += gridDim.x * 64u) {
        const uint32_t slot = in[q];
        const uint32_t req = fbits(S.p_rad[slot].w) >> 16;
        PathState st;
        st.rng.state = fbits(S.p_dir[slot].w);
        EventState ev;
        ev.idx = xyz(S.e_idx[slot]);
        bool map = false;
        for_each_request<kBatched>(true, req, [&](uint32_t r) {
            map = reqs[r].scene.u->lighting[3] == 2u;
            if (map) stage_shadow_map(reqs[r].scene, st, ev);
        });
        if (map) {
            S.e_map[slot] = f4(ev.map_dir, ev.map_pdf);
            S.e_maplt[slot] = f4(ev.map_light, 0.0f);
            S.p_dir[slot].w = bitsf(st.rng.state);
        }
    }
}

// Stage 4: next-event estimate, next direction, roulette (for the paths whose
// shadow rays are done).
template <bool kBatched>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_wf_scatter(const PtRequest* __restrict__ reqs, WfSlots S, WfQueues Q,
                                                   const uint32_t* __restrict__ in, uint32_t cur) {
    const uint32_t nxt = cur ^ 1u;
    const uint32_t n = Q.counts[kWfSC];
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t q = base + __lane_id();
        const bool active = q < n;
        const uint32_t slot = active ? in[q] : 0u;
        bool more = false;
        PathState st;
        SlotRef ref{0u, 0u, 0u};
        EventState ev;
        if (active) {
            load_path(S, slot, st, ref);
            load_event(S, slot, ev);
            ev.n = xyz(S.e_n[slot]);
            ev.light_dir = xyz(S.e_light[slot]);
            ev.light_t = xyz(S.e_lt[slot]);
        }
        for_each_request<kBatched>(active, ref.req, [&](uint32_t r) {
            const PtRequest& R = reqs[r];
            const FrameConsts fc = frame_consts(R.scene.u);
            ev.map_pdf = 0.0f;
            if (fc.env_map) {
                const float4 m = S.e_map[slot], ml = S.e_maplt[slot];
                ev.map_dir = xyz(m), ev.map_pdf = m.w, ev.map_light = xyz(ml);
            }
            more = stage_scatter(R.scene, fc, st, ev) && path_continues(fc, st);
        });
        if (active) {
            store_path(S, slot, st, ref);
            if (more) S.p_flight[slot].y = kWfNewFlight;
        }
        wave_push(active && more, slot, Q.ff[nxt], &Q.counts[kWfFF0 + nxt]);
        wave_push(active && !more, slot, Q.end[nxt], &Q.counts[kWfEND0 + nxt]);
    }
}

// Where a path waits between two iterations: a flight to begin or resume
// (ff), a suspended shadow ray (sh) or its end (end, for regen).
enum TailPhase : uint32_t { kFlight, kShadow, kScatter, kEnded };

// Runs the path in `slot`, waiting at `phase`, to its end, then the rest of its
// chunk, in one lane: the stages run in the order the iterations would run
// them, with the same functions, so the path draws the same random numbers and
// the chunk sums get the same additions in the same order as with iterations
// to the end. The chunk's sums are stored and the slot is left idle. False for
// an idle slot.
template <bool kGhosts = true>
__device__ __forceinline__ bool run_to_end(const PtRequest& R, const WfSlots& S, uint32_t slot, uint32_t phase) {
    const DevScene& sc = R.scene;
    const FrameConsts fc = frame_consts(sc.u);
    PathState st;
    SlotRef ref;
    load_path(S, slot, st, ref);
    if (ref.item == kNone) return false;  // an idle slot
    float4 c0 = S.c0[slot], c1 = S.c1[slot], c2 = S.c2[slot];
    bool resumed = phase == kFlight && S.p_flight[slot].y != kWfNewFlight;
    EventState ev;
    if (phase == kShadow) {  // as k_wf_shadow, without a bound
        load_event(S, slot, ev);
        ev.n = xyz(S.e_n[slot]);
        ev.light_dir = xyz(S.e_light[slot]);
        V3 result = xyz(S.e_lt[slot]);
        const uint2 fl = S.p_flight[slot];
        Flight f = flight_begin(sc.u, ev.idx, ev.light_dir);
        f.tt = bitsf(fl.x), f.i = int(fl.y);
        transmittance_steps(sc, f, result, st.rng, 4096);
        ev.light_t = result;
        phase = kScatter;
    }
    for (;;) {
        if (phase == kFlight) {  // k_wf_extend, then k_wf_shade and k_wf_shadow
            Flight f = flight_begin(sc.u, st.origin, st.dir);
            if (resumed) {
                const uint2 fl = S.p_flight[slot];
                f.tt = bitsf(fl.x), f.i = int(fl.y);
                resumed = false;
            }
            if (stage_extend_steps<kGhosts>(sc, fc, st, ev, f, 4096) == kFlowHit) {
                stage_shade(sc, fc, st, ev);
                if (st.bounce == 0u) {
                    c1 = make_float4(c1.x + st.albedo.x, c1.y + st.albedo.y, c1.z + st.albedo.z, c1.w + 0.0f);
                    c2 = make_float4(c2.x + st.normal.x, c2.y + st.normal.y, c2.z + st.normal.z, c2.w + st.depth);
                }
                stage_shadow(sc, st, ev);
                phase = kScatter;
            } else {
                if (st.bounce == 0u) {  // background: pt.comp:434-438
                    c0 = make_float4(c0.x + 0.f, c0.y + 0.f, c0.z + 0.f, c0.w + 1.0f);
                    c1 = make_float4(c1.x + 1.0f, c1.y + 1.0f, c1.z + 1.0f, c1.w + 0.f);
                    c2 = make_float4(c2.x + 0.f, c2.y + 0.f, c2.z + 0.f, c2.w + 1e4f);
                }
                phase = kEnded;
            }
        }
        if (phase == kScatter) {  // k_wf_shadow_map, k_wf_scatter
            ev.map_pdf = 0.0f;
            if (fc.env_map) stage_shadow_map(sc, st, ev);
            phase = stage_scatter(sc, fc, st, ev) && path_continues(fc, st) ? kFlight : kEnded;
            continue;
        }
        // k_wf_regen: the path's radiance into its chunk's sums (pt.comp:546-551),
        // then the chunk's next path or its partial sums.
        const V3 r = path_end(st);
        const float lum = dot(r, V3(0.2126f, 0.7152f, 0.0722f));
        c0 = make_float4(c0.x + r.x, c0.y + r.y, c0.z + r.z, c0.w);
        c1.w += lum * lum;
        ref.k++;
        uint32_t px, py, ci, pi, path_index, n_paths;
        item_coords(R, ref.item, ref.k, px, py, ci, pi, path_index, n_paths);
        if (ref.k >= n_paths) {
            store_partial(R, ci, pi, c0, c1, c2);
            break;
        }
        path_begin(sc, fc, px, py, path_index, st);
        phase = kFlight;
    }
    S.p_spec[slot] = make_float4(0.f, 0.f, 0.f, bitsf(kNone));  // the slot is idle
    return true;
}

// The tail of a single-request frame, once every work item is handed out and
// few paths remain: each path still in flight runs to its end in one launch,
// a lane per slot. Entries: the ended paths of end[cur] (regen's turn), the
// suspended shadow rays of sh[cur] and the flights of ff[cur] (new or suspended).
// Held to two waves per SIMD (256 registers, ~170 spilled, against 393 unbounded):
// each path's chain is latency-bound either way, and twice the resident lanes let
// the tail start at 128 Ki live paths instead of 64 Ki (PT-1 205.9 against 210.9
// ms per frame).
template <bool kGhosts>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_wf_finish(const PtRequest* __restrict__ reqs, WfSlots S, WfQueues Q,
                                                  uint32_t cur) {
    const uint32_t n_end = Q.counts[kWfEND0 + cur], n_sh = Q.counts[kWfSH0 + cur];
    const uint32_t n = n_end + n_sh + Q.counts[kWfFF0 + cur];
    for (uint32_t q = blockIdx.x * 64u + __lane_id(); q < n; q += gridDim.x * 64u) {
        uint32_t phase, slot;
        if (q < n_end) slot = Q.end[cur][q], phase = kEnded;
        else if (q < n_end + n_sh) slot = Q.sh[cur][q - n_end], phase = kShadow;
        else slot = Q.ff[cur][q - n_end - n_sh], phase = kFlight;
        run_to_end<kGhosts>(reqs[0], S, slot, phase);
    }
}

// Served engine, between two iterations: the paths of request-table entry `e`
// leave the wavefront. Its flights on ff[cur] and suspended shadow rays on
// sh[cur] have their queue entries set to kNone (padding, which the stages and
// the queue sort skip); its ended paths on end[cur] stay for regen. Without
// `drop`, up to `cap` of them are listed for k_wf_tail (slot | phase << 30;
// beyond `cap` they stay queued). With `drop` (a cancelled request), their
// slots are left idle on end[cur] and their work items counted in `done`.
__global__ void __launch_bounds__(256) k_wf_tail_gather(WfSlots S, WfQueues Q, uint32_t cur, uint32_t e, bool drop,
                                                        uint32_t* __restrict__ list, uint32_t* __restrict__ count,
                                                        uint32_t cap, uint32_t* __restrict__ done) {
    const uint32_t n_sh = Q.counts[kWfSH0 + cur];
    const uint32_t n = n_sh + Q.counts[kWfFF0 + cur];
    const uint32_t lane = __lane_id();
    for (uint32_t base = blockIdx.x * 256u + (threadIdx.x & ~63u); base < n; base += gridDim.x * 256u) {
        const uint32_t q = base + lane;
        uint32_t* at = nullptr;
        uint32_t slot = kNone;
        if (q < n) {
            at = q < n_sh ? &Q.sh[cur][q] : &Q.ff[cur][q - n_sh];
            slot = *at;
        }
        const bool mine = slot != kNone && S.req[slot] == e;
        if (drop) {
            if (mine) {
                *at = kNone;
                S.p_spec[slot] = make_float4(0.f, 0.f, 0.f, bitsf(kNone));
            }
            wave_push(mine, slot, Q.end[cur], &Q.counts[kWfEND0 + cur]);
            const uint64_t m = __ballot(mine);
            if (m && int(lane) == __ffsll((unsigned long long)m) - 1) atomicAdd(done, uint32_t(__popcll(m)));
        } else {
            const uint32_t i = wave_take(mine, count);
            if (mine && i < cap) {
                list[i] = slot | ((q < n_sh ? uint32_t(kShadow) : uint32_t(kFlight)) << 30);
                *at = kNone;
            }
        }
    }
}

// Served engine: the paths k_wf_tail_gather listed for entry `e` run to their
// ends (run_to_end: the frame is the same as with iterations), each finished
// work item counted in `done`. Runs on its own stream beside the iterations,
// which no longer reach these slots.
__global__ void __launch_bounds__(64) k_wf_tail(const PtRequest* __restrict__ reqs, WfSlots S, uint32_t e,
                                                const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                                                uint32_t cap, uint32_t* __restrict__ done) {
    const PtRequest& R = reqs[e];
    const uint32_t n = min(*count, cap);
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t q = base + __lane_id();
        bool finished = false;
        if (q < n) finished = run_to_end(R, S, list[q] & 0x3fffffffu, list[q] >> 30);
        const uint64_t m = __ballot(finished);
        if (m && int(__lane_id()) == __ffsll((unsigned long long)m) - 1) atomicAdd(done, uint32_t(__popcll(m)));
    }
}

// Served engine, once k_wf_tail has ended: its slots, idle now, back on
// end[cur], where regen gives them work or parks them.
__global__ void __launch_bounds__(256) k_wf_tail_return(WfQueues Q, uint32_t cur, const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ count, uint32_t cap) {
    const uint32_t n = min(*count, cap);
    for (uint32_t base = blockIdx.x * 256u + (threadIdx.x & ~63u); base < n; base += gridDim.x * 256u) {
        const uint32_t q = base + __lane_id();
        wave_push(q < n, q < n ? list[q] & 0x3fffffffu : 0u, Q.end[cur], &Q.counts[kWfEND0 + cur]);
    }
}

// The chunk sums of every pixel of a request, added in chunk order to the
// accumulation and guide images; the running mean to the HDR image (pt.comp:554-560).
__global__ void k_wf_resolve(PtRequest R) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R.n_px) return;
    const uint32_t tile = p >> 6, w = p & 63u;
    const uint32_t px = (tile % R.tiles_x) * 8u + (w & 7u), py = (tile / R.tiles_x) * 8u + (w >> 3);
    if (px >= R.width || py >= R.height) return;
    float4 s[3];
    for (int j = 0; j < 3; j++) s[j] = R.partial[size_t(j) * R.n_px + p];
    for (uint32_t c = 1; c < R.n_chunks; c++)
        for (int j = 0; j < 3; j++) {
            const float4 v = R.partial[(size_t(c) * 3u + j) * R.n_px + p];
            s[j] = make_float4(s[j].x + v.x, s[j].y + v.y, s[j].z + v.z, s[j].w + v.w);
        }
    const size_t i = size_t(py) * R.width + px;
    float4* out[3] = {R.accum, R.albedo, R.normal};
    for (int j = 0; j < 3; j++) {
        if (R.first_path != 0u) {
            const float4 a = out[j][i];
            s[j] = make_float4(a.x + s[j].x, a.y + s[j].y, a.z + s[j].z, a.w + s[j].w);
        }
        out[j][i] = s[j];
    }
    const float n = float(R.first_path + R.samples);
    const __half h[4] = {__float2half_rn(s[0].x / n), __float2half_rn(s[0].y / n), __float2half_rn(s[0].z / n),
                         __float2half_rn(s[0].w / n)};
    uint2 word;
    word.x = uint32_t(__half_as_ushort(h[0])) | (uint32_t(__half_as_ushort(h[1])) << 16);
    word.y = uint32_t(__half_as_ushort(h[2])) | (uint32_t(__half_as_ushort(h[3])) << 16);
    reinterpret_cast<uint2*>(R.hdr)[i] = word;
}

template <class K> int waves_for(K kernel, int cus, int per_cu_cap) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, 64, 0) != hipSuccess || blocks < 1) blocks = 4;
    return cus * std::min(blocks, per_cu_cap);
}

}  // namespace

void wf_init(const WfSlots& S, const WfQueues& Q, uint32_t slots, hipStream_t st) {
    hipLaunchKernelGGL(k_wf_init, dim3((std::max(slots, uint32_t(kWfCounts)) + 255) / 256), dim3(256), 0, st, S, Q,
                       slots);
}

WfGrid wf_grid(int cus) {
    WfGrid g;
    g.single = WfStageGrid{waves_for(k_wf_regen<false>, cus, 32), waves_for(k_wf_extend<false>, cus, 32),
                           waves_for(k_wf_shade<false>, cus, 32), waves_for(k_wf_shadow<false>, cus, 32),
                           waves_for(k_wf_scatter<false>, cus, 32)};
    g.extend_no_ghosts = waves_for(k_wf_extend<false, false>, cus, 32);
    g.extend_isolate = waves_for(k_wf_extend<false, false, false>, cus, 32);
    g.batched_extend_isolate = waves_for(k_wf_extend<true, false, false>, cus, 32);
    // The served regen's occupancy bounds the batched regen grid too (it runs in its place).
    g.batched = WfStageGrid{std::min(waves_for(k_wf_regen<true>, cus, 32), waves_for(k_wf_regen_served, cus, 32)),
                            waves_for(k_wf_extend<true>, cus, 32), waves_for(k_wf_shade<true>, cus, 32),
                            waves_for(k_wf_shadow<true>, cus, 32), waves_for(k_wf_scatter<true>, cus, 32)};
    g.sort = cus * 2;
    // Measured at PT-1 (512 x 512 x 400, 1024^2, 64 paths): 16 / 32 gave 220
    // Msamples/s against 124 unsuspended, 8-24 / 16-48 within 3 %; with the
    // three-wave free-flight kernel of isolate mode the autotuner picked 24 / 48
    // (210.7 against 218.8 ms per frame), and with the two-wave tail and 8M slots
    // 24 / 32 is faster still (201.8 against 203.5 ms).
    g.flight_steps = 24;
    g.shadow_steps = 32;
    g.tail_paths = 1u << 17;
    return g;
}

namespace {

// Extend over ff[cur], then shade, shadow and scatter over ev; with `sort`,
// each queue binned by request first.
template <bool kBatched>
void wf_stages(const PtRequest* reqs, const WfSlots& S, const WfQueues& Q, uint32_t cur, const WfGrid& grid, bool sort,
               bool env_map, bool ghosts, bool density, hipStream_t st) {
    const WfStageGrid& g = kBatched ? grid.batched : grid.single;
    const uint32_t* ff = Q.ff[cur];
    const uint32_t* n_ff = Q.counts + kWfFF0 + cur;
    const uint32_t* ev = Q.ev;
    const uint32_t* n_ev = Q.counts + kWfEV;
    if (sort) {
        uint32_t* hist = Q.bins;
        hipLaunchKernelGGL(k_wf_bin_count, dim3(grid.sort), dim3(256), 0, st, ff, n_ff, S.req, hist);
        hipLaunchKernelGGL(k_wf_bin_scatter, dim3(grid.sort), dim3(256), 0, st, ff, n_ff, S.req, hist,
                           hist + kWfBins, Q.sorted, Q.counts + kWfSORTFF);
        ff = Q.sorted;
        n_ff = Q.counts + kWfSORTFF;
    }
    // Engines over several requests keep the ghost branch while any request reads the
    // density: their ghost-free form (k_wf_extend<true, false, true>) served Batched-32 in
    // 0.91 s against 0.89 s.
    if (!ghosts && !density)
        hipLaunchKernelGGL((k_wf_extend<kBatched, false, false>),
                           dim3(kBatched ? grid.batched_extend_isolate : grid.extend_isolate), dim3(64), 0, st, reqs, S,
                           Q, ff, n_ff, cur, grid.flight_steps);
    else if (!kBatched && !ghosts)
        hipLaunchKernelGGL((k_wf_extend<false, false>), dim3(grid.extend_no_ghosts), dim3(64), 0, st, reqs, S, Q, ff,
                           n_ff, cur, grid.flight_steps);
    else
        hipLaunchKernelGGL(k_wf_extend<kBatched>, dim3(g.extend), dim3(64), 0, st, reqs, S, Q, ff, n_ff, cur,
                           grid.flight_steps);
    if (sort) {
        uint32_t* hist = Q.bins + 2 * kWfBins;
        hipLaunchKernelGGL(k_wf_bin_count, dim3(grid.sort), dim3(256), 0, st, ev, n_ev, S.req, hist);
        hipLaunchKernelGGL(k_wf_bin_scatter, dim3(grid.sort), dim3(256), 0, st, ev, n_ev, S.req, hist,
                           hist + kWfBins, Q.sorted, Q.counts + kWfSORTEV);
        ev = Q.sorted;
        n_ev = Q.counts + kWfSORTEV;
    }
    hipLaunchKernelGGL(k_wf_shade<kBatched>, dim3(g.shade), dim3(64), 0, st, reqs, S, Q, ev, n_ev);
    hipLaunchKernelGGL(k_wf_shadow<kBatched>, dim3(g.shadow), dim3(64), 0, st, reqs, S, Q, ev, n_ev, cur,
                       grid.shadow_steps);
    if (env_map) hipLaunchKernelGGL(k_wf_shadow_map<kBatched>, dim3(g.shadow), dim3(64), 0, st, reqs, S, Q, Q.sc);
    hipLaunchKernelGGL(k_wf_scatter<kBatched>, dim3(g.scatter), dim3(64), 0, st, reqs, S, Q, Q.sc, cur);
}

}  // namespace

void wf_iteration(const PtRequest* reqs, uint32_t n_reqs, uint32_t total_items, const WfSlots& S, const WfQueues& Q,
                  uint32_t cur, const WfGrid& g, bool env_map, bool ghosts, bool density, hipStream_t st) {
    if (n_reqs > 1) {
        hipLaunchKernelGGL(k_wf_regen<true>, dim3(g.batched.regen), dim3(64), 0, st, reqs, n_reqs, total_items, S, Q, cur);
        wf_stages<true>(reqs, S, Q, cur, g, true, env_map, ghosts, density, st);  // n_reqs <= kWfBins (WfEngine::run)
    } else {
        hipLaunchKernelGGL(k_wf_regen<false>, dim3(g.single.regen), dim3(64), 0, st, reqs, n_reqs, total_items, S, Q, cur);
        wf_stages<false>(reqs, S, Q, cur, g, false, env_map, ghosts, density, st);
    }
}

void wf_init_served(const WfSlots& S, const WfQueues& Q, uint32_t slots, hipStream_t st) {
    hipLaunchKernelGGL(k_wf_init_served, dim3((std::max(slots, uint32_t(kWfCounts)) + 255) / 256), dim3(256), 0, st, S,
                       Q, slots);
}

void wf_iteration_served(const PtRequest* reqs, const WfActive& A, const WfServed& V, const WfSlots& S,
                         const WfQueues& Q, uint32_t cur, const WfGrid& g, bool env_map, bool isolate,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_wf_regen_served, dim3(g.batched.regen), dim3(64), 0, st, reqs, A, V, S, Q, cur);
    wf_stages<true>(reqs, S, Q, cur, g, true, env_map, !isolate, !isolate, st);
}

void wf_finish(const PtRequest* reqs, const WfSlots& S, const WfQueues& Q, uint32_t cur, uint32_t paths, bool ghosts,
               hipStream_t st) {
    const dim3 grid(std::max(1u, (paths + 63u) / 64u));
    if (ghosts) hipLaunchKernelGGL(k_wf_finish<true>, grid, dim3(64), 0, st, reqs, S, Q, cur);
    else hipLaunchKernelGGL(k_wf_finish<false>, grid, dim3(64), 0, st, reqs, S, Q, cur);
}

void wf_tail_gather(const WfSlots& S, const WfQueues& Q, uint32_t cur, uint32_t entry, bool drop, uint32_t* list,
                    uint32_t* count, uint32_t cap, uint32_t* done, uint32_t slots, hipStream_t st) {
    const uint32_t blocks = std::max(1u, std::min((slots + 255u) / 256u, 4096u));
    hipLaunchKernelGGL(k_wf_tail_gather, dim3(blocks), dim3(256), 0, st, S, Q, cur, entry, drop, list, count, cap, done);
}

void wf_tail(const PtRequest* reqs, const WfSlots& S, uint32_t entry, const uint32_t* list, const uint32_t* count,
             uint32_t cap, uint32_t* done, hipStream_t st) {
    hipLaunchKernelGGL(k_wf_tail, dim3(std::max(1u, (cap + 63u) / 64u)), dim3(64), 0, st, reqs, S, entry, list, count,
                       cap, done);
}

void wf_tail_return(const WfQueues& Q, uint32_t cur, const uint32_t* list, const uint32_t* count, uint32_t cap,
                    hipStream_t st) {
    hipLaunchKernelGGL(k_wf_tail_return, dim3(std::max(1u, (cap + 255u) / 256u)), dim3(256), 0, st, Q, cur, list,
                       count, cap);
}

void wf_resolve(const PtRequest& r, hipStream_t st) {
    hipLaunchKernelGGL(k_wf_resolve, dim3((r.n_px + 255) / 256), dim3(256), 0, st, r);
}

}  // namespace vvr
