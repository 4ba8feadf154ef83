// The serving loop: many requests of many resident volumes on one device.
//
// One thread owns the device work, all of it on one stream. Each pass of its loop:
//   1. admits queued requests while the HBM budget, the request table and the
//      volumes allow (a volume whose label masks / majorant must be rebuilt for a
//      request is rebuilt only while no admitted request reads it);
//   2. launches the ray-marched requests admitted since the last pass as one
//      batched kernel, each followed by its read-back (ahead of the next
//      path-tracing batch, so they wait for at most one short batch);
//   3. runs a short batch of wavefront iterations for the path-traced requests
//      and reads back every increment's item counters after it; an increment
//      whose items are all handed out and whose paths in flight are few has
//      those paths taken out of the queues and run to their ends in one launch
//      on a second stream, beside the iterations (a tail launch), so that it
//      does not wait for its longest paths at one stage per iteration;
//   4. resolves, in order, each request's increments whose items have all
//      finished, reads the newest frame back, and starts a request's next
//      increment as soon as the current one has handed out all its items (two
//      increments of a request may be in flight: one in its tail, one filling
//      the slots), so a progressive request always offers work;
//   5. turns read-backs that have landed into completion records.
// A cancelled request stops handing out items at once and its paths in flight
// are dropped from the queues after the batches already queued.
// Requests that arrive meanwhile join at the next pass: they are admitted as
// memory and table entries free up, and their paths share the iterations of
// the requests already in flight (continuous batching). Work items are handed
// out in priority order: the increment with the lowest index first, then the
// oldest request, so new requests reach their first frame quickly.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "renderer_impl.hpp"

namespace vvr {
namespace {

constexpr uint32_t kTable = 256;    // request-table entries (path-traced increments in flight)
static_assert(kTable <= kWfBins, "the queue sort bins paths by request-table entry");
constexpr uint32_t kMaxJobs = 256;  // requests admitted at once
constexpr uint32_t kDvrRing = 8;    // batched ray-marching job tables in flight
constexpr uint32_t kInFlight = 2;   // increments of one request in flight

uint64_t now_ns() {
    return uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
                        std::chrono::steady_clock::now().time_since_epoch())
                        .count());
}

template <class T> T* halloc(size_t count, const char* what) {
    void* p = nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        throw Error{VVR_ENOMEM, std::string("cannot allocate pinned ") + what};
    }
    return static_cast<T*>(p);
}

void hfree(void* p) {
    if (p) (void)hipHostFree(p);
}

struct Job {
    uint64_t id = 0, tag = 0;
    vvr_renderer* vol = nullptr;
    bool holds = false;  // holds `vol` (hold_renderer) until the request is done with it
    vvr_frame_params params{};
    uint32_t width = 0, height = 0;
    bool pt = false;
    bool lit = false;  // ray marched with shadow or ambient-occlusion volumes
    // path tracing: spp_total paths per pixel in n_inc increments of `increment`
    uint32_t spp_total = 0, increment = 0, n_inc = 0;
    uint32_t inc_started = 0;
    std::deque<int> entries;  // table entries of the increments in flight, in increment order
    uint32_t spp_resolved = 0;
    std::atomic<uint32_t> spp_delivered{0};  // of the newest frame on the host
    uint64_t t_submit = 0, t_admit = 0;
    enum State { kQueued, kRunning, kDone };
    std::atomic<State> state{kQueued};
    bool read_pending = false;  // a frame read-back is in flight
    uint32_t read_spp = 0;      // paths per pixel of that frame
    int status = VVR_OK;
    std::string error;
    std::atomic<bool> cancelled{false};
    bool draining = false;   // cancelled: no new items, waiting for those handed out
    uint64_t drain_seq = 0;  // first batch that no longer hands out its items
    // device and pinned memory
    size_t bytes = 0;  // charged to the HBM budget
    int uslot = -1;    // uniform block
    float4 *accum = nullptr, *albedo = nullptr, *normal = nullptr, *partial = nullptr;
    float4* dn_scratch = nullptr;  // a-trous denoiser images (4 per pixel), when the request asks for it
    uint16_t* hdr = nullptr;
    // Pinned copies of the newest frames, two in turn: read-backs land in the
    // buffer that is not the front one, and become the front once complete.
    // Fields below are guarded by the server's mutex.
    uint16_t* h_hdr[2] = {nullptr, nullptr};
    int front = -1;              // the buffer holding the newest delivered frame
    int landing = -1;            // the buffer the read-back in flight writes
    int host_readers[2] = {0, 0};  // vvr_server_read_hdr calls copying out of each buffer
    hipEvent_t read_ev = nullptr;
    hipEvent_t vol_ready = nullptr;  // the renderer's stream at submission: work recorded before it
};

// One increment of a path-traced request in the wavefront engine.
struct Entry {
    Job* job = nullptr;
    uint32_t n_items = 0;
    uint64_t start_seq = 0;    // first batch whose counter read-back sees its reset counters
    bool handing_out = false;  // on the active list
    bool complete = false;     // every item finished
    uint64_t touched = 0;      // order in which it first handed out items (0: not yet)
    int tail = -1;             // the tail launch finishing its paths, while it runs
    bool tailed = false;       // it had one
};

size_t chunk_sum_bytes(const Job& j) {  // one increment's chunk sums
    return size_t(j.increment) * 3u * ((j.width + 7) / 8) * ((j.height + 7) / 8) * 64u * sizeof(float4);
}

// A finished request's pinned frames and events.
void free_host_side(Job* j) {
    for (auto& b : j->h_hdr) hfree(b), b = nullptr;
    if (j->read_ev) (void)hipEventDestroy(j->read_ev);
    if (j->vol_ready) (void)hipEventDestroy(j->vol_ready);
    j->read_ev = j->vol_ready = nullptr;
}

bool denoised(const Job& j) { return j.pt && j.params.path_tracing.denoiser == VVR_DENOISE_ATROUS; }
bool learned(const Job& j) { return j.pt && j.params.path_tracing.denoiser == VVR_DENOISE_LEARNED; }

size_t request_bytes(const Job& j) {
    const size_t px = size_t(j.width) * j.height;
    if (!j.pt) return px * 8;
    return px * (8 + 3 * 16) + std::min(j.n_inc, kInFlight) * chunk_sum_bytes(j) + (denoised(j) ? px * 4 * 16 : 0);
}

}  // namespace
}  // namespace vvr

struct vvr_server {
    int device = 0;
    vvr_server_config cfg{};
    hipStream_t stream = nullptr;  // every launch and copy of the server, in order

    // wavefront engine
    WfEngine engine;               // slot arrays and queues
    PtRequest* d_table = nullptr;
    PtRequest* h_table = nullptr;  // pinned staging, one per table entry
    uint32_t* d_ctr = nullptr;     // per entry, kWfCounterStride apart: head block, then done block
    // Pinned read-backs of both counter blocks, two in turn: the pass after a
    // batch is queued before the host waits for the batch's counters.
    uint32_t* h_ctr_buf[2] = {nullptr, nullptr};
    hipEvent_t ctr_ev[2] = {nullptr, nullptr};
    uint64_t ctr_seq[2] = {0, 0};    // the batch each read-back follows
    bool ctr_pending[2] = {false, false};
    int ctr_next = 0;
    const uint32_t* h_ctr = nullptr;  // the newest landed read-back
    uint64_t batch_seq = 0;           // wavefront batches launched
    uint64_t snap_seq = 0;            // the batch h_ctr follows
    uint64_t touch_seq = 0;           // increments that have begun handing out items
    uint32_t cur = 0, idle_in = 0;
    bool revive = false;
    Entry table[kTable];
    // Tail launches. An increment's paths advance one stage per iteration, so
    // once its items are all handed out it would wait for its longest paths
    // over many iterations; when few remain, they leave the queues and run to
    // their ends in one launch on tail_stream, beside the iterations.
    struct Tail {
        int entry = -1;            // the table entry it finishes (-1: free)
        uint32_t* list = nullptr;  // device: tail_cap listed slots, then the count
        uint32_t* count = nullptr;
        hipEvent_t gathered = nullptr, finished = nullptr;
    };
    static constexpr uint32_t kTails = 4;
    Tail tails[kTails];
    hipStream_t tail_stream = nullptr;
    uint32_t tail_cap = 0;  // paths one tail takes, at most (0: no tails)
    uint32_t active[kWfMaxActive] = {};
    uint32_t n_active = 0;
    uint32_t entries_in_use = 0;

    // ray marching: the passes' buffers (launch_dvr_passes, one request at a time,
    // grown to the largest frame seen), and the batched one-loop kernel's job tables
    // for frames that may draw ghosted labels
    uint8_t* dvr_mem = nullptr;
    size_t dvr_px = 0, dvr_tiles_cap = 0;
    DvrPasses dvr_passes{};
    DvrJob* d_jobs[kDvrRing] = {};
    DvrJob* h_jobs[kDvrRing] = {};
    hipEvent_t jobs_ev[kDvrRing] = {};
    bool jobs_used[kDvrRing] = {};
    uint32_t jobs_next = 0;

    // uniforms: one block per admitted request
    vvr_uniforms* d_uniforms = nullptr;
    vvr_uniforms* h_uniforms = nullptr;
    std::vector<int> free_uslots;

    std::map<vvr_renderer*, int> readers;        // admitted requests per volume
    std::map<vvr_renderer*, int> light_readers;  // of them, ray-marched ones reading its lighting volumes

    // requests
    std::mutex mu;
    std::condition_variable cv_submit, cv_done, cv_read;
    std::deque<Job*> queue;  // submitted, not admitted
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    std::vector<Job*> running;  // admitted, not finished
    std::deque<vvr_completion> completions;
    uint64_t next_id = 1;
    bool stopping = false;
    std::string fatal;
    vvr_server_stats stats{};
    std::thread worker;

    explicit vvr_server(int dev) : engine(dev) {}

    void loop();
    void admit(std::vector<Job*>& fresh_dvr);
    bool start_increment(Job* j);
    void launch_dvr(const std::vector<Job*>& batch);
    void pt_batch();
    void check_entries();
    void start_tails();
    void reap_tails();
    void advance_jobs();
    void stop_cancelled();
    void poll_reads();
    void read_back(Job* j, uint32_t spp);
    void drop_active(uint32_t e);
    void free_entry(int e);
    void finish(Job* j, int status, const std::string& msg, const vvr_completion* last = nullptr);
    void fail_all(int code, const std::string& msg);
    // The request is done with its renderer. A renderer destroyed meanwhile is
    // deleted here by its last request, after the server's work on it.
    void let_go(Job* j) {
        if (!j->holds) return;
        j->holds = false;
        if (let_go_renderer(j->vol)) {
            (void)hipStreamSynchronize(stream);
            delete j->vol;
        }
    }
    void release_memory(Job* j);
    vvr_completion record(const Job* j, int status, uint32_t spp, bool final_frame) const {
        vvr_completion c{};
        c.id = j->id, c.tag = j->tag, c.status = status, c.spp = spp, c.final_frame = final_frame ? 1u : 0u;
        c.submit_ns = j->t_submit, c.admit_ns = j->t_admit, c.frame_ns = now_ns();
        return c;
    }
};

void vvr_server::release_memory(Job* j) {
    dfree(j->accum), dfree(j->albedo), dfree(j->normal), dfree(j->partial), dfree(j->hdr), dfree(j->dn_scratch);
    j->accum = j->albedo = j->normal = j->partial = j->dn_scratch = nullptr;
    j->hdr = nullptr;
    if (j->uslot >= 0) free_uslots.push_back(j->uslot);
    j->uslot = -1;
    {
        std::lock_guard<std::mutex> lk(mu);
        stats.hbm_in_use -= j->bytes;
    }
    j->bytes = 0;
    if (j->vol) readers[j->vol]--;
    if (j->vol && j->lit) light_readers[j->vol]--;
}

// Ends a request: frees its device memory, marks it done and publishes its
// final completion (`last`, or an error record) in the same step, so a caller
// that sees the completion finds the request finished.
void vvr_server::finish(Job* j, int status, const std::string& msg, const vvr_completion* last) {
    release_memory(j);
    let_go(j);
    std::lock_guard<std::mutex> lk(mu);
    j->state = Job::kDone;
    j->status = status;
    j->error = msg;
    running.erase(std::remove(running.begin(), running.end(), j), running.end());
    completions.push_back(last ? *last : record(j, status, 0, true));
    if (last) stats.frames++;
    stats.completed++;
    cv_done.notify_all();
}

// Admits queued requests in order while they fit; a request that does not fit
// yet stays queued and does not block smaller ones behind it.
void vvr_server::admit(std::vector<Job*>& fresh_dvr) {
    std::vector<Job*> candidates;
    {
        std::lock_guard<std::mutex> lk(mu);
        candidates.assign(queue.begin(), queue.end());
    }
    for (Job* j : candidates) {
        // Under the lock from here to its removal from the queue: a request
        // cancelled meanwhile has left the queue, and its renderer may be gone.
        std::unique_lock<std::mutex> lk(mu);
        const auto qi = std::find(queue.begin(), queue.end(), j);
        if (qi == queue.end()) continue;
        const size_t need = request_bytes(*j);
        if (running.size() >= cfg.max_requests || free_uslots.empty()) break;
        if (stats.hbm_in_use + need > cfg.hbm_budget && !running.empty()) continue;
        if (j->pt && (entries_in_use >= kTable || n_active >= uint32_t(kWfMaxActive))) continue;
        vvr_renderer* v = j->vol;
        vvr_uniforms packed;
        pack_uniforms(v->geo, j->params, v->tf.bucket_max, v->has_labels, 0, j->width, j->height, j->id, &packed,
                      v->env_peak());
        // A derived volume is rebuilt only while no admitted request reads it:
        // masks and majorant wait for every reader of the volume, the lighting
        // volumes for the ray-marched requests that read them.
        if (v->needs_opacity_rebuild(j->params) && readers[v] > 0) continue;
        if (!j->pt && v->needs_lighting_rebuild(j->params, packed) && light_readers[v] > 0) continue;
        const bool rebuild = v->needs_prepare(j->params, packed);
        queue.erase(qi);
        lk.unlock();
        try {
            j->uslot = free_uslots.back();
            free_uslots.pop_back();
            readers[v]++;
            j->lit = !j->pt && (j->params.lighting.shadows || j->params.lighting.ambient_occlusion);
            if (j->lit) light_readers[v]++;
            const size_t px = size_t(j->width) * j->height;
            j->hdr = dalloc<uint16_t>(4 * px, "request image");
            if (j->pt) {
                j->accum = dalloc<float4>(px, "request accumulation");
                j->albedo = dalloc<float4>(px, "request albedo guide");
                j->normal = dalloc<float4>(px, "request normal guide");
                j->partial = reinterpret_cast<float4*>(
                    dalloc<uint8_t>(std::min(j->n_inc, kInFlight) * chunk_sum_bytes(*j), "request chunk sums"));
                if (denoised(*j)) j->dn_scratch = dalloc<float4>(4 * px, "request denoiser images");
            }
            for (auto& b : j->h_hdr) b = halloc<uint16_t>(4 * px, "frame read-back");
            if (!j->read_ev) check(hipEventCreateWithFlags(&j->read_ev, hipEventDisableTiming), "hipEventCreate");
            // Whatever the caller recorded on the renderer before submitting may
            // still read or write its volumes: the server's work waits for it.
            check(hipStreamWaitEvent(stream, j->vol_ready, 0), "hipStreamWaitEvent");
            j->bytes = need;
            {
                std::lock_guard<std::mutex> lk(mu);
                stats.hbm_in_use += need;
                stats.hbm_peak = std::max(stats.hbm_peak, stats.hbm_in_use);
                running.push_back(j);
            }
            vvr_uniforms& u = h_uniforms[j->uslot];
            u = packed;
            vvr_uniforms* du = d_uniforms + j->uslot;
            check(hipMemcpyAsync(du, &u, sizeof(u), hipMemcpyHostToDevice, stream), "uniform upload");
            if (rebuild) v->prepare(j->params, u, du, stream);
            j->t_admit = now_ns();
            j->state = Job::kRunning;
            if (j->pt) start_increment(j);
            else fresh_dvr.push_back(j);
        } catch (const Error& e) {
            finish(j, e.code, e.msg);
        }
    }
}

// Starts a request's next increment in a free table entry: fresh counters, its
// half of the chunk-sum ring, on the active list. False when no entry is free.
bool vvr_server::start_increment(Job* j) {
    int e = -1;
    for (uint32_t k = 0; k < kTable; k++)
        if (!table[k].job) {
            e = int(k);
            break;
        }
    if (e < 0 || n_active >= uint32_t(kWfMaxActive)) return false;
    const uint32_t k = j->inc_started;
    const uint32_t first = k * j->increment;
    const uint32_t n = std::min(j->increment, j->spp_total - first);
    PtRequest q{};
    q.scene = j->vol->scene(d_uniforms + j->uslot);
    q.accum = j->accum, q.albedo = j->albedo, q.normal = j->normal, q.hdr = j->hdr;
    WfEngine::shape(q, j->width, j->height, first, n, 1);
    q.partial =
        reinterpret_cast<float4*>(reinterpret_cast<uint8_t*>(j->partial) + (k % kInFlight) * chunk_sum_bytes(*j));
    q.item_base = 0;
    h_table[e] = q;
    check(hipMemcpyAsync(d_table + e, h_table + e, sizeof(PtRequest), hipMemcpyHostToDevice, stream), "request upload");
    check(hipMemsetAsync(d_ctr + e * kWfCounterStride, 0, 4, stream), "counter reset");
    check(hipMemsetAsync(d_ctr + (kTable + e) * kWfCounterStride, 0, 4, stream), "counter reset");
    Entry& en = table[e];
    en.job = j;
    en.n_items = q.n_items;
    en.start_seq = batch_seq;  // counter read-backs from earlier batches predate the reset
    en.handing_out = true;
    en.complete = false;
    active[n_active++] = uint32_t(e);
    entries_in_use++;
    j->entries.push_back(e);
    j->inc_started++;
    revive = true;
    return true;
}

void vvr_server::drop_active(uint32_t e) {
    for (uint32_t i = 0; i < n_active; i++)
        if (active[i] == e) {
            active[i] = active[--n_active];
            return;
        }
}

void vvr_server::free_entry(int e) {
    drop_active(uint32_t(e));
    table[e] = Entry{};
    entries_in_use--;
}

// Ray-marched requests: in passes (the march as parallel segments, as a
// renderer's own frames) one after another on the server's stream, through
// buffers grown to the largest frame; frames that may draw ghosted labels, or
// whose buffers would pass 4 GiB, in one batched launch of the one-loop kernel.
// Both give the same frame bit for bit.
void vvr_server::launch_dvr(const std::vector<Job*>& jobs) {
    std::vector<Job*> batch;
    for (Job* j : jobs) {
        const vvr_uniforms& hu = h_uniforms[j->uslot];
        const size_t px = size_t(j->width) * j->height, tiles = dvr_tiles(j->width, j->height);
        if (!dvr_in_passes(hu) || dvr_passes_bytes(j->width, j->height) > (size_t(4) << 30)) {
            batch.push_back(j);
            continue;
        }
        if (px > dvr_px || tiles > dvr_tiles_cap) {
            // Kernels queued on the stream may still use the old buffers.
            // (Like the path slots, the server's own buffers are not charged to the
            // requests' HBM budget.)
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            const size_t want_px = std::max(px, dvr_px), want_tiles = std::max(tiles, dvr_tiles_cap);
            dfree(dvr_mem);
            dvr_mem = nullptr;
            dvr_px = dvr_tiles_cap = 0;  // until the new buffers exist
            dvr_mem = dalloc<uint8_t>(dvr_passes_bytes_for(want_px, want_tiles), "server ray-marching passes");
            dvr_passes = carve_dvr_passes(dvr_mem, want_px, want_tiles);
            dvr_px = want_px, dvr_tiles_cap = want_tiles;
        }
        launch_dvr_passes(j->vol->scene(d_uniforms + j->uslot), dvr_passes, j->hdr, j->width, j->height, j->vol->cus,
                          dvr_segmented(hu), stream);
        check(hipGetLastError(), "ray marching passes");
        read_back(j, 0);
    }
    for (size_t b0 = 0; b0 < batch.size(); b0 += 64) {
        const size_t nb = std::min<size_t>(64, batch.size() - b0);
        const uint32_t k = jobs_next;
        jobs_next = (jobs_next + 1) % kDvrRing;
        if (jobs_used[k]) check(hipEventSynchronize(jobs_ev[k]), "job table reuse");
        uint32_t tiles = 0;
        for (size_t i = 0; i < nb; i++) {
            Job* j = batch[b0 + i];
            DvrJob& d = h_jobs[k][i];
            d.scene = j->vol->scene(d_uniforms + j->uslot);
            d.hdr = j->hdr;
            d.width = j->width, d.height = j->height;
            d.tiles_x = (j->width + 7) / 8;
            d.tile_base = tiles;
            tiles += d.tiles_x * ((j->height + 7) / 8);
        }
        check(hipMemcpyAsync(d_jobs[k], h_jobs[k], nb * sizeof(DvrJob), hipMemcpyHostToDevice, stream), "job upload");
        launch_dvr_batched(d_jobs[k], uint32_t(nb), tiles, stream);
        check(hipGetLastError(), "batched ray marching");
        check(hipEventRecord(jobs_ev[k], stream), "hipEventRecord");
        jobs_used[k] = true;
        for (size_t i = 0; i < nb; i++) read_back(batch[b0 + i], 0);
    }
}

// Copies a request's newest frame to its pinned buffer, after every earlier
// launch on the stream (and after any earlier read-back into the same buffer).
void vvr_server::read_back(Job* j, uint32_t spp) {
    const size_t bytes = size_t(j->width) * j->height * 8;
    int target;
    {
        // The buffer that is not the front one; a caller still copying out of it
        // (it was the front when the caller started) finishes first.
        std::unique_lock<std::mutex> lk(mu);
        target = j->front < 0 ? 0 : j->front ^ 1;
        cv_read.wait(lk, [&] { return j->host_readers[target] == 0; });
        j->landing = target;
    }
    check(hipMemcpyAsync(j->h_hdr[target], j->hdr, bytes, hipMemcpyDeviceToHost, stream), "frame read-back");
    check(hipEventRecord(j->read_ev, stream), "hipEventRecord");
    j->read_pending = true;
    j->read_spp = spp;
}

// A short batch of wavefront iterations over the increments handing out work.
// Priority: increments that have begun handing out items finish first, in the
// order they began, so that an increment is never left half done while others
// run (a request cancelled then would abandon the paths it already traced);
// then the lowest increment index (a new request's first frame before later
// frames of others), then the oldest request.
void vvr_server::pt_batch() {
    auto key = [&](uint32_t e) {
        const Job* j = table[e].job;
        const uint64_t t = table[e].touched;
        return std::make_tuple(t ? t : ~uint64_t(0), j->inc_started, j->id);
    };
    std::sort(active, active + n_active, [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
    // The first increment in this order receives this batch's first items.
    if (n_active && !table[active[0]].touched) table[active[0]].touched = ++touch_seq;
    WfActive A{};
    A.n = n_active;
    for (uint32_t i = 0; i < n_active; i++) A.r[i] = uint16_t(active[i]);
    WfServed V{};
    V.head = d_ctr;
    V.done = d_ctr + kTable * kWfCounterStride;
    // The light-map stage runs while any increment with paths in flight uses a map.
    // The free flights drop the density fetches and the ghost branch while every such increment
    // is in isolate mode and draws no ghost.
    bool env_map = false, isolate = true;
    for (uint32_t e = 0; e < kTable; e++) {
        if (!table[e].job || table[e].complete) continue;
        const vvr_uniforms& hu = h_uniforms[table[e].job->uslot];
        env_map = env_map || hu.lighting[3] == 2u;
        isolate = isolate && !frame_reads_density(hu) && !labels_may_draw_ghosts(hu);
    }
    for (uint32_t it = 0; it < cfg.iterations; it++) {
        V.idle_in = idle_in;
        V.revive = revive ? 1u : 0u;
        if (revive)
            check(hipMemsetAsync(engine.Q.counts + kWfIDLE0 + (idle_in ^ 1u), 0, 4, stream), "idle queue reset");
        wf_iteration_served(d_table, A, V, engine.S, engine.Q, cur, engine.grid, env_map, isolate, stream);
        check(hipGetLastError(), "wavefront iteration");
        if (revive) idle_in ^= 1u;
        revive = false;
        cur ^= 1u;
    }
    batch_seq++;
    std::lock_guard<std::mutex> lk(mu);
    stats.iterations += cfg.iterations;
}

// The counters have landed: increments that handed out all their items leave
// the active list; increments whose items all finished are complete.
void vvr_server::check_entries() {
    const uint32_t* head = h_ctr;
    const uint32_t* done = h_ctr + kTable * kWfCounterStride;
    for (uint32_t e = 0; e < kTable; e++) {
        Entry& en = table[e];
        if (!en.job || snap_seq < en.start_seq) continue;
        // Items went to it while an increment ahead of it ran out (the host
        // learns this from the counters only).
        if (!en.touched && head[e * kWfCounterStride] > 0) en.touched = ++touch_seq;
        if (en.handing_out && head[e * kWfCounterStride] >= en.n_items) {
            en.handing_out = false;
            drop_active(e);
        }
        // Complete once every item finished and no tail still writes its sums
        // (a reaped tail's work precedes later launches on the stream).
        if (!en.complete && en.tail < 0 && done[e * kWfCounterStride] >= en.n_items) en.complete = true;
    }
}

// Increments whose items are all handed out and whose paths in flight are
// few: those paths leave the queues after the batches queued so far and run to
// their ends on tail_stream (the counters say how many are left; a batch queued
// since only lowers that).
void vvr_server::start_tails() {
    if (!tail_cap) return;
    const uint32_t* done = h_ctr + kTable * kWfCounterStride;
    for (uint32_t e = 0; e < kTable; e++) {
        Entry& en = table[e];
        if (!en.job || en.handing_out || en.complete || en.tailed || en.job->draining || snap_seq < en.start_seq)
            continue;
        const uint32_t left = en.n_items - std::min(done[e * kWfCounterStride], en.n_items);
        if (left == 0 || left > tail_cap) continue;
        uint32_t k = 0;
        while (k < kTails && tails[k].entry >= 0) k++;
        if (k == kTails) return;
        Tail& t = tails[k];
        uint32_t* d_done = d_ctr + (kTable + e) * kWfCounterStride;
        check(hipMemsetAsync(t.count, 0, 4, stream), "tail count reset");
        wf_tail_gather(engine.S, engine.Q, cur, e, false, t.list, t.count, tail_cap, d_done, engine.slots, stream);
        check(hipGetLastError(), "tail gather");
        check(hipEventRecord(t.gathered, stream), "hipEventRecord");
        check(hipStreamWaitEvent(tail_stream, t.gathered, 0), "hipStreamWaitEvent");
        wf_tail(d_table, engine.S, e, t.list, t.count, tail_cap, d_done, tail_stream);
        check(hipGetLastError(), "tail launch");
        check(hipEventRecord(t.finished, tail_stream), "hipEventRecord");
        t.entry = int(e);
        en.tail = int(k);
        en.tailed = true;
    }
}

// Tails that have ended: the stream's later work waits for them (their sums,
// their slots), and their slots go back to the iterations.
void vvr_server::reap_tails() {
    for (uint32_t k = 0; k < kTails; k++) {
        Tail& t = tails[k];
        if (t.entry < 0) continue;
        const hipError_t q = hipEventQuery(t.finished);
        if (q == hipErrorNotReady) continue;
        check(q, "tail launch");
        check(hipStreamWaitEvent(stream, t.finished, 0), "hipStreamWaitEvent");
        wf_tail_return(engine.Q, cur, t.list, t.count, tail_cap, stream);
        check(hipGetLastError(), "tail return");
        table[t.entry].tail = -1;
        t.entry = -1;
    }
}

// Per path-traced request: resolve finished increments in order and read the
// newest frame back; start the next increment once none in flight is still
// handing out items; end cancelled requests once their paths have drained.
void vvr_server::advance_jobs() {
    const uint32_t* head = h_ctr;
    const uint32_t* done = h_ctr + kTable * kWfCounterStride;
    std::vector<Job*> snapshot;
    {
        std::lock_guard<std::mutex> lk(mu);
        snapshot = running;
    }
    for (Job* j : snapshot) {
        if (!j->pt) continue;
        if (j->draining) {
            bool drained = j->entries.empty() || snap_seq >= j->drain_seq;
            for (int e : j->entries) {
                const uint32_t handed = std::min(head[e * kWfCounterStride], table[e].n_items);
                drained = drained && done[e * kWfCounterStride] >= handed && table[e].tail < 0;
            }
            if (drained && !j->read_pending) {
                uint64_t abandoned = 0;
                for (int e : j->entries) {
                    abandoned += std::min(done[e * kWfCounterStride], table[e].n_items);
                    free_entry(e);
                }
                {
                    std::lock_guard<std::mutex> lk(mu);
                    stats.pt_paths_abandoned += abandoned;
                }
                j->entries.clear();
                const uint32_t spp = j->spp_delivered;
                const vvr_completion c = record(j, spp > 0 ? VVR_OK : VVR_ECANCELED, spp, true);
                finish(j, c.status, spp > 0 ? "" : "cancelled", &c);
            }
            continue;
        }
        bool resolved = false;
        while (!j->entries.empty() && table[j->entries.front()].complete) {
            const int e = j->entries.front();
            j->entries.pop_front();
            wf_resolve(h_table[e], stream);
            check(hipGetLastError(), "resolve");
            const uint32_t n = h_table[e].samples;
            free_entry(e);
            j->spp_resolved += n;
            resolved = true;
            std::lock_guard<std::mutex> lk(mu);
            stats.pt_paths += uint64_t(j->width) * j->height * n;
        }
        if (resolved) {
            // The denoised frame of everything resolved so far, as the renderer
            // denoises after each frame (renderer.rs:1795-1797).
            if (denoised(*j)) {
                launch_denoise(j->accum, j->albedo, j->normal, j->dn_scratch, j->hdr, j->width, j->height,
                               float(j->spp_resolved), stream);
                check(hipGetLastError(), "denoiser");
            } else if (learned(*j)) {  // the network attached to the request's renderer
                unet_denoise_frame(j->vol->unet, j->accum, j->albedo, j->normal, float(j->spp_resolved), j->width,
                                   j->height, j->hdr, stream);
                check(hipGetLastError(), "learned denoiser");
            }
            read_back(j, j->spp_resolved);
        }
        const bool busy = std::any_of(j->entries.begin(), j->entries.end(), [&](int e) { return table[e].handing_out; });
        if (!j->cancelled && !busy && j->inc_started < j->n_inc && j->entries.size() < kInFlight) start_increment(j);
    }
}

// Cancelled path-traced requests stop handing out work at once; advance_jobs
// ends them when the paths already handed out have finished.
void vvr_server::stop_cancelled() {
    std::vector<Job*> snapshot;
    {
        std::lock_guard<std::mutex> lk(mu);
        snapshot = running;
    }
    for (Job* j : snapshot) {
        if (!j->pt || !j->cancelled || j->draining) continue;
        j->draining = true;
        j->drain_seq = batch_seq;
        for (int e : j->entries) {
            table[e].handing_out = false;
            drop_active(uint32_t(e));
            // Its paths in flight are abandoned after the batches queued so far
            // (those on the end queue finish at the next regen).
            wf_tail_gather(engine.S, engine.Q, cur, uint32_t(e), true, nullptr, nullptr, 0,
                           d_ctr + (kTable + e) * kWfCounterStride, engine.slots, stream);
            check(hipGetLastError(), "dropping cancelled paths");
        }
    }
}

// Frames whose read-back has landed become completions.
void vvr_server::poll_reads() {
    std::vector<Job*> snapshot;
    {
        std::lock_guard<std::mutex> lk(mu);
        snapshot = running;
    }
    for (Job* j : snapshot) {
        if (!j->read_pending) continue;
        const hipError_t q = hipEventQuery(j->read_ev);
        if (q == hipErrorNotReady) continue;
        check(q, "frame read-back");
        j->read_pending = false;
        {
            std::lock_guard<std::mutex> lk(mu);
            j->front = j->landing;
            j->landing = -1;
        }
        j->spp_delivered = j->read_spp;
        const bool last = !j->pt || j->read_spp >= j->spp_total;
        const vvr_completion c = record(j, VVR_OK, j->read_spp, last);
        if (last) {
            finish(j, VVR_OK, "", &c);
        } else {
            std::lock_guard<std::mutex> lk(mu);
            completions.push_back(c);
            stats.frames++;
            cv_done.notify_all();
        }
    }
}

void vvr_server::loop() {
    try {
        check(hipSetDevice(device), "hipSetDevice");
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                if (running.empty() && queue.empty()) {
                    if (stopping) break;
                    cv_submit.wait(lk, [&] { return stopping || !queue.empty(); });
                    if (stopping && queue.empty()) break;
                }
            }
            std::vector<Job*> fresh_dvr;
            admit(fresh_dvr);
            if (!fresh_dvr.empty()) launch_dvr(fresh_dvr);
            stop_cancelled();
            if (entries_in_use > 0) {
                // One short batch per pass, queued before the host waits for the
                // previous batch's counters, so the device runs this batch while
                // the host acts on the last one (ray-marched requests admitted
                // meanwhile wait for at most two batches).
                pt_batch();
                const int b = ctr_next;
                ctr_next ^= 1;
                check(hipMemcpyAsync(h_ctr_buf[b], d_ctr, 2 * kTable * kWfCounterStride * 4, hipMemcpyDeviceToHost,
                                     stream),
                      "counter read-back");
                check(hipEventRecord(ctr_ev[b], stream), "hipEventRecord");
                ctr_seq[b] = batch_seq - 1;
                ctr_pending[b] = true;
                const int w = ctr_pending[b ^ 1] ? (b ^ 1) : b;
                check(hipEventSynchronize(ctr_ev[w]), "counter read-back");
                ctr_pending[w] = false;
                h_ctr = h_ctr_buf[w];
                snap_seq = ctr_seq[w];
                check_entries();
                reap_tails();
                start_tails();
            } else {
                for (int k = 0; k < 2; k++)
                    if (ctr_pending[k]) {  // no entries left: the read-back only needs to land
                        check(hipEventSynchronize(ctr_ev[k]), "counter read-back");
                        ctr_pending[k] = false;
                    }
                bool reading = false;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    for (Job* j : running) reading |= j->read_pending;
                }
                if (reading) std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            advance_jobs();
            poll_reads();
        }
    } catch (const Error& e) {
        fail_all(e.code, e.msg);
    } catch (const std::bad_alloc&) {
        fail_all(VVR_ENOMEM, "host allocation failed");
    } catch (...) {
        fail_all(VVR_EDEVICE, "unexpected internal error");
    }
}

// The loop failed: every admitted and queued request ends with the error (and
// can be released), and later submissions are refused.
void vvr_server::fail_all(int code, const std::string& msg) {
    (void)hipStreamSynchronize(stream);
    if (tail_stream) (void)hipStreamSynchronize(tail_stream);
    std::vector<Job*> snap;
    {
        std::lock_guard<std::mutex> lk(mu);
        fatal = msg;
        snap = running;
    }
    for (Job* j : snap) finish(j, code, msg);
    std::lock_guard<std::mutex> lk(mu);
    for (Job* j : queue) {
        let_go(j);
        j->state = Job::kDone;
        j->status = code;
        j->error = msg;
        completions.push_back(record(j, code, 0, true));
        stats.completed++;
    }
    queue.clear();
    cv_done.notify_all();
}

namespace {

void create_server(vvr_ctx* ctx, const vvr_server_config* cfg, vvr_server** out) {
    if (!ctx || !out) throw Error{VVR_EINVAL, "null argument"};
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto s = std::make_unique<vvr_server>(ctx->device);
    s->device = ctx->device;
    vvr_server_config c = cfg ? *cfg : vvr_server_config{};
    if (!c.hbm_budget) c.hbm_budget = uint64_t(64) << 30;
    if (!c.max_requests || c.max_requests > kMaxJobs) c.max_requests = kMaxJobs;
    if (!c.iterations) c.iterations = 2;
    if (c.iterations > 64) throw Error{VVR_EINVAL, "iterations per pass must be 1..64"};
    s->cfg = c;
    // The served engine's defaults, as measured on the serving trace and Batched-32: 4M
    // slots, tracking runs of 24 / 48.
    s->engine.set_shape(vvr_pt_shape{c.slots ? c.slots : 1u << 22, c.flight_steps ? c.flight_steps : 24u,
                                     c.shadow_steps ? c.shadow_steps : 48u});
    check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    s->engine.reserve_slots(s->engine.max_slots);
    s->engine.reserve(1, 1);
    s->d_table = dalloc<PtRequest>(kTable, "request table");
    s->h_table = halloc<PtRequest>(kTable, "request table staging");
    s->d_ctr = dalloc<uint32_t>(2 * kTable * kWfCounterStride, "request counters");
    check(hipMemsetAsync(s->d_ctr, 0, 2 * kTable * kWfCounterStride * 4, s->stream), "counter reset");
    for (int k = 0; k < 2; k++) {
        s->h_ctr_buf[k] = halloc<uint32_t>(2 * kTable * kWfCounterStride, "counter read-back");
        std::memset(s->h_ctr_buf[k], 0, 2 * kTable * kWfCounterStride * 4);
        check(hipEventCreateWithFlags(&s->ctr_ev[k], hipEventDisableTiming), "hipEventCreate");
    }
    s->h_ctr = s->h_ctr_buf[0];
    for (uint32_t k = 0; k < kDvrRing; k++) {
        s->d_jobs[k] = dalloc<DvrJob>(64, "ray-marching jobs");
        s->h_jobs[k] = halloc<DvrJob>(64, "ray-marching job staging");
        check(hipEventCreateWithFlags(&s->jobs_ev[k], hipEventDisableTiming), "hipEventCreate");
    }
    // Served tails run k_wf_tail (one wave per SIMD): 64 Ki paths measured best on the
    // serving trace (256 Ki cost a quarter of the throughput); VVR_PT_TAIL overrides.
    s->tail_cap = std::getenv("VVR_PT_TAIL") ? s->engine.grid.tail_paths
                                             : std::min(s->engine.grid.tail_paths, 1u << 16);
    if (s->tail_cap) {
        check(hipStreamCreateWithFlags(&s->tail_stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& t : s->tails) {
            t.list = dalloc<uint32_t>(size_t(s->tail_cap) + 64, "tail lists");
            t.count = t.list + s->tail_cap;
            check(hipEventCreateWithFlags(&t.gathered, hipEventDisableTiming), "hipEventCreate");
            check(hipEventCreateWithFlags(&t.finished, hipEventDisableTiming), "hipEventCreate");
        }
    }
    s->d_uniforms = dalloc<vvr_uniforms>(kMaxJobs, "request uniforms");
    s->h_uniforms = halloc<vvr_uniforms>(kMaxJobs, "request uniform staging");
    for (int i = int(kMaxJobs) - 1; i >= 0; i--) s->free_uslots.push_back(i);
    wf_init_served(s->engine.S, s->engine.Q, s->engine.slots, s->stream);
    check(hipGetLastError(), "wavefront init");
    check(hipStreamSynchronize(s->stream), "server set-up");
    vvr_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->loop(); });
    *out = raw;
}

// One trial of config `c`: the requests submitted at once to a server of their
// own, timed from the first submission to the last request's final frame.
double trial_seconds(vvr_ctx* ctx, const vvr_server_config& c, const vvr_request* rq, uint32_t n) {
    vvr_server* raw = nullptr;
    create_server(ctx, &c, &raw);
    std::unique_ptr<vvr_server, void (*)(vvr_server*)> s(raw, vvr_server_destroy);
    std::vector<uint64_t> ids(n);
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < n; i++) {
        const int e = vvr_server_submit(s.get(), &rq[i], &ids[i]);
        if (e != VVR_OK) throw Error{e, vvr_last_error()};
    }
    uint32_t left = n;
    std::vector<vvr_completion> buf(64);
    while (left > 0) {
        uint32_t got = 0;
        const int e = vvr_server_poll(s.get(), buf.data(), uint32_t(buf.size()), &got, 100);
        if (e != VVR_OK) throw Error{e, vvr_last_error()};
        for (uint32_t k = 0; k < got; k++) {
            if (buf[k].status != VVR_OK) throw Error{buf[k].status, "a trial request failed"};
            if (buf[k].final_frame) left--;
        }
        if (std::chrono::steady_clock::now() - t0 > std::chrono::minutes(10))
            throw Error{VVR_EDEVICE, "autotune trial did not finish within 10 minutes"};
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (uint64_t id : ids) (void)vvr_server_release(s.get(), id);
    return sec;
}

// The server's engine shape for a request mix: trials of the requests over the
// path slots (1, 2, 4 and 8 Mi at two iterations per pass), then over the
// iterations per pass (1, 2, 4, 8) at the best slot count; the fastest shape
// goes into *cfg, whose other fields are kept. A first trial at *cfg readies the
// volumes (masks, majorants) so that every trial does the same work.
void autotune_server(vvr_ctx* ctx, const vvr_request* rq, uint32_t n, vvr_server_config* cfg) {
    if (!ctx || !rq || !cfg || n == 0) throw Error{VVR_EINVAL, "autotune needs a context, requests and a config"};
    vvr_server_config best = *cfg;
    (void)trial_seconds(ctx, best, rq, n);
    double best_s = 1e300;
    for (uint32_t mi : {1u, 2u, 4u, 8u}) {
        vvr_server_config c = *cfg;
        c.slots = mi << 20;
        c.iterations = 2;
        const double t = trial_seconds(ctx, c, rq, n);
        if (t < best_s) best_s = t, best = c;
    }
    const uint32_t slots = best.slots;
    for (uint32_t it : {1u, 4u, 8u}) {
        vvr_server_config c = *cfg;
        c.slots = slots;
        c.iterations = it;
        const double t = trial_seconds(ctx, c, rq, n);
        if (t < best_s) best_s = t, best = c;
    }
    cfg->slots = best.slots;
    cfg->iterations = best.iterations;
}

}  // namespace

extern "C" {

int vvr_server_autotune(vvr_ctx* ctx, const vvr_request* requests, uint32_t n, vvr_server_config* config) {
    return guarded([&] { autotune_server(ctx, requests, n, config); });
}

int vvr_server_create(vvr_ctx* ctx, const vvr_server_config* config, vvr_server** out) {
    return guarded([&] { create_server(ctx, config, out); });
}

int vvr_server_submit(vvr_server* s, const vvr_request* rq, uint64_t* id) {
    return guarded([&] {
        if (!s || !rq || !rq->volume || !rq->params) throw Error{VVR_EINVAL, "null argument"};
        if (!rq->width || !rq->height || rq->width > 16384 || rq->height > 16384)
            throw Error{VVR_EINVAL, "image size must be 1..16384 per axis"};
        validate_frame(*rq->params);
        if (rq->params->path_tracing_enabled && rq->params->path_tracing.denoiser == VVR_DENOISE_LEARNED &&
            !rq->volume->unet)
            throw Error{VVR_EINVAL, "no learned denoiser is attached to the request's renderer"};
        auto j = std::make_unique<Job>();
        j->vol = rq->volume;
        j->params = *rq->params;
        j->width = rq->width, j->height = rq->height;
        j->tag = rq->tag;
        j->pt = rq->params->path_tracing_enabled != 0;
        if (j->pt) {
            j->spp_total = std::max(rq->params->path_tracing.samples_per_frame, 1u);
            j->increment = s->cfg.pt_increment ? std::min(s->cfg.pt_increment, j->spp_total) : j->spp_total;
            j->n_inc = (j->spp_total + j->increment - 1) / j->increment;
            // Work-item counters are 32-bit and may run past an increment's end
            // by the slot count before the increment leaves the active list.
            const uint64_t px = uint64_t((rq->width + 7) / 8) * ((rq->height + 7) / 8) * 64u;
            if (px * j->increment >= (1ull << 31))
                throw Error{VVR_ELIMIT, "one increment exceeds 2^31 paths: lower pt_increment"};
        }
        if (size_t need = request_bytes(*j); need > s->cfg.hbm_budget)
            throw Error{VVR_ELIMIT, "request needs " + std::to_string(need) + " bytes, over the server's HBM budget"};
        check(hipSetDevice(s->device), "hipSetDevice");
        check(hipEventCreateWithFlags(&j->vol_ready, hipEventDisableTiming), "hipEventCreate");
        check(hipEventRecord(j->vol_ready, j->vol->stream), "hipEventRecord");
        j->t_submit = now_ns();
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) throw Error{VVR_ESTATE, "server is stopping"};
        if (!s->fatal.empty()) throw Error{VVR_EDEVICE, "server failed: " + s->fatal};
        j->id = s->next_id++;
        if (id) *id = j->id;
        hold_renderer(j->vol);
        j->holds = true;
        s->queue.push_back(j.get());
        s->jobs[j->id] = std::move(j);
        s->stats.submitted++;
        s->cv_submit.notify_all();
    });
}

int vvr_server_poll(vvr_server* s, vvr_completion* out, uint32_t cap, uint32_t* n, uint32_t wait_ms) {
    return guarded([&] {
        if (!s || (!out && cap)) throw Error{VVR_EINVAL, "null argument"};
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->completions.empty() && wait_ms)
            s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->completions.empty(); });
        uint32_t k = 0;
        while (k < cap && !s->completions.empty()) {
            out[k++] = s->completions.front();
            s->completions.pop_front();
        }
        if (n) *n = k;
    });
}

int vvr_server_read_hdr(vvr_server* s, uint64_t id, float* rgba, size_t n_floats) {
    return guarded([&] {
        if (!s || !rgba) throw Error{VVR_EINVAL, "null argument"};
        Job* j = nullptr;
        int f = -1;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto it = s->jobs.find(id);
            if (it == s->jobs.end()) throw Error{VVR_EINVAL, "unknown request"};
            j = it->second.get();
            const bool delivered = j->pt ? j->spp_delivered > 0 : (j->state == Job::kDone && j->status == VVR_OK);
            if (!delivered || j->front < 0) {
                if (j->state == Job::kDone && j->status != VVR_OK)
                    throw Error{j->status, "the request failed: " + j->error};
                throw Error{VVR_ESTATE, "the request has no delivered frame"};
            }
            if (n_floats != size_t(j->width) * j->height * 4)
                throw Error{VVR_EINVAL, "buffer holds " + std::to_string(n_floats) + " floats, the frame " +
                                            std::to_string(size_t(j->width) * j->height * 4)};
            f = j->front;
            j->host_readers[f]++;  // no read-back lands in this buffer until the copy below is done
        }
        const uint16_t* src = j->h_hdr[f];
        for (size_t i = 0; i < n_floats; i++) rgba[i] = f16_bits_to_f32(src[i]);
        std::lock_guard<std::mutex> lk(s->mu);
        j->host_readers[f]--;
        s->cv_read.notify_all();
    });
}

int vvr_server_cancel(vvr_server* s, uint64_t id) {
    return guarded([&] {
        if (!s) throw Error{VVR_EINVAL, "null argument"};
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) throw Error{VVR_EINVAL, "unknown request"};
        Job* j = it->second.get();
        j->cancelled = true;
        auto q = std::find(s->queue.begin(), s->queue.end(), j);
        if (q != s->queue.end()) {
            s->queue.erase(q);
            s->let_go(j);
            j->state = Job::kDone;
            j->status = VVR_ECANCELED;
            j->error = "cancelled";
            s->completions.push_back(s->record(j, VVR_ECANCELED, 0, true));
            s->stats.completed++;
            s->cv_done.notify_all();
        }
    });
}

int vvr_server_release(vvr_server* s, uint64_t id) {
    return guarded([&] {
        if (!s) throw Error{VVR_EINVAL, "null argument"};
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) throw Error{VVR_EINVAL, "unknown request"};
        Job* j = it->second.get();
        if (j->state != Job::kDone) throw Error{VVR_ESTATE, "the request has not finished"};
        if (j->host_readers[0] || j->host_readers[1]) throw Error{VVR_ESTATE, "the request's frame is being read"};
        free_host_side(j);
        s->jobs.erase(it);
    });
}

int vvr_server_get_stats(vvr_server* s, vvr_server_stats* out) {
    return guarded([&] {
        if (!s || !out) throw Error{VVR_EINVAL, "null argument"};
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->stats;
    });
}

void vvr_server_destroy(vvr_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        // Requests not admitted yet are dropped.
        for (Job* j : s->queue) {
            j->state = Job::kDone;
            s->let_go(j);  // never admitted: no work of the server reads its renderer
        }
        s->queue.clear();
        s->cv_submit.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    if (s->tail_stream) (void)hipStreamSynchronize(s->tail_stream);
    for (auto& kv : s->jobs) {
        Job* j = kv.second.get();
        if (j->state == Job::kRunning) s->release_memory(j);
        s->let_go(j);
        free_host_side(j);
    }
    dfree(s->d_table), dfree(s->d_ctr), dfree(s->d_uniforms), dfree(s->dvr_mem);
    hfree(s->h_table), hfree(s->h_uniforms), hfree(s->h_ctr_buf[0]), hfree(s->h_ctr_buf[1]);
    for (auto& e : s->ctr_ev)
        if (e) (void)hipEventDestroy(e);
    for (uint32_t k = 0; k < kDvrRing; k++) {
        dfree(s->d_jobs[k]);
        hfree(s->h_jobs[k]);
        if (s->jobs_ev[k]) (void)hipEventDestroy(s->jobs_ev[k]);
    }
    for (auto& t : s->tails) {
        dfree(t.list);
        if (t.gathered) (void)hipEventDestroy(t.gathered);
        if (t.finished) (void)hipEventDestroy(t.finished);
    }
    if (s->tail_stream) (void)hipStreamDestroy(s->tail_stream);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"
