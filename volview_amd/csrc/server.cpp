// The serving loop: many requests of many resident volumes on one device.
//
// One thread owns the device work. Each pass of its loop:
//   1. admits queued requests while the HBM budget, the request table and the
//      volumes allow (a volume whose label masks / majorant must be rebuilt for a
//      request is rebuilt only while no admitted request reads it);
//   2. launches the ray-marched requests admitted since the last pass as one
//      batched kernel on the ray-marching stream, each followed by its read-back;
//   3. runs a batch of wavefront iterations for the path-traced requests on the
//      path-tracing stream (all of them draw work items side by side), and reads
//      back the requests' finished-item counters two batches behind;
//   4. resolves each path-traced increment whose items are all finished and
//      reads its frame back; queues the request's next increment, if any;
//   5. turns read-backs that have landed into completion records.
// Requests that arrive meanwhile join at the next pass: they are admitted as
// memory and table entries free up, and their paths share the iterations of
// the requests already in flight (continuous batching).
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "renderer_impl.hpp"

namespace vvr {
namespace {

constexpr uint32_t kTable = 256;       // request-table entries (path-traced requests in flight)
constexpr uint32_t kMaxJobs = 256;     // requests admitted at once
constexpr uint32_t kBatch = 4;         // wavefront iterations per launch batch
constexpr uint32_t kDvrRing = 8;       // batched ray-marching job tables in flight

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
    vvr_frame_params params{};
    uint32_t width = 0, height = 0;
    bool pt = false;
    uint32_t spp_total = 0, increment = 0, spp_done = 0, spp_running = 0;
    uint64_t t_submit = 0, t_admit = 0;
    enum State { kQueued, kRunning, kReading, kDone } state = kQueued;
    bool final_read = false;    // the read-back in flight is the last frame
    int status = VVR_OK;
    std::string error;
    // device and pinned memory
    size_t bytes = 0;            // charged to the HBM budget
    int uslot = -1;              // uniform block
    float4 *accum = nullptr, *albedo = nullptr, *normal = nullptr, *partial = nullptr;
    uint16_t* hdr = nullptr;
    uint16_t* h_hdr = nullptr;   // pinned copy of the latest frame
    hipEvent_t read_ev = nullptr;
    int entry = -1;              // request-table entry while path tracing
    uint32_t n_items = 0;        // of the running increment
    uint64_t start_seq = 0;      // first wavefront batch that sees the increment's counters
};

}  // namespace
}  // namespace vvr

struct vvr_server {
    int device = 0;
    vvr_server_config cfg{};
    hipStream_t s_pt = nullptr, s_dvr = nullptr;

    // wavefront engine
    WfEngine engine;  // slot arrays and queues
    PtRequest* d_table = nullptr;
    PtRequest* h_table = nullptr;  // pinned staging, one entry per table entry
    uint32_t* d_ctr = nullptr;     // head[kTable], done[kTable]
    uint32_t* h_done[2] = {nullptr, nullptr};
    hipEvent_t done_ev[2] = {nullptr, nullptr};
    bool done_pending[2] = {false, false};
    uint64_t batch_seq = 0;        // wavefront batches launched
    uint64_t snap_seq[2] = {0, 0}; // the batch each counter read-back follows
    uint32_t cur = 0, idle_in = 0;
    bool revive = false;
    Job* table[kTable] = {};
    uint32_t active[kWfMaxActive] = {};
    uint32_t n_active = 0;

    // ray marching
    DvrJob* d_jobs[kDvrRing] = {};
    DvrJob* h_jobs[kDvrRing] = {};
    hipEvent_t jobs_ev[kDvrRing] = {};
    bool jobs_used[kDvrRing] = {};
    uint32_t jobs_next = 0;

    // uniforms: one block per admitted request
    vvr_uniforms* d_uniforms = nullptr;
    vvr_uniforms* h_uniforms = nullptr;
    std::vector<int> free_uslots;

    // volumes: admitted requests per renderer, and the event after its last rebuild
    std::map<vvr_renderer*, int> readers;

    // requests
    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    std::deque<Job*> queue;                               // submitted, not admitted
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    std::vector<Job*> running;                            // admitted, not finished
    std::deque<vvr_completion> completions;
    uint64_t next_id = 1;
    bool stopping = false;
    std::string fatal;
    vvr_server_stats stats{};
    std::thread worker;

    explicit vvr_server(int dev) : engine(dev) {}

    void loop();
    void admit(std::vector<Job*>& fresh_dvr);
    void start_increment(Job* j);
    void launch_dvr(const std::vector<Job*>& batch);
    void pt_batch();
    void check_increments(int k);
    void poll_reads();
    void finish(Job* j, int status, const std::string& msg);
    void release_memory(Job* j);
    bool idle_locked() const { return queue.empty() && running.empty(); }
};

namespace vvr {
namespace {

size_t frame_bytes(const Job& j, uint32_t chunk_count) {
    const size_t px = size_t(j.width) * j.height;
    const size_t tiles = size_t((j.width + 7) / 8) * ((j.height + 7) / 8) * 64u;
    if (!j.pt) return px * 8;
    return px * (8 + 3 * 16) + size_t(chunk_count) * 3u * tiles * 16u;
}

}  // namespace
}  // namespace vvr

void vvr_server::release_memory(Job* j) {
    dfree(j->accum), dfree(j->albedo), dfree(j->normal), dfree(j->partial), dfree(j->hdr);
    j->accum = j->albedo = j->normal = j->partial = nullptr;
    j->hdr = nullptr;
    if (j->uslot >= 0) free_uslots.push_back(j->uslot);
    j->uslot = -1;
    {
        std::lock_guard<std::mutex> lk(mu);
        stats.hbm_in_use -= j->bytes;
    }
    j->bytes = 0;
    if (j->vol) readers[j->vol]--;
}

void vvr_server::finish(Job* j, int status, const std::string& msg) {
    release_memory(j);
    std::lock_guard<std::mutex> lk(mu);
    j->state = Job::kDone;
    j->status = status;
    j->error = msg;
    running.erase(std::remove(running.begin(), running.end(), j), running.end());
    if (status != VVR_OK) {
        vvr_completion c{};
        c.id = j->id, c.tag = j->tag, c.status = status, c.final_frame = 1;
        c.submit_ns = j->t_submit, c.admit_ns = j->t_admit, c.frame_ns = now_ns();
        completions.push_back(c);
    }
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
        const size_t need = frame_bytes(*j, j->pt ? (j->increment + 0) : 0);
        if (running.size() >= cfg.max_requests || free_uslots.empty()) break;
        if (stats.hbm_in_use + need > cfg.hbm_budget && !running.empty()) continue;
        if (j->pt && n_active >= uint32_t(kWfMaxActive)) continue;
        int entry = -1;
        if (j->pt) {
            for (uint32_t e = 0; e < kTable; e++)
                if (!table[e]) {
                    entry = int(e);
                    break;
                }
            if (entry < 0) continue;
        }
        vvr_renderer* v = j->vol;
        const OpacityKey key = opacity_key(j->params.labels, v->generation, v->has_labels, v->geo.min_spacing);
        const bool rebuild = !v->mask_key || !v->majorant_key || *v->majorant_key != key ||
                             (key.mode != VVR_LABELS_OFF && *v->mask_key != key);
        if (rebuild && readers[v] > 0) continue;  // wait until no admitted request reads this volume
        {
            std::lock_guard<std::mutex> lk(mu);
            queue.erase(std::find(queue.begin(), queue.end(), j));
        }
        try {
            j->uslot = free_uslots.back();
            free_uslots.pop_back();
            readers[v]++;
            const size_t px = size_t(j->width) * j->height;
            j->hdr = dalloc<uint16_t>(4 * px, "request image");
            if (j->pt) {
                j->accum = dalloc<float4>(px, "request accumulation");
                j->albedo = dalloc<float4>(px, "request albedo guide");
                j->normal = dalloc<float4>(px, "request normal guide");
                const size_t tiles = size_t((j->width + 7) / 8) * ((j->height + 7) / 8) * 64u;
                j->partial = dalloc<float4>(size_t(j->increment) * 3u * tiles, "request chunk sums");
            }
            j->h_hdr = halloc<uint16_t>(4 * px, "frame read-back");
            if (!j->read_ev) check(hipEventCreateWithFlags(&j->read_ev, hipEventDisableTiming), "hipEventCreate");
            j->bytes = need;
            {
                std::lock_guard<std::mutex> lk(mu);
                stats.hbm_in_use += need;
                stats.hbm_peak = std::max(stats.hbm_peak, stats.hbm_in_use);
                running.push_back(j);
            }
            vvr_uniforms& u = h_uniforms[j->uslot];
            pack_uniforms(v->geo, j->params, v->tf.bucket_max, v->has_labels, 0, j->width, j->height, j->id, &u);
            vvr_uniforms* du = d_uniforms + j->uslot;
            check(hipMemcpyAsync(du, &u, sizeof(u), hipMemcpyHostToDevice, s_dvr), "uniform upload");
            if (rebuild) {
                v->prepare(j->params, du, s_dvr);
                // Path tracing reads the rebuilt volumes too.
                hipEvent_t ev;
                check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
                check(hipEventRecord(ev, s_dvr), "hipEventRecord");
                check(hipStreamWaitEvent(s_pt, ev, 0), "hipStreamWaitEvent");
                (void)hipEventDestroy(ev);
            }
            j->t_admit = now_ns();
            j->state = Job::kRunning;
            if (j->pt) {
                // The uniform block was written on the ray-marching stream.
                hipEvent_t ev;
                check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
                check(hipEventRecord(ev, s_dvr), "hipEventRecord");
                check(hipStreamWaitEvent(s_pt, ev, 0), "hipStreamWaitEvent");
                (void)hipEventDestroy(ev);
                j->entry = entry;
                table[entry] = j;
                start_increment(j);
            } else {
                fresh_dvr.push_back(j);
            }
        } catch (const Error& e) {
            finish(j, e.code, e.msg);
        }
    }
}

// The next increment of a path-traced request: a table entry with fresh counters.
void vvr_server::start_increment(Job* j) {
    const uint32_t first = j->spp_done;
    const uint32_t n = std::min(j->increment, j->spp_total - first);
    PtRequest q{};
    q.scene = j->vol->scene(d_uniforms + j->uslot);
    q.accum = j->accum, q.albedo = j->albedo, q.normal = j->normal, q.hdr = j->hdr;
    WfEngine::shape(q, j->width, j->height, first, n, 1);
    q.partial = j->partial;
    q.item_base = 0;
    const int e = j->entry;
    h_table[e] = q;
    check(hipMemcpyAsync(d_table + e, h_table + e, sizeof(PtRequest), hipMemcpyHostToDevice, s_pt), "request upload");
    check(hipMemsetAsync(d_ctr + e, 0, 4, s_pt), "counter reset");
    check(hipMemsetAsync(d_ctr + kTable + e, 0, 4, s_pt), "counter reset");
    j->n_items = q.n_items;
    j->spp_running = n;
    j->start_seq = batch_seq;  // counter snapshots from earlier batches predate the reset
    active[n_active++] = uint32_t(e);
    revive = true;
}

void vvr_server::launch_dvr(const std::vector<Job*>& batch) {
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
        check(hipMemcpyAsync(d_jobs[k], h_jobs[k], nb * sizeof(DvrJob), hipMemcpyHostToDevice, s_dvr), "job upload");
        launch_dvr_batched(d_jobs[k], uint32_t(nb), tiles, s_dvr);
        check(hipGetLastError(), "batched ray marching");
        check(hipEventRecord(jobs_ev[k], s_dvr), "hipEventRecord");
        jobs_used[k] = true;
        for (size_t i = 0; i < nb; i++) {
            Job* j = batch[b0 + i];
            const size_t bytes = size_t(j->width) * j->height * 8;
            check(hipMemcpyAsync(j->h_hdr, j->hdr, bytes, hipMemcpyDeviceToHost, s_dvr), "frame read-back");
            check(hipEventRecord(j->read_ev, s_dvr), "hipEventRecord");
            j->final_read = true;
            j->state = Job::kReading;
        }
    }
}

// A batch of wavefront iterations over the active requests, then the
// finished-item counters into read-back slot `k`.
void vvr_server::pt_batch() {
    WfActive A{};
    A.n = n_active;
    for (uint32_t i = 0; i < n_active; i++) A.r[i] = uint16_t(active[i]);
    WfServed V{};
    V.head = d_ctr;
    V.done = d_ctr + kTable;
    for (uint32_t it = 0; it < kBatch; it++) {
        V.idle_in = idle_in;
        V.revive = revive ? 1u : 0u;
        if (revive)
            check(hipMemsetAsync(engine.Q.counts + kWfIDLE0 + (idle_in ^ 1u), 0, 4, s_pt), "idle queue reset");
        wf_iteration_served(d_table, A, V, engine.S, engine.Q, cur, engine.grid, s_pt);
        check(hipGetLastError(), "wavefront iteration");
        if (revive) idle_in ^= 1u;
        revive = false;
        cur ^= 1u;
    }
    batch_seq++;
    std::lock_guard<std::mutex> lk(mu);
    stats.iterations += kBatch;
}

// Read-back slot `k` (finished-item counters) has landed: resolve and read back
// every increment whose items are all finished; queue the next increments.
void vvr_server::check_increments(int k) {
    const uint32_t* done = h_done[k];
    std::vector<Job*> ended;
    for (uint32_t i = 0; i < n_active; i++) {
        Job* j = table[active[i]];
        if (j && snap_seq[k] >= j->start_seq && done[active[i]] >= j->n_items) ended.push_back(j);
    }
    for (Job* j : ended) {
        const uint32_t e = uint32_t(j->entry);
        for (uint32_t i = 0; i < n_active; i++)
            if (active[i] == e) {
                active[i] = active[--n_active];
                break;
            }
        wf_resolve(h_table[e], s_pt);
        check(hipGetLastError(), "resolve");
        j->spp_done += j->spp_running;
        {
            std::lock_guard<std::mutex> lk(mu);
            stats.pt_paths += uint64_t(j->width) * j->height * j->spp_running;
        }
        const size_t bytes = size_t(j->width) * j->height * 8;
        // A previous frame's read-back into the same pinned buffer has landed
        // (increments are read back in order on this stream).
        check(hipMemcpyAsync(j->h_hdr, j->hdr, bytes, hipMemcpyDeviceToHost, s_pt), "frame read-back");
        check(hipEventRecord(j->read_ev, s_pt), "hipEventRecord");
        j->final_read = j->spp_done >= j->spp_total;
        j->state = Job::kReading;
        if (j->final_read) {
            table[e] = nullptr;
            j->entry = -1;
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
        if (j->state != Job::kReading) continue;
        const hipError_t q = hipEventQuery(j->read_ev);
        if (q == hipErrorNotReady) continue;
        check(q, "frame read-back");
        vvr_completion c{};
        c.id = j->id, c.tag = j->tag, c.status = VVR_OK;
        c.spp = j->pt ? j->spp_done : 0u;
        c.final_frame = j->final_read ? 1u : 0u;
        c.submit_ns = j->t_submit, c.admit_ns = j->t_admit, c.frame_ns = now_ns();
        {
            std::lock_guard<std::mutex> lk(mu);
            completions.push_back(c);
            stats.frames++;
            cv_done.notify_all();
        }
        if (j->final_read) {
            finish(j, VVR_OK, "");
        } else {
            j->state = Job::kRunning;
            start_increment(j);
        }
    }
}

void vvr_server::loop() {
    try {
        check(hipSetDevice(device), "hipSetDevice");
        int slot_k = 0;  // read-back slot of the next batch
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
            if (n_active > 0) {
                pt_batch();
                check(hipMemcpyAsync(h_done[slot_k], d_ctr + kTable, kTable * 4, hipMemcpyDeviceToHost, s_pt),
                      "counter read-back");
                snap_seq[slot_k] = batch_seq - 1;
                check(hipEventRecord(done_ev[slot_k], s_pt), "hipEventRecord");
                done_pending[slot_k] = true;
                slot_k ^= 1;
                // The other slot's counters are one batch older: wait for them.
                if (done_pending[slot_k]) {
                    check(hipEventSynchronize(done_ev[slot_k]), "counter read-back");
                    done_pending[slot_k] = false;
                    check_increments(slot_k);
                }
            } else {
                // Nothing to trace: drain counter read-backs still in flight.
                for (int k = 0; k < 2; k++)
                    if (done_pending[k]) {
                        check(hipEventSynchronize(done_ev[k]), "counter read-back");
                        done_pending[k] = false;
                    }
                bool reading = false;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    for (Job* j : running) reading |= j->state == Job::kReading;
                }
                if (reading) std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            poll_reads();
        }
    } catch (const Error& e) {
        std::lock_guard<std::mutex> lk(mu);
        fatal = e.msg;
        for (Job* j : running) {
            vvr_completion c{};
            c.id = j->id, c.tag = j->tag, c.status = e.code, c.final_frame = 1;
            completions.push_back(c);
        }
        cv_done.notify_all();
    }
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
    s->cfg = c;
    s->engine.max_slots = c.slots ? c.slots : s->engine.max_slots;
    check(hipStreamCreateWithFlags(&s->s_pt, hipStreamNonBlocking), "hipStreamCreate");
    check(hipStreamCreateWithFlags(&s->s_dvr, hipStreamNonBlocking), "hipStreamCreate");
    s->engine.reserve_slots(s->engine.max_slots);
    s->engine.reserve(1, 1);
    s->d_table = dalloc<PtRequest>(kTable, "request table");
    s->h_table = halloc<PtRequest>(kTable, "request table staging");
    s->d_ctr = dalloc<uint32_t>(2 * kTable, "request counters");
    check(hipMemsetAsync(s->d_ctr, 0, 2 * kTable * 4, s->s_pt), "counter reset");
    for (int k = 0; k < 2; k++) {
        s->h_done[k] = halloc<uint32_t>(kTable, "counter read-back");
        check(hipEventCreateWithFlags(&s->done_ev[k], hipEventDisableTiming), "hipEventCreate");
    }
    for (uint32_t k = 0; k < kDvrRing; k++) {
        s->d_jobs[k] = dalloc<DvrJob>(64, "ray-marching jobs");
        s->h_jobs[k] = halloc<DvrJob>(64, "ray-marching job staging");
        check(hipEventCreateWithFlags(&s->jobs_ev[k], hipEventDisableTiming), "hipEventCreate");
    }
    s->d_uniforms = dalloc<vvr_uniforms>(kMaxJobs, "request uniforms");
    s->h_uniforms = halloc<vvr_uniforms>(kMaxJobs, "request uniform staging");
    for (int i = int(kMaxJobs) - 1; i >= 0; i--) s->free_uslots.push_back(i);
    wf_init_served(s->engine.S, s->engine.Q, s->engine.slots, s->s_pt);
    check(hipGetLastError(), "wavefront init");
    check(hipStreamSynchronize(s->s_pt), "server set-up");
    vvr_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->loop(); });
    *out = raw;
}

}  // namespace

extern "C" {

int vvr_server_create(vvr_ctx* ctx, const vvr_server_config* config, vvr_server** out) {
    return guarded([&] { create_server(ctx, config, out); });
}

int vvr_server_submit(vvr_server* s, const vvr_request* rq, uint64_t* id) {
    return guarded([&] {
        if (!s || !rq || !rq->volume || !rq->params) throw Error{VVR_EINVAL, "null argument"};
        if (!rq->width || !rq->height || rq->width > 16384 || rq->height > 16384)
            throw Error{VVR_EINVAL, "image size must be 1..16384 per axis"};
        validate_frame(*rq->params);
        auto j = std::make_unique<Job>();
        j->vol = rq->volume;
        j->params = *rq->params;
        j->width = rq->width, j->height = rq->height;
        j->tag = rq->tag;
        j->pt = rq->params->path_tracing_enabled != 0;
        if (j->pt) {
            j->spp_total = std::max(rq->params->path_tracing.samples_per_frame, 1u);
            j->increment = s->cfg.pt_increment ? std::min(s->cfg.pt_increment, j->spp_total) : j->spp_total;
        }
        if (size_t need = frame_bytes(*j, j->increment); need > s->cfg.hbm_budget)
            throw Error{VVR_ELIMIT, "request needs " + std::to_string(need) + " bytes, over the server's HBM budget"};
        j->t_submit = now_ns();
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) throw Error{VVR_ESTATE, "server is stopping"};
        if (!s->fatal.empty()) throw Error{VVR_EDEVICE, "server failed: " + s->fatal};
        j->id = s->next_id++;
        if (id) *id = j->id;
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

int vvr_server_read_hdr(vvr_server* s, uint64_t id, float* rgba) {
    return guarded([&] {
        if (!s || !rgba) throw Error{VVR_EINVAL, "null argument"};
        Job* j = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto it = s->jobs.find(id);
            if (it == s->jobs.end()) throw Error{VVR_EINVAL, "unknown request"};
            j = it->second.get();
            if (!(j->state == Job::kDone || (j->pt && j->spp_done > 0 && j->state == Job::kRunning)))
                throw Error{VVR_ESTATE, "the request has no delivered frame"};
            if (j->status != VVR_OK) throw Error{j->status, "the request failed: " + j->error};
        }
        const size_t n = size_t(j->width) * j->height * 4;
        for (size_t i = 0; i < n; i++) rgba[i] = f16_bits_to_f32(j->h_hdr[i]);
    });
}

int vvr_server_release(vvr_server* s, uint64_t id) {
    return guarded([&] {
        if (!s) throw Error{VVR_EINVAL, "null argument"};
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) throw Error{VVR_EINVAL, "unknown request"};
        if (it->second->state != Job::kDone) throw Error{VVR_ESTATE, "the request has not finished"};
        hfree(it->second->h_hdr);
        if (it->second->read_ev) (void)hipEventDestroy(it->second->read_ev);
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
        for (Job* j : s->queue) j->state = Job::kDone;
        s->queue.clear();
        s->cv_submit.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->s_pt);
    (void)hipStreamSynchronize(s->s_dvr);
    for (auto& kv : s->jobs) {
        Job* j = kv.second.get();
        if (j->state != Job::kDone) s->release_memory(j);
        hfree(j->h_hdr);
        if (j->read_ev) (void)hipEventDestroy(j->read_ev);
    }
    dfree(s->d_table), dfree(s->d_ctr), dfree(s->d_uniforms);
    hfree(s->h_table), hfree(s->h_uniforms);
    for (int k = 0; k < 2; k++) {
        hfree(s->h_done[k]);
        if (s->done_ev[k]) (void)hipEventDestroy(s->done_ev[k]);
    }
    for (uint32_t k = 0; k < kDvrRing; k++) {
        dfree(s->d_jobs[k]);
        hfree(s->h_jobs[k]);
        if (s->jobs_ev[k]) (void)hipEventDestroy(s->jobs_ev[k]);
    }
    if (s->s_pt) (void)hipStreamDestroy(s->s_pt);
    if (s->s_dvr) (void)hipStreamDestroy(s->s_dvr);
    delete s;
}

}  // extern "C"
