// server.hip -- a serving loop over the build and lookup path: requests arrive one by one,
// builds run on concurrent stream lanes, lookups are batched continuously into one launch, and
// the DAGs stay resident in device memory under an HBM budget.
//
// The reference has no server: its renderer builds at start-up and evaluates once per frame
// (src/DeferredRenderer.cpp:221-230, 333).  Here requests carry device buffers:
//   build  (depth, N, zTileIndex, zTileNum)  -> a resident DAG, named by the request's ticket
//   lookup (DAG ticket, NDC points, count)   -> one visibility byte per point
// Admission: a build is started only when the budget holds the resident DAGs, the reserved
// bytes of builds in flight, the lanes' scratch and this build's scratch growth plus an upper
// bound of its DAG (pre-merge word count); those come from the exact node counts computed at
// the start of every build.  Unpinned DAGs are evicted least recently used first to make room;
// a DAG is pinned while a lookup on it is queued or running.
// Serial mode processes every request in arrival order on one stream, one at a time: the plain
// path, for comparison.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "abi.hpp"

namespace cpvs {
namespace {

constexpr uint32_t kMaxServerLanes = 16;
constexpr uint32_t kMaxSegs = 4096;              // lookup requests per batch
constexpr uint64_t kMaxBatchQueries = 1ull << 28;  // queries per batch

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Req {
    uint64_t ticket = 0;
    int kind = 0;  // 0 build, 1 lookup
    const float* depth = nullptr;
    uint32_t N = 0, t = 0, T = 1;
    uint64_t dag = 0;
    const float* ndc = nullptr;
    uint32_t stride = 3;
    uint64_t count = 0;
    uint8_t* vis = nullptr;
    int status = -1;  // -1 pending, else a CPVS status
    std::string error;
    double t_submit = 0.0, t_done = 0.0;
};

struct Entry {
    cpvs_dag* dag = nullptr;
    uint64_t bytes = 0;
    uint64_t last_use = 0;
    uint32_t pins = 0;
};

int levels_of(u32 N) {
    int n = 1;
    while ((1u << (n - 1)) < N) ++n;
    return n;
}

}  // namespace
}  // namespace cpvs

using namespace cpvs;

struct cpvs_server {
    cpvs_ctx* ctx = nullptr;
    bool serial = false;
    uint64_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_work;  // new requests or stop
    std::condition_variable cv_done;  // a request finished
    std::condition_variable cv_mem;   // memory was released
    std::deque<Req*> builds, lookups, fifo;  // fifo: serial mode
    std::unordered_map<uint64_t, Req*> reqs;
    std::unordered_map<uint64_t, Entry> store;
    uint64_t resident_bytes = 0, reserved_bytes = 0, scratch_bytes = 0;
    uint64_t next_ticket = 1, clock = 0;
    uint64_t n_builds = 0, n_lookups = 0, n_batches = 0, n_evicted = 0, n_queries = 0, n_trims = 0;
    bool stop = false;
    std::vector<Worker*> lanes;
    Worker* lk = nullptr;          // lookup lane
    LookupSeg* h_segs = nullptr;   // pinned staging of one batch's segment table
    LookupSeg* d_segs = nullptr;
    std::vector<std::thread> threads;
    std::vector<uint64_t> lane_scratch;  // arena bytes held by each lane
};

namespace {

// Called with s->mu held.  Evicts unpinned DAGs, least recently used first, until `need` more
// bytes fit; false if they cannot.  The evicted DAGs go to `victims`, to be destroyed once the
// lock is released.  No lookup pins a victim and its build has completed, so no queued work
// reads its words: they return to the context's buffer cache at once, with no device
// synchronisation, and the next build's words are taken from there.
bool make_room(cpvs_server* s, uint64_t need, std::vector<cpvs_dag*>& victims) {
    while (s->resident_bytes + s->reserved_bytes + s->scratch_bytes + need > s->budget) {
        auto victim = s->store.end();
        for (auto it = s->store.begin(); it != s->store.end(); ++it)
            if (it->second.pins == 0 && (victim == s->store.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == s->store.end()) return false;
        victims.push_back(victim->second.dag);
        s->resident_bytes -= victim->second.bytes;
        s->store.erase(victim);
        ++s->n_evicted;
    }
    return true;
}

void destroy_all(std::vector<cpvs_dag*>& v) {
    for (cpvs_dag* d : v) dag_destroy_idle(d);
    v.clear();
}

struct AdmitState {
    cpvs_server* s;
    uint32_t lane;
    uint64_t reserved = 0;   // bytes reserved by this build (DAG bound + scratch growth)
    uint64_t dag_bound = 0;  // the part of it that bounds the DAG's word buffer
};

// Reserves `need` bytes for lane a (replacing what it reserved before), evicting or waiting for
// memory; CPVS_EBUDGET when they can never fit.
int reserve_bytes(AdmitState* a, uint64_t need) {
    cpvs_server* s = a->s;
    std::vector<cpvs_dag*> victims;
    std::unique_lock<std::mutex> lk(s->mu);
    s->reserved_bytes -= a->reserved;  // a retry replaces the first request
    a->reserved = 0;
    int rc = CPVS_OK;
    for (;;) {
        if (make_room(s, need, victims)) break;
        // Nothing evictable: wait for lookups to unpin or other builds to finish.  That cannot
        // happen in serial mode (this thread is the only one) or when nothing is in flight.
        bool pinned = false;
        for (auto& kv : s->store) pinned = pinned || kv.second.pins;
        if (s->serial || (s->reserved_bytes == 0 && !pinned)) {
            set_error("server: build needs %llu bytes; budget %llu, resident %llu (%s), lane scratch %llu",
                      (unsigned long long)need, (unsigned long long)s->budget, (unsigned long long)s->resident_bytes,
                      pinned ? "pinned by queued lookups" : "none evictable", (unsigned long long)s->scratch_bytes);
            rc = CPVS_EBUDGET;
            break;
        }
        s->cv_mem.wait(lk);
    }
    if (rc == CPVS_OK) {
        a->reserved = need;
        s->reserved_bytes += need;
    }
    lk.unlock();
    destroy_all(victims);
    return rc;
}

int admit_build(void* user, const Arena& arena, u64 scratch_bytes, u64 dag_bytes_max) {
    AdmitState* a = static_cast<AdmitState*>(user);
    const uint64_t growth = scratch_bytes > arena.cap ? scratch_bytes + (scratch_bytes >> 3) - arena.cap : 0;
    a->dag_bound = dag_bytes_max;
    return reserve_bytes(a, growth + dag_bytes_max);
}

// A build's DAG words: a cached buffer when one fits, else new memory.  Device memory of the
// server is its resident DAGs, the reservations of builds in flight (this one's included, which
// bounds these words), the lanes' arenas and the context's cached buffers; before new memory is
// taken the cache is trimmed so that this sum stays within the budget (hipFree waits for the
// device, so this happens only when no cached buffer fits and the budget is tight).
void* alloc_dag_words(void* user, size_t bytes) {
    AdmitState* a = static_cast<AdmitState*>(user);
    cpvs_server* s = a->s;
    BufferCache& bc = s->ctx->buffers;
    if (void* p = bc.take_cached(bytes, a->dag_bound)) return p;  // within the reservation
    uint64_t used;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        used = s->resident_bytes + s->reserved_bytes + s->scratch_bytes;
    }
    const uint64_t keep = used < s->budget ? s->budget - used : 0;
    if (bc.cached_bytes() > keep) {
        bc.trim(keep);
        std::lock_guard<std::mutex> lk(s->mu);
        ++s->n_trims;
    }
    return bc.take(bytes);
}

void finish(cpvs_server* s, Req* r, int status) {
    r->status = status;
    if (status != CPVS_OK) r->error = cpvs_last_error();
    r->t_done = now_ms();
}

// One build on lane `li` (pyramid in the lane's level arena, admission, DAG into the store).
int run_build(cpvs_server* s, uint32_t li, Req* r) {
    Worker& w = *s->lanes[li];
    cpvs_pyramid p;
    p.ctx = s->ctx;
    p.N = r->N;
    p.n = levels_of(r->N);
    p.depth = const_cast<float*>(r->depth);
    p.borrowed = true;
    AdmitState a{s, li, 0};
    int rc = CPVS_OK;
    // The lane's pyramid-level arena grows before the build's admission: reserve that growth first.
    // Without leafmasks the octree reads pyramid levels 1 and 2 as well.
    const bool all_levels = s->ctx->leafmasks.load() == 0;
    const size_t lv_bytes = pyramid_levels_bytes(r->N, all_levels);
    if (lv_bytes > w.levels.cap) rc = reserve_bytes(&a, lv_bytes + (lv_bytes >> 3) - w.levels.cap);
    if (rc == CPVS_OK) rc = pyramid_build_levels(&p, all_levels, w.stream, &w.levels);
    if (rc == CPVS_OK && a.reserved) {
        std::lock_guard<std::mutex> lk(s->mu);  // the growth is lane scratch from here on
        s->reserved_bytes -= a.reserved;
        a.reserved = 0;
        s->scratch_bytes -= s->lane_scratch[li];
        s->lane_scratch[li] = w.arena.cap + w.levels.cap;
        s->scratch_bytes += s->lane_scratch[li];
    }
    Admit adm{admit_build, &a, alloc_dag_words};
    cpvs_dag* d = nullptr;
    if (rc == CPVS_OK) rc = build_dag(s->ctx, w.arena, w.pinned, s->ctx->merge_mode, &p, r->t, r->T, w.stream, &d, &adm);
    std::lock_guard<std::mutex> lk(s->mu);
    s->reserved_bytes -= a.reserved;
    s->scratch_bytes -= s->lane_scratch[li];
    s->lane_scratch[li] = w.arena.cap + w.levels.cap;
    s->scratch_bytes += s->lane_scratch[li];
    if (rc == CPVS_OK) {
        Entry e;
        e.dag = d;
        e.bytes = s->ctx->buffers.size_of(d->d_words);  // the buffer, which may exceed the words
        e.last_use = ++s->clock;
        s->store[r->ticket] = e;
        s->resident_bytes += e.bytes;
    }
    ++s->n_builds;
    finish(s, r, rc);
    s->cv_mem.notify_all();
    s->cv_done.notify_all();
    return rc;
}

void build_loop(cpvs_server* s, uint32_t li) {
    (void)hipSetDevice(s->ctx->device);
    for (;;) {
        Req* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->builds.empty(); });
            if (s->builds.empty()) return;  // stopping
            r = s->builds.front();
            s->builds.pop_front();
        }
        run_build(s, li, r);
    }
}

// Runs one batch of lookups (segments already filled in s->h_segs) and completes them.
int run_lookups(cpvs_server* s, std::vector<Req*>& batch, const std::vector<DagDesc>& descs) {
    uint32_t blocks = 0;
    for (size_t k = 0; k < batch.size(); ++k) {
        Req* r = batch[k];
        LookupSeg& g = s->h_segs[k];
        g.d = descs[k];
        g.ndc = r->ndc;
        g.vis = r->vis;
        g.stride = r->stride;
        g.count = (u32)r->count;
        g.block0 = blocks;
        g.pad = 0;
        blocks += (u32)((r->count + 255) / 256);
    }
    const hipStream_t st = s->lk->stream;
    hipError_t e = hipMemcpyAsync(s->d_segs, s->h_segs, sizeof(LookupSeg) * batch.size(), hipMemcpyHostToDevice, st);
    int rc = (e == hipSuccess) ? launch_segments_eval(s->d_segs, (u32)batch.size(), blocks, st)
                               : hip_fail(e, "server lookup table copy", __FILE__, __LINE__);
    if (rc == CPVS_OK) {
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = hip_fail(e, "server lookup", __FILE__, __LINE__);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (Req* r : batch) {
        auto it = s->store.find(r->dag);
        if (it != s->store.end()) --it->second.pins;
        s->n_queries += r->count;
        ++s->n_lookups;
        finish(s, r, rc);
    }
    ++s->n_batches;
    s->cv_mem.notify_all();
    s->cv_done.notify_all();
    return rc;
}

// Continuous batching: every lookup pending when the lookup lane frees up goes into the next
// launch (up to kMaxSegs requests and kMaxBatchQueries points).
void lookup_loop(cpvs_server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::vector<Req*> batch;
    std::vector<DagDesc> descs;
    for (;;) {
        batch.clear();
        descs.clear();
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->lookups.empty(); });
            if (s->lookups.empty()) return;  // stopping
            uint64_t q = 0;
            while (!s->lookups.empty() && batch.size() < kMaxSegs) {
                Req* r = s->lookups.front();
                if (!batch.empty() && q + r->count > kMaxBatchQueries) break;
                s->lookups.pop_front();
                q += r->count;
                batch.push_back(r);
                descs.push_back(dag_desc(s->store.at(r->dag).dag));  // pinned since submission
            }
        }
        run_lookups(s, batch, descs);
    }
}

// Serial mode: one request at a time, in arrival order, on lane 0's stream.
void serial_loop(cpvs_server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::vector<Req*> one(1);
    std::vector<DagDesc> desc(1);
    for (;;) {
        Req* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->fifo.empty(); });
            if (s->fifo.empty()) return;
            r = s->fifo.front();
            s->fifo.pop_front();
            if (r->kind == 1) desc[0] = dag_desc(s->store.at(r->dag).dag);  // pinned since submission
        }
        if (r->kind == 0) {
            run_build(s, 0, r);
        } else {
            one[0] = r;
            run_lookups(s, one, desc);
        }
    }
}

void shutdown(cpvs_server* s) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->threads)
        if (t.joinable()) t.join();
    s->threads.clear();
}

}  // namespace

extern "C" {

int cpvs_server_create(cpvs_ctx* ctx, uint32_t lanes, uint64_t hbm_budget_bytes, int serial, cpvs_server** out) {
    CHECK_ARG(ctx && out, "server_create: NULL argument");
    CHECK_ARG(lanes >= 1 && lanes <= kMaxServerLanes, "server_create: %u build lanes (1..%u)", lanes, kMaxServerLanes);
    CHECK_ARG(hbm_budget_bytes > 0, "server_create: zero budget");
    GUARD_BEGIN
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    cpvs_server* s = new cpvs_server();
    s->ctx = ctx;
    s->serial = serial != 0;
    s->budget = hbm_budget_bytes;
    {  // memory the process already holds (other contexts' scratch, caller buffers) is not there to give
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const uint64_t avail = free_b > (1ull << 30) ? free_b - (1ull << 30) : 0;
            if (avail < s->budget) s->budget = avail ? avail : 1;
        } else {
            (void)hipGetLastError();
        }
    }
    const uint32_t nl = s->serial ? 1 : lanes;
    int rc = CPVS_OK;
    for (uint32_t i = 0; i < nl && rc == CPVS_OK; ++i) {
        Worker* w = new Worker();
        rc = w->create();
        if (rc) delete w;
        else s->lanes.push_back(w);
    }
    s->lane_scratch.assign(s->lanes.size(), 0);
    if (rc == CPVS_OK) {
        if (s->serial) {
            s->lk = s->lanes[0];
        } else {
            s->lk = new Worker();
            rc = s->lk->create();
            if (rc) {
                delete s->lk;
                s->lk = nullptr;
            }
        }
    }
    if (rc == CPVS_OK) {
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&s->h_segs), sizeof(LookupSeg) * kMaxSegs, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&s->d_segs, sizeof(LookupSeg) * kMaxSegs);
        if (e != hipSuccess) rc = hip_fail(e, "server_create buffers", __FILE__, __LINE__);
    }
    if (rc == CPVS_OK) {
        try {
            if (s->serial) {
                s->threads.emplace_back(serial_loop, s);
            } else {
                for (uint32_t i = 0; i < nl; ++i) s->threads.emplace_back(build_loop, s, i);
                s->threads.emplace_back(lookup_loop, s);
            }
        } catch (...) {
            set_error("server_create: could not start threads");
            rc = CPVS_EINTERNAL;
        }
    }
    if (rc != CPVS_OK) {
        const std::string msg = cpvs_last_error();
        cpvs_server_destroy(s);
        set_error("%s", msg.c_str());
        return rc;
    }
    *out = s;
    return CPVS_OK;
    GUARD_END
}

int cpvs_server_submit_build(cpvs_server* s, const float* depth, uint32_t N, uint32_t zTileIndex, uint32_t zTileNum,
                             uint64_t* ticket) {
    CHECK_ARG(s && depth && ticket, "server_submit_build: NULL argument");
    CHECK_ARG(is_pow2(N) && N >= 16 && N <= 65536, "server_submit_build: N=%u must be a power of two in [16, 65536]", N);
    CHECK_ARG(zTileNum >= 1 && zTileIndex < zTileNum, "server_submit_build: zTileIndex %u must be < zTileNum %u",
              zTileIndex, zTileNum);
    CHECK_ARG((uint64_t)N * zTileNum <= (1ull << 21), "server_submit_build: N*zTileNum exceeds 2^21");
    GUARD_BEGIN
    Req* r = new Req();
    r->kind = 0;
    r->depth = depth;
    r->N = N;
    r->t = zTileIndex;
    r->T = zTileNum;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->ticket = s->next_ticket++;
        r->t_submit = now_ms();
        s->reqs[r->ticket] = r;
        (s->serial ? s->fifo : s->builds).push_back(r);
    }
    s->cv_work.notify_all();
    *ticket = r->ticket;
    return CPVS_OK;
    GUARD_END
}

int cpvs_server_submit_lookup(cpvs_server* s, uint64_t dag, const float* ndc, uint32_t stride, uint64_t count,
                              uint8_t* vis, uint64_t* ticket) {
    CHECK_ARG(s && ticket && ((ndc && vis) || !count), "server_submit_lookup: NULL argument");
    CHECK_ARG(stride >= 3, "server_submit_lookup: stride %u < 3", stride);
    CHECK_ARG(count <= 0xFFFFFF00ull, "server_submit_lookup: %llu points exceed one request", (unsigned long long)count);
    GUARD_BEGIN
    Req* r = new Req();
    r->kind = 1;
    r->dag = dag;
    r->ndc = ndc;
    r->stride = stride;
    r->count = count;
    r->vis = vis;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->store.find(dag);
        if (it == s->store.end()) {
            delete r;
            set_error("server_submit_lookup: DAG %llu is not resident", (unsigned long long)dag);
            return CPVS_ESTATE;
        }
        ++it->second.pins;
        it->second.last_use = ++s->clock;
        r->ticket = s->next_ticket++;
        r->t_submit = now_ms();
        s->reqs[r->ticket] = r;
        (s->serial ? s->fifo : s->lookups).push_back(r);
    }
    s->cv_work.notify_all();
    *ticket = r->ticket;
    return CPVS_OK;
    GUARD_END
}

int cpvs_server_wait(cpvs_server* s, uint64_t ticket, double* latency_ms) {
    CHECK_ARG(s, "server_wait: NULL server");
    Req* r = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        CHECK_ARG(it != s->reqs.end(), "server_wait: unknown ticket %llu", (unsigned long long)ticket);
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->status >= 0; });
        s->reqs.erase(it);
    }
    const int st = r->status;
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    if (st != CPVS_OK) set_error("%s", r->error.c_str());
    delete r;
    return st;
}

int cpvs_server_poll(cpvs_server* s, uint64_t ticket, int* done) {
    CHECK_ARG(s && done, "server_poll: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    CHECK_ARG(it != s->reqs.end(), "server_poll: unknown ticket %llu", (unsigned long long)ticket);
    *done = it->second->status >= 0 ? 1 : 0;
    return CPVS_OK;
}

int cpvs_server_dag(cpvs_server* s, uint64_t dag, const cpvs_dag** out) {
    CHECK_ARG(s && out, "server_dag: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->store.find(dag);
    if (it == s->store.end()) {
        set_error("server_dag: DAG %llu is not resident", (unsigned long long)dag);
        return CPVS_ESTATE;
    }
    *out = it->second.dag;
    return CPVS_OK;
}

int cpvs_server_release(cpvs_server* s, uint64_t dag) {
    CHECK_ARG(s, "server_release: NULL server");
    cpvs_dag* d = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->store.find(dag);
        if (it == s->store.end()) {
            set_error("server_release: DAG %llu is not resident", (unsigned long long)dag);
            return CPVS_ESTATE;
        }
        if (it->second.pins) {
            set_error("server_release: DAG %llu has lookups in flight", (unsigned long long)dag);
            return CPVS_ESTATE;
        }
        d = it->second.dag;
        s->resident_bytes -= it->second.bytes;
        s->store.erase(it);
    }
    dag_destroy_idle(d);  // unpinned: no lookup reads it
    s->cv_mem.notify_all();
    return CPVS_OK;
}

int cpvs_server_resident(cpvs_server* s, uint64_t* ids, uint32_t cap, uint32_t* count) {
    CHECK_ARG(s && count, "server_resident: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *count = (uint32_t)s->store.size();
    uint32_t k = 0;
    if (ids)
        for (auto& kv : s->store)
            if (k < cap) ids[k++] = kv.first;
    return CPVS_OK;
}

int cpvs_server_stats(cpvs_server* s, uint64_t* out, uint32_t cap) {
    CHECK_ARG(s && out, "server_stats: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t v[] = {s->n_builds, s->n_lookups, s->n_batches, s->n_queries, s->n_evicted,
                          (uint64_t)s->store.size(), s->resident_bytes, s->scratch_bytes, s->reserved_bytes, s->budget,
                          s->n_trims, (uint64_t)s->ctx->buffers.cached_bytes()};
    for (uint32_t i = 0; i < cap && i < sizeof(v) / sizeof(v[0]); ++i) out[i] = v[i];
    return CPVS_OK;
}

int cpvs_server_destroy(cpvs_server* s) {
    if (!s) return CPVS_OK;
    hipSetDevice(s->ctx->device);
    shutdown(s);
    for (auto& kv : s->reqs) delete kv.second;  // requests never waited for
    s->reqs.clear();
    for (auto& kv : s->store) cpvs_dag_destroy(kv.second.dag);
    s->store.clear();
    for (Worker* w : s->lanes) {
        w->destroy();
        delete w;
    }
    if (s->lk && !s->serial) {
        s->lk->destroy();
        delete s->lk;
    }
    if (s->h_segs) hipHostFree(s->h_segs);
    if (s->d_segs) hipFree(s->d_segs);
    delete s;
    return CPVS_OK;
}

}  // extern "C"
