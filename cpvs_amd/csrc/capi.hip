// capi.hip -- the extern "C" boundary (include/cpvs.h): handles, argument checks, errors.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "abi.hpp"
#include "internal.hpp"
#include "scan.hpp"

namespace cpvs {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

thread_local Prof* t_prof = nullptr;

void Prof::resolve() {
    std::vector<Rec> recs;
    {
        std::lock_guard<std::mutex> lk(mu);
        recs.swap(pending);
    }
    std::vector<std::pair<const char*, float>> done;
    done.reserve(recs.size());
    for (auto& r : recs) {
        float ms = 0.0f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess)
            done.emplace_back(r.name, ms);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    std::lock_guard<std::mutex> lk(mu);
    for (auto& d : done) {
        auto& e = acc[d.first];
        e.first += d.second;
        e.second += 1;
    }
}

void Prof::reset() {
    resolve();
    std::lock_guard<std::mutex> lk(mu);
    acc.clear();
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return CPVS_EHIP;
}

int Arena::reserve(size_t bytes, hipStream_t s, bool* moved) {
    used = 0;
    if (moved) *moved = false;
    if (bytes <= cap) return CPVS_OK;
    if (fixed) {
        set_error("build: %zu bytes of scratch needed, the fixed scratch holds %zu", bytes, cap);
        return CPVS_EINVAL;
    }
    if (moved) *moved = true;
    if (base) {
        CPVS_HIP_CHECK(hipStreamSynchronize(s));
        CPVS_HIP_CHECK(hipFree(base));
        base = nullptr;
        cap = 0;
    }
    size_t want = bytes + (bytes >> 3);
    if (hipMalloc(&base, want) != hipSuccess) {
        base = nullptr;
        (void)hipGetLastError();
        set_error("scratch arena: hipMalloc of %zu bytes failed", want);
        return CPVS_ENOMEM;
    }
    cap = want;
    return CPVS_OK;
}

void Arena::release() {
    if (base) hipFree(base);
    base = nullptr;
    cap = used = 0;
}

}  // namespace cpvs

using namespace cpvs;



extern "C" {

const char* cpvs_last_error(void) { return g_err; }

int cpvs_ctx_create(int device, cpvs_ctx** out) {
    CHECK_ARG(out, "ctx_create: out is NULL");
    GUARD_BEGIN
    int ndev = 0;
    CPVS_HIP_CHECK(hipGetDeviceCount(&ndev));
    CHECK_ARG(device >= 0 && device < ndev, "ctx_create: device %d of %d", device, ndev);
    CPVS_HIP_CHECK(hipSetDevice(device));
    cpvs_ctx* c = new cpvs_ctx();
    c->device = device;
    c->buffers.device = device;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->num_cus = cus;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    // The no-allocation build's fixed scratch on this device (its address, found once per device).
    int rc0 = noalloc_setup(device);
    if (rc0 != CPVS_OK) {
        hipStreamDestroy(c->stream);
        delete c;
        return rc0;
    }
    c->pinned_words = kPinnedWords;
    e = hipHostMalloc(reinterpret_cast<void**>(&c->pinned), kPinnedBytes, kPinnedFlags);
    if (e != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return hip_fail(e, "hipHostMalloc", __FILE__, __LINE__);
    }
    *out = c;
    return CPVS_OK;
    GUARD_END
}

namespace cpvs {

namespace {
// Best fit of at most `most` bytes, from one of the cache's lists; mu held.
void* fit_from(std::multimap<size_t, void*>& m, size_t bytes, size_t most, size_t* got) {
    auto it = m.lower_bound(bytes);
    if (it == m.end() || it->first > most) return nullptr;
    void* p = it->second;
    *got = it->first;
    m.erase(it);
    return p;
}
}  // namespace

void* BufferCache::take_cached(size_t bytes, size_t limit) {
    bytes = bytes ? bytes : 1;
    const size_t most = std::min(2 * bytes, std::max(limit, bytes));
    std::unique_lock<std::mutex> lk(mu);
    size_t got = 0;
    void* p = fit_from(free_, bytes, most, &got);
    if (!p && !pending_.empty()) {
        auto it = pending_.lower_bound(bytes);
        if (it != pending_.end() && it->first <= most) {
            // Every buffer freed so far becomes idle with one device synchronisation (those freed
            // while it runs stay pending: their readers may have been queued after it began).
            std::multimap<size_t, void*> parked;
            parked.swap(pending_);
            lk.unlock();
            (void)hipSetDevice(device);
            (void)hipDeviceSynchronize();
            lk.lock();
            ++device_syncs;
            free_.insert(parked.begin(), parked.end());
            p = fit_from(free_, bytes, most, &got);
        }
    }
    if (!p) return nullptr;
    live_[p] = got;
    cached -= got;
    return p;
}

void* BufferCache::take(size_t bytes) {
    bytes = bytes ? bytes : 1;
    if (void* p = take_cached(bytes)) return p;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        release();  // the cached buffers may be what is missing
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    std::lock_guard<std::mutex> lk(mu);
    live_[p] = bytes;
    return p;
}

void BufferCache::park(void* p, bool idle) {
    if (!p) return;
    bool drop = false;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = live_.find(p);
        const size_t bytes = (it != live_.end()) ? it->second : 0;
        if (it != live_.end()) live_.erase(it);
        if (bytes && cached + bytes <= kMaxBytes) {
            (idle ? free_ : pending_).emplace(bytes, p);
            cached += bytes;
        } else {
            drop = true;
        }
    }
    if (drop) {
        (void)hipSetDevice(device);
        (void)hipFree(p);  // waits for the device: no queued work still reads p
    }
}

void BufferCache::give(void* p) { park(p, false); }
void BufferCache::give_idle(void* p) { park(p, true); }

size_t BufferCache::size_of(void* p) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = live_.find(p);
    return it == live_.end() ? 0 : it->second;
}

size_t BufferCache::cached_bytes() {
    std::lock_guard<std::mutex> lk(mu);
    return cached;
}

size_t BufferCache::trim(size_t keep) {
    std::vector<void*> drop;
    size_t freed = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto* m : {&free_, &pending_}) {
            while (cached > keep && !m->empty()) {
                auto it = std::prev(m->end());  // largest first
                drop.push_back(it->second);
                cached -= it->first;
                freed += it->first;
                m->erase(it);
            }
        }
        if (!drop.empty()) ++trims;
    }
    (void)hipSetDevice(device);
    for (void* p : drop) (void)hipFree(p);  // hipFree waits for the device
    return freed;
}

void BufferCache::release() {
    std::lock_guard<std::mutex> lk(mu);
    (void)hipSetDevice(device);
    for (auto* m : {&free_, &pending_}) {
        for (auto& kv : *m) (void)hipFree(kv.second);
        m->clear();
    }
    cached = 0;
}

}  // namespace cpvs

int cpvs_ctx_destroy(cpvs_ctx* ctx) {
    if (!ctx) return CPVS_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ctx->buffers.release();
    ctx->arena.release();
    for (cpvs::Worker* w : ctx->lanes) {
        w->destroy();
        delete w;
    }
    ctx->lanes.clear();
    if (ctx->pinned) hipHostFree(ctx->pinned);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return CPVS_OK;
}

int cpvs_ctx_synchronize(cpvs_ctx* ctx) {
    CHECK_ARG(ctx, "ctx is NULL");
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    CPVS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return CPVS_OK;
}

int cpvs_ctx_profile(cpvs_ctx* ctx, int mode) {
    CHECK_ARG(ctx, "ctx_profile: NULL ctx");
    if (mode < 0) ctx->prof.reset();
    else ctx->prof_on = mode != 0;
    return CPVS_OK;
}

int cpvs_ctx_profile_read(cpvs_ctx* ctx, char* buf, uint64_t cap) {
    CHECK_ARG(ctx && buf && cap, "ctx_profile_read: NULL argument");
    const auto acc = ctx->prof.snapshot();
    std::string out;
    char line[256];
    for (const auto& kv : acc) {
        snprintf(line, sizeof(line), "%s\t%.6f\t%ld\n", kv.first.c_str(), kv.second.first, kv.second.second);
        out += line;
    }
    if (out.size() + 1 > cap) {
        set_error("ctx_profile_read: %zu bytes needed", out.size() + 1);
        return CPVS_EINVAL;
    }
    std::memcpy(buf, out.c_str(), out.size() + 1);
    return CPVS_OK;
}

int cpvs_ctx_set_merge(cpvs_ctx* ctx, int mode) {
    CHECK_ARG(ctx, "ctx_set_merge: NULL ctx");
    CHECK_ARG(mode == CPVS_MERGE_TABLE || mode == CPVS_MERGE_SORT, "ctx_set_merge: mode %d", mode);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->merge_mode = mode;
    return CPVS_OK;
}

int cpvs_ctx_set_leafmasks(cpvs_ctx* ctx, int on) {
    CHECK_ARG(ctx, "ctx_set_leafmasks: NULL ctx");
    CHECK_ARG(on == 0 || on == 1, "ctx_set_leafmasks: %d (0 or 1)", on);
    ctx->leafmasks.store(on);
    return CPVS_OK;
}

int cpvs_ctx_set_hash_bits(cpvs_ctx* ctx, uint32_t leaf_bits, uint32_t inner_bits) {
    CHECK_ARG(ctx, "ctx_set_hash_bits: NULL ctx");
    CHECK_ARG(leaf_bits < 32 && inner_bits < 32, "ctx_set_hash_bits: %u, %u (0 .. 31)", leaf_bits, inner_bits);
    ctx->leaf_hash_bits.store(leaf_bits);
    ctx->inner_hash_bits.store(inner_bits);
    return CPVS_OK;
}

int cpvs_ctx_leafmasks(const cpvs_ctx* ctx, int* on) {
    CHECK_ARG(ctx && on, "ctx_leafmasks: NULL argument");
    *on = ctx->leafmasks.load();
    return CPVS_OK;
}

int cpvs_dag_leafmasks(const cpvs_dag* d, int* on) {
    CHECK_ARG(d && on, "dag_leafmasks: NULL argument");
    *on = d->leafmasks ? 1 : 0;
    return CPVS_OK;
}

int cpvs_dag_merge_info(const cpvs_dag* d, int* mode, uint64_t* key_collisions) {
    CHECK_ARG(d, "dag_merge_info: NULL dag");
    if (mode) *mode = d->merge_mode;
    if (key_collisions) *key_collisions = d->key_collisions;
    return CPVS_OK;
}

int cpvs_ctx_stream(cpvs_ctx* ctx, void** stream) {
    CHECK_ARG(ctx && stream, "ctx_stream: NULL argument");
    *stream = ctx->stream;
    return CPVS_OK;
}

// ---------------------------------------------------------------------------------------
// Pyramid
// ---------------------------------------------------------------------------------------

int cpvs_pyramid_build(cpvs_ctx* ctx, const float* depth, int depth_on_device, uint32_t N, uint32_t flags, void* stream,
                       cpvs_pyramid** out) {
    CHECK_ARG(ctx && depth && out, "pyramid_build: NULL argument");
    CHECK_ARG(is_pow2(N) && N >= 2 && N <= 65536, "pyramid_build: N=%u must be a power of two in [2, 65536]", N);
    CHECK_ARG(!(flags & CPVS_PYR_BORROW_DEPTH) || depth_on_device, "pyramid_build: BORROW_DEPTH needs device depth");
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);
    ProfInstall prof(ctx->prof_on ? &ctx->prof : nullptr);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = pick_stream(ctx, stream);
    cpvs_pyramid* p = new cpvs_pyramid();
    p->ctx = ctx;
    p->N = N;
    int n = 1;
    while ((1u << (n - 1)) < N) ++n;
    p->n = n;
    const size_t bytes = sizeof(float) * (size_t)N * N;
    if (flags & CPVS_PYR_BORROW_DEPTH) {
        p->depth = const_cast<float*>(depth);
        p->borrowed = true;
    } else {
        if (hipMalloc(&p->depth, bytes) != hipSuccess) {
            (void)hipGetLastError();
            delete p;
            set_error("pyramid_build: hipMalloc of %zu bytes failed", bytes);
            return CPVS_ENOMEM;
        }
        const hipError_t e = hipMemcpyAsync(p->depth, depth, bytes,
                                            depth_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            hipFree(p->depth);
            delete p;
            return hip_fail(e, "pyramid depth copy", __FILE__, __LINE__);
        }
    }
    // Without leafmasks the octree reads levels 1 and 2 as well.
    int rc = pyramid_build_levels(p, (flags & CPVS_PYR_ALL_LEVELS) != 0 || ctx->leafmasks.load() == 0, s);
    // A borrowed device depth map on the context's own stream needs no wait here: builds on that
    // stream run after the levels, and everything else waits for them (pyramid_wait).  A copied
    // input is waited for, so the caller may reuse its buffer, and so is a caller's stream, which
    // the caller may destroy.
    p->pending = p->borrowed && s == ctx->stream;
    if (rc == CPVS_OK && !p->pending) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = hip_fail(e, "pyramid sync", __FILE__, __LINE__);
    }
    if (rc != CPVS_OK) {
        if (!p->borrowed && p->depth) hipFree(p->depth);
        if (p->levels && p->levels_owned) ctx->buffers.give(p->levels);
        delete p;
        return rc;
    }
    *out = p;
    return CPVS_OK;
    GUARD_END
}

int cpvs_pyramid_num_levels(const cpvs_pyramid* p, int* out) {
    CHECK_ARG(p && out, "pyramid_num_levels: NULL argument");
    *out = p->n;
    return CPVS_OK;
}

int cpvs_pyramid_get(const cpvs_pyramid* p, uint32_t level, uint32_t x, uint32_t y, float* mn, float* mx) {
    CHECK_ARG(p && mn && mx, "pyramid_get: NULL argument");
    CHECK_ARG((int)level < p->n, "pyramid_get: level %u >= %d", level, p->n);
    const uint32_t w = p->N >> level;
    CHECK_ARG(x < w && y < w, "pyramid_get: (%u,%u) outside level %u of width %u", x, y, level, w);
    CPVS_HIP_CHECK(hipSetDevice(p->ctx->device));
    CPVS_HIP_CHECK(pyramid_wait(p, nullptr));
    if (level == 0) {
        float v;
        CPVS_HIP_CHECK(hipMemcpy(&v, p->depth + (size_t)y * p->N + x, sizeof(float), hipMemcpyDeviceToHost));
        *mn = *mx = v;
        return CPVS_OK;
    }
    if (!p->lv[level]) {
        set_error("pyramid_get: level %u not stored (build with CPVS_PYR_ALL_LEVELS)", level);
        return CPVS_ESTATE;
    }
    float2 v;
    CPVS_HIP_CHECK(hipMemcpy(&v, p->lv[level] + (size_t)y * w + x, sizeof(float2), hipMemcpyDeviceToHost));
    *mn = v.x;
    *mx = v.y;
    return CPVS_OK;
}

int cpvs_pyramid_copy_level(const cpvs_pyramid* p, uint32_t level, float* host, uint64_t cap) {
    CHECK_ARG(p && host, "pyramid_copy_level: NULL argument");
    CHECK_ARG((int)level < p->n, "pyramid_copy_level: level %u >= %d", level, p->n);
    const size_t w = p->N >> level;
    CPVS_HIP_CHECK(hipSetDevice(p->ctx->device));
    CPVS_HIP_CHECK(pyramid_wait(p, nullptr));
    if (level == 0) {
        CHECK_ARG(cap >= w * w, "pyramid_copy_level: capacity %llu < %zu", (unsigned long long)cap, w * w);
        CPVS_HIP_CHECK(hipMemcpy(host, p->depth, sizeof(float) * w * w, hipMemcpyDeviceToHost));
        return CPVS_OK;
    }
    if (!p->lv[level]) {
        set_error("pyramid_copy_level: level %u not stored (build with CPVS_PYR_ALL_LEVELS)", level);
        return CPVS_ESTATE;
    }
    CHECK_ARG(cap >= 2 * w * w, "pyramid_copy_level: capacity %llu < %zu", (unsigned long long)cap, 2 * w * w);
    CPVS_HIP_CHECK(hipMemcpy(host, p->lv[level], sizeof(float2) * w * w, hipMemcpyDeviceToHost));
    return CPVS_OK;
}

int cpvs_pyramid_destroy(cpvs_pyramid* p) {
    if (!p) return CPVS_OK;
    hipSetDevice(p->ctx->device);
    if (!p->borrowed && p->depth) hipFree(p->depth);
    if (p->levels && p->levels_owned) p->ctx->buffers.give(p->levels);
    delete p;
    return CPVS_OK;
}

// ---------------------------------------------------------------------------------------
// Build
// ---------------------------------------------------------------------------------------

static int check_tiles(const cpvs_pyramid* p, uint32_t t, uint32_t T) {
    // CompressedShadow(uint numLevels) asserts numLevels > 3 (src/CompressedShadow.cpp:44-47).
    CHECK_ARG(p->n >= 4, "build: a %ux%u map has %d levels; the octree needs N >= 8", p->N, p->N, p->n);
    CHECK_ARG(T >= 1 && t < T, "zTileIndex %u must be < zTileNum %u", t, T);
    CHECK_ARG((uint64_t)p->N * T <= (1ull << 21), "N*zTileNum = %llu exceeds 2^21", (unsigned long long)p->N * T);
    return CPVS_OK;
}

int cpvs_estimate_build(cpvs_ctx* ctx, const cpvs_pyramid* p, uint32_t t, uint32_t T, uint64_t* nodes,
                        uint64_t* scratch_bytes, uint64_t* max_dag_bytes) {
    CHECK_ARG(ctx && p, "estimate_build: NULL argument");
    int rc = check_tiles(p, t, T);
    if (rc) return rc;
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    CPVS_HIP_CHECK(pyramid_wait(p, s));
    CountLayout lay;
    const int n = p->n, R = n - 2;
    const bool leafm = n >= 5 && ctx->leafmasks.load() != 0;
    rc = pyramid_node_counts(p, t, T, leafm, pinned_partials(ctx->pinned), nullptr, s, &lay);
    if (rc) return rc;
    CPVS_HIP_CHECK(hipStreamSynchronize(s));
    const int bottom = leafm ? 2 : 0;
    u64 cnt[kMaxLevels] = {};
    sum_node_counts(lay, pinned_partials(ctx->pinned), cnt);
    cnt[R] = 1;
    u64 cmax = 1, per_level = 0, words = 0;
    for (int L = bottom; L <= R; ++L) {
        const u64 c = cnt[L] ? cnt[L] : 1;
        cmax = c > cmax ? c : cmax;
        per_level += arena_bytes(c, 8) + 5 * arena_bytes(c, 4);
        // Pre-merge words: mask + one pointer per child (two words per partial slice at leaves).
        words += cnt[L] + ((L > bottom) ? cnt[L - 1] : 0) + ((bottom == 2 && L == 2) ? 16 * cnt[L] : 0);
    }
    const u64 c2 = (bottom == 2) ? (cnt[2] ? cnt[2] : 1) : 0;
    const u64 ncol = (bottom == 2) ? (u64)(p->N >> 3) * (p->N >> 3) : 0;
    const u64 leaf = (bottom == 2) ? arena_bytes(c2 * 8, 8) + 2 * arena_bytes(c2, 4) +
                                         arena_bytes(ncol, 16)
                                   : 0;
    u64 slots = 1024;
    while (slots < cmax + cmax / 2 + 1) slots <<= 1;
    const u64 merge = (ctx->merge_mode == CPVS_MERGE_TABLE)
                          ? arena_bytes(slots, 8)
                          : 4 * arena_bytes(cmax, 4) + arena_bytes(sort_scratch_words((u32)cmax), 4);
    const u64 temps = 2 * arena_bytes(cmax, 4) + merge + arena_bytes(scan_scratch_words((u32)(cmax > ncol ? cmax : ncol)), 4) +
                      (1 << 20);
    if (nodes)
        for (int L = 0; L < n; ++L) nodes[L] = cnt[L];
    // The words are staged at their pre-merge bound before the exact allocation.
    if (scratch_bytes) *scratch_bytes = per_level + leaf + temps + arena_bytes(words + 17, 4);
    if (max_dag_bytes) *max_dag_bytes = 4 * words;
    return CPVS_OK;
    GUARD_END
}

int cpvs_build(cpvs_ctx* ctx, const cpvs_pyramid* p, uint32_t t, uint32_t T, void* stream, cpvs_dag** out) {
    CHECK_ARG(ctx && p && out, "build: NULL argument");
    CHECK_ARG(p->ctx == ctx, "build: pyramid belongs to another context");
    int rc = check_tiles(p, t, T);
    if (rc) return rc;
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);
    ProfInstall prof(ctx->prof_on ? &ctx->prof : nullptr);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = pick_stream(ctx, stream);
    CPVS_HIP_CHECK(pyramid_wait(p, s));
    return build_dag(ctx, ctx->arena, ctx->pinned, ctx->merge_mode, p, t, T, s, out);
    GUARD_END
}

int cpvs_dag_info(const cpvs_dag* d, uint32_t* num_levels, uint64_t* num_words, int* total_vis) {
    CHECK_ARG(d, "dag_info: NULL dag");
    if (num_levels) *num_levels = (uint32_t)d->n;
    if (num_words) *num_words = d->words;
    if (total_vis) *total_vis = (d->root_mask == 0x5555u) ? CPVS_VISIBLE : (d->root_mask == 0u) ? CPVS_SHADOW : CPVS_PARTIAL;
    return CPVS_OK;
}

int cpvs_dag_level_stats(const cpvs_dag* d, uint64_t* nodes, uint64_t* uniques, int cap) {
    CHECK_ARG(d, "dag_level_stats: NULL dag");
    for (int L = 0; L < cap && L < kMaxLevels; ++L) {
        if (nodes) nodes[L] = d->nodes[L];
        if (uniques) uniques[L] = d->uniques[L];
    }
    return CPVS_OK;
}

int cpvs_dag_copy_words(const cpvs_dag* d, uint32_t* host, uint64_t cap) {
    CHECK_ARG(d && host, "dag_copy_words: NULL argument");
    CHECK_ARG(cap >= d->words, "dag_copy_words: capacity %llu < %llu words", (unsigned long long)cap,
              (unsigned long long)d->words);
    CPVS_HIP_CHECK(hipSetDevice(d->ctx->device));
    CPVS_HIP_CHECK(hipMemcpy(host, d->d_words, sizeof(uint32_t) * d->words, hipMemcpyDeviceToHost));
    return CPVS_OK;
}

int cpvs_dag_device_words(const cpvs_dag* d, const uint32_t** dptr) {
    CHECK_ARG(d && dptr, "dag_device_words: NULL argument");
    *dptr = d->d_words;
    return CPVS_OK;
}

int cpvs_dag_traverse(cpvs_dag* d, const float* ndc, uint64_t count, int try_leaf, int* vis) {
    CHECK_ARG(d && (ndc || !count) && (vis || !count), "dag_traverse: NULL argument");
    // Leafmask DAGs are only walkable with leafmasks and plain ones (numLevels < 5, or built without
    // leafmasks) without (src/CompressedShadow.cpp:20-27,404-463); the other combinations read leaf
    // payload as pointers in the reference.
    CHECK_ARG((try_leaf != 0) == d->leafmasks, "dag_traverse: tryLeafmasks=%d invalid for a DAG %s leafmasks", try_leaf,
              d->leafmasks ? "with" : "without");
    if (!count) return CPVS_OK;
    GUARD_BEGIN
    cpvs_ctx* ctx = d->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    float* d_ndc = nullptr;
    int* d_out = nullptr;
    CPVS_HIP_CHECK(hipMalloc(&d_ndc, sizeof(float) * 3 * count));
    if (hipMalloc(&d_out, sizeof(int) * count) != hipSuccess) {
        hipFree(d_ndc);
        set_error("dag_traverse: hipMalloc failed");
        return CPVS_ENOMEM;
    }
    int rc = CPVS_OK;
    hipError_t e = hipMemcpyAsync(d_ndc, ndc, sizeof(float) * 3 * count, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        rc = launch_traverse_single(d->d_words, (u32)d->words, d->n, d_ndc, count, try_leaf, d_out, s);
        if (rc == CPVS_OK) {
            e = hipMemcpyAsync(vis, d_out, sizeof(int) * count, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
        }
    }
    hipFree(d_ndc);
    hipFree(d_out);
    if (e != hipSuccess) return hip_fail(e, "dag_traverse", __FILE__, __LINE__);
    return rc;
    GUARD_END
}

}  // extern "C"

namespace cpvs {
void dag_destroy_idle(cpvs_dag* d) {
    if (!d) return;
    if (d->d_words) d->ctx->buffers.give_idle(d->d_words);
    delete d;
}
}  // namespace cpvs

extern "C" {

int cpvs_dag_set_top_levels(cpvs_dag* d, int levels) {
    CHECK_ARG(d, "dag_set_top_levels: NULL DAG");
    CHECK_ARG(levels >= -1 && levels <= kMaxTopLevels, "dag_set_top_levels: %d levels (-1..%d)", levels, kMaxTopLevels);
    // As a container's table over one cell (cpvs_container_copy_to_gpu): at most the DAG's inner
    // levels above the last; automatic: the largest within kTopTableBytes, none below 3 levels.
    const int lead = d->n - 1 - (d->leafmasks ? 3 : 1);
    int K = (levels < 0) ? std::min(lead, kMaxTopLevels) : std::min(levels, lead);
    if (levels < 0) {
        while (K > 0 && (sizeof(u64) << (3 * K)) > kTopTableBytes) --K;
        if (K < 3) K = 0;
    }
    // A root that is all lit or all shadow is answered before any walk; offsets must stay below
    // the table's sentinels.
    if (d->root_mask == 0u || d->root_mask == 0x5555u || d->words >= 0xFFFFFFFDull) K = 0;
    cpvs_ctx* ctx = d->ctx;
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    u64* nt = nullptr;
    if (K > 0) {
        nt = static_cast<u64*>(ctx->buffers.take(sizeof(u64) << (3 * K)));
        if (!nt) {
            set_error("dag_set_top_levels: allocation of a %d-level table failed", K);
            return CPVS_ENOMEM;
        }
        ProfInstall prof(ctx->prof_on ? &ctx->prof : nullptr);
        int rc = launch_build_top(d->d_words, (u32)d->words, nullptr, 1u, d->n, (u32)K, nt, ctx->stream);
        if (rc == CPVS_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = CPVS_EHIP;
        if (rc) {
            ctx->buffers.give(nt);
            if (rc == CPVS_EHIP) set_error("dag_set_top_levels: building the table failed");
            return rc;
        }
    }
    if (d->d_top) d->retired_tops.push_back(d->d_top);
    d->d_top = nt;
    d->top_levels = (u32)K;
    return CPVS_OK;
    GUARD_END
}

int cpvs_dag_top_levels(const cpvs_dag* d, int* levels) {
    CHECK_ARG(d && levels, "dag_top_levels: NULL argument");
    *levels = d->d_top ? (int)d->top_levels : 0;
    return CPVS_OK;
}

int cpvs_dag_destroy(cpvs_dag* d) {
    if (!d) return CPVS_OK;
    hipSetDevice(d->ctx->device);
    if (d->d_words) d->ctx->buffers.give(d->d_words);
    if (d->d_top) d->ctx->buffers.give(d->d_top);
    for (u64* t : d->retired_tops) d->ctx->buffers.give(t);
    delete d;
    return CPVS_OK;
}

// ---------------------------------------------------------------------------------------
// Container
// ---------------------------------------------------------------------------------------

int cpvs_container_create(cpvs_ctx* ctx, uint32_t length, cpvs_container** out) {
    CHECK_ARG(ctx && out, "container_create: NULL argument");
    CHECK_ARG(is_pow2(length) && length <= 64, "container_create: length %u must be a power of two <= 64", length);
    GUARD_BEGIN
    cpvs_container* c = new cpvs_container();
    c->ctx = ctx;
    c->length = length;
    c->dags.assign((size_t)length * length * length, nullptr);
    *out = c;
    return CPVS_OK;
    GUARD_END
}

int cpvs_container_set(cpvs_container* c, cpvs_dag* dag, uint32_t x, uint32_t y, uint32_t z) {
    CHECK_ARG(c, "container_set: NULL container");
    CHECK_ARG(x < c->length && y < c->length && z < c->length, "container_set: (%u,%u,%u) outside length %u", x, y, z,
              c->length);
    if (c->freed_cpu) {
        set_error("container_set: container was moved to the GPU");
        return CPVS_ESTATE;
    }
    const size_t i = ((size_t)z * c->length + y) * c->length + x;
    if (c->dags[i] && c->dags[i] != dag) cpvs_dag_destroy(c->dags[i]);
    c->dags[i] = dag;
    return CPVS_OK;
}

int cpvs_container_get(const cpvs_container* c, uint32_t x, uint32_t y, uint32_t z, const cpvs_dag** out) {
    CHECK_ARG(c && out, "container_get: NULL argument");
    CHECK_ARG(x < c->length && y < c->length && z < c->length, "container_get: (%u,%u,%u) outside length %u", x, y, z,
              c->length);
    *out = c->freed_cpu ? nullptr : c->dags[((size_t)z * c->length + y) * c->length + x];
    return CPVS_OK;
}

int cpvs_container_set_filter_size(cpvs_container* c, uint32_t size) {
    CHECK_ARG(c, "container_set_filter_size: NULL container");
    c->filter_size = size;
    return CPVS_OK;
}

int cpvs_container_set_top_levels(cpvs_container* c, int levels) {
    CHECK_ARG(c, "container_set_top_levels: NULL container");
    CHECK_ARG(levels >= -1 && levels <= kMaxTopLevels, "container_set_top_levels: %d levels (-1..%d)", levels,
              kMaxTopLevels);
    CHECK_ARG(levels < 0 || ((u64)c->dags.size() << (3 * levels)) * sizeof(u64) <= kTopTableMaxBytes,
              "container_set_top_levels: %d levels over %zu cells exceed %llu bytes", levels, c->dags.size(),
              (unsigned long long)kTopTableMaxBytes);
    c->top_request = levels;
    return CPVS_OK;
}

int cpvs_container_top_levels(const cpvs_container* c, int* levels) {
    CHECK_ARG(c && levels, "container_top_levels: NULL argument");
    *levels = (int)c->top_levels;
    return CPVS_OK;
}

int cpvs_container_copy_to_gpu(cpvs_container* c) {
    CHECK_ARG(c, "container_copy_to_gpu: NULL container");
    if (c->freed_cpu) {
        set_error("container_copy_to_gpu: DAGs were already freed");
        return CPVS_ESTATE;
    }
    GUARD_BEGIN
    const size_t cells = c->dags.size();
    int dl = -1;
    u64 total = 0;
    std::vector<u32> grid(cells);
    for (size_t i = 0; i < cells; ++i) {
        const cpvs_dag* d = c->dags[i];
        if (!d) {
            set_error("container_copy_to_gpu: slot %zu is empty", i);
            return CPVS_ESTATE;
        }
        if (dl < 0) dl = d->n;
        if (d->n != dl) {
            set_error("container_copy_to_gpu: DAG %zu has %d levels, DAG 0 has %d", i, d->n, dl);
            return CPVS_EINVAL;
        }
        // One lookup walks every cell (traverse.cs:8-9 compiles LEAFMASKS in or out).
        if (d->leafmasks != c->dags[0]->leafmasks) {
            set_error("container_copy_to_gpu: DAG %zu %s leafmasks, DAG 0 %s", i, d->leafmasks ? "has" : "has no",
                      c->dags[0]->leafmasks ? "has" : "has none");
            return CPVS_EINVAL;
        }
        // createTopLevelGrid (src/CompressedShadowContainer.cpp:71-91).
        grid[i] = (d->root_mask == 0u) ? 0xFFFFFFFFu : (d->root_mask == 0x5555u) ? 0xFFFFFFFEu : (u32)total;
        total += d->words;
    }
    CHECK_ARG(total < 0xFFFFFFFEull, "container: %llu combined words exceed 32-bit offsets", (unsigned long long)total);
    cpvs_ctx* ctx = c->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if (c->d_dag) hipFree(c->d_dag);
    if (c->d_grid) hipFree(c->d_grid);
    if (c->d_top) hipFree(c->d_top);
    c->d_dag = nullptr;
    c->d_grid = nullptr;
    c->d_top = nullptr;
    c->top_levels = 0;
    if (hipMalloc(&c->d_dag, sizeof(u32) * (total ? total : 1)) != hipSuccess ||
        hipMalloc(&c->d_grid, sizeof(u32) * cells) != hipSuccess) {
        (void)hipGetLastError();
        set_error("container_copy_to_gpu: hipMalloc failed");
        return CPVS_ENOMEM;
    }
    // combineDAGs (src/CompressedShadowContainer.cpp:52-69): concatenation in slot order.
    u64 off = 0;
    for (size_t i = 0; i < cells; ++i) {
        const cpvs_dag* d = c->dags[i];
        CPVS_HIP_CHECK(hipMemcpyAsync(c->d_dag + off, d->d_words, sizeof(u32) * d->words, hipMemcpyDeviceToDevice, s));
        off += d->words;
    }
    CPVS_HIP_CHECK(hipMemcpyAsync(c->d_grid, grid.data(), sizeof(u32) * cells, hipMemcpyHostToDevice, s));
    // The lookups' top table (launch_build_top): K levels below the grid cell, no more than the
    // DAGs' inner levels above the last (level 3 with leafmasks, 1 without) nor kMaxTopLevels, and
    // within kTopTableBytes; skipped below 3 levels, where it saves little.
    const int lead = dl - 1 - (c->dags[0]->leafmasks ? 3 : 1);
    int K = (c->top_request < 0) ? std::min(lead, kMaxTopLevels) : std::min(c->top_request, lead);
    if (c->top_request < 0) {
        while (K > 0 && ((u64)cells << (3 * K)) * sizeof(u64) > kTopTableBytes) --K;
        if (K < 3) K = 0;
    }
    if (K > 0 && total < 0xFFFFFFFDull) {
        const u64 entries = (u64)cells << (3 * K);
        if (hipMalloc(&c->d_top, sizeof(u64) * entries) != hipSuccess) {
            (void)hipGetLastError();
            set_error("container_copy_to_gpu: hipMalloc of the %llu-entry top table failed",
                      (unsigned long long)entries);
            return CPVS_ENOMEM;
        }
        ProfInstall prof(ctx->prof_on ? &ctx->prof : nullptr);
        int rc = launch_build_top(c->d_dag, (u32)total, c->d_grid, (u32)cells, dl, (u32)K, c->d_top, s);
        if (rc) return rc;
        c->top_levels = (u32)K;
    }
    CPVS_HIP_CHECK(hipStreamSynchronize(s));
    c->dag_words = total;
    c->dag_levels = dl;
    c->leafmasks = c->dags[0]->leafmasks;
    int gl = 0;
    while ((1u << gl) < c->length) ++gl;
    c->grid_levels = gl;  // (uint)log8(length^3) (src/CompressedShadowContainer.cpp:42)
    c->on_gpu = true;
    return CPVS_OK;
    GUARD_END
}

int cpvs_container_free_on_cpu(cpvs_container* c) {
    CHECK_ARG(c, "container_free_on_cpu: NULL container");
    for (auto& d : c->dags) {
        cpvs_dag_destroy(d);
        d = nullptr;
    }
    c->freed_cpu = true;
    return CPVS_OK;
}

int cpvs_container_move_to_gpu(cpvs_container* c) {
    int rc = cpvs_container_copy_to_gpu(c);
    if (rc) return rc;
    return cpvs_container_free_on_cpu(c);
}

int cpvs_container_info(const cpvs_container* c, int* dag_levels, int* grid_levels, uint64_t* num_words) {
    CHECK_ARG(c, "container_info: NULL container");
    if (dag_levels) *dag_levels = c->dag_levels;
    if (grid_levels) *grid_levels = c->grid_levels;
    if (num_words) *num_words = c->dag_words;
    return CPVS_OK;
}

int cpvs_container_copy_words(const cpvs_container* c, uint32_t* hw, uint64_t capw, uint32_t* hg, uint64_t capg) {
    CHECK_ARG(c, "container_copy_words: NULL container");
    if (!c->on_gpu) {
        set_error("container_copy_words: not on the GPU yet");
        return CPVS_ESTATE;
    }
    CPVS_HIP_CHECK(hipSetDevice(c->ctx->device));
    if (hw) {
        CHECK_ARG(capw >= c->dag_words, "container_copy_words: word capacity too small");
        CPVS_HIP_CHECK(hipMemcpy(hw, c->d_dag, sizeof(u32) * c->dag_words, hipMemcpyDeviceToHost));
    }
    if (hg) {
        CHECK_ARG(capg >= c->dags.size(), "container_copy_words: grid capacity too small");
        CPVS_HIP_CHECK(hipMemcpy(hg, c->d_grid, sizeof(u32) * c->dags.size(), hipMemcpyDeviceToHost));
    }
    return CPVS_OK;
}

static Mat4 to_mat(const float* m) {
    Mat4 M;
    static const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    std::memcpy(M.m, m ? m : I, sizeof(M.m));
    return M;
}

int cpvs_container_evaluate(cpvs_container* c, const float* pos, uint32_t stride, uint64_t count, const float* m,
                            uint8_t* vis, void* stream) {
    CHECK_ARG(c && (pos || !count) && (vis || !count), "container_evaluate: NULL argument");
    CHECK_ARG(stride == 3 || stride == 4, "container_evaluate: stride %u must be 3 or 4", stride);
    if (!c->on_gpu) {
        set_error("container_evaluate: copyToGPU first");
        return CPVS_ESTATE;
    }
    CPVS_HIP_CHECK(hipSetDevice(c->ctx->device));
    ProfInstall prof(c->ctx->prof_on ? &c->ctx->prof : nullptr);
    return launch_evaluate(c->d_dag, (u32)c->dag_words, c->d_grid, c->dag_levels, c->grid_levels, c->leafmasks, c->d_top,
                           c->top_levels, pos, stride, count,
                           to_mat(m), vis, pick_stream(c->ctx, stream));
}

int cpvs_container_evaluate_pcf(cpvs_container* c, const float* pos, uint32_t stride, uint64_t count, const float* m,
                                float* vis, void* stream) {
    CHECK_ARG(c && (pos || !count) && (vis || !count), "container_evaluate_pcf: NULL argument");
    CHECK_ARG(stride == 3 || stride == 4, "container_evaluate_pcf: stride %u must be 3 or 4", stride);
    CHECK_ARG(c->filter_size >= 1 && c->filter_size <= kMaxFilterSize, "container_evaluate_pcf: filter size %u (1..%u)",
              c->filter_size, kMaxFilterSize);
    if (!c->on_gpu) {
        set_error("container_evaluate_pcf: copyToGPU first");
        return CPVS_ESTATE;
    }
    CPVS_HIP_CHECK(hipSetDevice(c->ctx->device));
    ProfInstall prof(c->ctx->prof_on ? &c->ctx->prof : nullptr);
    return launch_evaluate_pcf(c->d_dag, (u32)c->dag_words, c->d_grid, c->dag_levels, c->grid_levels, c->leafmasks, c->d_top,
                               c->top_levels, pos, stride, count,
                               to_mat(m), c->filter_size, vis, pick_stream(c->ctx, stream));
}

int cpvs_container_evaluate_image(cpvs_container* c, const float* pos, uint32_t width, uint32_t height, const float* m,
                                  uint8_t* vis, void* stream) {
    CHECK_ARG(c && pos && vis, "container_evaluate_image: NULL argument");
    if (!c->on_gpu) {
        set_error("container_evaluate_image: copyToGPU first");
        return CPVS_ESTATE;
    }
    CPVS_HIP_CHECK(hipSetDevice(c->ctx->device));
    ProfInstall prof(c->ctx->prof_on ? &c->ctx->prof : nullptr);
    return launch_evaluate_image(c->d_dag, (u32)c->dag_words, c->d_grid, c->dag_levels, c->grid_levels, c->leafmasks,
                                 c->d_top, c->top_levels, pos, width, height,
                                 to_mat(m), vis, pick_stream(c->ctx, stream));
}

int cpvs_container_tune_top_levels(cpvs_container* c, const float* pos, uint32_t stride, uint64_t count,
                                   const float* m, uint8_t* vis, const int* cands, uint32_t ncand, uint32_t reps,
                                   int* best, double* ms) {
    CHECK_ARG(c && pos && vis && count && cands && ncand && best, "container_tune_top_levels: NULL or empty argument");
    CHECK_ARG(reps >= 1 && reps <= 100, "container_tune_top_levels: reps %u must be in [1, 100]", reps);
    if (c->freed_cpu) {
        set_error("container_tune_top_levels: the DAGs were freed (free_on_cpu): nothing to rebuild from");
        return CPVS_ESTATE;
    }
    const int keep = c->top_request;
    for (uint32_t k = 0; k < ncand; ++k) {
        const int rc = cpvs_container_set_top_levels(c, cands[k]);
        c->top_request = keep;
        if (rc) return rc;
    }
    GUARD_BEGIN
    CPVS_HIP_CHECK(hipSetDevice(c->ctx->device));
    hipStream_t s = c->ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CPVS_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        set_error("container_tune_top_levels: hipEventCreate failed");
        return CPVS_EHIP;
    }
    int rc = CPVS_OK;
    float best_ms = 0.0f;
    uint32_t pick = 0;
    for (uint32_t k = 0; k < ncand && rc == CPVS_OK; ++k) {
        c->top_request = cands[k];
        rc = cpvs_container_copy_to_gpu(c);
        // one warm-up, then the best of reps timed lookups
        if (rc == CPVS_OK) rc = cpvs_container_evaluate(c, pos, stride, count, m, vis, s);
        float t = 0.0f;
        for (uint32_t r = 0; r < reps && rc == CPVS_OK; ++r) {
            float one = 0.0f;
            if (hipEventRecord(e0, s) != hipSuccess) rc = CPVS_EHIP;
            if (rc == CPVS_OK) rc = cpvs_container_evaluate(c, pos, stride, count, m, vis, s);
            if (rc == CPVS_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                  hipEventElapsedTime(&one, e0, e1) != hipSuccess))
                rc = CPVS_EHIP;
            if (rc == CPVS_OK && (r == 0 || one < t)) t = one;
        }
        if (rc == CPVS_OK) {
            if (ms) ms[k] = t;
            if (k == 0 || t < best_ms) {
                best_ms = t;
                pick = k;
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc == CPVS_EHIP) set_error("container_tune_top_levels: event timing failed");
    c->top_request = (rc == CPVS_OK) ? cands[pick] : keep;
    if (rc == CPVS_OK) {
        rc = cpvs_container_copy_to_gpu(c);
        *best = (int)pick;
    }
    return rc;
    GUARD_END
}

int cpvs_container_destroy(cpvs_container* c) {
    if (!c) return CPVS_OK;
    hipSetDevice(c->ctx->device);
    for (auto* d : c->dags) cpvs_dag_destroy(d);
    if (c->d_dag) hipFree(c->d_dag);
    if (c->d_grid) hipFree(c->d_grid);
    if (c->d_top) hipFree(c->d_top);
    delete c;
    return CPVS_OK;
}

// ---------------------------------------------------------------------------------------
// Helpers
// ---------------------------------------------------------------------------------------

int cpvs_device_alloc(cpvs_ctx* ctx, uint64_t bytes, void** dptr) {
    CHECK_ARG(ctx && dptr, "device_alloc: NULL argument");
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        set_error("device_alloc: hipMalloc of %llu bytes failed", (unsigned long long)bytes);
        return CPVS_ENOMEM;
    }
    return CPVS_OK;
}

int cpvs_device_free(cpvs_ctx* ctx, void* dptr) {
    CHECK_ARG(ctx, "device_free: NULL ctx");
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    if (dptr) CPVS_HIP_CHECK(hipFree(dptr));
    return CPVS_OK;
}

int cpvs_memcpy_h2d(cpvs_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    CHECK_ARG(ctx && dst && src, "memcpy_h2d: NULL argument");
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    CPVS_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return CPVS_OK;
}

int cpvs_memcpy_d2d(cpvs_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream) {
    CHECK_ARG(ctx && dst && src, "memcpy_d2d: NULL argument");
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    CPVS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, pick_stream(ctx, stream)));
    return CPVS_OK;
}

int cpvs_device_copy(cpvs_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream) {
    CHECK_ARG(ctx && dst && src, "device_copy: NULL argument");
    CHECK_ARG(bytes % 16 == 0, "device_copy: %llu bytes is not a multiple of 16", (unsigned long long)bytes);
    CHECK_ARG(((uintptr_t)dst | (uintptr_t)src) % 16 == 0, "device_copy: pointers must be 16-byte aligned");
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    return launch_device_copy(dst, src, bytes, pick_stream(ctx, stream));
}

int cpvs_memcpy_d2h(cpvs_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    CHECK_ARG(ctx && dst && src, "memcpy_d2h: NULL argument");
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    CPVS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    CPVS_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return CPVS_OK;
}

int cpvs_event_record(cpvs_ctx* ctx, void* stream, void** event) {
    CHECK_ARG(ctx && event, "event_record: NULL argument");
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    hipEvent_t e;
    CPVS_HIP_CHECK(hipEventCreate(&e));
    CPVS_HIP_CHECK(hipEventRecord(e, pick_stream(ctx, stream)));
    *event = e;
    return CPVS_OK;
}

int cpvs_event_elapsed_ms(void* start, void* stop, float* ms) {
    CHECK_ARG(start && stop && ms, "event_elapsed_ms: NULL argument");
    CPVS_HIP_CHECK(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
    CPVS_HIP_CHECK(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)));
    return CPVS_OK;
}

int cpvs_event_destroy(void* event) {
    if (event) hipEventDestroy(reinterpret_cast<hipEvent_t>(event));
    return CPVS_OK;
}

int cpvs_gen_depth(cpvs_ctx* ctx, float* d_out, uint32_t fullN, uint64_t seed, uint32_t nboxes, int noise, uint32_t x0,
                   uint32_t y0, uint32_t tileN, void* stream) {
    CHECK_ARG(ctx && d_out, "gen_depth: NULL argument");
    CHECK_ARG(fullN >= 1 && tileN >= 1 && (uint64_t)x0 + tileN <= fullN && (uint64_t)y0 + tileN <= fullN,
              "gen_depth: tile [%u,%u)+%u outside %u", x0, y0, tileN, fullN);
    GUARD_BEGIN
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    return launch_gen_depth(d_out, fullN, seed, nboxes, noise, x0, y0, tileN, pick_stream(ctx, stream));
    GUARD_END
}

int cpvs_gen_queries_uniform(cpvs_ctx* ctx, float* d_out, uint64_t seed, uint64_t count, uint32_t stride, void* stream) {
    CHECK_ARG(ctx && (d_out || !count), "gen_queries_uniform: NULL argument");
    CHECK_ARG(stride == 3 || stride == 4, "gen_queries_uniform: stride %u must be 3 or 4", stride);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    return launch_gen_queries_uniform(d_out, seed, 0, 1, count, stride, pick_stream(ctx, stream));
}

int cpvs_gen_queries_uniform_range(cpvs_ctx* ctx, float* d_out, uint64_t seed, uint64_t first, uint64_t step,
                                   uint64_t count, uint32_t stride, void* stream) {
    CHECK_ARG(ctx && (d_out || !count), "gen_queries_uniform_range: NULL argument");
    CHECK_ARG(stride == 3 || stride == 4, "gen_queries_uniform_range: stride %u must be 3 or 4", stride);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    return launch_gen_queries_uniform(d_out, seed, first, step, count, stride, pick_stream(ctx, stream));
}

int cpvs_gen_queries_gbuffer(cpvs_ctx* ctx, float* d_out, const float* d_depth, uint32_t N, uint64_t seed, uint32_t W,
                             uint32_t H, void* stream) {
    CHECK_ARG(ctx && d_out && d_depth, "gen_queries_gbuffer: NULL argument");
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    return launch_gen_queries_gbuffer(d_out, d_depth, N, seed, W, H, pick_stream(ctx, stream));
}

}  // extern "C"
