// batch.hip -- many maps at once: concurrent build lanes, the lane-count tuner, and DAG sets
// (many DAGs queried in one launch).
//
// The reference builds maps one after another (src/DeferredRenderer.cpp:150-187, one
// CompressedShadow::create per tile and z-tile).  One build is a chain of ~60 short launches
// with three host readbacks; small maps leave the GPU idle between them.  A batch runs L lanes,
// each a host thread with its own stream, scratch arena, pyramid-level arena and pinned staging,
// pulling maps from one queue in decreasing size (so a lane's arenas grow on its first map and
// are reused after).  Every map goes through the same kernels as cpvs_build, so the DAG words
// are those of a plain build.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "abi.hpp"

namespace cpvs {

int Worker::create() {
    CPVS_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    const hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&pinned), kPinnedBytes, kPinnedFlags);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(stream);
        stream = nullptr;
        pinned = nullptr;
        return hip_fail(e, "hipHostMalloc", __FILE__, __LINE__);
    }
    return CPVS_OK;
}

int Worker::forest_staging(size_t bytes) {
    if (bytes <= fpin_bytes) return CPVS_OK;
    CPVS_HIP_CHECK(hipStreamSynchronize(stream));  // no copy still reads the old staging
    if (fpin) (void)hipHostFree(fpin);
    fpin = nullptr;
    fpin_bytes = 0;
    const size_t want = bytes + (bytes >> 2);
    CPVS_HIP_CHECK(hipHostMalloc(&fpin, want, kPinnedFlags));
    fpin_bytes = want;
    return CPVS_OK;
}

void Worker::destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    arena.release();
    levels.release();
    if (fpin) (void)hipHostFree(fpin);
    fpin = nullptr;
    fpin_bytes = 0;
    if (pinned) (void)hipHostFree(pinned);
    if (stream) (void)hipStreamDestroy(stream);
    pinned = nullptr;
    stream = nullptr;
}

namespace {

constexpr uint32_t kMaxLanes = 16;
constexpr uint32_t kDefaultLanes = 4;
constexpr uint32_t kDefaultForestMtexels = 256;

int levels_of(u32 N) {
    int n = 1;
    while ((1u << (n - 1)) < N) ++n;
    return n;
}

int check_item(const cpvs_batch_item& it, uint32_t i) {
    CHECK_ARG(it.depth, "batch item %u: NULL depth", i);
    CHECK_ARG(is_pow2(it.N) && it.N >= 8 && it.N <= 65536, "batch item %u: N=%u must be a power of two in [8, 65536]", i,
              it.N);
    CHECK_ARG(it.zTileNum >= 1 && it.zTileIndex < it.zTileNum, "batch item %u: zTileIndex %u must be < zTileNum %u", i,
              it.zTileIndex, it.zTileNum);
    CHECK_ARG((uint64_t)it.N * it.zTileNum <= (1ull << 21), "batch item %u: N*zTileNum exceeds 2^21", i);
    return CPVS_OK;
}

// MinMaxHierarchy of one depth map on lane w (levels in the lane's arena).
int build_pyramid(cpvs_ctx* ctx, Worker& w, const cpvs_batch_item& it, cpvs_pyramid* p) {
    p->ctx = ctx;
    p->N = it.N;
    p->n = levels_of(it.N);
    p->depth = const_cast<float*>(it.depth);
    p->borrowed = true;
    return pyramid_build_levels(p, ctx->leafmasks.load() == 0, w.stream, &w.levels);
}

int ensure_lanes(cpvs_ctx* ctx, uint32_t lanes) {
    while (ctx->lanes.size() < lanes) {
        Worker* w = new Worker();
        const int rc = w->create();
        if (rc) {
            delete w;
            return rc;
        }
        ctx->lanes.push_back(w);
    }
    return CPVS_OK;
}

// Runs the batch on `lanes` lanes; the calling thread is lane 0.  Caller holds ctx->mu.
int run_batch(cpvs_ctx* ctx, const cpvs_batch_item* items, uint32_t count, uint32_t lanes, cpvs_dag** out,
              double* done_ms) {
    for (uint32_t i = 0; i < count; ++i) out[i] = nullptr;
    if (!count) return CPVS_OK;
    int rc = ensure_lanes(ctx, lanes);
    if (rc) return rc;
    // Items of one depth map (same pointer and N: the z-tiles of a tile) form a group, built by one
    // lane from one pyramid, as the reference's loop builds one MinMaxHierarchy per tile
    // (src/DeferredRenderer.cpp:150-187).  Groups go out largest map first.
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const u64 sa = (u64)items[a].N * items[a].N, sb = (u64)items[b].N * items[b].N;
        if (sa != sb) return sa > sb;
        if (items[a].depth != items[b].depth) return std::less<const float*>()(items[a].depth, items[b].depth);
        return false;
    });
    std::vector<uint32_t> group_start;  // group g is order[group_start[g], group_start[g + 1])
    for (uint32_t j = 0; j < count; ++j) {
        const cpvs_batch_item& a = items[order[j]];
        if (j == 0 || a.depth != items[order[j - 1]].depth || a.N != items[order[j - 1]].N) group_start.push_back(j);
    }
    const uint32_t ngroups = (uint32_t)group_start.size();
    group_start.push_back(count);
    std::atomic<uint32_t> next{0};
    std::atomic<bool> failed{false};
    std::mutex emu;
    int first_rc = CPVS_OK;
    std::string first_msg;
    const int merge_mode = ctx->merge_mode;
    const auto t0 = std::chrono::steady_clock::now();

    auto record = [&](int lrc, const char* what) {
        std::lock_guard<std::mutex> lk(emu);
        if (first_rc == CPVS_OK) {
            first_rc = lrc;
            first_msg = what;
        }
        failed.store(true);
    };
    auto body = [&](Worker* w) {
        if (hipSetDevice(ctx->device) != hipSuccess) {
            record(CPVS_EHIP, "batch: hipSetDevice failed on a lane thread");
            return;
        }
        while (!failed.load(std::memory_order_relaxed)) {
            const uint32_t g = next.fetch_add(1, std::memory_order_relaxed);
            if (g >= ngroups) break;
            cpvs_pyramid p;
            bool ok = true;
            for (uint32_t j = group_start[g]; ok && j < group_start[g + 1]; ++j) {
                const uint32_t i = order[j];
                cpvs_dag* d = nullptr;
                int lrc;
                try {
                    lrc = (j == group_start[g]) ? build_pyramid(ctx, *w, items[i], &p) : CPVS_OK;
                    if (lrc == CPVS_OK)
                        lrc = build_dag(ctx, w->arena, w->pinned, merge_mode, &p, items[i].zTileIndex, items[i].zTileNum,
                                        w->stream, &d);
                } catch (const std::bad_alloc&) {
                    set_error("host allocation failed");
                    lrc = CPVS_ENOMEM;
                } catch (...) {
                    set_error("internal exception");
                    lrc = CPVS_EINTERNAL;
                }
                if (lrc != CPVS_OK) {
                    char msg[600];
                    snprintf(msg, sizeof(msg), "batch item %u: %s", i, cpvs_last_error());
                    record(lrc, msg);
                    ok = false;
                    break;
                }
                out[i] = d;
                if (done_ms)
                    done_ms[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
            if (!ok) break;
        }
        (void)hipStreamSynchronize(w->stream);
    };

    std::vector<std::thread> threads;
    threads.reserve(lanes);
    try {
        for (uint32_t k = 1; k < lanes; ++k) threads.emplace_back(body, ctx->lanes[k]);
    } catch (...) {
        failed.store(true);
        for (auto& t : threads) t.join();
        set_error("batch: could not start %u lane threads", lanes - 1);
        rc = CPVS_EINTERNAL;
        for (uint32_t i = 0; i < count; ++i) {
            cpvs_dag_destroy(out[i]);
            out[i] = nullptr;
        }
        return rc;
    }
    body(ctx->lanes[0]);
    for (auto& t : threads) t.join();
    if (first_rc != CPVS_OK) {
        for (uint32_t i = 0; i < count; ++i) {
            cpvs_dag_destroy(out[i]);
            out[i] = nullptr;
        }
        set_error("%s", first_msg.c_str());
        return first_rc;
    }
    return CPVS_OK;
}

// ---------------------------------------------------------------------------------------
// Forests: many maps through one launch chain (forest.hip).  Maps are grouped by depth map (the
// z-tiles of one map share its pyramid and stay in one forest), largest first, and packed into
// forests of at most forest_mtexels x 2^20 texels of work (N^2 / zTileNum per item).  Lanes take
// forests from one queue; each builds the pyramids of its forest's maps into its level arena,
// then the forest.  A forest whose levels outgrow 32-bit words is split in two and built again.
// ---------------------------------------------------------------------------------------
int run_forests(cpvs_ctx* ctx, const cpvs_batch_item* items, uint32_t count, const cpvs_batch_config& cfg, cpvs_dag** out,
                double* done_ms) {
    for (uint32_t i = 0; i < count; ++i) out[i] = nullptr;
    if (!count) return CPVS_OK;
    const uint32_t lanes = cfg.lanes;
    int rc = ensure_lanes(ctx, lanes);
    if (rc) return rc;
    std::vector<uint32_t> order;  // leafmask maps, grouped by depth map, largest first
    std::vector<uint32_t> single;  // N < 16: no leafmask levels, built one by one
    for (uint32_t i = 0; i < count; ++i) (items[i].N >= 16 ? order : single).push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        if (items[a].N != items[b].N) return items[a].N > items[b].N;
        if (items[a].depth != items[b].depth) return std::less<const float*>()(items[a].depth, items[b].depth);
        return false;
    });
    struct Group {
        uint32_t begin, end;  // order[begin, end): the items of one depth map
        u64 work;
    };
    std::vector<Group> groups;
    for (uint32_t j = 0; j < (uint32_t)order.size(); ++j) {
        const cpvs_batch_item& a = items[order[j]];
        if (j == 0 || a.depth != items[order[j - 1]].depth || a.N != items[order[j - 1]].N)
            groups.push_back(Group{j, j, 0});
        groups.back().end = j + 1;
        groups.back().work += (u64)a.N * a.N / a.zTileNum;
    }
    const u64 budget = (u64)std::max<uint32_t>(cfg.forest_mtexels, 1) << 20;
    std::vector<std::pair<uint32_t, uint32_t>> forests;  // groups [first, second)
    for (uint32_t g = 0; g < (uint32_t)groups.size();) {
        uint32_t e = g + 1;
        u64 w = groups[g].work, n = groups[g].end - groups[g].begin;
        while (e < groups.size() && w + groups[e].work <= budget && n + (groups[e].end - groups[e].begin) <= 65535) {
            w += groups[e].work;
            n += groups[e].end - groups[e].begin;
            ++e;
        }
        forests.push_back({g, e});
        g = e;
    }
    ForestOpts fo;
    fo.merge_mode = ctx->merge_mode;
    fo.leaf_cols = cfg.leaf_cols;
    fo.leaf_waves = cfg.leaf_waves;
    fo.graph = cfg.graph != 0;

    std::atomic<uint32_t> next{0};
    std::atomic<bool> failed{false};
    std::mutex emu;
    int first_rc = CPVS_OK;
    std::string first_msg;
    const auto t0 = std::chrono::steady_clock::now();
    auto record = [&](int lrc, const char* what) {
        std::lock_guard<std::mutex> lk(emu);
        if (first_rc == CPVS_OK) {
            first_rc = lrc;
            first_msg = what;
        }
        failed.store(true);
    };
    auto stamp = [&](uint32_t i) {
        if (done_ms) done_ms[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    // Builds fin[0 .. m) (items idx[0 .. m)) as one forest, halving it while it is too large for
    // 32-bit offsets or its scratch cannot be allocated (smaller forests need less).
    std::function<int(Worker*, std::vector<ForestIn>&, std::vector<uint32_t>&, uint32_t, uint32_t)> build_range =
        [&](Worker* w, std::vector<ForestIn>& fin, std::vector<uint32_t>& idx, uint32_t a, uint32_t b) -> int {
        std::vector<cpvs_dag*> dags(b - a, nullptr);
        int lrc = build_forest(ctx, *w, fin.data() + a, b - a, fo, dags.data());
        if (lrc == kForestTooLarge || lrc == CPVS_ENOMEM) {
            if (b - a == 1) {  // one DAG alone: the single build
                cpvs_dag* d = nullptr;
                lrc = build_dag(ctx, w->arena, w->pinned, fo.merge_mode, fin[a].p, fin[a].t, fin[a].T, w->stream, &d);
                if (lrc == CPVS_OK) {
                    out[idx[a]] = d;
                    stamp(idx[a]);
                }
                return lrc;
            }
            const uint32_t mid = a + (b - a) / 2;
            lrc = build_range(w, fin, idx, a, mid);
            return lrc ? lrc : build_range(w, fin, idx, mid, b);
        }
        if (lrc == CPVS_OK)
            for (uint32_t k = a; k < b; ++k) {
                out[idx[k]] = dags[k - a];
                stamp(idx[k]);
            }
        return lrc;
    };
    auto body = [&](Worker* w, bool first) {
        if (hipSetDevice(ctx->device) != hipSuccess) {
            record(CPVS_EHIP, "batch: hipSetDevice failed on a lane thread");
            return;
        }
        try {
            while (!failed.load(std::memory_order_relaxed)) {
                const uint32_t f = next.fetch_add(1, std::memory_order_relaxed);
                if (f >= forests.size()) break;
                const uint32_t g0 = forests[f].first, g1 = forests[f].second;
                size_t lv_bytes = 0;
                for (uint32_t g = g0; g < g1; ++g) lv_bytes += pyramid_levels_bytes(items[order[groups[g].begin]].N, false);
                int lrc = w->levels.reserve(lv_bytes, w->stream);
                std::vector<cpvs_pyramid> pyr(g1 - g0);
                std::vector<ForestIn> fin;
                std::vector<uint32_t> idx;
                for (uint32_t g = g0; g < g1 && lrc == CPVS_OK; ++g) {
                    cpvs_pyramid& p = pyr[g - g0];
                    const cpvs_batch_item& it = items[order[groups[g].begin]];
                    p.ctx = ctx;
                    p.N = it.N;
                    p.n = levels_of(it.N);
                    p.depth = const_cast<float*>(it.depth);
                    p.borrowed = true;
                    lrc = pyramid_build_levels(&p, false, w->stream, &w->levels, true);
                    for (uint32_t j = groups[g].begin; j < groups[g].end; ++j) {
                        fin.push_back(ForestIn{&p, items[order[j]].zTileIndex, items[order[j]].zTileNum});
                        idx.push_back(order[j]);
                    }
                }
                if (lrc == CPVS_OK) lrc = build_range(w, fin, idx, 0, (uint32_t)fin.size());
                if (lrc != CPVS_OK) {
                    char msg[600];
                    snprintf(msg, sizeof(msg), "batch forest %u: %s", f, cpvs_last_error());
                    record(lrc, msg);
                    break;
                }
            }
            // Maps without leafmask levels, one by one on the first lane.
            for (size_t k = 0; first && k < single.size() && !failed.load(); ++k) {
                const uint32_t i = single[k];
                cpvs_pyramid p;
                int lrc = build_pyramid(ctx, *w, items[i], &p);
                cpvs_dag* d = nullptr;
                if (lrc == CPVS_OK)
                    lrc = build_dag(ctx, w->arena, w->pinned, fo.merge_mode, &p, items[i].zTileIndex, items[i].zTileNum,
                                    w->stream, &d);
                if (lrc != CPVS_OK) {
                    char msg[600];
                    snprintf(msg, sizeof(msg), "batch item %u: %s", i, cpvs_last_error());
                    record(lrc, msg);
                    break;
                }
                out[i] = d;
                stamp(i);
            }
        } catch (const std::bad_alloc&) {
            record(CPVS_ENOMEM, "batch: host allocation failed");
        } catch (...) {
            record(CPVS_EINTERNAL, "batch: internal exception");
        }
        (void)hipStreamSynchronize(w->stream);
    };
    std::vector<std::thread> threads;
    threads.reserve(lanes);
    try {
        for (uint32_t k = 1; k < lanes; ++k) threads.emplace_back(body, ctx->lanes[k], false);
    } catch (...) {
        failed.store(true);
        for (auto& t : threads) t.join();
        set_error("batch: could not start %u lane threads", lanes - 1);
        return CPVS_EINTERNAL;
    }
    body(ctx->lanes[0], true);
    for (auto& t : threads) t.join();
    if (first_rc != CPVS_OK) {
        for (uint32_t i = 0; i < count; ++i) {
            cpvs_dag_destroy(out[i]);
            out[i] = nullptr;
        }
        set_error("%s", first_msg.c_str());
        return first_rc;
    }
    return CPVS_OK;
}

// The configuration a batch runs with: the caller's, or the context's with defaults filled in.
cpvs_batch_config effective(const cpvs_ctx* ctx, const cpvs_batch_config* cfg, uint32_t lanes) {
    cpvs_batch_config c = cfg ? *cfg : ctx->batch;
    if (!cfg && !c.lanes && !c.forest_mtexels && !c.leaf_cols) {  // never set: the defaults
        c.forest_mtexels = kDefaultForestMtexels;
    }
    if (!c.lanes) c.lanes = kDefaultLanes;
    if (lanes) c.lanes = lanes;
    if (c.leaf_cols != 4) c.leaf_cols = 8;
    if (c.leaf_waves != 2) c.leaf_waves = 1;
    c.graph = c.graph ? 1u : 0u;
    return c;
}

int check_config(const cpvs_batch_config& c) {
    CHECK_ARG(c.lanes <= kMaxLanes, "batch config: %u lanes (at most %u)", c.lanes, kMaxLanes);
    CHECK_ARG(c.leaf_cols == 0 || c.leaf_cols == 4 || c.leaf_cols == 8, "batch config: leaf_cols %u (4 or 8)", c.leaf_cols);
    CHECK_ARG(c.leaf_waves <= 2, "batch config: leaf_waves %u (1 or 2)", c.leaf_waves);
    CHECK_ARG(c.graph <= 1, "batch config: graph %u (0 or 1)", c.graph);
    return CPVS_OK;
}

// Forests build leafmask DAGs only: without leafmasks the batch runs one map per launch chain.
int run_config(cpvs_ctx* ctx, const cpvs_batch_item* items, uint32_t count, const cpvs_batch_config& c, cpvs_dag** out,
               double* done_ms) {
    return (c.forest_mtexels && ctx->leafmasks.load()) ? run_forests(ctx, items, count, c, out, done_ms)
                                                       : run_batch(ctx, items, count, c.lanes, out, done_ms);
}

}  // namespace
}  // namespace cpvs

using namespace cpvs;

struct cpvs_dagset {
    cpvs_ctx* ctx = nullptr;
    uint32_t count = 0;
    DagDesc* d_desc = nullptr;
};

extern "C" {

int cpvs_build_batch(cpvs_ctx* ctx, const cpvs_batch_item* items, uint32_t count, uint32_t lanes, cpvs_dag** out,
                     double* done_ms) {
    CHECK_ARG(ctx && (items || !count) && (out || !count), "build_batch: NULL argument");
    CHECK_ARG(lanes <= kMaxLanes, "build_batch: %u lanes (at most %u)", lanes, kMaxLanes);
    for (uint32_t i = 0; i < count; ++i) {
        const int rc = check_item(items[i], i);
        if (rc) return rc;
    }
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    // Work already queued on the context stream (e.g. the depth generator) comes first.
    CPVS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return run_config(ctx, items, count, effective(ctx, nullptr, lanes), out, done_ms);
    GUARD_END
}

int cpvs_build_batch_config(cpvs_ctx* ctx, const cpvs_batch_item* items, uint32_t count, const cpvs_batch_config* cfg,
                            cpvs_dag** out, double* done_ms) {
    CHECK_ARG(ctx && (items || !count) && (out || !count), "build_batch_config: NULL argument");
    if (cfg) {
        const int rc = check_config(*cfg);
        if (rc) return rc;
    }
    for (uint32_t i = 0; i < count; ++i) {
        const int rc = check_item(items[i], i);
        if (rc) return rc;
    }
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    CPVS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return run_config(ctx, items, count, effective(ctx, cfg, 0), out, done_ms);
    GUARD_END
}

int cpvs_ctx_set_batch_config(cpvs_ctx* ctx, const cpvs_batch_config* cfg) {
    CHECK_ARG(ctx && cfg, "ctx_set_batch_config: NULL argument");
    const int rc = check_config(*cfg);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->batch = effective(ctx, cfg, 0);
    return CPVS_OK;
}

int cpvs_ctx_batch_config(cpvs_ctx* ctx, cpvs_batch_config* cfg) {
    CHECK_ARG(ctx && cfg, "ctx_batch_config: NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    *cfg = effective(ctx, nullptr, 0);
    return CPVS_OK;
}

int cpvs_ctx_set_batch_lanes(cpvs_ctx* ctx, uint32_t lanes) {
    CHECK_ARG(ctx, "ctx_set_batch_lanes: NULL ctx");
    CHECK_ARG(lanes <= kMaxLanes, "ctx_set_batch_lanes: %u lanes (at most %u)", lanes, kMaxLanes);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->batch = effective(ctx, nullptr, lanes);
    return CPVS_OK;
}

int cpvs_ctx_release_scratch(cpvs_ctx* ctx, uint64_t* freed) {
    CHECK_ARG(ctx, "ctx_release_scratch: NULL ctx");
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);  // builds and batches hold it while they run
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    CPVS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    uint64_t n = ctx->arena.cap;
    ctx->arena.release();
    for (Worker* w : ctx->lanes) {
        CPVS_HIP_CHECK(hipStreamSynchronize(w->stream));
        n += w->arena.cap + w->levels.cap;
        w->arena.release();
        w->levels.release();
    }
    n += ctx->buffers.trim(0);
    if (freed) *freed = n;
    return CPVS_OK;
    GUARD_END
}

int cpvs_ctx_device_name(cpvs_ctx* ctx, char* buf, uint64_t cap) {
    CHECK_ARG(ctx && buf && cap, "ctx_device_name: NULL argument");
    hipDeviceProp_t prop;
    CPVS_HIP_CHECK(hipGetDeviceProperties(&prop, ctx->device));
    snprintf(buf, cap, "%s %s %dCU", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return CPVS_OK;
}

int cpvs_ctx_batch_lanes(cpvs_ctx* ctx, uint32_t* lanes) {
    CHECK_ARG(ctx && lanes, "ctx_batch_lanes: NULL argument");
    *lanes = effective(ctx, nullptr, 0).lanes;
    return CPVS_OK;
}

}  // extern "C"

namespace {
// Times each candidate configuration on the caller's batch (one warm-up run each, which sizes the
// lanes' arenas, then the best of reps) and stores the fastest as the context's.  ctx->mu held.
int tune_configs(cpvs_ctx* ctx, const cpvs_batch_item* items, uint32_t count, const std::vector<cpvs_batch_config>& cands,
                 uint32_t reps, uint32_t* best, double* ms) {
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    CPVS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::vector<cpvs_dag*> out(count, nullptr);
    auto destroy_all = [&]() {
        for (auto& d : out) {
            cpvs_dag_destroy(d);
            d = nullptr;
        }
    };
    double best_ms = 0.0;
    uint32_t best_k = 0;
    for (uint32_t k = 0; k < (uint32_t)cands.size(); ++k) {
        int rc = run_config(ctx, items, count, cands[k], out.data(), nullptr);  // warm-up
        destroy_all();
        if (rc) return rc;
        double m = 0.0;
        for (uint32_t r = 0; r < reps; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            rc = run_config(ctx, items, count, cands[k], out.data(), nullptr);
            const double e = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            destroy_all();
            if (rc) return rc;
            m = (r == 0) ? e : std::min(m, e);
        }
        if (ms) ms[k] = m;
        if (k == 0 || m < best_ms) {
            best_ms = m;
            best_k = k;
        }
    }
    ctx->batch = cands[best_k];
    *best = best_k;
    return CPVS_OK;
}
}  // namespace

extern "C" {

int cpvs_tune_batch(cpvs_ctx* ctx, const cpvs_batch_item* items, uint32_t count, const uint32_t* candidates,
                    uint32_t ncand, uint32_t reps, uint32_t* best, double* ms) {
    CHECK_ARG(ctx && items && count && candidates && ncand && best, "tune_batch: NULL or empty argument");
    CHECK_ARG(reps >= 1 && reps <= 100, "tune_batch: reps %u must be in [1, 100]", reps);
    for (uint32_t k = 0; k < ncand; ++k)
        CHECK_ARG(candidates[k] >= 1 && candidates[k] <= kMaxLanes, "tune_batch: candidate %u lanes", candidates[k]);
    for (uint32_t i = 0; i < count; ++i) {
        const int rc = check_item(items[i], i);
        if (rc) return rc;
    }
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::vector<cpvs_batch_config> cands(ncand);
    for (uint32_t k = 0; k < ncand; ++k) cands[k] = effective(ctx, nullptr, candidates[k]);
    uint32_t bk = 0;
    const int rc = tune_configs(ctx, items, count, cands, reps, &bk, ms);
    if (rc) return rc;
    *best = candidates[bk];
    return CPVS_OK;
    GUARD_END
}

int cpvs_tune_batch_config(cpvs_ctx* ctx, const cpvs_batch_item* items, uint32_t count, const cpvs_batch_config* cands,
                           uint32_t ncand, uint32_t reps, uint32_t* best, double* ms) {
    CHECK_ARG(ctx && items && count && cands && ncand && best, "tune_batch_config: NULL or empty argument");
    CHECK_ARG(reps >= 1 && reps <= 100, "tune_batch_config: reps %u must be in [1, 100]", reps);
    for (uint32_t k = 0; k < ncand; ++k) {
        const int rc = check_config(cands[k]);
        if (rc) return rc;
    }
    for (uint32_t i = 0; i < count; ++i) {
        const int rc = check_item(items[i], i);
        if (rc) return rc;
    }
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::vector<cpvs_batch_config> c(ncand);
    for (uint32_t k = 0; k < ncand; ++k) c[k] = effective(ctx, &cands[k], 0);
    return tune_configs(ctx, items, count, c, reps, best, ms);
    GUARD_END
}

int cpvs_dagset_create(cpvs_ctx* ctx, const cpvs_dag* const* dags, uint32_t count, cpvs_dagset** out) {
    CHECK_ARG(ctx && dags && count && out, "dagset_create: NULL or empty argument");
    std::vector<DagDesc> h(count);
    for (uint32_t i = 0; i < count; ++i) {
        const cpvs_dag* d = dags[i];
        CHECK_ARG(d, "dagset_create: DAG %u is NULL", i);
        CHECK_ARG(d->ctx == ctx, "dagset_create: DAG %u belongs to another context", i);
        // One lookup for the whole set, as one shader (traverse.cs:8-9 compiles LEAFMASKS in or out).
        CHECK_ARG(d->leafmasks == dags[0]->leafmasks, "dagset_create: DAG %u %s leafmasks, DAG 0 %s", i,
                  d->leafmasks ? "has" : "has no", dags[0]->leafmasks ? "has" : "has none");
        CHECK_ARG(d->words < 0xFFFFFFFEull, "dagset_create: DAG %u exceeds 32-bit offsets", i);
        h[i] = dag_desc(d);
    }
    GUARD_BEGIN
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    cpvs_dagset* s = new cpvs_dagset();
    s->ctx = ctx;
    s->count = count;
    if (hipMalloc(&s->d_desc, sizeof(DagDesc) * count) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        set_error("dagset_create: hipMalloc failed");
        return CPVS_ENOMEM;
    }
    const hipError_t e = hipMemcpy(s->d_desc, h.data(), sizeof(DagDesc) * count, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(s->d_desc);
        delete s;
        return hip_fail(e, "dagset_create copy", __FILE__, __LINE__);
    }
    *out = s;
    return CPVS_OK;
    GUARD_END
}

int cpvs_dagset_evaluate(cpvs_dagset* s, const float* ndc, uint32_t stride, uint64_t count, const uint32_t* ids,
                         uint8_t* vis, void* stream) {
    CHECK_ARG(s && ((ndc && vis) || !count), "dagset_evaluate: NULL argument");
    CHECK_ARG(stride >= 3, "dagset_evaluate: stride %u < 3", stride);
    CPVS_HIP_CHECK(hipSetDevice(s->ctx->device));
    return launch_dagset_eval(s->d_desc, s->count, DagDesc{}, ids, ndc, stride, count, vis, pick_stream(s->ctx, stream));
}

int cpvs_dagset_destroy(cpvs_dagset* s) {
    if (!s) return CPVS_OK;
    hipSetDevice(s->ctx->device);
    if (s->d_desc) hipFree(s->d_desc);
    delete s;
    return CPVS_OK;
}

int cpvs_dag_evaluate(const cpvs_dag* d, const float* ndc, uint32_t stride, uint64_t count, uint8_t* vis, void* stream) {
    CHECK_ARG(d && ((ndc && vis) || !count), "dag_evaluate: NULL argument");
    CHECK_ARG(stride >= 3, "dag_evaluate: stride %u < 3", stride);
    CHECK_ARG(d->words < 0xFFFFFFFEull, "dag_evaluate: DAG exceeds 32-bit offsets");
    CPVS_HIP_CHECK(hipSetDevice(d->ctx->device));
    return launch_dagset_eval(nullptr, 0, dag_desc(d), nullptr, ndc, stride, count, vis, pick_stream(d->ctx, stream));
}

}  // extern "C"
