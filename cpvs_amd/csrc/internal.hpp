// internal.hpp -- host-side objects behind the opaque C-ABI handles and the kernel launchers
// each translation unit provides.
#pragma once

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cpvs.h"
#include "common.hpp"

namespace cpvs {

constexpr int kMaxLevels = 24;  // N <= 2^21 texels per side
constexpr uint32_t kMaxFilterSize = 64;  // PCF kernel width limit (64 x 64 lookups per point)
// A container's lookup top table: at most this many levels, and (chosen automatically) at most
// kTopTableBytes, one XCD's L2, so that lookups in any order find it cached; an explicit request
// (coherent queries gain from more levels) may take up to kTopTableMaxBytes.
constexpr int kMaxTopLevels = 9;
constexpr uint64_t kTopTableBytes = 4ull << 20;
constexpr uint64_t kTopTableMaxBytes = 1ull << 30;

void set_error(const char* fmt, ...);

// Bump allocator over one device buffer, reset per build.  Grows (after a stream sync) when a
// build needs more than it holds.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    bool fixed = false;  // a caller's buffer of `cap` bytes: never grown (reserve fails beyond it)
    // Ensure capacity and reset `used`; *moved (optional) tells whether the buffer was reallocated
    // (its contents are then lost).
    int reserve(size_t bytes, hipStream_t s, bool* moved = nullptr);
    template <class T>
    T* take(size_t count) {
        const size_t bytes = (count * sizeof(T) + 255) & ~(size_t)255;
        char* p = base + used;
        used += bytes;
        return reinterpret_cast<T*>(p);
    }
    void release();
};

inline size_t arena_bytes(size_t count, size_t elem) { return (count * elem + 255) & ~(size_t)255; }

// Pinned host staging per stream lane: kPinnedWords u64 words of readback staging, then
// kPinnedPartials u32 node-count partials (one per count workgroup, written by the kernel).
constexpr size_t kPinnedWords = 256;
constexpr size_t kPinnedPartials = 65536;
constexpr size_t kPinnedBytes = sizeof(u64) * kPinnedWords + sizeof(u32) * kPinnedPartials;
inline u32* pinned_partials(u64* pinned) { return reinterpret_cast<u32*>(pinned + kPinnedWords); }
// Kernels store node counts and meta words straight into the staging, and the host reads them
// after a stream synchronize.  Coherent host memory keeps those stores out of the GPU's caches
// (the default, non-coherent kind may hold them in L2 past the kernel's end).
constexpr unsigned kPinnedFlags = hipHostMallocCoherent | hipHostMallocMapped;

// One build lane of a batch: a stream with its own scratch arena, pyramid-level arena and
// pinned readback staging.  Lanes run concurrently from their own host threads.
struct Worker {
    hipStream_t stream = nullptr;
    Arena arena;   // build scratch
    Arena levels;  // pyramid levels of the item being built (of all maps of a forest)
    u64* pinned = nullptr;
    void* fpin = nullptr;  // a forest's pinned staging (count partials, tables, readback), grown as needed
    size_t fpin_bytes = 0;
    int create();  // on the current device
    void destroy();
    int forest_staging(size_t bytes);  // ensures fpin holds bytes (waits for the stream first)
};

// forest.hip: many DAGs built as one forest with one launch chain (see there).
struct ForestIn {
    const cpvs_pyramid* p;  // levels >= 3 stored, resident until the build returns
    u32 t, T;               // zTileIndex, zTileNum
};
struct ForestOpts {
    int merge_mode = CPVS_MERGE_TABLE;
    u32 leaf_cols = 8;    // leaf kernel: level-3 columns per wave (4 or 8)
    u32 leaf_waves = 1;   // leaf kernel: waves per workgroup (1 or 2)
    bool graph = false;   // capture the level chain between the two readbacks in a hipGraph
};
constexpr int kForestTooLarge = -1002;  // a level's words or nodes exceed 32 bits: split the forest
// Builds the DAGs of items[0 .. M) on w's stream; out[i] receives item i's DAG.  Pyramids must
// be complete on w.stream's timeline.  Returns kForestTooLarge when the forest must be split.
int build_forest(cpvs_ctx* ctx, Worker& w, const ForestIn* items, u32 M, const ForestOpts& o, cpvs_dag** out);

}  // namespace cpvs

namespace cpvs {
// Freed device buffers of DAG words and pyramid levels kept for reuse: a hipMalloc of a few MB costs
// about 14 us of host time, during which a build's GPU work waits.  Best fit within twice the
// request; at most kMaxBytes are kept (beyond that a buffer is released with hipFree, which waits
// for the device).
// A freed buffer may still be read by work queued on any stream, so give() parks it as pending
// and the free itself waits for nothing.  A pending buffer is handed out again only after a
// device synchronisation that follows its free: take() performs one when no ready buffer fits and
// a pending one would, which readies every pending buffer at once.  give_idle() is for buffers
// the caller knows no queued work reads (a server's evicted DAG: no lookup pins it, its build has
// completed); they are ready at once.
struct BufferCache {
    static constexpr size_t kMaxBytes = size_t(2) << 30;
    std::mutex mu;
    std::multimap<size_t, void*> free_;     // size -> ready buffer
    std::multimap<size_t, void*> pending_;  // size -> freed buffer not yet known idle
    std::unordered_map<void*, size_t> live_;  // handed-out buffer -> its size
    size_t cached = 0;   // bytes in free_ and pending_
    uint64_t device_syncs = 0, trims = 0;
    int device = 0;
    void* take(size_t bytes);            // nullptr when hipMalloc fails (even after emptying the cache)
    // A cached buffer of at most min(2 * bytes, limit) bytes, or nullptr; never allocates.
    void* take_cached(size_t bytes, size_t limit = ~size_t(0));
    void give(void* p);                  // p from take(); null is ignored
    void give_idle(void* p);             // p from take(), with no queued work reading it
    size_t size_of(void* p);             // bytes of a handed-out buffer (0 if unknown)
    size_t cached_bytes();
    size_t trim(size_t keep);            // hipFree cached buffers until at most `keep` bytes stay
    void release();                      // hipFree of every cached buffer
private:
    void park(void* p, bool idle);
};
}  // namespace cpvs

struct cpvs_ctx {
    int device = 0;
    cpvs_batch_config batch{};  // how batches run (cpvs_ctx_set_batch_config); lanes 0 = default
    cpvs::BufferCache buffers;
    int num_cus = 256;  // compute units of the device (persistent grids size themselves to it)
    hipStream_t stream = nullptr;
    cpvs::Arena arena;
    cpvs::u64* pinned = nullptr;  // host readback staging (pinned)
    size_t pinned_words = 0;
    std::mutex mu;
    bool prof_on = false;
    cpvs::Prof prof;
    int merge_mode = CPVS_MERGE_TABLE;
    // The reference's LEAFMASKS switch (src/CompressedShadow.cpp:17, shader/traverse.cs:8-9), on by
    // default: builds of maps of numLevels >= 5 end in 17-word leafmask nodes at level 2; off, they
    // go down to level 0 (cpvs_ctx_set_leafmasks).
    std::atomic<int> leafmasks{1};
    // Merge hashes of the table path cut to this many low bits (0: full 64-bit hashes), at the leaf
    // level and at the levels above it, so that distinct keys collide and the full-key compares are
    // exercised (cpvs_ctx_set_hash_bits).
    std::atomic<uint32_t> leaf_hash_bits{0}, inner_hash_bits{0};
    std::vector<cpvs::Worker*> lanes;  // batch build lanes, created on first use and kept
};

struct cpvs_pyramid {
    cpvs_ctx* ctx = nullptr;
    uint32_t N = 0;
    int n = 0;                 // numLevels including the root image
    float* depth = nullptr;    // N*N floats (owned unless borrowed)
    bool borrowed = false;
    bool levels_owned = true;  // false when the levels live in a batch lane's arena
    bool pending = false;      // the levels may still be in flight on the context's stream
    float2* levels = nullptr;  // one allocation holding every stored level
    float2* lv[cpvs::kMaxLevels] = {};  // per-level pointer (null when not stored)
};

struct cpvs_dag {
    cpvs_ctx* ctx = nullptr;
    int n = 0;                    // numLevels
    uint64_t words = 0;           // word count
    uint32_t* d_words = nullptr;  // device words (owned)
    uint32_t root_mask = 0;
    uint32_t N = 0, zTileIndex = 0, zTileNum = 1;
    uint64_t nodes[cpvs::kMaxLevels] = {};    // nodes per level before merging
    uint64_t uniques[cpvs::kMaxLevels] = {};  // nodes per level after merging
    int merge_mode = CPVS_MERGE_TABLE;        // merge path that produced the words
    uint32_t key_collisions = 0;              // 32-bit sort-key collisions resolved by scanning
    bool leafmasks = false;                   // level 2 holds leafmask nodes (lookups stop there)
    uint64_t* d_top = nullptr;                // lookup top table (cpvs_dag_set_top_levels), or null
    uint32_t top_levels = 0;
    std::vector<uint64_t*> retired_tops;      // replaced tables, which DAG sets made earlier still name
};

struct cpvs_container {
    cpvs_ctx* ctx = nullptr;
    uint32_t length = 1;
    std::vector<cpvs_dag*> dags;  // length^3 slots, index z*L^2 + y*L + x
    uint32_t filter_size = 1;
    uint32_t* d_dag = nullptr;   // combined words
    uint64_t dag_words = 0;
    uint32_t* d_grid = nullptr;  // length^3 entries
    int dag_levels = 0, grid_levels = 0;
    bool leafmasks = true;       // the DAGs' mode (all alike)
    uint64_t* d_top = nullptr;   // the lookups' top table (launch_build_top), or null
    uint32_t top_levels = 0;     // its levels (0: none)
    int top_request = -1;        // cpvs_container_set_top_levels: -1 automatic
    bool on_gpu = false;
    bool freed_cpu = false;
};

namespace cpvs {

// pyramid.hip.  Levels go to a new allocation owned by p, or, when `into` is given, to that
// arena (p then does not own them; the arena is reset, or with append the levels follow what it
// holds, within a reservation the caller made).
int pyramid_build_levels(cpvs_pyramid* p, bool store_all, hipStream_t s, Arena* into = nullptr, bool append = false);
// Exact node counts per level for z-tile t of T.  Each count workgroup writes its partial count
// to partials[block] (pinned host memory, kPinnedPartials entries; may be null), and
// sum_node_counts adds them up per level once the stream has reached that point: cnt[L] = nodes
// at octree level L.  partials[block_start[nlev]] receives the root's childmask.  col_info (may be null) receives, per level-3 texel (leafmask mode), its
// partial-cell count (.x) and first partial z cell within the z-tile (.z).
// Before work on stream s reads p's levels: a pyramid still in flight on the context's stream is
// waited for when s is another stream.
inline hipError_t pyramid_wait(const cpvs_pyramid* p, hipStream_t s) {
    return (p->pending && s != p->ctx->stream) ? hipStreamSynchronize(p->ctx->stream) : hipSuccess;
}

struct CountLayout {
    int nlev = 0;
    int level[kMaxLevels] = {};            // pyramid level of each block range
    u32 block_start[kMaxLevels + 1] = {};  // first block of each range
};
// leaf: the octree ends in leafmask nodes at level 2 (counts from level 3 up) rather than at level 0
// (counts from level 1 up, which needs pyramid levels 1 and 2).
// exact: count only the cells under a partial cell at every level above (slower; needed only
// for maps with NaN depths, whose plain counts can exceed the octree's nodes).
int pyramid_node_counts(const cpvs_pyramid* p, u32 t, u32 T, bool leaf, u32* partials, uint4* col_info, hipStream_t s,
                        CountLayout* layout = nullptr, bool exact = false);
void sum_node_counts(const CountLayout& lay, const u32* partials, u64* cnt);

// Admission hook of a build: called once the exact node counts are known and before any scratch
// or DAG memory is taken, with the scratch bytes the build needs (its arena grows only if it
// holds less) and an upper bound of the DAG bytes.  May block; a non-zero return aborts the
// build with that status.  Called again (replacing the first request) if the build retries.
struct Admit {
    int (*fn)(void* user, const Arena& arena, u64 scratch_bytes, u64 dag_bytes_max);
    void* user;
    // Allocates the DAG's words (within the admitted bound) when set; the context's buffer cache
    // otherwise.  nullptr = out of memory.
    void* (*alloc)(void* user, size_t bytes) = nullptr;
};

// Destroys a DAG whose words no queued work reads: its buffer is ready for reuse at once.
void dag_destroy_idle(cpvs_dag* d);
// Bytes of the stored pyramid levels of an N x N map (levels >= 3, or all with store_all).
size_t pyramid_levels_bytes(u32 N, bool store_all);

// A build's words written into a caller's device buffer instead of a new cpvs_dag.
struct DagOut {
    u32* words = nullptr;  // device, cap words
    u64 cap = 0;
    u64 nwords = 0;        // out: the DAG's word count
    u32 root_mask = 0;     // out
};

// build.hip: one DAG from a pyramid on stream s, scratch from A, readbacks staged in pinned
// (kPinnedWords words of pinned host memory).  With `into`, the words go to into->words (an
// error if they exceed into->cap) and *out is not touched.
int build_dag(cpvs_ctx* ctx, Arena& A, u64* pinned, int merge_mode, const cpvs_pyramid* p, u32 t, u32 T, hipStream_t s,
              cpvs_dag** out, const Admit* admit = nullptr, DagOut* into = nullptr);

// noalloc.hip: finds this device's fixed scratch (once per device); called by cpvs_ctx_create.
int noalloc_setup(int device);

// lookup.hip
struct Mat4 {
    float m[16];
};
// leaf: the DAGs end in leafmask nodes at level 2; 0: they go down to level 0.  top (may be null):
// the container's top table of topk levels (launch_build_top).
int launch_evaluate(const u32* dag, u32 words, const u32* grid, int dl, int gl, int leaf, const u64* top, u32 topk,
                    const float* pos, u32 stride, u64 count, const Mat4& m, uint8_t* vis, hipStream_t s);
int launch_evaluate_pcf(const u32* dag, u32 words, const u32* grid, int dl, int gl, int leaf, const u64* top, u32 topk,
                        const float* pos, u32 stride, u64 count, const Mat4& m, u32 f, float* vis, hipStream_t s);
int launch_evaluate_image(const u32* dag, u32 words, const u32* grid, int dl, int gl, int leaf, const u64* top, u32 topk,
                          const float* pos, u32 width, u32 height, const Mat4& m, uint8_t* vis, hipStream_t s);
// The top table of a container: cells x 8^K entries (the walk's state after K levels per cell);
// grid null: one DAG of its own (cells 1, base 0).
int launch_build_top(const u32* dag, u32 words, const u32* grid, u32 cells, int dl, u32 K, u64* top, hipStream_t s);
// One DAG of a set, walked as a one-cell container.  base is 0, or the grid sentinel
// 0xFFFFFFFF / 0xFFFFFFFE when the root is all shadow / all lit.
struct DagDesc {
    const u32* words;
    u32 nwords;
    u32 base;
    u32 dl;    // numLevels
    u32 leaf;  // 1: leafmask nodes at level 2; 0: the DAG goes down to level 0
    const u64* top;  // the DAG's lookup top table (launch_build_top, one cell), or null
    u32 topk;        // its levels
};
// The descriptor of a built DAG (createTopLevelGrid sentinels for a one-cell grid,
// src/CompressedShadowContainer.cpp:71-91).
inline DagDesc dag_desc(const cpvs_dag* d) {
    DagDesc r{};
    r.words = d->d_words;
    r.nwords = (u32)d->words;
    r.base = (d->root_mask == 0u) ? 0xFFFFFFFFu : (d->root_mask == 0x5555u) ? 0xFFFFFFFEu : 0u;
    r.dl = (u32)d->n;
    r.leaf = d->leafmasks ? 1u : 0u;
    r.top = d->d_top;
    r.topk = d->d_top ? d->top_levels : 0u;
    return r;
}
constexpr uint8_t kBadId = 0xFF;  // visibility written for a query whose DAG id is outside the set
int launch_dagset_eval(const DagDesc* set, u32 nset, const DagDesc& one, const u32* ids, const float* ndc, u32 stride,
                       u64 count, uint8_t* vis, hipStream_t s);
// Many lookup requests in one launch: request r's queries [0, count) use DAG d, read NDC from
// ndc (stride floats apart) and write vis.  Requests take whole 256-query blocks, the first at
// block0 (ascending).
struct LookupSeg {
    DagDesc d;
    const float* ndc;
    uint8_t* vis;
    u32 stride;
    u32 count;
    u32 block0;
    u32 pad;
};
int launch_segments_eval(const LookupSeg* segs, u32 nseg, u32 blocks, hipStream_t s);
int launch_traverse_single(const u32* dag, u32 words, int n, const float* ndc, u64 count, int try_leaf, int* out,
                           hipStream_t s);

// gen.hip
int launch_gen_depth(float* out, u32 fullN, u64 seed, u32 nboxes, int noise, u32 x0, u32 y0, u32 tileN, hipStream_t s);
int launch_gen_queries_uniform(float* out, u64 seed, u64 first, u64 step, u64 count, u32 stride, hipStream_t s);
int launch_device_copy(void* dst, const void* src, u64 bytes, hipStream_t s);
int launch_gen_queries_gbuffer(float* out, const float* depth, u32 N, u64 seed, u32 W, u32 H, hipStream_t s);

}  // namespace cpvs
