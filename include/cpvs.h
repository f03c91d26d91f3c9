/*
 * cpvs.h -- C ABI of cpvs_amd, an MI355X (gfx950) implementation of the precomputed voxelized
 * shadow build and lookup path of TobiasRp/cpvs.
 *
 * Every entry point returns a status code (CPVS_OK on success); no exception crosses the ABI.
 * cpvs_last_error() gives the calling thread's message for the last failure.  Handles are
 * opaque; one handle may be used from one thread at a time.  Streams are hipStream_t values
 * passed as void*; NULL selects the context's own stream.  Device pointers are HIP device
 * memory of the context's device.
 *
 * Each entry point names the reference interface it replaces (file:line in TobiasRp/cpvs).
 */
#ifndef CPVS_AMD_CPVS_H
#define CPVS_AMD_CPVS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPVS_OK 0
#define CPVS_EINVAL 1    /* bad argument (size not a power of two, index out of range, ...) */
#define CPVS_ENOMEM 2    /* device or host allocation failed */
#define CPVS_EHIP 3      /* a HIP runtime call failed */
#define CPVS_EINTERNAL 4 /* an internal consistency check failed */
#define CPVS_ESTATE 5    /* call not valid in the object's current state */
#define CPVS_EBUDGET 6   /* admission refused: estimate exceeds the memory budget */

/* CompressedShadow::NodeVisibility (src/CompressedShadow.h:22-26). */
#define CPVS_SHADOW 0
#define CPVS_VISIBLE 1
#define CPVS_PARTIAL 2

/* Subtree merge paths (cpvs_ctx_set_merge); both give the same words. */
#define CPVS_MERGE_TABLE 0 /* 64-bit key hash table, first occurrence by atomic min (default) */
#define CPVS_MERGE_SORT 1  /* LSD radix sort of (32-bit key hash, node index) */

/* cpvs_pyramid_build flags. */
#define CPVS_PYR_ALL_LEVELS 1u   /* also store levels 1 and 2 (read only by builds without leafmasks) */
#define CPVS_PYR_BORROW_DEPTH 2u /* depth is device memory the caller keeps alive; not copied */

typedef struct cpvs_ctx cpvs_ctx;
typedef struct cpvs_pyramid cpvs_pyramid;
typedef struct cpvs_dag cpvs_dag;
typedef struct cpvs_container cpvs_container;

/* Message of the calling thread's last failed call ("" if none). */
const char *cpvs_last_error(void);

/* Context: one HIP device, its own non-blocking stream and a scratch arena for builds.
 * (The reference has no counterpart: it runs on the host and the GL context.) */
int cpvs_ctx_create(int device, cpvs_ctx **out);
int cpvs_ctx_destroy(cpvs_ctx *ctx);
int cpvs_ctx_synchronize(cpvs_ctx *ctx);
int cpvs_ctx_stream(cpvs_ctx *ctx, void **stream);
/* Per-kernel timing: mode 1 brackets every kernel launched by this context's entry points with
 * HIP events on its stream, 0 stops, -1 clears the totals.  profile_read writes one line per
 * kernel name: "name<TAB>total_ms<TAB>launches".  (No reference counterpart: the reference
 * times the whole precompute with std::chrono, src/main.cpp:228-234.) */
int cpvs_ctx_profile(cpvs_ctx *ctx, int mode);
/* Selects how builds merge common subtrees, the replacement of mergeLevel's linear search
 * (src/CompressedShadowUtil.h:154-182): CPVS_MERGE_TABLE (default) or CPVS_MERGE_SORT.  Two
 * different keys that the table path cannot tell apart (equal hash tags in one probe chain) are
 * detected by the full-key compare, and the build is redone with the sort path. */
int cpvs_ctx_set_merge(cpvs_ctx *ctx, int mode);
/* Test hook with no counterpart in the reference (whose linear search compares whole keys): the
 * table path's 64-bit key hashes cut to their low bits (1 .. 31; 0, the default, keeps them
 * whole), leaf_bits at the leafmask level and inner_bits above it, so that distinct keys share
 * hashes.  The full-key compares must then detect every such collision and send the build to the
 * sort path, with the same words. */
int cpvs_ctx_set_hash_bits(cpvs_ctx *ctx, uint32_t leaf_bits, uint32_t inner_bits);
/* The reference's compile-time LEAFMASKS switch (#define LEAFMASKS, src/CompressedShadow.cpp:17;
 * shader/traverse.cs:8-9, MIN_LEVEL 34-38), as a run-time mode of the context, on (1) by default.
 * Off (0), builds of maps of numLevels >= 5 go down to level 0 with 1+k-word nodes (useLeafmasks
 * false, src/CompressedShadow.cpp:20-41), pyramids made afterwards keep levels 1 and 2, and
 * forests are not used (batches build map by map).  Every DAG records the mode it was built in;
 * lookups, traversals and the file format follow the DAG's mode, and a container or DAG set takes
 * DAGs of one mode only.  Set it before builds start: a build racing a change may fail with an
 * error. */
int cpvs_ctx_set_leafmasks(cpvs_ctx *ctx, int on);
int cpvs_ctx_leafmasks(const cpvs_ctx *ctx, int *on);
int cpvs_ctx_profile_read(cpvs_ctx *ctx, char *buf, uint64_t cap);

/* MinMaxHierarchy::MinMaxHierarchy(const ImageF&) (src/MinMaxHierarchy.h:23,
 * src/MinMaxHierarchy.cpp:9-97).  depth: N x N row-major float32, host memory or (with
 * depth_on_device) device memory.  N must be a power of two, 2 <= N <= 65536 (cpvs_build needs
 * N >= 8, src/CompressedShadow.cpp:44-47). */
int cpvs_pyramid_build(cpvs_ctx *ctx, const float *depth, int depth_on_device, uint32_t N, uint32_t flags,
                       void *stream, cpvs_pyramid **out);
/* MinMaxHierarchy::getNumLevels (src/MinMaxHierarchy.h:60-62). */
int cpvs_pyramid_num_levels(const cpvs_pyramid *p, int *out);
/* MinMaxHierarchy::getMin / getMax (src/MinMaxHierarchy.h:31-55); level 0 returns the depth
 * for both.  Synchronous device read. */
int cpvs_pyramid_get(const cpvs_pyramid *p, uint32_t level, uint32_t x, uint32_t y, float *min, float *max);
/* MinMaxHierarchy::getLevel (src/MinMaxHierarchy.h:67-72): level 0 as N*N floats, level L >= 1
 * as (N>>L)^2 interleaved (min, max) pairs.  Fails with CPVS_ESTATE for levels 1-2 unless the
 * pyramid was built with CPVS_PYR_ALL_LEVELS (or N < 16). */
int cpvs_pyramid_copy_level(const cpvs_pyramid *p, uint32_t level, float *host, uint64_t cap_floats);
int cpvs_pyramid_destroy(cpvs_pyramid *p);

/* Exact octree node counts per level before merging for z-tile zTileIndex of zTileNum
 * (nodes_per_level[L] for L < numLevels, may be NULL), the scratch bytes cpvs_build will use,
 * and an upper bound of the DAG's bytes.  Used for admission control; no reference
 * counterpart (the reference sizes vectors as it goes, src/CompressedShadow.cpp:87-169). */
int cpvs_estimate_build(cpvs_ctx *ctx, const cpvs_pyramid *p, uint32_t zTileIndex, uint32_t zTileNum,
                        uint64_t *nodes_per_level, uint64_t *scratch_bytes, uint64_t *max_dag_bytes);

/* CompressedShadow::create(const MinMaxHierarchy&, uint zTileIndex, uint zTileNum)
 * (src/CompressedShadow.h:48-49, src/CompressedShadow.cpp:49-59).  zTileNum is explicit
 * instead of the reference's process-global depthOffset (src/CompressedShadowUtil.cpp:10-14).
 * The DAG's words stay in device memory.  Returns after the build completed on the stream. */
int cpvs_build(cpvs_ctx *ctx, const cpvs_pyramid *p, uint32_t zTileIndex, uint32_t zTileNum, void *stream,
               cpvs_dag **out);

/* No counterpart in the reference: cpvs_build for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc or hipFree: its scratch is a fixed device array of cpvs_build_noalloc_scratch_bytes()
 * bytes per device (64 MiB, held from context creation until the process ends; enough for maps up
 * to 1024^2 and most of 2048^2), and the DAG's words go to d_words (device, cap_words words).
 * *num_words gets the word count, *root_mask (may be NULL) the root's childmask.  A map whose
 * scratch or words exceed those sizes fails with CPVS_EINVAL and a message naming both sizes.
 * The words equal cpvs_build's.  Calls on one device share its array: calls through one context
 * are serialised, calls through several contexts on one device race.  Returns after the build
 * completed on the stream. */
int cpvs_build_noalloc(cpvs_ctx *ctx, const cpvs_pyramid *p, uint32_t zTileIndex, uint32_t zTileNum, void *stream,
                       uint32_t *d_words, uint64_t cap_words, uint64_t *num_words, uint32_t *root_mask);
size_t cpvs_build_noalloc_scratch_bytes(void);
/* CompressedShadow::getNumLevels / getDAG().size() / getTotalVisibility
 * (src/CompressedShadow.h:69-77, src/CompressedShadow.cpp:66-72). */
int cpvs_dag_info(const cpvs_dag *d, uint32_t *num_levels, uint64_t *num_words, int *total_visibility);
/* Node counts per level before and after merging (cap entries each, may be NULL). */
int cpvs_dag_level_stats(const cpvs_dag *d, uint64_t *nodes, uint64_t *uniques, int cap);
/* Merge path that produced the DAG and the 32-bit sort-key collisions it resolved. */
int cpvs_dag_merge_info(const cpvs_dag *d, int *mode, uint64_t *key_collisions);
/* 1 when level 2 holds leafmask nodes (numLevels >= 5 built with leafmasks), else 0. */
int cpvs_dag_leafmasks(const cpvs_dag *d, int *on);
/* CompressedShadow::getDAG (src/CompressedShadow.h:75-77): copy the words to host memory. */
int cpvs_dag_copy_words(const cpvs_dag *d, uint32_t *host, uint64_t cap_words);
int cpvs_dag_device_words(const cpvs_dag *d, const uint32_t **device_words);
/* CompressedShadow::traverse(vec3 ndc, bool tryLeafmasks) (src/CompressedShadow.cpp:404-463)
 * for count NDC points (3 floats each, host memory), evaluated on the device; vis[i] gets
 * CPVS_SHADOW, CPVS_VISIBLE or CPVS_PARTIAL. */
int cpvs_dag_traverse(cpvs_dag *d, const float *ndc, uint64_t count, int try_leafmasks, int *vis);
int cpvs_dag_destroy(cpvs_dag *d);

/* CompressedShadowContainer(uint length) (src/CompressedShadowContainer.h:21-25): length^3
 * DAG slots.  The single-shadow constructor (:28-32) is create(length=1) + set(dag,0,0,0). */
int cpvs_container_create(cpvs_ctx *ctx, uint32_t length, cpvs_container **out);
/* CompressedShadowContainer::set (src/CompressedShadowContainer.h:34-39); takes ownership of
 * dag (a DAG already in the slot is destroyed). */
int cpvs_container_set(cpvs_container *c, cpvs_dag *dag, uint32_t x, uint32_t y, uint32_t z);
/* CompressedShadowContainer::get (src/CompressedShadowContainer.h:41-46); *out is NULL after
 * free_on_cpu, as in the reference. */
int cpvs_container_get(const cpvs_container *c, uint32_t x, uint32_t y, uint32_t z, const cpvs_dag **out);
/* setFilterSize (src/CompressedShadowContainer.h:71-73); stored, unused by the lookup, as in
 * the reference (shader/traverse.cs:16-17). */
int cpvs_container_set_filter_size(cpvs_container *c, uint32_t size);
/* copyToGPU (src/CompressedShadowContainer.cpp:31-46): combineDAGs (:52-69) and
 * createTopLevelGrid (:71-91) on the device.  Grid sentinels are 0xFFFFFFFF (shadow) and
 * 0xFFFFFFFE (visible), the values shader/traverse.cs:85-88 tests.  Every slot holds a DAG of the
 * same numLevels and leafmask mode (one traverse.cs walks them all). */
int cpvs_container_copy_to_gpu(cpvs_container *c);
/* No counterpart in the reference (its lookup walks every level, shader/traverse.cs:89-133).
 * copy_to_gpu also builds a top table: per grid cell, the walk's state after its first K levels,
 * which the lookups read in place of walking them (same results).  levels -1 (default): K
 * automatic (the most levels within 4 MiB, one XCD's L2); 0: no table; 1..9: K = min(levels, the DAGs' inner
 * levels), a table of at most 1 GiB.  Takes effect at the next copy_to_gpu.  top_levels reports
 * the K of the table built (0: none). */
int cpvs_container_set_top_levels(cpvs_container *c, int levels);
int cpvs_container_top_levels(const cpvs_container *c, int *levels);
/* Tunes the top table on the caller's lookups (no counterpart in the reference): for each
 * candidate levels value (as cpvs_container_set_top_levels takes), copy_to_gpu and the best of
 * reps timed cpvs_container_evaluate calls after one warm-up on the context stream.  Keeps the
 * fastest (set and built), returns its index in *best; ms[k] (optional) gets candidate k's time.
 * Needs the DAGs on the host side (before free_on_cpu). */
int cpvs_container_tune_top_levels(cpvs_container *c, const float *positions, uint32_t stride_floats, uint64_t count,
                                   const float *light_view_proj, uint8_t *vis, const int *candidates,
                                   uint32_t ncand, uint32_t reps, int *best, double *ms);
/* freeOnCPU / moveToGPU (src/CompressedShadowContainer.h:55-68). */
int cpvs_container_free_on_cpu(cpvs_container *c);
int cpvs_container_move_to_gpu(cpvs_container *c);
int cpvs_container_info(const cpvs_container *c, int *dag_levels, int *grid_levels, uint64_t *num_words);
int cpvs_container_copy_words(const cpvs_container *c, uint32_t *host_words, uint64_t cap_words,
                              uint32_t *host_grid, uint64_t cap_grid);
/* CompressedShadowContainer::evaluate (src/CompressedShadowContainer.cpp:93-124) running the
 * lookup of shader/traverse.cs:135-149: positions (device, stride_floats = 3 or 4 floats per
 * point) are multiplied by light_view_proj (16 floats, column-major, host memory; NULL =
 * identity), divided by w, quantised and looked up; vis (device) gets 0 or 1 per point. */
int cpvs_container_evaluate(cpvs_container *c, const float *positions, uint32_t stride_floats, uint64_t count,
                            const float *light_view_proj, uint8_t *vis, void *stream);
/* Percentage-closer filtering with the container's filter size f (setFilterSize; the reference
 * stores it and leaves it unused, shader/traverse.cs:16-17): vis (device, one float per point)
 * gets the fraction of lit leaf cells among the f x f paths (px+dx, py+dy, pz), dx, dy in
 * [-(f-1)/2, f/2], clamped to the grid.  f = 1 gives 0.0 / 1.0 as evaluate.  f in 1..64. */
int cpvs_container_evaluate_pcf(cpvs_container *c, const float *positions, uint32_t stride_floats, uint64_t count,
                                const float *light_view_proj, float *vis, void *stream);
/* The same over a width x height rgba32f position image (row-major, 4 floats per texel), as the
 * reference's Texture2D arguments; vis is a width x height byte image. */
int cpvs_container_evaluate_image(cpvs_container *c, const float *positions_rgba, uint32_t width, uint32_t height,
                                  const float *light_view_proj, uint8_t *vis, void *stream);
int cpvs_container_destroy(cpvs_container *c);

/* ---- DAG byte format ----------------------------------------------------------------------------
 * CompressedShadow::getDAG's word array (src/CompressedShadow.h:75-77) behind a 64-byte versioned
 * header (magic "CPVSDAG", version 1, numLevels, N, zTileIndex, zTileNum, root mask, word count,
 * checksum); loading checks the header, the checksum and the DAG structure before any upload.
 * serialize with buf NULL returns the size in *size. */
int cpvs_dag_serialize(const cpvs_dag *dag, void *buf, uint64_t cap, uint64_t *size);
/* Checks a serialized DAG without a GPU: header, checksum and structure (root at word 0, levels
 * contiguous, every pointer the start of a next-level node, every node referenced, no trailing
 * words).  Optional outputs: numLevels, word count, nodes per level (cap entries). */
int cpvs_dag_check(const void *buf, uint64_t size, uint32_t *num_levels, uint64_t *words, uint64_t *level_nodes,
                   int cap);
int cpvs_dag_deserialize(cpvs_ctx *ctx, const void *buf, uint64_t size, cpvs_dag **out);
int cpvs_dag_save(const cpvs_dag *dag, const char *path);
int cpvs_dag_load(cpvs_ctx *ctx, const char *path, cpvs_dag **out);

/* ---- Batches: many maps built concurrently, many DAGs queried in one launch ------------------
 * No counterpart in the reference, which builds tiles one after another
 * (src/DeferredRenderer.cpp:150-187); each map is built exactly as cpvs_build builds it. */

/* One map of a batch: N x N float depth in device memory (borrowed for the call). */
typedef struct cpvs_batch_item {
    const float *depth;
    uint32_t N;          /* power of two in [8, 65536] */
    uint32_t zTileIndex; /* as CompressedShadow::create (src/CompressedShadow.h:48-49) */
    uint32_t zTileNum;
    uint32_t reserved;   /* 0 */
} cpvs_batch_item;

/* MinMaxHierarchy + CompressedShadow::create for every item, on `lanes` concurrent streams
 * (0 = the context's tuned or default count; at most 16).  out[i] receives item i's DAG (words
 * identical to cpvs_build's).  done_ms (optional, host) receives each item's completion time in
 * ms since the call began.  On error no DAG is returned and out[] is all NULL. */
int cpvs_build_batch(cpvs_ctx *ctx, const cpvs_batch_item *items, uint32_t count, uint32_t lanes, cpvs_dag **out,
                     double *done_ms);
int cpvs_ctx_set_batch_lanes(cpvs_ctx *ctx, uint32_t lanes); /* 0 = default */
int cpvs_ctx_batch_lanes(cpvs_ctx *ctx, uint32_t *lanes);    /* the count a batch with lanes = 0 uses */
/* "<name> <arch> <CUs>CU" of the context's device (keys the lane-count tuning cache). */
int cpvs_ctx_device_name(cpvs_ctx *ctx, char *buf, uint64_t cap);
/* Frees the context's build scratch, its batch lanes' scratch and pyramid-level arenas, and its
 * cache of freed DAG and pyramid buffers (each grows to the largest build or forest it served and
 * is otherwise kept for reuse); *freed (may be NULL) receives the bytes.  Waits for the context's
 * streams; call it between builds (it takes the lock builds and batches hold).  No counterpart in
 * the reference, which keeps no device scratch. */
int cpvs_ctx_release_scratch(cpvs_ctx *ctx, uint64_t *freed);
/* Times the batch once per candidate lane count (best of reps runs, DAGs discarded), stores the
 * fastest as the context's batch lane count and returns it in *best; ms[k] (optional) receives
 * candidate k's time. */
int cpvs_tune_batch(cpvs_ctx *ctx, const cpvs_batch_item *items, uint32_t count, const uint32_t *candidates,
                    uint32_t ncand, uint32_t reps, uint32_t *best, double *ms);

/* How a batch runs.  forest_mtexels = 0: each map is its own launch chain (one cpvs_build per map),
 * `lanes` chains at once.  forest_mtexels > 0: maps are built as forests -- many DAGs through one
 * launch chain, every octree level one array over all of them, all counts from one launch and one
 * readback -- each forest holding maps of at most forest_mtexels x 2^20 texels of work (N^2 /
 * zTileNum per item; the z-tiles of one map stay in one forest and share its pyramid), `lanes`
 * forests at once.  leaf_cols (4 or 8) and leaf_waves (1 or 2) shape the forest's leaf kernel
 * (level-3 columns per wave, waves per workgroup); graph = 1 captures each forest's level chain in
 * a hipGraph and launches it as one.  Words are identical to cpvs_build's in every configuration.
 * Maps with N < 16 (no leafmask levels) are built one by one. */
typedef struct cpvs_batch_config {
    uint32_t lanes;           /* 1..16 */
    uint32_t forest_mtexels;  /* 0 = one map per launch chain */
    uint32_t leaf_cols;       /* 4 or 8 */
    uint32_t leaf_waves;      /* 1 or 2 */
    uint32_t graph;           /* 0 or 1 */
    uint32_t reserved[3];     /* 0 */
} cpvs_batch_config;
/* The configuration batches with lanes = 0 use (initially: 4 lanes of forests of 256 Mtexels,
 * 8 columns per wave, 1 wave per workgroup, no graph). */
int cpvs_ctx_set_batch_config(cpvs_ctx *ctx, const cpvs_batch_config *cfg);
int cpvs_ctx_batch_config(cpvs_ctx *ctx, cpvs_batch_config *cfg);
/* cpvs_build_batch with an explicit configuration (NULL: the context's). */
int cpvs_build_batch_config(cpvs_ctx *ctx, const cpvs_batch_item *items, uint32_t count, const cpvs_batch_config *cfg,
                            cpvs_dag **out, double *done_ms);
/* The autotuner: times the batch once per candidate configuration (best of reps runs after one
 * warm-up of each, DAGs discarded), stores the fastest as the context's configuration and returns
 * its index in *best; ms[k] (optional) receives candidate k's time. */
int cpvs_tune_batch_config(cpvs_ctx *ctx, const cpvs_batch_item *items, uint32_t count, const cpvs_batch_config *cands,
                           uint32_t ncand, uint32_t reps, uint32_t *best, double *ms);

/* A set of DAGs of one leafmask mode (cpvs_dag_leafmasks), each looked up as a one-DAG container
 * (CompressedShadowContainer(unique_ptr<CompressedShadow>), src/CompressedShadowContainer.h:27-32)
 * with an identity light matrix: NDC in, visibility out.  The DAGs are borrowed and must outlive
 * the set. */
typedef struct cpvs_dagset cpvs_dagset;
int cpvs_dagset_create(cpvs_ctx *ctx, const cpvs_dag *const *dags, uint32_t count, cpvs_dagset **out);
/* count queries (device NDC, stride >= 3 floats apart); query i uses DAG ids[i] (device u32), or
 * DAG i mod count when ids is NULL.  vis (device) gets 0/1, or 255 for an id outside the set. */
int cpvs_dagset_evaluate(cpvs_dagset *s, const float *ndc, uint32_t stride, uint64_t count, const uint32_t *ids,
                         uint8_t *vis, void *stream);
int cpvs_dagset_destroy(cpvs_dagset *s);
/* The same for one DAG. */
int cpvs_dag_evaluate(const cpvs_dag *dag, const float *ndc, uint32_t stride, uint64_t count, uint8_t *vis,
                      void *stream);
/* No counterpart in the reference: a lookup top table of the DAG's own, as a one-cell container
 * builds (cpvs_container_set_top_levels; levels -1 automatic, 0 none, 1..9), read by
 * cpvs_dag_evaluate, by DAG sets created after the call and by server lookups; results are the
 * same with any setting.  Built on the context stream, synchronously; call it while no lookup on
 * the DAG is queued.  A replaced table is kept until the DAG is destroyed (DAG sets created earlier
 * name it).  Its bytes are not counted in a server's budget.  top_levels reports the levels built. */
int cpvs_dag_set_top_levels(cpvs_dag *d, int levels);
int cpvs_dag_top_levels(const cpvs_dag *d, int *levels);

/* ---- Serving: a request loop over builds and lookups -----------------------------------------
 * No counterpart in the reference (its renderer builds at start-up and evaluates once per frame,
 * src/DeferredRenderer.cpp:221-230,333).  Requests carry device buffers, borrowed until the
 * request completes.  Builds run on `lanes` concurrent streams; every lookup pending when the
 * lookup stream frees up goes into one launch; DAGs stay resident, named by their build's ticket,
 * within hbm_budget_bytes (resident DAGs + builds in flight, each reserving its scratch growth
 * and an upper bound of its DAG from the exact node counts).  The budget is capped at creation to
 * the device memory then free, less 1 GiB (cpvs_server_stats reports the budget in force).  Least recently used DAGs without
 * queued lookups are evicted to admit a build; a build that cannot fit fails with CPVS_EBUDGET.
 * Resident bytes are the DAGs' buffer sizes; the lanes' pyramid-level and scratch arenas count as
 * lane scratch, and their growth is admitted before it happens.  An evicted or released DAG's
 * buffer returns to the context's buffer cache at once, with no device synchronisation (no lookup
 * pins it), and later builds take their words from there; when no cached buffer fits, the cache is
 * trimmed (hipFree) so that resident + reserved + scratch + cached bytes stay within the budget.
 * Work a caller queues itself on a DAG from cpvs_server_dag must therefore complete before that
 * DAG is released or can be evicted.
 * serial != 0: one request at a time in arrival order on one stream (the plain path). */
typedef struct cpvs_server cpvs_server;
int cpvs_server_create(cpvs_ctx *ctx, uint32_t lanes, uint64_t hbm_budget_bytes, int serial, cpvs_server **out);
int cpvs_server_submit_build(cpvs_server *s, const float *depth, uint32_t N, uint32_t zTileIndex, uint32_t zTileNum,
                             uint64_t *ticket);
/* dag: the ticket of a completed build whose DAG is resident (else CPVS_ESTATE). */
int cpvs_server_submit_lookup(cpvs_server *s, uint64_t dag, const float *ndc, uint32_t stride, uint64_t count,
                              uint8_t *vis, uint64_t *ticket);
/* Blocks until the request is done; returns its status; latency_ms = completion - submission. */
int cpvs_server_wait(cpvs_server *s, uint64_t ticket, double *latency_ms);
int cpvs_server_poll(cpvs_server *s, uint64_t ticket, int *done);
/* The resident DAG of a build (valid until released or evicted). */
int cpvs_server_dag(cpvs_server *s, uint64_t dag, const cpvs_dag **out);
int cpvs_server_release(cpvs_server *s, uint64_t dag);
int cpvs_server_resident(cpvs_server *s, uint64_t *ids, uint32_t cap, uint32_t *count);
/* builds, lookups, lookup launches, queries, evictions, resident DAGs, resident bytes, lane scratch
 * bytes, reserved bytes, budget, cache trims, the context's cached bytes (the first cap of them). */
int cpvs_server_stats(cpvs_server *s, uint64_t *out, uint32_t cap);
/* Finishes every queued request, then frees the resident DAGs. */
int cpvs_server_destroy(cpvs_server *s);

/* Device memory and timing helpers for callers without their own HIP allocator. */
int cpvs_device_alloc(cpvs_ctx *ctx, uint64_t bytes, void **dptr);
int cpvs_device_free(cpvs_ctx *ctx, void *dptr);
int cpvs_memcpy_h2d(cpvs_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int cpvs_memcpy_d2h(cpvs_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int cpvs_memcpy_d2d(cpvs_ctx *ctx, void *dst, const void *src, uint64_t bytes, void *stream); /* asynchronous */
/* A plain copy kernel (16-byte aligned, bytes a multiple of 16), asynchronous: the measured HBM
 * bandwidth and the scale of HBM traffic counters come from it. */
int cpvs_device_copy(cpvs_ctx *ctx, void *dst, const void *src, uint64_t bytes, void *stream);
int cpvs_event_record(cpvs_ctx *ctx, void *stream, void **event);
int cpvs_event_elapsed_ms(void *start, void *stop, float *ms); /* synchronises on stop */
int cpvs_event_destroy(void *event);

/* Synthetic inputs (the reference renders scenes instead).
 * Depth tile [x0,x0+tileN)^2 of a fullN x fullN map: plane 0.6+0.1u, optional 4-octave value
 * noise of amplitude 0.05, nboxes boxes of side 16..128 and depth U[0.1,0.5). */
int cpvs_gen_depth(cpvs_ctx *ctx, float *d_out, uint32_t fullN, uint64_t seed, uint32_t nboxes, int noise,
                   uint32_t x0, uint32_t y0, uint32_t tileN, void *stream);
/* count uniform light-space points as NDC (stride 3 or 4; the 4th component is 1). */
int cpvs_gen_queries_uniform(cpvs_ctx *ctx, float *d_out, uint64_t seed, uint64_t count, uint32_t stride,
                             void *stream);
/* Points first, first+step, ... (count of them) of the same stream as cpvs_gen_queries_uniform. */
int cpvs_gen_queries_uniform_range(cpvs_ctx *ctx, float *d_out, uint64_t seed, uint64_t first, uint64_t step,
                                   uint64_t count, uint32_t stride, void *stream);
/* 2*W*H G-buffer points (float4 NDC) over the N x N depth map d_depth: per pixel one point on
 * the surface and one jittered by U(-2,2) voxels in z. */
int cpvs_gen_queries_gbuffer(cpvs_ctx *ctx, float *d_out, const float *d_depth, uint32_t N, uint64_t seed,
                             uint32_t W, uint32_t H, void *stream);

#ifdef __cplusplus
}
#endif
#endif
