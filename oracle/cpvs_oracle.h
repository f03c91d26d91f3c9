/*
 * cpvs_oracle.h -- CPU restatement of the TobiasRp/cpvs build and lookup path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (cpvs_amd/) links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker / CPU baseline.
 *
 * Every function cites the reference file:line it restates.  Paths are relative
 * to the reference repository root.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_pyramid_*           MinMaxTest known answers (test/MinMaxTest.cpp:41-90)
 *                          + closed form (exact 2x2 min/max, checked with numpy)
 *   or_create_childmask    createChildmask known answer 0x88aa
 *                          (test/CompressedShadowUtilTest.cpp:14-19)
 *   or_create_leaf_node    brute-force midpoint compare on every voxel
 *   or_child_coordinates   test/CompressedShadowUtilTest.cpp:29-65
 *   or_merge_level         test/CompressedShadowUtilTest.cpp:83-97
 *   or_create              traverse known answers (test/CompressedShadowTest.cpp:50-152),
 *                          brute-force midpoint theorem on every voxel of small maps,
 *                          the independent structural DAG builder (oracle/structural.py),
 *                          and faithful == hashed merge word-for-word
 *   or_traverse_cpu        test/CompressedShadowTest.cpp:50-152 + brute force
 *   or_traverse_cs         brute force (no reference test exercises traverse.cs)
 *   or_gen_*               committed fixture checksums (tests/golden/)
 */
#ifndef CPVS_ORACLE_H
#define CPVS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of pyramid levels including the root image (MinMaxHierarchy.h:60-62). */
int or_num_levels(int N);
/* Floats needed for pyramid levels 1..n-1, (min,max) interleaved (MinMaxHierarchy.cpp:9-27). */
size_t or_pyramid_floats(int N);
/* Offset (in floats) of level L >= 1 inside the pyramid buffer. */
size_t or_pyramid_level_offset(int N, int L);
/* Build levels 1..n-1 of the min/max pyramid of the N x N image D (MinMaxHierarchy.cpp:9-97). */
void or_pyramid_build(const float *D, int N, float *P);
/* getMin / getMax (MinMaxHierarchy.h:31-55); level 0 returns the depth for both. */
float or_pyramid_min(const float *D, const float *P, int N, int L, int x, int y);
float or_pyramid_max(const float *D, const float *P, int N, int L, int x, int y);

/* cs::createChildmask (CompressedShadowUtil.cpp:20-54), zTileNum passed explicitly. */
uint32_t or_create_childmask(const float *D, const float *P, int N, int level,
                             int ox, int oy, int oz, uint32_t zTileNum);
/* cs::createChildmask1x1x8 (CompressedShadowUtil.cpp:80-99): returns the 16-bit mask,
 * writes 0..8 partial leafmasks to masks[] and their count to *count. */
uint32_t or_create_leaf_node(const float *D, int N, int ox, int oy, int oz,
                             uint32_t zTileNum, uint64_t *masks, int *count);
/* cs::getChildCoordinates (CompressedShadowUtil.cpp:101-117); returns count, xyz triples. */
int or_child_coordinates(uint32_t childmask, int px, int py, int pz, int *out_xyz);
/* cs::mergeLevel (CompressedShadowUtil.h:154-182) on a level of nodeSize-word nodes.
 * mapping[i] = new relative offset of old node i (in words); returns the unique count.
 * mode 0 = the reference's linear search, mode 1 = hash table (same output). */
uint32_t or_merge_level(const uint32_t *level_words, size_t num_nodes, uint32_t nodeSize,
                        uint32_t *out_words, uint32_t *mapping, int mode);

/* CompressedShadow::create (CompressedShadow.cpp:49-59).  Returns a malloc'd word
 * array (free with or_free) and its length; numLevels in *num_levels.
 * merge_mode: 0 = faithful O(n*u) mergeLevel, 1 = hashed (identical words).
 * Returns 0 on success, nonzero on bad arguments / allocation failure. */
/* The reference's compile-time LEAFMASKS switch (CompressedShadow.cpp:17): on by default. */
void or_set_leafmasks(int on);
int or_create(const float *D, const float *P, int N, uint32_t zTileIndex, uint32_t zTileNum,
              int merge_mode, uint32_t **out_words, size_t *out_len, int *num_levels);
void or_free(void *p);

/* Number of SVO nodes per level before merging, from the last create() call of
 * this thread (levels 0..n-2); used by tests to check node counts. */
void or_last_node_counts(uint64_t *out, int n);

/* CompressedShadow::getTotalVisibility (CompressedShadow.cpp:66-72). */
int or_total_visibility(const uint32_t *words);

/* CompressedShadow::traverse (CompressedShadow.cpp:404-463) for one NDC point.
 * Returns 0 SHADOW, 1 VISIBLE, 2 PARTIAL. */
int or_traverse_cpu(const uint32_t *words, int numLevels, const float ndc[3], int tryLeafmasks);

/* traverse.cs:41-133 transcription for one integer path (grid + dag). */
int or_traverse_path_cs(const uint32_t *dag, const uint32_t *grid, int dagLevels,
                        int gridLevels, int px, int py, int pz);
/* traverse.cs:43-48 getPathFromNDC with R = 1 << (dagLevels+gridLevels-1); the
 * path is clamped to [0, R-1] (out-of-range input is undefined in the reference). */
void or_path_from_ndc(const float ndc[3], int dagLevels, int gridLevels, int *path);
/* traverse.cs:135-149 main() over count positions (stride floats apart), matrix m
 * column-major (NULL = identity).  vis[i] = 0/1. */
void or_evaluate(const uint32_t *dag, const uint32_t *grid, int dagLevels, int gridLevels,
                 const float *pos, uint32_t stride, size_t count, const float *m,
                 uint8_t *vis, int nthreads);

/* CompressedShadowContainer::createTopLevelGrid (CompressedShadowContainer.cpp:71-91)
 * with the sentinels traverse.cs tests (0xFFFFFFFF shadow, 0xFFFFFFFE visible). */
void or_top_level_grid(const uint32_t *const *dags, const size_t *lens, size_t count,
                       uint32_t *grid);

/* Brute-force leaf-resolution visibility: float(pz)+0.5f <= D[py][px]*float(N*T). */
int or_brute_visible(const float *D, int N, uint32_t zTileNum, int px, int py, int pz);

/* Synthetic generator (the build's own input spec; the reference renders scenes). */
void or_gen_depth_tile(int fullN, uint64_t seed, int nboxes, int noise, int x0, int y0,
                       int tileN, float *out);
void or_gen_queries_uniform(uint64_t seed, size_t count, uint32_t stride, float *out);
void or_gen_queries_gbuffer(const float *D, int N, uint64_t seed, int W, int H, float *out);

#ifdef __cplusplus
}
#endif
#endif
