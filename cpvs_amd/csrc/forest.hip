// forest.hip -- many CompressedShadow::create calls as one forest with one launch chain.
//
// The reference builds one DAG after another (src/DeferredRenderer.cpp:150-187: every z-tile of
// every tile, each create() a separate pass over its own vectors).  One build here is a chain of
// about 60 launches and three host readbacks; a small map keeps the GPU busy for only a few of
// them.  A forest builds M DAGs (maps of any sizes, and z-tiles of one map sharing its pyramid)
// with the launches of one build:
//   * every octree level L is one array over all DAGs, DAG-major: DAG f's level-L nodes form one
//     segment, in its own breadth-first order.  DAGs are ordered by root level, highest first; a
//     level then lists the children emitted from the level above (whose DAGs come first) followed
//     by the roots at that level, so one scan per level emits every DAG's children in place;
//   * a node's coordinate carries its DAG's index (fpack), through which classification and the
//     leaf kernel find the DAG's pyramid, size and z-tile;
//   * mask words carry the DAG's index in their upper 16 bits, so the merge kernels of the single
//     build (hash tables, key compares, ranks) serve the forest unchanged and never merge across
//     DAGs; a level's ranks and word offsets are global, and each DAG's level base and local
//     offsets follow from the ranks and offsets at its segment starts (k_forest_bases);
//   * the counts of all DAGs come from one launch (k_forest_counts) and one readback, and the
//     chain between the two readbacks can be captured in a hipGraph (ForestOpts::graph).
// The words of each DAG equal those of its own cpvs_build (the same keys, ranks and layout).
#include <algorithm>
#include <cstring>
#include <vector>

#include "build_parts.hpp"

namespace cpvs {

// ---------------------------------------------------------------------------------------
// Device tables.
// ---------------------------------------------------------------------------------------

// One DAG of the forest.
struct FItem {
    const float* depth;            // the map (level 0)
    const float2* lv[kMaxLevels];  // pyramid levels >= 3
    u32 N, T, t, R;                // side, zTileNum, zTileIndex, root level (numLevels - 2)
    u32 w3, w3_shift, col_base;    // leaf columns: N/8 per side, first column record
    u32 pad;
};

// Node coordinates at levels >= 3 with the DAG's index: x, y < 2^13 (N <= 65536), z < 2^18
// (N * zTileNum <= 2^21), DAG < 2^20.
__host__ __device__ inline u64 fpack(u32 x, u32 y, u32 z, u32 g) {
    return (u64)x | ((u64)y << 13) | ((u64)z << 26) | ((u64)g << 44);
}
__device__ inline u32 fx(u64 c) { return (u32)(c & 0x1FFF); }
__device__ inline u32 fy(u64 c) { return (u32)((c >> 13) & 0x1FFF); }
__device__ inline u32 fz(u64 c) { return (u32)((c >> 26) & 0x3FFFF); }
__device__ inline u32 fg(u64 c) { return (u32)(c >> 44); }

constexpr u32 kNone = 0xFFFFFFFFu;
constexpr u32 kMaxForest = 65535;  // DAG indices live in the upper 16 bits of mask words
constexpr u32 kKL = kMaxLevels;    // row length of the per-(DAG, level) tables

// ---------------------------------------------------------------------------------------
// 1. Counts of every DAG in one launch.  Record k covers pyramid level `level` of DAG `item`, from
//    block block0 on (kCountTile texels per block).
// ---------------------------------------------------------------------------------------
struct FCountRec {
    u32 block0, item, level, pad;
};
constexpr u32 kFCountPer = 8;
constexpr u32 kFCountTile = 256 * kFCountPer;

// exact: only the cells under a partial cell at every level above (reachable_cells; maps with NaN
// depths, whose plain counts can exceed the emitted nodes).
__global__ __launch_bounds__(256) void k_forest_counts(const FCountRec* __restrict__ recs, u32 nrec,
                                                       const FItem* __restrict__ items, u32* __restrict__ partials,
                                                       uint4* __restrict__ colinfo, u32 root_slot0, u32 exact) {
    __shared__ u32 s_wave[4];
    u32 lo = 0, hi = nrec;  // recs[lo].block0 <= blockIdx.x < recs[hi].block0
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (recs[mid].block0 <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const FCountRec r = recs[lo];
    const FItem& it = items[r.item];
    const int L = (int)r.level;
    const u32 w = it.N >> L;
    const float h = (float)((it.N >> L) * it.T);
    const int zlo = (int)(it.t * (it.N >> L)), zhi = (int)((it.t + 1) * (it.N >> L));
    const float2* lv = it.lv[L];
    uint4* colc = (L == 3) ? colinfo + it.col_base : nullptr;
    const u32 base = (blockIdx.x - r.block0) * kFCountTile;
    u32 acc = 0;
#pragma unroll
    for (u32 k = 0; k < kFCountPer; ++k) {
        const u32 j = base + k * 256 + threadIdx.x;
        if (j < w * w) {
            int s;
            const u32 c = exact ? reachable_cells(it.lv, j % w, j / w, L, (int)it.R, it.N, it.T, zlo, zhi, &s)
                                : partial_cells(lv[j], h, zlo, zhi, &s);
            if (colc) colc[j] = make_uint4(c, 0u, (u32)max(s, 0), r.item);
            acc += c;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        if ((u32)L == it.R) {
            // The root (0, 0, 2t): its childmask, as classify_node computes it.
            u32 m = 0;
#pragma unroll
            for (u32 ch = 0; ch < 8; ++ch) {
                const float2 v = lv[ch & 3];
                m |= classify_cell(2 * it.t + (ch >> 2), v.x * h, v.y * h) << (2 * ch);
            }
            partials[root_slot0 + r.item] = m;
        }
    }
}

// ---------------------------------------------------------------------------------------
// 2. Roots, classification and emission.
// ---------------------------------------------------------------------------------------
struct FLevels {
    u64* coord[kMaxLevels];
    u32* mask[kMaxLevels];
    u32* cstart[kMaxLevels];
    u32* rank[kMaxLevels];
    u32* uniq[kMaxLevels];
    u32* off[kMaxLevels];
};

// Root of DAG g at position seg[g][R] of level R (0, 0, 2t); the metadata words zeroed.
__global__ void k_forest_roots(const FItem* __restrict__ items, u32 M, const u32* __restrict__ seg, FLevels lv,
                               u32* __restrict__ meta) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMetaWords) meta[i] = 0u;
    if (i >= M) return;
    const FItem& it = items[i];
    lv.coord[it.R][seg[(size_t)i * kKL + it.R]] = fpack(0, 0, 2 * it.t, i);
}

// createChildmask of a forest node: the single build's classification on the node's own pyramid,
// with the DAG's index in the mask's upper half.  The index is clamped to the forest's M items (as
// classify_node clamps the cell): a node past its level's emitted children (counts that exceed
// them, which the host then reports) still reads inside the forest's arrays.
__device__ inline u32 fclassify(u64 c, int L, const FItem* __restrict__ items, u32 M) {
    const u32 g = min(fg(c), M - 1);
    const FItem& it = items[g];
    return classify_node(pack_xyz(fx(c), fy(c), fz(c)), it.lv[L], it.depth, it.N, L, it.T) | (g << 16);
}

__global__ __launch_bounds__(256) void k_forest_classify(const u64* __restrict__ coord, u32 count,
                                                         const FItem* __restrict__ items, u32 M, int L,
                                                         u32* __restrict__ mask) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) mask[i] = fclassify(coord[i], L, items, M);
}

struct FClassify {
    const u64* coord;
    const FItem* items;
    u32 M;
    int L;
    __device__ u32 operator()(u32 i) const { return fclassify(coord[i], L, items, M); }
};
struct FClassifyCountIn {
    FClassify cls;
    __device__ u32 operator()(u32 i) const { return num_children(cls(i)); }
};

// The scan's output stage: first child index, and the children in (parent, child index) order --
// coordinates for inner children, the column map for leaves (getChildCoordinates,
// src/CompressedShadowUtil.cpp:101-117).
struct FEmitOut {
    const u32* mask;
    const u64* coord;
    u32* cstart;
    u64* child_coord;       // inner children
    u32 child_cap;          // emitted nodes of the child level; writes beyond it are dropped
    const FItem* items;
    u32 M;                  // items in the forest
    const uint4* colinfo;   // leaf children: column records ...
    u32* colbfs;            // ... and (column slot -> breadth-first index), when set
    u32* mask_out = nullptr;  // when set, the node is classified here and its mask stored
    FClassify cls = {};
    struct Pre {
        u32 m;
        u64 c;
    };
    __device__ Pre load(u32 i) const { return Pre{mask_out ? cls(i) : mask[i], coord[i]}; }
    __device__ void operator()(u32 i, u32 first, u32 k, Pre pre) const {
        cstart[i] = first;
        if (mask_out) mask_out[i] = pre.m;
        if (!k) return;
        const u32 m = pre.m;
        const u64 c = pre.c;
        const u32 x = fx(c), y = fy(c), z = fz(c), g = min(fg(c), M - 1);
        u32 j = first;
        if (colbfs) {
            const FItem& it = items[g];
            const uint4* info = colinfo + it.col_base;
            const u32 xc = min(x, it.w3 - 2) & ~1u, yc = min(y, it.w3 - 2);  // inside the DAG's columns
            uint4 ci[4];
#pragma unroll
            for (u32 q = 0; q < 4; ++q) {
                ci[q] = make_uint4(0u, 0u, 0u, 0u);
                if (m & ((2u << (2 * q)) | (2u << (2 * (q + 4))))) ci[q] = info[(yc + (q >> 1)) * it.w3 + xc + (q & 1)];
            }
#pragma unroll
            for (u32 ch = 0; ch < 8; ++ch) {
                if (!(m & (2u << (2 * ch)))) continue;
                const uint4 r = ci[ch & 3];
                const u32 slot = r.y + (z + (ch >> 2) - r.z);
                if (j < child_cap && slot < child_cap) colbfs[slot] = j;
                ++j;
            }
            return;
        }
#pragma unroll
        for (u32 ch = 0; ch < 8; ++ch) {
            if (!(m & (2u << (2 * ch)))) continue;
            if (j < child_cap) child_coord[j] = fpack(2 * (x + (ch & 1)), 2 * (y + ((ch >> 1) & 1)), 2 * (z + (ch >> 2)), g);
            ++j;
        }
    }
};

// ---------------------------------------------------------------------------------------
// 3. Leaves: the single build's column waves (build.hip, k_leaf_waves) over the forest's list of
//    columns with leaf nodes.  A column's record names its DAG (.w), whose map, size and z-tile
//    give the column's depth block and scale.  kCols columns per wave, kWaves waves per group.
// ---------------------------------------------------------------------------------------
template <u32 kCols, u32 kWaves>
__global__ __launch_bounds__(64 * kWaves) void k_forest_leaf(const uint4* __restrict__ colinfo, const u32* __restrict__ colbfs,
                                                             const FItem* __restrict__ items, u32 nleaf,
                                                             u32* __restrict__ mask_out, u64* __restrict__ leafm,
                                                             u32* __restrict__ runhead,
                                                             u64* __restrict__ tab, u32 tab_slots,
                                                             const u32* __restrict__ col_list, const u32* __restrict__ n_list,
                                                             u32 list_cap) {
    __shared__ int s_f[kWaves][kCols][64];
    __shared__ u32 s_b[kWaves][kCols][64];
    __shared__ uint8_t s_chg[kWaves][128];  // each wave's changed-node flags (leaf_column)
    const u32 lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 c0 = (blockIdx.x * kWaves + wave) * kCols;
    const u32 nvalid = min(*n_list, list_cap);
    const u32 clast = nvalid ? nvalid - 1 : 0u;
    const u32 my_col = nvalid ? col_list[min(c0 + (lane & (kCols - 1)), clast)] : 0u;
    uint4 ci = colinfo[my_col];
    if (lane >= kCols || c0 + lane >= nvalid) ci = make_uint4(0u, 0u, 0u, 0u);
    float d[kCols];
    float HT[kCols];
#pragma unroll
    for (u32 k = 0; k < kCols; ++k) {
        const u32 c = __builtin_amdgcn_readlane(my_col, k);
        const FItem& it = items[__builtin_amdgcn_readlane(ci.w, k)];
        const u32 local = c - it.col_base;
        const u32 x3 = local & (it.w3 - 1), y3 = local >> it.w3_shift;
        HT[k] = (float)(it.N * it.T);
        const float* blk = it.depth + ((size_t)(8 * y3) * it.N + 8 * x3);  // uniform base, per-lane offset
        d[k] = 0.0f;
        const u32 loff = 4u * ((lane >> 3) * it.N + (lane & 7));  // bytes, below 2^32 (N <= 65536)
        if (__builtin_amdgcn_readlane(ci.x, k) != 0u)
            d[k] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(blk) + loff);
    }
    u32 b[kCols];
#pragma unroll
    for (u32 k = 0; k < kCols; ++k) {
        const u32 cnt = __builtin_amdgcn_readlane(ci.x, k), st = __builtin_amdgcn_readlane(ci.y, k);
        const u32* run = colbfs + (cnt ? st : 0u);
        b[k] = run[(lane < cnt) ? lane : 0u];
    }
    // The columns' minimum and maximum f from their DAGs' level-3 pyramid texels (lane k < kCols:
    // column k's minimum, lane k + kCols its maximum), or from the texels for a block with a NaN.
    int fext;
    {
        const uint4 cj = colinfo[my_col];  // lanes k and k + kCols read column k's record
        const FItem& it = items[cj.w];
        const float2 ext = it.lv[3][my_col - it.col_base];
        fext = leaf_f(((lane & kCols) ? ext.y : ext.x) * (float)(it.N * it.T));
    }
    u32 nan_cols = 0;
#pragma unroll
    for (u32 k = 0; k < kCols; ++k) {
        const float D = d[k] * HT[k];
        if (__ballot(!(D == D))) nan_cols |= 1u << k;
        s_f[wave][k][lane] = leaf_f(D);
        s_b[wave][k][lane] = b[k];
    }

    // The wave's columns with leaf nodes (ci is zero for lanes past the list).
    for (u32 todo = (u32)__ballot(ci.x != 0u); todo; todo &= todo - 1u) {
        const u32 k = (u32)__builtin_ctz(todo);
        const u32 cnt = __builtin_amdgcn_readlane(ci.x, k);
        const u32 base = __builtin_amdgcn_readlane(ci.y, k);
        const int zs = (int)__builtin_amdgcn_readlane(ci.z, k);
        const u32 gtag = __builtin_amdgcn_readlane(ci.w, k) << 16;
        const int f = s_f[wave][k][lane];
        int fmin = __builtin_amdgcn_readlane(fext, k), fmax = __builtin_amdgcn_readlane(fext, k + kCols);
        if ((nan_cols >> k) & 1u) {
            fmin = __ockl_wfred_min_i32(f);
            fmax = __ockl_wfred_max_i32(f);
        }
        leaf_column(f, fmin, fmax, zs, cnt, s_b[wave][k][lane], colbfs + base, gtag, nleaf,
                    kWaves == 1 ? s_chg[0] : s_chg[wave], lane, mask_out, leafm, runhead);
    }
    clear_table(tab, tab_slots);
}

// ---------------------------------------------------------------------------------------
// 4. Compress: each DAG's level sizes and bases from the global ranks and offsets at its segment
//    starts, the readback, and the words.
// ---------------------------------------------------------------------------------------

// Host readback layout (u32 words after the metadata block): per DAG its word count and root
// mask, then per (DAG, level) its unique nodes.
struct FBasesOut {
    u32* meta_host;  // kMetaWords
    u32* total;      // M
    u32* root;       // M
    u32* uniq;       // M * kKL
};

__global__ void k_forest_bases(const FItem* __restrict__ items, u32 M, FLevels lv, const u32* __restrict__ seg,
                               const u32* __restrict__ nxt, u32* __restrict__ base_off, u32* __restrict__ B,
                               const u32* __restrict__ meta, int Lmin, FBasesOut ho) {
    const u32 g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < M) {
        const FItem& it = items[g];
        unsigned long long acc = 0;
        for (int L = (int)it.R; L >= Lmin; --L) {
            const size_t k = (size_t)g * kKL + L;
            const u32 s = seg[k];
            u32 S = 0, U = 0, o0 = 0;
            if (s != kNone) {
                // Ranks bounded by the level's unique nodes: a rank past the level's emitted children
                // was never written (the host then rejects the forest's words).
                const u32 Ul = meta[kMetaU + L], rmax = Ul ? Ul - 1 : 0u;
                const u32 r0 = min(lv.rank[L][s], rmax);
                o0 = lv.off[L][r0];
                const u32 e = nxt[k];
                const u32 r1 = (e != kNone) ? min(lv.rank[L][e], rmax) : Ul;
                const u32 o1 = (e != kNone) ? lv.off[L][r1] : meta[kMetaS + L];
                S = o1 - o0;
                U = r1 - r0;
            }
            base_off[k] = o0;
            B[k] = (acc > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (u32)acc;
            acc += S;
            ho.uniq[k] = U;
        }
        ho.total[g] = (acc > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (u32)acc;
        ho.root[g] = lv.mask[it.R][seg[(size_t)g * kKL + it.R]] & 0xFFFFu;
    }
    if (blockIdx.x == 0) {
        __syncthreads();
        for (u32 i = threadIdx.x; i < kMetaWords; i += blockDim.x) ho.meta_host[i] = meta[i];
    }
    __threadfence_system();
}

struct FWriteArgs {
    int nlev;
    u32 first[kMaxLevels + 1];  // first block of each level's range
    int level[kMaxLevels];
    u32 U[kMaxLevels];          // unique nodes of the level (all DAGs)
};

// Unique node r of level L (global rank order) into its DAG's buffer: [mask, absolute offsets of
// the partial children] or [mask, lo, hi, ...] at the leaf level (src/CompressedShadow.cpp:326-392).
__global__ __launch_bounds__(kNodeThreads) void k_forest_write(FWriteArgs a, FLevels lv, const u64* __restrict__ leafm,
                                                                const u32* __restrict__ base_off, const u32* __restrict__ B,
                                                                u32* const* __restrict__ outp) {
    int k = 0;
    while (k + 1 < a.nlev && blockIdx.x >= a.first[k + 1]) ++k;
    const int L = a.level[k];
    const u32 r = (blockIdx.x - a.first[k]) * blockDim.x + threadIdx.x;
    if (r >= a.U[k]) return;
    const u32 i = lv.uniq[L][r];
    const u32 m = lv.mask[L][i];
    const u32 g = m >> 16;
    const size_t gk = (size_t)g * kKL + L;
    const u32 o = B[gk] + (lv.off[L][r] - base_off[gk]);
    u32* out = outp[g];
    out[o] = m & 0xFFFFu;
    const u32 kc = num_children(m);
    if (!kc) return;
    if (L == 2) {
        const u64* p = leafm + (size_t)i * 8;  // a first occurrence heads its run: its own row
        u64 v[8];
#pragma unroll
        for (u32 j = 0; j < 8; ++j) v[j] = (j < kc) ? p[j] : 0ull;
#pragma unroll
        for (u32 j = 0; j < 8; ++j)
            if (j < kc) {
                out[o + 1 + 2 * j] = (u32)v[j];
                out[o + 2 + 2 * j] = (u32)(v[j] >> 32);
            }
        return;
    }
    u32 cr[8];
    load_child_ranks(lv.cstart[L], lv.rank[L - 1], i, kc, cr);
    const u32 cb = B[gk - 1], co0 = base_off[gk - 1];
    u32 co[8];
#pragma unroll
    for (u32 j = 0; j < 8; ++j) co[j] = (j < kc) ? lv.off[L - 1][cr[j]] : 0u;
#pragma unroll
    for (u32 j = 0; j < 8; ++j)
        if (j < kc) out[o + 1 + j] = cb + (co[j] - co0);
}

// ---------------------------------------------------------------------------------------
// Host orchestration.
// ---------------------------------------------------------------------------------------
namespace {

inline dim3 grid_of(u64 n, u32 block) { return dim3((u32)std::max<u64>(1, (n + block - 1) / block)); }

constexpr int kRetrySort = -1010;
constexpr int kRetryFull = -1011;
constexpr int kRetryExact = -1012;  // the counts exceed the emitted nodes (a map with NaN depths)

// Launches the leaf kernel in the shape the options name.
void launch_leaf(const ForestOpts& o, u32 units_cols, hipStream_t s, const uint4* colinfo, const u32* colbfs,
                 const FItem* items, u32 nleaf, u32* mask, u64* leafm, u32* runhead, u64* tab,
                 u32 slots, const u32* clist, const u32* nclist, u32 cap) {
    const u32 per = o.leaf_cols * o.leaf_waves;
    const dim3 grid((units_cols + per - 1) / per);
    if (o.leaf_cols == 4 && o.leaf_waves == 1)
        CPVS_LAUNCH("forest_leaf", (k_forest_leaf<4, 1>), grid, dim3(64), 0, s, colinfo, colbfs, items, nleaf, mask, leafm,
                    runhead, tab, slots, clist, nclist, cap);
    else if (o.leaf_cols == 4)
        CPVS_LAUNCH("forest_leaf", (k_forest_leaf<4, 2>), grid, dim3(128), 0, s, colinfo, colbfs, items, nleaf, mask, leafm,
                    runhead, tab, slots, clist, nclist, cap);
    else if (o.leaf_waves == 1)
        CPVS_LAUNCH("forest_leaf", (k_forest_leaf<8, 1>), grid, dim3(64), 0, s, colinfo, colbfs, items, nleaf, mask, leafm,
                    runhead, tab, slots, clist, nclist, cap);
    else
        CPVS_LAUNCH("forest_leaf", (k_forest_leaf<8, 2>), grid, dim3(128), 0, s, colinfo, colbfs, items, nleaf, mask, leafm,
                    runhead, tab, slots, clist, nclist, cap);
}

// Pinned staging of one forest build.
struct Stage {
    u32* partials;  // nblocks + M (root masks)
    FItem* items;   // M
    FCountRec* recs;
    u32* seg;       // M * kKL
    u32* nxt;       // M * kKL
    u32** ptrs;     // M output pointers
    u32* back;      // readback: kMetaWords, then totals (M), roots (M), uniques (M * kKL)
};

template <class T>
T* carve(char*& p, size_t n) {
    T* r = reinterpret_cast<T*>(p);
    p += (n * sizeof(T) + 255) & ~(size_t)255;
    return r;
}

int forest_once(cpvs_ctx* ctx, Worker& w, const ForestIn* in, u32 M, const ForestOpts& o, int merge_mode, bool small_leaf,
                bool exact, cpvs_dag** out) {
    const hipStream_t s = w.stream;
    const bool use_table = merge_mode == CPVS_MERGE_TABLE;

    // --- order (root level, highest first) and the item table --------------------------------
    std::vector<u32> ord(M);
    for (u32 i = 0; i < M; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](u32 a, u32 b) { return in[a].p->n > in[b].p->n; });
    std::vector<FItem> hi(M);
    u64 ncol = 0;
    int Rmax = 0;
    u32 nrec = 0, nblocks = 0;
    for (u32 f = 0; f < M; ++f) {
        const ForestIn& x = in[ord[f]];
        FItem& it = hi[f];
        std::memset(&it, 0, sizeof(it));
        it.depth = x.p->depth;
        for (int L = 0; L < kMaxLevels; ++L) it.lv[L] = x.p->lv[L];
        it.N = x.p->N;
        it.T = x.T;
        it.t = x.t;
        it.R = (u32)(x.p->n - 2);
        it.w3 = x.p->N >> 3;
        it.w3_shift = (u32)__builtin_ctz(it.w3);
        it.col_base = (u32)ncol;
        ncol += (u64)it.w3 * it.w3;
        Rmax = std::max(Rmax, (int)it.R);
        nrec += it.R - 2;
    }
    if (ncol >= 0xFFFFFFF0ull) return kForestTooLarge;
    std::vector<FCountRec> recs;
    recs.reserve(nrec);
    for (u32 f = 0; f < M; ++f)
        for (u32 L = 3; L <= hi[f].R; ++L) {
            const u64 w2 = (u64)(hi[f].N >> L) * (hi[f].N >> L);
            recs.push_back(FCountRec{nblocks, f, L, 0});
            nblocks += (u32)((w2 + kFCountTile - 1) / kFCountTile);
        }

    // --- pinned staging -------------------------------------------------------------------------
    auto stage_bytes = [&](size_t nb) {
        return arena_bytes(nb + M, 4) + arena_bytes(M, sizeof(FItem)) + arena_bytes(recs.size(), sizeof(FCountRec)) +
               2 * arena_bytes((size_t)M * kKL, 4) + arena_bytes(M, sizeof(u32*)) +
               arena_bytes(kMetaWords + 2 * (size_t)M + (size_t)M * kKL, 4);
    };
    int rc = w.forest_staging(stage_bytes(nblocks));
    if (rc) return rc;
    Stage st;
    {
        char* p = static_cast<char*>(w.fpin);
        st.partials = carve<u32>(p, nblocks + M);
        st.items = carve<FItem>(p, M);
        st.recs = carve<FCountRec>(p, recs.size());
        st.seg = carve<u32>(p, (size_t)M * kKL);
        st.nxt = carve<u32>(p, (size_t)M * kKL);
        st.ptrs = carve<u32*>(p, M);
        st.back = carve<u32>(p, kMetaWords + 2 * (size_t)M + (size_t)M * kKL);
    }
    std::memcpy(st.items, hi.data(), sizeof(FItem) * M);
    std::memcpy(st.recs, recs.data(), sizeof(FCountRec) * recs.size());

    // --- arena, part 1: what the counts write and read ----------------------------------------
    Arena& A = w.arena;
    const size_t head = arena_bytes(ncol, sizeof(uint4)) + arena_bytes(kMetaWords, 4) + arena_bytes(M, sizeof(FItem)) +
                        arena_bytes(recs.size(), sizeof(FCountRec)) + 4 * arena_bytes((size_t)M * kKL, 4) +
                        arena_bytes(M, sizeof(u32*));
    rc = A.reserve(head, s);
    if (rc) return rc;
    auto take_head = [&](uint4** colinfo, u32** meta, FItem** items, FCountRec** drecs, u32** seg, u32** nxt, u32** boff,
                         u32** B, u32*** ptrs) {
        *colinfo = A.take<uint4>(ncol);
        *meta = A.take<u32>(kMetaWords);
        *items = A.take<FItem>(M);
        *drecs = A.take<FCountRec>(recs.size());
        *seg = A.take<u32>((size_t)M * kKL);
        *nxt = A.take<u32>((size_t)M * kKL);
        *boff = A.take<u32>((size_t)M * kKL);
        *B = A.take<u32>((size_t)M * kKL);
        *ptrs = A.take<u32*>(M);
    };
    uint4* colinfo;
    u32 *meta, *seg, *nxt, *boff, *B;
    FItem* items;
    FCountRec* drecs;
    u32** dptrs;
    take_head(&colinfo, &meta, &items, &drecs, &seg, &nxt, &boff, &B, &dptrs);
    auto run_counts = [&](u32* partials) -> int {
        CPVS_HIP_CHECK(hipMemcpyAsync(items, st.items, sizeof(FItem) * M, hipMemcpyHostToDevice, s));
        CPVS_HIP_CHECK(hipMemcpyAsync(drecs, st.recs, sizeof(FCountRec) * recs.size(), hipMemcpyHostToDevice, s));
        CPVS_LAUNCH("forest_counts", k_forest_counts, dim3(nblocks), dim3(256), 0, s, drecs, (u32)recs.size(), items, partials,
                    colinfo, nblocks, exact ? 1u : 0u);
        CPVS_HIP_CHECK(hipGetLastError());
        return CPVS_OK;
    };
    rc = run_counts(st.partials);
    if (rc) return rc;
    CPVS_HIP_CHECK(hipStreamSynchronize(s));

    // --- exact counts and the forest's layout -----------------------------------------------------
    std::vector<u64> cnt((size_t)M * kKL, 0);  // nodes of DAG f at level L
    for (size_t k = 0; k < recs.size(); ++k) {
        const u32 b1 = (k + 1 < recs.size()) ? recs[k + 1].block0 : nblocks;
        u64 c = 0;
        for (u32 b = recs[k].block0; b < b1; ++b) c += st.partials[b];
        cnt[(size_t)recs[k].item * kKL + recs[k].level - 1] = c;  // partial level-L cells = level L-1 nodes
    }
    for (u32 f = 0; f < M; ++f) cnt[(size_t)f * kKL + hi[f].R] = 1;
    u64 C[kMaxLevels] = {}, E[kMaxLevels] = {};
    for (u32 f = 0; f < M; ++f)
        for (int L = 2; L <= (int)hi[f].R; ++L) {
            const u64 c = cnt[(size_t)f * kKL + L];
            C[L] += c;
            if ((int)hi[f].R > L) E[L] += c;
        }
    int Lmin = Rmax;  // the lowest level with nodes (levels between may be empty: DAGs ending early)
    for (int L = Rmax - 1; L >= 2; --L)
        if (C[L]) Lmin = L;
    for (int L = 2; L <= Rmax; ++L) {
        const u64 per = (L == 2) ? 17 : 9;
        if (C[L] * per >= 0xFFFFFFF0ull) return kForestTooLarge;
    }
    // Segments: DAG f's level-L nodes start at seg[f][L]; nxt = the next DAG's start at L.
    for (int L = 2; L <= Rmax; ++L) {
        u64 pos = 0;
        u32 last = kNone;
        for (u32 f = 0; f < M; ++f) {
            const size_t k = (size_t)f * kKL + L;
            st.seg[k] = kNone;
            st.nxt[k] = kNone;
            if ((int)hi[f].R < L) continue;
            const u64 c = cnt[k];
            if (!c) continue;
            st.seg[k] = (u32)pos;
            if (last != kNone) st.nxt[(size_t)last * kKL + L] = (u32)pos;
            last = f;
            pos += c;
        }
    }
    for (u32 f = 0; f < M; ++f)
        for (int L = 0; L < 2; ++L) st.seg[(size_t)f * kKL + L] = st.nxt[(size_t)f * kKL + L] = kNone;
    for (u32 f = 0; f < M; ++f)
        for (int L = Rmax + 1; L < kMaxLevels; ++L) st.seg[(size_t)f * kKL + L] = st.nxt[(size_t)f * kKL + L] = kNone;

    // --- scratch ----------------------------------------------------------------------------------
    u32 cmax = 1;
    for (int L = Lmin; L <= Rmax; ++L) cmax = std::max<u32>(cmax, (u32)C[L]);
    const u32 S = table_slots(cmax);
    const u32 scan_n = (u32)std::max<u64>(ncol, cmax);
    const size_t c2 = C[2] ? C[2] : 1;
    size_t need = head;
    for (int L = Lmin; L <= Rmax; ++L) {
        const size_t c = C[L] ? C[L] : 1;
        need += (L > 2 ? arena_bytes(c, 8) + arena_bytes(c, 4) : 0) + 4 * arena_bytes(c, 4);  // no coords, starts at leaves
    }
    need += arena_bytes(c2 * 8, 8) + 2 * arena_bytes(c2, 4);  // leafmask rows, run heads, column map
    need += arena_bytes(ncol, 4) + arena_bytes(1, 4);
    need += 2 * arena_bytes(cmax, 4);
    if (use_table) need += arena_bytes(S, 8);
    else need += 4 * arena_bytes(cmax, 4) + arena_bytes(sort_scratch_words(cmax), 4);
    need += arena_bytes(scan_scratch_words(scan_n), 4);
    need += 1 << 20;
    bool moved = false;
    rc = A.reserve(need, s, &moved);
    if (rc) return rc;
    take_head(&colinfo, &meta, &items, &drecs, &seg, &nxt, &boff, &B, &dptrs);
    if (moved) {
        rc = run_counts(st.partials);  // the column records were lost with the old arena
        if (rc) return rc;
    }
    FLevels lv{};
    for (int L = Lmin; L <= Rmax; ++L) {
        const size_t c = C[L] ? C[L] : 1;
        lv.coord[L] = (L > 2) ? A.take<u64>(c) : nullptr;
        lv.mask[L] = A.take<u32>(c);
        lv.cstart[L] = (L > 2) ? A.take<u32>(c) : nullptr;
        lv.rank[L] = A.take<u32>(c);
        lv.uniq[L] = A.take<u32>(c);
        lv.off[L] = A.take<u32>(c);
    }
    u64* leafm = A.take<u64>(c2 * 8);
    u32* runhead = A.take<u32>(c2);
    u32* colbfs = A.take<u32>(c2);
    u32* clist = A.take<u32>(ncol ? ncol : 1);
    u32* nclist = A.take<u32>(1);
    u32* rep = A.take<u32>(cmax);
    u32* tslot = A.take<u32>(cmax);
    u32* segs = tslot;
    u64* tab = use_table ? A.take<u64>(S) : nullptr;
    u32 *skey = nullptr, *sval = nullptr, *ktmp = nullptr, *vtmp = nullptr, *sort_scratch = nullptr;
    if (!use_table) {
        skey = A.take<u32>(cmax);
        sval = A.take<u32>(cmax);
        ktmp = A.take<u32>(cmax);
        vtmp = A.take<u32>(cmax);
        sort_scratch = A.take<u32>(sort_scratch_words(cmax));
    }
    u32* scan_scratch = A.take<u32>(scan_scratch_words(scan_n));
    if (A.used > A.cap) {
        set_error("forest: scratch layout exceeds its reservation (%zu > %zu)", A.used, A.cap);
        return CPVS_EINTERNAL;
    }
    CPVS_HIP_CHECK(hipMemcpyAsync(seg, st.seg, sizeof(u32) * M * kKL, hipMemcpyHostToDevice, s));
    CPVS_HIP_CHECK(hipMemcpyAsync(nxt, st.nxt, sizeof(u32) * M * kKL, hipMemcpyHostToDevice, s));

    // --- the level chain (optionally captured) ----------------------------------------------------
    hipGraph_t graph = nullptr;
    const bool capture = o.graph && !t_prof;  // per-kernel timing records events: launch directly then
    if (capture) CPVS_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    {
        if (ncol)
            device_scan(ColPairIn{colinfo}, ColListOut{colinfo, clist, nclist, (u32)ncol}, (u32)ncol, OpSum64{}, nullptr,
                        scan_scratch, s);
        CPVS_LAUNCH("forest_roots", k_forest_roots, grid_of(std::max<u64>(M, kMetaWords), 256), dim3(256), 0, s, items, M, seg,
                    lv, meta);
        // Top-down: classification and emission, level by level.
        for (int L = Rmax; L >= std::max(Lmin, 3); --L) {
            const u32 c = (u32)C[L];
            if (!c) continue;
            const bool leaf_kids = (L == 3);
            FEmitOut emit{lv.mask[L], lv.coord[L], lv.cstart[L], leaf_kids ? nullptr : lv.coord[L - 1],
                          (u32)(leaf_kids ? C[2] : E[L - 1]), items, M, colinfo, leaf_kids ? colbfs : nullptr};
            if (scan_prefetches(c)) {
                const FClassify cls{lv.coord[L], items, M, L};
                emit.mask_out = lv.mask[L];
                emit.cls = cls;
                device_scan(FClassifyCountIn{cls}, emit, c, OpSum{}, meta + kMetaC + L, scan_scratch, s);
            } else {
                CPVS_LAUNCH("forest_classify", k_forest_classify, grid_of(c, 256), dim3(256), 0, s, lv.coord[L], c, items, M,
                            L, lv.mask[L]);
                device_scan(ChildCountIn{lv.mask[L]}, emit, c, OpSum{}, meta + kMetaC + L, scan_scratch, s);
            }
        }
        const u32 slots2 = use_table ? leaf_table_slots((u32)C[2], small_leaf) : 0u;
        if (C[2]) {
            const u32 bound = (u32)std::min<u64>(ncol, C[2]);
            launch_leaf(o, bound, s, colinfo, colbfs, items, (u32)C[2], lv.mask[2], leafm, runhead, tab, slots2,
                        clist, nclist, bound);
        }
        // Bottom-up merge over every level with nodes (roots included: a root is alone in its DAG).
        int prev = -1;  // the level merged before (its ranks are the next level's child ranks)
        for (int L = Lmin; L <= Rmax; ++L) {
            const u32 c = (u32)C[L];
            if (!c) {
                prev = -1;
                continue;
            }
            const bool is_leaf = (L == 2);
            const u32* child_rank = (L > 2 && C[L - 1]) ? lv.rank[L - 1] : nullptr;
            const u32* rh = is_leaf ? runhead : nullptr;
            u64* rpart = use_table ? reinterpret_cast<u64*>(scan_scratch) : nullptr;
            const u32 rrounds = scan_rounds(c);
            const u32 nzero = (rpart && rrounds > 1) ? ceil_div<u32>(c, rrounds * kScanThreads) : 0u;
            if (use_table) {
                const u32 Sl = is_leaf ? slots2 : table_slots(c);
                if (!is_leaf && prev < 0) CPVS_LAUNCH("table_clear", k_table_clear, grid_of(Sl, 256), dim3(256), 0, s, tab, Sl);
                if (is_leaf) {
                    CPVS_LAUNCH("insert_leaf", k_insert_leaf, dim3(ceil_div<u32>(c, kLeafWindow)), dim3(kLeafInsertThreads), 0, s,
                                lv.mask[2], leafm, runhead, c, tab, Sl, tslot, meta + kMetaFull, rpart, nzero,
                                ctx->leaf_hash_bits.load());
                    CPVS_LAUNCH("resolve", k_resolve<LeafEq>, grid_of(c, kNodeThreads), dim3(kNodeThreads), 0, s, tslot, tab,
                                runhead, c, LeafEq{lv.mask[2], leafm, nullptr, c}, rep, meta + kMetaColl, lv.mask[2], 1u,
                                rpart, rrounds);
                } else {
                    CPVS_LAUNCH("insert_inner", k_insert_inner, grid_of(c, kInsertBlock), dim3(kInsertBlock), 0, s, lv.mask[L],
                                lv.cstart[L], child_rank, c, tab, Sl, tslot, meta + kMetaFull, rpart, nzero,
                                ctx->inner_hash_bits.load());
                    CPVS_LAUNCH("resolve", k_resolve<InnerEq>, grid_of(c, kNodeThreads), dim3(kNodeThreads), 0, s, tslot, tab,
                                nullptr, c, InnerEq{lv.mask[L], lv.cstart[L], child_rank}, rep, meta + kMetaColl, lv.mask[L],
                                0u, rpart, rrounds);
                }
                device_scan_from_partials(FirstSizeIn{rep, lv.mask[L], is_leaf ? 1u : 0u},
                                          FirstSizeOut{lv.rank[L], lv.uniq[L], lv.off[L], meta + kMetaU + L, meta + kMetaS + L,
                                                       c},
                                          c, OpSum64{}, nullptr, rpart, s);
            } else {
                if (is_leaf)
                    CPVS_LAUNCH("sortkeys", k_sortkeys_leaf, grid_of(c, 256), dim3(256), 0, s, lv.mask[2], leafm, runhead, c,
                                skey, sval);
                else
                    CPVS_LAUNCH("sortkeys", k_sortkeys_inner, grid_of(c, 256), dim3(256), 0, s, lv.mask[L], lv.cstart[L],
                                child_rank, c, skey, sval);
                radix_sort_pairs(skey, sval, ktmp, vtmp, c, 32, sort_scratch, s);
                device_scan(SegHeadIn{skey}, SegStartOut{segs}, c, OpMax{}, nullptr, scan_scratch, s);
                if (is_leaf)
                    CPVS_LAUNCH("representative", k_representative<LeafEq>, grid_of(c, 256), dim3(256), 0, s, sval, segs, c,
                                LeafEq{lv.mask[2], leafm, runhead, c}, rep, meta + kMetaSortColl);
                else
                    CPVS_LAUNCH("representative", k_representative<InnerEq>, grid_of(c, 256), dim3(256), 0, s, sval, segs, c,
                                InnerEq{lv.mask[L], lv.cstart[L], child_rank}, rep, meta + kMetaSortColl);
                rh = nullptr;
                device_scan(FirstSizeIn{rep, lv.mask[L], is_leaf ? 1u : 0u},
                            FirstSizeOut{lv.rank[L], lv.uniq[L], lv.off[L], meta + kMetaU + L, meta + kMetaS + L, c}, c,
                            OpSum64{}, nullptr, scan_scratch, s);
            }
            const u32 S_next = (use_table && L + 1 <= Rmax && C[L + 1]) ? table_slots((u32)C[L + 1]) : 0u;
            CPVS_LAUNCH("rank_copies", k_rank_copies, grid_of(c, kNodeThreads), dim3(kNodeThreads), 0, s, rep, rh, c,
                        lv.rank[L], tab, S_next);
            prev = L;
        }
        FBasesOut ho;
        ho.meta_host = st.back;
        ho.total = st.back + kMetaWords;
        ho.root = ho.total + M;
        ho.uniq = ho.root + M;
        CPVS_LAUNCH("forest_bases", k_forest_bases, grid_of(M, 256), dim3(256), 0, s, items, M, lv, seg, nxt, boff, B, meta,
                    Lmin, ho);
    }
    hipError_t e = hipGetLastError();
    if (capture) {
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = ec;
        hipGraphExec_t exec = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphLaunch(exec, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    } else if (e == hipSuccess) {
        e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) return hip_fail(e, "forest level chain", __FILE__, __LINE__);

    const u32* hm = st.back;
    // Counts first: a level whose count exceeds its emitted children merged unwritten nodes too.
    for (int L = std::max(Lmin, 3); L <= Rmax; ++L) {
        if (!C[L]) continue;
        const u64 want = (L == 3) ? C[2] : E[L - 1];
        if (hm[kMetaC + L] != want) {
            if (!exact) return kRetryExact;  // cells under a NaN texel were counted: count exactly
            set_error("forest: level %d emitted %u children, expected %llu", L, hm[kMetaC + L], (unsigned long long)want);
            return CPVS_EINTERNAL;
        }
    }
    if (use_table && hm[kMetaFull] != 0) return small_leaf ? kRetryFull : kRetrySort;
    if (use_table && hm[kMetaColl] != 0) return kRetrySort;
    const u32* totals = st.back + kMetaWords;
    const u32* roots = totals + M;
    const u32* uq = roots + M;
    for (u32 f = 0; f < M; ++f)
        if (totals[f] >= 0xFFFFFFFEu) {
            set_error("forest: DAG %u exceeds 32-bit offsets", ord[f]);
            return CPVS_EINVAL;
        }

    // --- the DAGs' buffers and words --------------------------------------------------------------
    std::vector<u32*> bufs(M, nullptr);
    auto free_bufs = [&]() {
        for (u32* b : bufs) ctx->buffers.give(b);
    };
    for (u32 f = 0; f < M; ++f) {
        bufs[f] = static_cast<u32*>(ctx->buffers.take(sizeof(u32) * std::max<u32>(totals[f], 1)));
        if (!bufs[f]) {
            free_bufs();
            set_error("forest: hipMalloc of %u DAG words failed", totals[f]);
            return CPVS_ENOMEM;
        }
        st.ptrs[f] = bufs[f];
    }
    e = hipMemcpyAsync(dptrs, st.ptrs, sizeof(u32*) * M, hipMemcpyHostToDevice, s);
    FWriteArgs wa{};
    u32 wblocks = 0;
    for (int L = Rmax; L >= Lmin; --L) {
        const u32 U = hm[kMetaU + L];
        if (!C[L] || !U) continue;
        wa.level[wa.nlev] = L;
        wa.U[wa.nlev] = U;
        wa.first[wa.nlev++] = wblocks;
        wblocks += ceil_div<u32>(U, kNodeThreads);
    }
    wa.first[wa.nlev] = wblocks;
    if (e == hipSuccess && wblocks) {
        CPVS_LAUNCH("forest_write", k_forest_write, dim3(wblocks), dim3(kNodeThreads), 0, s, wa, lv, leafm, boff, B,
                    dptrs);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        free_bufs();
        return hip_fail(e, "forest write", __FILE__, __LINE__);
    }
    for (u32 f = 0; f < M; ++f) {
        const ForestIn& x = in[ord[f]];
        cpvs_dag* d = new cpvs_dag();
        d->ctx = ctx;
        d->n = x.p->n;
        d->words = totals[f];
        d->d_words = bufs[f];
        d->root_mask = roots[f];
        d->N = x.p->N;
        d->zTileIndex = x.t;
        d->zTileNum = x.T;
        d->merge_mode = merge_mode;
        d->key_collisions = hm[kMetaSortColl];
        d->leafmasks = true;
        const int R = (int)hi[f].R;
        for (int L = 0; L < kMaxLevels; ++L) {
            d->nodes[L] = (L >= 2 && L <= R) ? cnt[(size_t)f * kKL + L] : 0;
            d->uniques[L] = (L >= Lmin && L <= R) ? uq[(size_t)f * kKL + L] : 0;
        }
        out[ord[f]] = d;
    }
    return CPVS_OK;
}

}  // namespace

int build_forest(cpvs_ctx* ctx, Worker& w, const ForestIn* items, u32 M, const ForestOpts& o, cpvs_dag** out) {
    for (u32 i = 0; i < M; ++i) out[i] = nullptr;
    if (!M) return CPVS_OK;
    if (M > kMaxForest) return kForestTooLarge;
    if (!ctx->leafmasks.load()) {
        set_error("forest: forests build leafmask DAGs; the context builds without leafmasks");
        return CPVS_EINVAL;
    }
    for (u32 i = 0; i < M; ++i) {
        const cpvs_pyramid* p = items[i].p;
        if (p->n < 5 || p->N > 65536 || (u64)p->N * items[i].T > (1ull << 21) || items[i].t >= items[i].T) {
            set_error("forest: item %u (N=%u, z-tile %u of %u) is not a leafmask map within the limits", i, p->N,
                      items[i].t, items[i].T);
            return CPVS_EINVAL;
        }
        for (int L = 3; L < p->n; ++L)
            if (!p->lv[L]) {
                set_error("forest: item %u lacks pyramid level %d", i, L);
                return CPVS_EINTERNAL;
            }
    }
    // As build_dag: a full table, a key collision and counts above the emitted nodes each switch to
    // their fallback once.
    int mode = o.merge_mode, rc = CPVS_OK;
    bool small_leaf = true, exact = false;
    for (int attempt = 0; attempt < 4; ++attempt) {
        rc = forest_once(ctx, w, items, M, o, mode, small_leaf, exact, out);
        if (rc == kRetryFull && small_leaf) small_leaf = false;
        else if (rc == kRetrySort && mode != CPVS_MERGE_SORT) mode = CPVS_MERGE_SORT, small_leaf = false;
        else if (rc == kRetryExact && !exact) exact = true;
        else break;
    }
    if (rc == kRetryFull || rc == kRetrySort || rc == kRetryExact) {
        set_error("forest: no consistent build after every fallback (status %d)", rc);
        return CPVS_EINTERNAL;
    }
    return rc;
}

}  // namespace cpvs
