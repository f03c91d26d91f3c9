// build.hip -- CompressedShadow::create on the device (src/CompressedShadow.cpp:49-392).
//
// The reference builds an uncompressed 9/17-word octree on the host, merges equal subtrees
// level by level with a linear search, drops the holes and relayouts the array.  Here every
// octree level is a set of compact per-node arrays, processed as:
//
//   1. BFS classification, root -> level 3 (or 0): one thread per node reads its 2x2 pyramid
//      texels and writes the 16-bit child mask; an exclusive scan of the partial-child counts
//      gives each node's first child, and the scan's output stage writes the children's
//      coordinates in (parent, child index) order -- the reference's breadth-first order.
//   2. Leaves (level 2): each wave takes 8 of the 8x8 texel columns (level-3 texels), lane =
//      texel, one column at a time.  The leaf nodes of a column are its partial level-3 cells,
//      contiguous in z; a node equal to its z-neighbour (the common case along walls) is copied,
//      not recomputed.  The level-3 scan records where each (column, z) leaf sits in
//      breadth-first order.
//   3. Merge, bottom-up from the leaves to the root's children.  Key = (mask, partial leafmasks)
//      or (mask, merged ranks of the partial children).  Default (a): the key's 64-bit hash goes
//      into an open-addressing table whose 64-bit slots hold a hash tag and the smallest node
//      index with it, so every node finds the first occurrence of its key; the full key is
//      compared against it, and a collision re-runs the build with (b).  (b): a stable LSD radix
//      sort of (32-bit hash, node index), segment starts by a max-scan and a full-key compare
//      against the segment head (a mismatch falls back to a scan of the segment).  Either way a
//      scan of first-occurrence flags gives each node's rank = order of first occurrence, as
//      mergeLevel's linear search yields (src/CompressedShadowUtil.h:154-182).
//   4. Compress: per level a scan of unique node sizes (1+k words, 1+2k at the leaf level),
//      level bases root first, then one pass writes [mask, absolute child offsets] or
//      [mask, lo, hi, ...] (src/CompressedShadow.cpp:326-392).
#include <algorithm>
#include <cstdio>
#include <cstring>

#include "build_parts.hpp"

namespace cpvs {

__global__ __launch_bounds__(256) void k_classify(const u64* __restrict__ coord, u32 count, const float2* __restrict__ lv,
                                                  const float* __restrict__ depth, u32 N, int L, u32 T,
                                                  u32* __restrict__ mask_out) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    mask_out[i] = classify_node(coord[i], lv, depth, N, L, T);
}

// ---------------------------------------------------------------------------------------
// 2. Leaves: createChildmask1x1x8 + createLeafmask (src/CompressedShadowUtil.cpp:59-99).
//
// Slice s of leaf node z3 (depth z = 8*z3 + s) is lit at texel t iff float(z) + 0.5 <= d_t*H
// (absoluteVisible, H = N*T), i.e. iff u_t >= s with f_t = floor(d_t*H - 0.5) and
// u_t = f_t - 8*z3.  With one texel per lane, slice s's leafmask is B_s = ballot(u >= s): eight
// compares give the whole node.  The B_s are nested (B_0 >= B_1 >= ...), so the slices with
// B_s all ones (fully lit) come first, then the partial ones, then those with B_s = 0 (shadow).
// Everything past the compares is wave-uniform scalar work: the childmask, the key hash over
// (mask, partial leafmasks), and the lane placement of the leafmasks for the store.
// Node z3 equals node z3-1 of the same column unless some u_t lies in [-8, 6].
// ---------------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------------
// The leaf kernel: kLeafCols level-3 columns per wave, one texel per lane.
// * All loads of a wave's columns are issued together (column info, then the depth blocks and
//   the first 64 breadth-first indices of each column); the columns are then walked one after the
//   other with no memory wait.  A column of more than 64 nodes (box walls) loads the indices of
//   its later 64-node chunks in the loop.
// * Which nodes differ from their predecessor follows from the texels alone.  With
//   h_t = f_t + 8 - 8*zs, node q changes iff some h_t - 8q lies in [0, 14]: q = floor(h_t/8), or
//   q = floor(h_t/8) - 1 when h_t mod 8 != 7.  Each texel flags at most two nodes in LDS and a
//   ballot over the flags read back gives a chunk's changed nodes at once; only those (the run
//   heads, 1.3 M of 6.7 M leaves at 8192^2) get compares.
// * The kernel is bound by instruction issue: the per-run-head work is wave-uniform, and the
//   scalar unit is one per CU, shared by its four SIMDs.  What cut it: the changed nodes from LDS
//   flags and one ballot instead of two wave OR-reductions (0.262 -> 0.244 ms), lane inserts by
//   v_writelane and run heads propagated head by head (0.239 ms), the wave's columns walked by a
//   bitmask of those with nodes (0.224 ms), slice counts from per-chunk offsets (0.217 ms), and
//   simpler flag addressing (0.209 ms).  The vector units are now the busier (73%, scalar 57%).
// * Only a run's first node stores its mask and leafmasks: the merge and the final write read
//   those of first occurrences only, and a key's first occurrence is always the first node of a
//   run (breadth-first order grows with z inside a column).  Every node stores the breadth-first
//   index of its run's first node.
// * A run head's leafmasks go to the row at its own breadth-first index (8 slots per node), so the
//   leaf insert, the key compares and the write find them at the node's index, with no slot map.
// * Measured on MI355X (earlier forms of the kernel): 8 columns per wave 0.31 ms, 4 columns 0.32
//   ms, 16 columns 0.42 ms; one
//   column per one-wave workgroup (loads issued per column, run heads found by a ballot per node)
//   0.55 ms; one lane per node over LDS-staged columns 0.51 ms (64 texels of work per node).
//   Workgroups of 1, 2, 4 and 8 waves: 0.308, 0.315, 0.320 and 0.335 ms.
// Workgroups are dealt to the XCDs in turn (block b to XCD b mod kXcds), each XCD with its own L2.
// XCD x takes the contiguous units [x*unit_chunk, (x+1)*unit_chunk): neighbouring columns, whose
// 8-texel depth rows share 128-byte lines, are then read through one L2.
// ---------------------------------------------------------------------------------------

// kList (several z-tiles): the wave's columns are col_list[c0 .. c0+7], the z-tile's columns with
// leaf nodes in column order, *n_list of them; ncol (at most the leaf node count) bounds the grid.
// Otherwise the wave's columns are c0 .. c0+7 of all ncol.
template <bool kList>
__global__ __launch_bounds__(64 * kLeafWaves) void k_leaf_waves(LeafColumns lc, u32 ncol, u32 nleaf,
                                                                const float* __restrict__ depth, u32 N, u32 T,
                                                                const float2* __restrict__ lv3,
                                                                u32* __restrict__ mask_out, u64* __restrict__ leafm,
                                                                u32* __restrict__ runhead,
                                                                u32 unit_chunk,
                                                                u64* __restrict__ tab, u32 tab_slots,
                                                                const u32* __restrict__ col_list,
                                                                const u32* __restrict__ n_list) {
    // Per lane: f_t and the breadth-first index of node `lane` of each of the wave's columns (a lane
    // only ever reads back its own entries, so the arrays need no barrier).
    __shared__ int s_f[kLeafWaves][kLeafCols][64];
    __shared__ u32 s_b[kLeafWaves][kLeafCols][64];
    // A chunk's changed-node flags (leaf_column): byte j + 1 for node q0 + j, byte 65 a sink.
    __shared__ uint8_t s_chg[kLeafWaves][128];
    // The wave index is uniform across the wave; saying so keeps the column bases in scalar registers.
    const u32 lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Workgroups are dealt to the XCDs in turn; XCD x takes the contiguous units [x*unit_chunk, ...).
    // A column list fills only its first units, so there consecutive units go to different XCDs.
    const u32 unit = kList ? blockIdx.x : (blockIdx.x % kXcds) * unit_chunk + blockIdx.x / kXcds;
    const u32 c0 = (unit * kLeafWaves + wave) * kLeafCols;
    // Every load below is unconditional (addresses clamped into range, values masked afterwards), so
    // the compiler counts them exactly and waits only for what each step uses.
    const u32 nvalid = kList ? min(*n_list, ncol) : ncol;  // columns of this launch
    const u32 clast = kList ? (nvalid ? nvalid - 1 : 0u) : ncol - 1;
    // lane k < kLeafCols: column c0 + k of the launch (list entries past nvalid are never read)
    const u32 my_slot = min(c0 + (lane & (kLeafCols - 1)), clast);
    const u32 my_col = kList ? (nvalid ? col_list[my_slot] : 0u) : my_slot;
    uint4 ci = lc.info[my_col];
    if (lane >= kLeafCols || c0 + lane >= nvalid) ci = make_uint4(0u, 0u, 0u, 0u);
    // Addresses are a wave-uniform base (the column's block, its run of indices) plus a per-lane
    // offset computed once, so every load takes its address from scalar registers and one vector
    // offset, with no per-column vector arithmetic.
    // Byte offset of texel (lane & 7, lane >> 3) of a column's block: below 2^32 (N <= 2^21), so a
    // load takes it as its 32-bit vector offset from the block's scalar base.
    const u32 loff = 4u * ((lane >> 3) * N + (lane & 7));
    float d[kLeafCols];
    // Without a list, a wave's columns are c0 .. c0+7 of one row of level-3 texels when the row
    // holds a multiple of 8 columns: their blocks are 32 bytes apart, one base and immediate offsets.
    const bool row_run = !kList && lc.w3 >= kLeafCols && c0 + kLeafCols <= ncol;
    if (row_run) {
        const u32 x3 = c0 & (lc.w3 - 1), y3 = c0 >> lc.w3_shift;
        const char* blk = reinterpret_cast<const char*>(depth + ((size_t)(8 * y3) * N + 8 * x3)) + loff;
#pragma unroll
        for (u32 k = 0; k < kLeafCols; ++k) {
            const float* p = reinterpret_cast<const float*>(blk + 32 * k);
            if (T > 1) {
                d[k] = 0.0f;
                if (__builtin_amdgcn_readlane(ci.x, k) != 0u) d[k] = *p;
            } else {
                d[k] = __builtin_nontemporal_load(p);
            }
        }
    } else if (T > 1) {
        // A column with no leaf node in this z-tile (most columns of a tile) is not read: its f
        // values are never used.  With one z-tile every column has nodes, and the check would only
        // cost issue slots.
#pragma unroll
        for (u32 k = 0; k < kLeafCols; ++k) {
            const u32 c = kList ? __builtin_amdgcn_readlane(my_col, k) : min(c0 + k, clast);
            const u32 x3 = c & (lc.w3 - 1), y3 = c >> lc.w3_shift;
            const float* blk = depth + ((size_t)(8 * y3) * N + 8 * x3);
            d[k] = 0.0f;
            if (__builtin_amdgcn_readlane(ci.x, k) != 0u) d[k] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(blk) + loff);
        }
    } else {
#pragma unroll
        for (u32 k = 0; k < kLeafCols; ++k) {
            const u32 c = kList ? __builtin_amdgcn_readlane(my_col, k) : min(c0 + k, clast);
            const u32 x3 = c & (lc.w3 - 1), y3 = c >> lc.w3_shift;
            const float* blk = depth + ((size_t)(8 * y3) * N + 8 * x3);
            // Nontemporal: the map's last read (the merge kernels that follow keep more of the L2).
            d[k] = __builtin_nontemporal_load(reinterpret_cast<const float*>(reinterpret_cast<const char*>(blk) + loff));
        }
    }
    u32 b[kLeafCols];
#pragma unroll
    for (u32 k = 0; k < kLeafCols; ++k) {
        const u32 cnt = __builtin_amdgcn_readlane(ci.x, k), st = __builtin_amdgcn_readlane(ci.y, k);
        const u32* run = lc.bfs + (cnt ? st : 0u);  // st + lane < nleaf for lane < cnt: the counts sum to it
        b[k] = run[(lane < cnt) ? lane : 0u];
    }
    const float HT = (float)(N * T);
    // The columns' minimum and maximum f from their level-3 pyramid texels (the 8x8 blocks' depth
    // minimum and maximum): lane k < kLeafCols holds column k's, lane k + kLeafCols its maximum.
    // A block holding a NaN takes them from its texels instead (the pyramid keeps a NaN only as a
    // first operand, so its extremes do not tell).
    const float2 ext = lv3[my_col];
    const int fext = leaf_f(((lane & kLeafCols) ? ext.y : ext.x) * HT);
    u32 nan_cols = 0;
#pragma unroll
    for (u32 k = 0; k < kLeafCols; ++k) {
        const float D = d[k] * HT;
        if (__ballot(!(D == D))) nan_cols |= 1u << k;
        s_f[wave][k][lane] = leaf_f(D);
        s_b[wave][k][lane] = b[k];
    }

    // The wave's columns with leaf nodes (ci is zero for lanes past the launch's columns).
    for (u32 todo = (u32)__ballot(ci.x != 0u); todo; todo &= todo - 1u) {
        const u32 k = (u32)__builtin_ctz(todo);
        const u32 cnt = __builtin_amdgcn_readlane(ci.x, k);
        const u32 base = __builtin_amdgcn_readlane(ci.y, k);
        const int zs = (int)__builtin_amdgcn_readlane(ci.z, k);
        const int f = s_f[wave][k][lane];
        int fmin = __builtin_amdgcn_readlane(fext, k), fmax = __builtin_amdgcn_readlane(fext, k + kLeafCols);
        if ((nan_cols >> k) & 1u) {
            fmin = __ockl_wfred_min_i32(f);
            fmax = __ockl_wfred_max_i32(f);
        }
        // (with one wave per workgroup its flags sit at a constant LDS address)
        leaf_column(f, fmin, fmax, zs, cnt, s_b[wave][k][lane], lc.bfs + base, 0u, nleaf,
                    kLeafWaves == 1 ? s_chg[0] : s_chg[wave], lane, mask_out, leafm, runhead);
    }
    // The leaf level's merge table is cleared here, after all the work (stores issued earlier
    // would be waited for with the loads): this kernel leaves HBM bandwidth to spare.
    clear_table(tab, tab_slots);
}

// ---------------------------------------------------------------------------------------
// 3. Merge.
// ---------------------------------------------------------------------------------------
// --- (a) hash table --------------------------------------------------------------------------

// One 64-bit word per slot: the key hash's high 32 bits (its tag) above the smallest node index
// with that hash, so a probe touches one cache line and a new key costs one atomic.  ~0 marks an
// empty slot (node indices stay below 2^32 - 1).  Two keys whose tags are equal and whose probe
// chains meet share a slot; the full-key compare in k_resolve then sends the build down the
// sorting path, as for equal 64-bit hashes.

__global__ __launch_bounds__(256) void k_table_clear(u64* __restrict__ tab, u32 S) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S) tab[i] = kEmptySlot;
}

__global__ __launch_bounds__(kLeafInsertThreads) void k_insert_leaf(const u32* __restrict__ mask, const u64* __restrict__ leafm,
                                                                    const u32* __restrict__ runhead, u32 n,
                                                                    u64* __restrict__ tab, u32 S, u32* __restrict__ tslot,
                                                                    u32* __restrict__ full, u64* __restrict__ zero,
                                                                    u32 nzero, u32 hbits) {
    __shared__ u64 gh[kLeafSlots];
    __shared__ u32 gi[kLeafSlots];  // smallest node index of the slot's hash, then its global slot
    __shared__ unsigned short glist[kLeafWindow];  // LDS slots in use
    __shared__ u32 gcount;
    clear_partials(zero, nzero);
    for (u32 j = threadIdx.x; j < kLeafSlots; j += blockDim.x) {
        gh[j] = 0ull;  // 0 marks an empty LDS slot
        gi[j] = kNoRep;
    }
    if (threadIdx.x == 0) gcount = 0;
    const u32 i0 = blockIdx.x * kLeafWindow + threadIdx.x;
    u32 rh[kLeafPer];
#pragma unroll
    for (u32 k = 0; k < kLeafPer; ++k) {
        const u32 i = i0 + k * kLeafInsertThreads;
        rh[k] = (i < n) ? runhead[i] : kNoRep;
    }
    // The key hash of each run's first node (the run's smallest index), from its mask and
    // leafmasks (its row at its own index): computed here, lane-parallel, rather than wave-uniform
    // in the leaf kernel.
    u32 hm[kLeafPer];
#pragma unroll
    for (u32 k = 0; k < kLeafPer; ++k) {
        const u32 i = i0 + k * kLeafInsertThreads;
        hm[k] = (rh[k] == i) ? mask[i] : 0u;
    }
    u64 h[kLeafPer];
#pragma unroll
    for (u32 k = 0; k < kLeafPer; ++k) {
        const u32 i = i0 + k * kLeafInsertThreads;
        h[k] = (rh[k] == i) ? weaken_hash(leaf_key_hash(hm[k], leafm + (size_t)i * 8), hbits) : 0ull;
        if (h[k] == 0ull && rh[k] == i) h[k] = 1ull;
    }
    __syncthreads();
    u32 ls[kLeafPer];
#pragma unroll
    for (u32 k = 0; k < kLeafPer; ++k) {
        const u32 i = i0 + k * kLeafInsertThreads;
        ls[k] = 0;
        if (rh[k] != i) continue;
        u32 sl = (u32)(h[k] ^ (h[k] >> 29)) & (kLeafSlots - 1);
        for (;;) {
            const u64 cur = atomicCAS(reinterpret_cast<unsigned long long*>(gh + sl), 0ull, (unsigned long long)h[k]);
            if (cur == 0ull) {
                glist[atomicAdd(&gcount, 1u)] = (unsigned short)sl;
                break;
            }
            if (cur == h[k]) break;
            sl = (sl + 1) & (kLeafSlots - 1);
        }
        atomicMin(gi + sl, i);
        ls[k] = sl;
    }
    __syncthreads();
    const u32 m = gcount;  // at most kLeafWindow
    for (u32 j = threadIdx.x; j < m; j += blockDim.x) {
        const u32 sl = glist[j];
        gi[sl] = table_insert(tab, S, gh[sl], gi[sl], full);
    }
    __syncthreads();
#pragma unroll
    for (u32 k = 0; k < kLeafPer; ++k) {
        const u32 i = i0 + k * kLeafInsertThreads;
        if (rh[k] == i) tslot[i] = gi[ls[k]];
    }
}

__global__ __launch_bounds__(kInsertBlock) void k_insert_inner(const u32* __restrict__ mask, const u32* __restrict__ cstart,
                                                               const u32* __restrict__ child_rank, u32 n, u64* __restrict__ tab,
                                                               u32 S, u32* __restrict__ tslot, u32* __restrict__ full,
                                                               u64* __restrict__ zero, u32 nzero, u32 hbits) {
    clear_partials(zero, nzero);
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < n;
    const u64 h = act ? weaken_hash(inner_key_hash(mask, cstart, child_rank, i), hbits) : 0ull;
    const u32 slot = table_insert_group(tab, S, h, i, act, full);
    if (act) tslot[i] = slot;
}

// --- (b) radix sort ----------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_sortkeys_leaf(const u32* __restrict__ mask, const u64* __restrict__ leafm,
                                                       const u32* __restrict__ runhead,
                                                       u32 n, u32* __restrict__ skey, u32* __restrict__ sval) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u32 h = runhead[i];  // the key is the run's first node's
    h = (h < n) ? h : i;
    skey[i] = fold32(leaf_key_hash(mask[h], leafm + (size_t)h * 8));
    sval[i] = i;
}

__global__ __launch_bounds__(256) void k_sortkeys_inner(const u32* __restrict__ mask, const u32* __restrict__ cstart,
                                                        const u32* __restrict__ child_rank, u32 n, u32* __restrict__ skey,
                                                        u32* __restrict__ sval) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    skey[i] = fold32(inner_key_hash(mask, cstart, child_rank, i));
    sval[i] = i;
}

// --- ranks -------------------------------------------------------------------------------------

// Copies take their first occurrence's rank.  The merge table of the next level (tab_slots > 0)
// is cleared here too: this level's last table read (k_resolve) ran before.
__global__ __launch_bounds__(kNodeThreads) void k_rank_copies(const u32* __restrict__ rep, const u32* __restrict__ runhead, u32 count,
                                                     u32* __restrict__ rank, u64* __restrict__ tab, u32 tab_slots) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        u32 r = rep[i];
        const u32 h = runhead ? runhead[i] : i;  // loaded with rep[i]: most leaf nodes need it
        if (r == kNoRep) r = (h < count) ? rep[h] : kNoRep;
        if (r != i && r < count) rank[i] = rank[r];
    }
    clear_table(tab, tab_slots);
}

__global__ void k_root_init(u32* rank, u32* uniq, u32* n_uniq) {
    rank[0] = 0;
    uniq[0] = 0;
    *n_uniq = 1;
}

// First launch after the scratch is laid out when no level is small: the root's coordinate, and
// the meta words zeroed (k_small_topdown does both otherwise).
__global__ void k_start(u64* root_coord, u64 root, u32* meta, u32 meta_words) {
    *root_coord = root;
    for (u32 i = 0; i < meta_words; ++i) meta[i] = 0u;
}

// ---------------------------------------------------------------------------------------
// 4. Compress.
// ---------------------------------------------------------------------------------------
// B[R] = 0, B[L-1] = B[L] + S[L]; total = B[bottom] + S[bottom] (levels absent have S = 0).  One
// thread; every S is loaded before any base is stored (stored one by one behind each load, the
// loads would wait for one another).
__device__ inline void level_bases(u32* meta, int R, int bottom) {
    u32 S[kMaxLevels];
#pragma unroll
    for (int L = 0; L < kMaxLevels; ++L) S[L] = meta[kMetaS + L];
    unsigned long long acc = 0;
#pragma unroll
    for (int L = kMaxLevels - 1; L >= 0; --L) {
        if (L <= R && L >= bottom) {
            meta[kMetaB + L] = (u32)acc;
            acc += S[L];
        }
    }
    meta[kMetaW] = (acc > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (u32)acc;
}

// The build's readback: the metadata block copied to pinned host memory by the whole workgroup,
// one word per thread (after a barrier, so that thread 0's bases are in).
__device__ inline void meta_to_host(const u32* meta, u32* host_meta) {
    __syncthreads();
    for (u32 i = threadIdx.x; i < kMetaWords; i += blockDim.x) host_meta[i] = meta[i];
    __threadfence_system();
}

__global__ __launch_bounds__(128) void k_level_bases(u32* __restrict__ meta, int R, int bottom,
                                                     const u32* __restrict__ root_mask, u32* __restrict__ host_meta) {
    if (threadIdx.x == 0) {
        meta[kMetaRoot] = *root_mask;
        level_bases(meta, R, bottom);
    }
    meta_to_host(meta, host_meta);
}

// Every level's words in one launch: once the offsets and bases are known the levels' words are
// independent, and a launch per level costs a launch each.  Level k takes the blocks
// [first[k], first[k+1]), one unique node per thread.
struct WriteLevel {
    const u32* uniq;
    const u32* mask;
    const u32* cstart;
    const u32* off;
    const u32* child_rank;  // null at the deepest level
    const u32* child_off;
    u32 child_n;            // nodes of the child level
    u32 U, base, child_base, leaf;
    int L;                  // the octree level (k_write_levels_dev reads U and the bases of meta)
};
struct WriteArgs {
    int nlev;
    u32 first[kMaxLevels + 1];
    WriteLevel lv[kMaxLevels];
    const u64* leafm;
    u32 nleaf;  // nodes of the leaf level (at least 1)
};

__global__ __launch_bounds__(kNodeThreads) void k_write_levels(WriteArgs a, u32* __restrict__ out) {
    int k = 0;
    while (k + 1 < a.nlev && blockIdx.x >= a.first[k + 1]) ++k;
    const WriteLevel& l = a.lv[k];
    const u32 r = (blockIdx.x - a.first[k]) * blockDim.x + threadIdx.x;
    if (r >= l.U) return;
    if (l.leaf)
        write_leaf_node(r, l.uniq, l.mask, a.leafm, a.nleaf, l.off, l.base, out);
    else
        write_inner_node(r, l.uniq, l.mask, l.cstart, l.off, l.base, l.child_rank, l.child_off, l.child_base, l.child_n,
                         out);
}

// The same with the unique counts and level bases read from the metadata block, so that it runs
// right behind the merge with no host readback between: a grid-stride loop over the levels'
// blocks (their number is known on the device only), the words into a staging area of the
// pre-merge word bound.
__global__ __launch_bounds__(kNodeThreads) void k_write_levels_dev(WriteArgs a, const u32* __restrict__ meta,
                                                                  u32* __restrict__ out, u64 cap) {
    __shared__ u32 s_first[kMaxLevels + 1], s_U[kMaxLevels], s_base[kMaxLevels], s_cbase[kMaxLevels];
    if (threadIdx.x == 0) {
        u32 acc = 0;
        for (int k = 0; k < a.nlev; ++k) {
            const int L = a.lv[k].L;
            s_U[k] = meta[kMetaU + L];
            s_base[k] = meta[kMetaB + L];
            s_cbase[k] = a.lv[k].child_rank ? meta[kMetaB + L - 1] : 0u;
            s_first[k] = acc;
            acc += (s_U[k] + kNodeThreads - 1) / kNodeThreads;
        }
        s_first[a.nlev] = acc;
    }
    __syncthreads();
    const u32 nb = s_first[a.nlev];
    for (u32 vb = blockIdx.x; vb < nb; vb += gridDim.x) {
        int k = 0;
        while (k + 1 < a.nlev && vb >= s_first[k + 1]) ++k;
        const WriteLevel& l = a.lv[k];
        const u32 r = (vb - s_first[k]) * blockDim.x + threadIdx.x;
        if (r >= s_U[k]) continue;
        // Offsets lie below the word count, itself below the staging's pre-merge bound; a node
        // that would not fit (only if the build were inconsistent, which the host then reports)
        // is not written.
        if ((u64)s_base[k] + l.off[r] + 17 > cap) continue;
        if (l.leaf)
            write_leaf_node(r, l.uniq, l.mask, a.leafm, a.nleaf, l.off, s_base[k], out);
        else
            write_inner_node(r, l.uniq, l.mask, l.cstart, l.off, s_base[k], l.child_rank, l.child_off, s_cbase[k],
                             l.child_n, out);
    }
}

// ---------------------------------------------------------------------------------------
// Host orchestration.
// ---------------------------------------------------------------------------------------
namespace {

struct Level {
    u32 count = 0;
    u64* coord = nullptr;
    u32* mask = nullptr;
    u32* cstart = nullptr;
    u32* rank = nullptr;
    u32* uniq = nullptr;
    u32* off = nullptr;
};

// ---------------------------------------------------------------------------------------
// Small levels in one workgroup.  The upper octree levels hold few nodes (at 8192^2 the levels
// above 6 hold a few thousand together), yet each costs about ten launches of the per-level
// path at several microseconds apiece.  A level of at most kSmallMax nodes is handled by one
// 1024-thread workgroup instead, all such levels in one launch per phase: top-down
// classification and child emission, bottom-up merge (hash table in LDS), and compress
// (size scans, level bases, words).  Results are those of the per-level path.
// ---------------------------------------------------------------------------------------
// Measured at 8192^2: a limit of 4096 nodes beats 2048 (1.29 against 1.31 ms per build) and 8192
// (1.34 ms: one workgroup then spends more on the 7.9 K-node level than its per-level launches do).
constexpr u32 kSmallMax = 4096;
constexpr int kSmallThreads = 1024;
constexpr u32 kSmallPer = kSmallMax / kSmallThreads;  // nodes per thread
constexpr u32 kSmallSlots = 2 * kSmallMax;             // LDS table of 8-byte slots: 64 KB

struct SmallArgs {
    int R, Ls, deepest;  // the small levels are [Ls, R]; deepest = lowest level with nodes
    u32 count[kMaxLevels];
    u64* coord[kMaxLevels];
    u32* mask[kMaxLevels];
    u32* cstart[kMaxLevels];
    u32* rank[kMaxLevels];
    u32* uniq[kMaxLevels];
    u32* off[kMaxLevels];
    const float2* lv[kMaxLevels];
    u32* meta;
    u32* host_meta;  // pinned host copy of meta, written when the compress step is done
    u32 sizes_in_merge;  // every larger level has its word offsets: k_small_merge finishes compress
};

// Exclusive sum over the workgroup, one value per thread; *total = the sum.
__device__ inline u32 small_scan(u32 v, u32* s_wave, u32* total) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const u32 inc = wave_inclusive_scan(v, OpSum{});
    if (lane == kWave - 1) s_wave[wid] = inc;
    __syncthreads();
    u32 wpre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kSmallThreads / kWave; ++w) {
        const u32 x = s_wave[w];
        if (w < wid) wpre += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return wpre + inc - v;
}

// Thread t owns nodes [t*E, t*E + E) of a level of c nodes (contiguous, so scans keep order).
__device__ inline u32 small_per_thread(u32 c) { return (c + kSmallThreads - 1) / kSmallThreads; }

// The small levels' cells stay in LDS (two buffers: a level and its children); only the children
// of the last small level, which the multi-workgroup path classifies, go to memory.  Each node's
// mask and cell are kept in registers from its classification to its emission.
__global__ __launch_bounds__(kSmallThreads) void k_small_topdown(SmallArgs a, const float* __restrict__ depth, u32 N, u32 T,
                                                                 u64 root, LeafColumns cols) {
    __shared__ u32 s_wave[kSmallThreads / kWave];
    __shared__ u64 s_coord[2][kSmallMax];
    if (threadIdx.x == 0) s_coord[a.R & 1][0] = root;  // create(): root offset (0, 0, 2*zTileIndex)
    for (u32 i = threadIdx.x; i < kMetaWords; i += kSmallThreads) a.meta[i] = 0u;  // the first launch to use them
    __syncthreads();
    for (int L = a.R; L >= a.Ls; --L) {
        const u32 c = a.count[L];  // at most kSmallMax
        const u32 E = small_per_thread(c), i0 = threadIdx.x * E;
        const u64* cell = s_coord[L & 1];
        u32 k[kSmallPer], m[kSmallPer], sum = 0;
        u64 cc[kSmallPer];
#pragma unroll
        for (u32 e = 0; e < kSmallPer; ++e) {
            k[e] = 0;
            m[e] = 0;
            cc[e] = 0;
            const u32 i = i0 + e;
            if (e < E && i < c) {
                cc[e] = cell[i];
                m[e] = classify_node(cc[e], a.lv[L], depth, N, L, T);
                a.mask[L][i] = m[e];
                k[e] = num_children(m[e]);
                sum += k[e];
            }
        }
        u32 total;
        u32 first = small_scan(sum, s_wave, &total);
        // Children on a small level (count at most kSmallMax) go to the other LDS buffer.
        u64* child = (L - 1 >= a.Ls) ? s_coord[(L - 1) & 1] : a.coord[L - 1];
        EmitChildrenOut emit{a.mask[L], nullptr, a.cstart[L], child, a.count[L - 1], (L == 3) ? cols : LeafColumns{}};
#pragma unroll
        for (u32 e = 0; e < kSmallPer; ++e) {
            const u32 i = i0 + e;
            if (e < E && i < c) {
                emit(i, first, k[e], EmitChildrenOut::Pre{m[e], cc[e]});
                first += k[e];
            }
        }
        if (threadIdx.x == 0) a.meta[kMetaC + L] = total;
        __syncthreads();  // the child level's coordinates are read next
    }
}

// The LDS table of the small levels: one 64-bit word per slot, as in the global table.
__device__ inline u32 lds_insert(u64* tab, u32 S, u64 h, u32 i) {
    const u32 tag = (u32)(h >> 32);
    const u64 mine = ((u64)tag << 32) | i;
    u32 slot = (u32)(h ^ (h >> 31)) & (S - 1);
    for (;;) {
        const u64 cur = atomicCAS(reinterpret_cast<unsigned long long*>(tab + slot), (unsigned long long)kEmptySlot,
                                  (unsigned long long)mine);
        if (cur == kEmptySlot) return slot;
        if ((u32)(cur >> 32) == tag) {
            atomicMin(reinterpret_cast<unsigned long long*>(tab + slot), (unsigned long long)mine);
            return slot;
        }
        slot = (slot + 1) & (S - 1);
    }
}

// Merge of the small levels Ls .. R-1 (the root is never merged), then the root's rank.  A node
// takes the smallest index with its slot's tag; if that node's key differs (a collision of tags
// or of 64-bit hashes), it looks for its first occurrence by comparing keys with every lower index.
// The small levels' ranks are also kept in LDS (two buffers: a level and its children), where the
// next level's keys, key compares and the rank copies read them.  A level's masks and child starts
// are loaded while the level below is merged, and staged in LDS for the keys and compares.
__global__ __launch_bounds__(kSmallThreads) void k_small_merge(SmallArgs a) {
    __shared__ u64 tab[kSmallSlots];
    __shared__ u32 s_wave[kSmallThreads / kWave];
    __shared__ u32 s_rank[2][kSmallMax];
    __shared__ u32 s_mask[kSmallMax];
    __shared__ u32 s_cstart[kSmallMax];
    u32 pm[kSmallPer], pc[kSmallPer];  // masks and child starts of the next level's nodes
    auto prefetch = [&](int L) {
        const u32 c = a.count[L], E = small_per_thread(c), i0 = threadIdx.x * E;
#pragma unroll
        for (u32 e = 0; e < kSmallPer; ++e) {
            const u32 i = i0 + e;
            pm[e] = 0;
            pc[e] = 0;
            if (e < E && i < c) {
                pm[e] = a.mask[L][i];
                pc[e] = a.cstart[L][i];
            }
        }
    };
    if (a.Ls < a.R) prefetch(a.Ls);
    for (int L = a.Ls; L < a.R; ++L) {
        const u32 c = a.count[L];  // at most kSmallMax
        u32 S = 64;
        while (S < 2 * c) S <<= 1;
        for (u32 j = threadIdx.x; j < S; j += kSmallThreads) tab[j] = kEmptySlot;
        {
            const u32 E = small_per_thread(c), i0 = threadIdx.x * E;
#pragma unroll
            for (u32 e = 0; e < kSmallPer; ++e) {
                const u32 i = i0 + e;
                if (e < E && i < c) {
                    s_mask[i] = pm[e];
                    s_cstart[i] = pc[e];
                }
            }
        }
        __syncthreads();
        if (L + 1 < a.R) prefetch(L + 1);
        // The first small level's children were ranked by the multi-workgroup path, in memory.
        const u32* child_rank = (L > a.deepest) ? ((L > a.Ls) ? s_rank[(L - 1) & 1] : a.rank[L - 1]) : nullptr;
        u32* rank_l = s_rank[L & 1];
        const u32 E = small_per_thread(c), i0 = threadIdx.x * E;
        u32 slot[kSmallPer];
#pragma unroll
        for (u32 e = 0; e < kSmallPer; ++e) {
            const u32 i = i0 + e;
            slot[e] = 0;
            if (e < E && i < c) slot[e] = lds_insert(tab, S, inner_key_hash(s_mask, s_cstart, child_rank, i), i);
        }
        __syncthreads();
        const InnerEq eq{s_mask, s_cstart, child_rank};
        // One scan of (first occurrences << 16 | their words) gives ranks and word offsets (a level
        // of at most kSmallMax nodes of at most 9 words keeps the low half below 2^16).
        u32 rep[kSmallPer], sz[kSmallPer], packed = 0;
#pragma unroll
        for (u32 e = 0; e < kSmallPer; ++e) {
            const u32 i = i0 + e;
            rep[e] = i;
            sz[e] = 0;
            if (e < E && i < c) {
                u32 r = (u32)tab[slot[e]];
                if (r != i && !eq(i, r)) {
                    r = i;
                    for (u32 j = 0; j < i; ++j)
                        if (eq(i, j)) {
                            r = j;
                            break;
                        }
                }
                rep[e] = r;
                if (r == i) {
                    sz[e] = 1u + num_children(s_mask[i]);
                    packed += (1u << 16) + sz[e];
                }
            }
        }
        u32 total;
        const u32 ex = small_scan(packed, s_wave, &total);
        u32 rk = ex >> 16, of = ex & 0xFFFFu;
#pragma unroll
        for (u32 e = 0; e < kSmallPer; ++e) {
            const u32 i = i0 + e;
            if (e < E && i < c && rep[e] == i) {
                a.rank[L][i] = rk;
                rank_l[i] = rk;
                a.uniq[L][rk] = i;
                if (a.sizes_in_merge) a.off[L][rk] = of;
                ++rk;
                of += sz[e];
            }
        }
        if (threadIdx.x == 0) {
            a.meta[kMetaU + L] = total >> 16;
            if (a.sizes_in_merge) a.meta[kMetaS + L] = total & 0xFFFFu;
        }
        __syncthreads();  // first occurrences ranked
#pragma unroll
        for (u32 e = 0; e < kSmallPer; ++e) {
            const u32 i = i0 + e;
            if (e < E && i < c && rep[e] != i) {
                const u32 r = rank_l[rep[e]];
                a.rank[L][i] = r;
                rank_l[i] = r;
            }
        }
        __syncthreads();  // the next level reads these ranks
    }
    if (threadIdx.x == 0) {
        a.rank[a.R][0] = 0;
        a.uniq[a.R][0] = 0;
        a.meta[kMetaU + a.R] = 1;
        if (a.sizes_in_merge) {
            // What k_small_compress does otherwise: the root's size, the level bases, the readback.
            const u32 root = a.mask[a.R][0];
            a.off[a.R][0] = 0;
            a.meta[kMetaS + a.R] = 1u + num_children(root);
            a.meta[kMetaRoot] = root;
            level_bases(a.meta, a.R, a.deepest);
        }
    }
    if (a.sizes_in_merge) meta_to_host(a.meta, a.host_meta);  // uniform: every thread reaches the barrier
}

// Size scans of the small levels (1 + k words per unique node), then the level bases of all
// levels (the per-level scans of the larger levels ran before).
__global__ __launch_bounds__(kSmallThreads) void k_small_compress(SmallArgs a) {
    __shared__ u32 s_wave[kSmallThreads / kWave];
    for (int L = a.R; L >= a.Ls; --L) {
        const u32 U = a.meta[kMetaU + L];
        const u32 E = small_per_thread(U), i0 = threadIdx.x * E;
        u32 sz[kSmallPer], sum = 0;
#pragma unroll
        for (u32 e = 0; e < kSmallPer; ++e) {
            const u32 r = i0 + e;
            sz[e] = (e < E && r < U) ? 1u + num_children(a.mask[L][a.uniq[L][r]]) : 0u;
            sum += sz[e];
        }
        u32 total;
        u32 ex = small_scan(sum, s_wave, &total);
#pragma unroll
        for (u32 e = 0; e < kSmallPer; ++e) {
            const u32 r = i0 + e;
            if (e < E && r < U) {
                a.off[L][r] = ex;
                ex += sz[e];
            }
        }
        if (threadIdx.x == 0) a.meta[kMetaS + L] = total;
    }
    __syncthreads();  // every level's size (thread 0 wrote them) before the bases
    if (threadIdx.x == 0) {
        a.meta[kMetaRoot] = a.mask[a.R][0];
        level_bases(a.meta, a.R, a.deepest);
    }
    meta_to_host(a.meta, a.host_meta);
}

// The DAG's word buffer: from the admission's allocator when it has one (a server keeps its budget
// there), else from the context's buffer cache.
void* alloc_words(cpvs_ctx* ctx, const Admit* admit, u64 words) {
    const size_t bytes = sizeof(u32) * words;
    return (admit && admit->alloc) ? admit->alloc(admit->user, bytes) : ctx->buffers.take(bytes);
}

inline dim3 grid1d(u32 n, u32 block = 256) { return dim3(ceil_div<u32>(n ? n : 1, block)); }

constexpr int kRetryWithSort = -1000;
constexpr int kRetryFullTable = -1001;
constexpr int kRetryExactCounts = -1002;  // the counts exceed the emitted nodes (a map with NaN depths)
static_assert(sizeof(u32) * kMetaWords <= sizeof(u64) * kPinnedWords, "meta readback fits the pinned staging");


// small_leaf: a leaf level of at least kSmallLeafTableMin nodes gets a merge table sized for a
// quarter of them (only a run's first node is inserted: a fifth of them at 8192^2, with 4% of the
// nodes distinct), which keeps the table's clear and its random accesses small; a table found too
// small (maps of mostly distinct leaves, such as white noise) is reported as kRetryFullTable and
// the build runs again.
__global__ void k_store_word(u32* dst, u32 v) { *dst = v; }

// The staged words into the DAG's own buffer (16-byte vectors where both are aligned).
__global__ __launch_bounds__(256) void k_copy_words(const u32* __restrict__ src, u32* __restrict__ dst, u64 n) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 i0 = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if ((((uintptr_t)src | (uintptr_t)dst) & 15) == 0) {
        const u64 n4 = n / 4;
        for (u64 i = i0; i < n4; i += stride) reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
        for (u64 i = 4 * n4 + i0; i < n; i += stride) dst[i] = src[i];
    } else {
        for (u64 i = i0; i < n; i += stride) dst[i] = src[i];
    }
}

int build_once(cpvs_ctx* ctx, Arena& A, u64* pinned, const cpvs_pyramid* p, u32 t, u32 T, hipStream_t s, int merge_mode,
               cpvs_dag** out, const Admit* admit, bool small_leaf, bool exact, DagOut* into) {
    const int n = p->n;
    const int R = n - 2;
    const bool leaf = n >= 5 && ctx->leafmasks != 0;  // useLeafmasks (src/CompressedShadow.cpp:20-27)
    const int bottom = leaf ? 2 : 0;
    const u32 N = p->N;
    const bool use_table = merge_mode == CPVS_MERGE_TABLE;

    // --- exact node counts (one readback) ---------------------------------------------
    const u32 w3 = leaf ? (N >> 3) : 0;
    const u32 ncol = w3 * w3;
    const size_t meta_bytes = arena_bytes(kMetaWords, sizeof(u32));
    const size_t col_bytes = leaf ? arena_bytes(ncol, sizeof(uint4)) : 0;
    int rc = A.reserve(meta_bytes + col_bytes, s);
    if (rc) return rc;
    uint4* colinfo = leaf ? A.take<uint4>(ncol) : nullptr;
    CountLayout lay;
    rc = pyramid_node_counts(p, t, T, leaf, pinned_partials(pinned), colinfo, s, &lay, exact);
    if (rc) return rc;
    CPVS_HIP_CHECK(hipStreamSynchronize(s));
    u64 cnt[kMaxLevels] = {};
    sum_node_counts(lay, pinned_partials(pinned), cnt);
    cnt[R] = 1;
    if (cnt[R - 1] == 0) {
        // No partial child of the root (a z-tile above or below every surface, a flat map): the
        // DAG is the root's childmask alone, written without any further kernel.
        const u32 word = pinned_partials(pinned)[lay.block_start[lay.nlev]];
        if (into) {  // a caller's buffer: the word is stored by a kernel, nothing is allocated
            if (into->cap < 1) {
                set_error("build: the DAG has 1 word, the output holds %llu", (unsigned long long)into->cap);
                return CPVS_EINVAL;
            }
            CPVS_LAUNCH("store_word", k_store_word, dim3(1), dim3(1), 0, s, into->words, word);
            CPVS_HIP_CHECK(hipGetLastError());
            CPVS_HIP_CHECK(hipStreamSynchronize(s));
            into->nwords = 1;
            into->root_mask = word;
            return CPVS_OK;
        }
        if (admit) {
            rc = admit->fn(admit->user, A, 0, sizeof(u32));
            if (rc) return rc;
        }
        u32* d_words = static_cast<u32*>(alloc_words(ctx, admit, 1));
        if (!d_words) {
            set_error("build: hipMalloc of one DAG word failed");
            return CPVS_ENOMEM;
        }
        const hipError_t e = hipMemcpy(d_words, &word, sizeof(u32), hipMemcpyHostToDevice);  // an idle buffer
        if (e != hipSuccess) {
            ctx->buffers.give(d_words);
            return hip_fail(e, "build root word", __FILE__, __LINE__);
        }
        cpvs_dag* d = new cpvs_dag();
        d->ctx = ctx;
        d->n = n;
        d->words = 1;
        d->d_words = d_words;
        d->root_mask = word;
        d->N = N;
        d->zTileIndex = t;
        d->zTileNum = T;
        d->merge_mode = merge_mode;
        d->leafmasks = leaf;
        d->nodes[R] = 1;
        d->uniques[R] = 1;
        *out = d;
        return CPVS_OK;
    }
    for (int L = bottom; L <= R; ++L) {
        if (cnt[L] > 0xFFFFFFF0ull) {
            set_error("build: level %d has %llu nodes (limit 2^32)", L, (unsigned long long)cnt[L]);
            return CPVS_EINVAL;
        }
    }

    // --- scratch sizing ----------------------------------------------------------------------
    u32 cmax = 1;
    for (int L = bottom; L < R; ++L) cmax = cnt[L] > cmax ? (u32)cnt[L] : cmax;
    const u32 S = table_slots(cmax);
    const u32 scan_n = (leaf && ncol > cmax) ? ncol : cmax;
    size_t need = meta_bytes + col_bytes;
    for (int L = bottom; L <= R; ++L) {
        const size_t c = cnt[L] ? cnt[L] : 1;
        need += arena_bytes(c, 8) + 5 * arena_bytes(c, 4);
    }
    const size_t c2 = leaf ? (cnt[2] ? cnt[2] : 1) : 1;
    if (leaf) need += arena_bytes(c2 * 8, 8) + 2 * arena_bytes(c2, 4);  // leafmask rows, run heads, column map
    need += 2 * arena_bytes(cmax, 4);
    const bool col_list = leaf && T > 1;  // compact the z-tile's columns with leaf nodes
    if (col_list) need += arena_bytes(ncol, 4) + arena_bytes(1, 4);
    if (use_table) need += arena_bytes(S, 8);
    else need += 4 * arena_bytes(cmax, 4) + arena_bytes(sort_scratch_words(cmax), 4);
    need += arena_bytes(scan_scratch_words(scan_n), 4);
    need += 1 << 20;
    // Pre-merge words bound the DAG: mask and child pointers per node, two words per partial slice
    // at the leaves.  The words are written into a staging area of that size right behind the
    // merge (no host readback before the write), then copied to an exact allocation; a caller's
    // output buffer (into) is written directly after the readback instead.
    u64 bound_words = 0;
    for (int L = bottom; L <= R; ++L)
        bound_words += cnt[L] + ((L > bottom) ? cnt[L - 1] : 0) + ((leaf && L == 2) ? 16 * cnt[L] : 0);
    const bool staged = !into;
    if (staged) need += arena_bytes(bound_words + 17, 4);  // 17 words of slack for the bound check
    if (admit) {
        rc = admit->fn(admit->user, A, need, 4 * bound_words);
        if (rc) return rc;
    }
    bool moved = false;
    rc = A.reserve(need, s, &moved);  // keeps the contents when large enough
    if (rc) return rc;
    colinfo = leaf ? A.take<uint4>(ncol) : nullptr;
    u32* meta = A.take<u32>(kMetaWords);
    if (leaf && moved) {  // the arena was reallocated: count the columns again
        rc = pyramid_node_counts(p, t, T, leaf, nullptr, colinfo, s, nullptr, exact);
        if (rc) return rc;
    }
    // meta is zeroed by the first kernel below (k_small_topdown or k_start).

    Level lv[kMaxLevels];
    for (int L = bottom; L <= R; ++L) {
        const size_t c = cnt[L] ? cnt[L] : 1;
        lv[L].count = (u32)cnt[L];
        lv[L].coord = A.take<u64>(c);
        lv[L].mask = A.take<u32>(c);
        lv[L].cstart = A.take<u32>(c);
        lv[L].rank = A.take<u32>(c);
        lv[L].uniq = A.take<u32>(c);
        lv[L].off = A.take<u32>(c);
    }
    u64* leafm = leaf ? A.take<u64>(c2 * 8) : nullptr;
    u32* runhead = leaf ? A.take<u32>(c2) : nullptr;
    u32* colbfs = leaf ? A.take<u32>(c2) : nullptr;
    u32* clist = col_list ? A.take<u32>(ncol) : nullptr;
    u32* nclist = col_list ? A.take<u32>(1) : nullptr;
    u32* rep = A.take<u32>(cmax);
    u32* tslot = A.take<u32>(cmax);  // table slot of each node, or segment starts when sorting
    u32* seg = tslot;
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
    u32* staging = staged ? A.take<u32>(bound_words + 17) : nullptr;
    if (A.used > A.cap) {
        set_error("build: scratch layout exceeds its reservation (%zu > %zu)", A.used, A.cap);
        return CPVS_EINTERNAL;
    }

    LeafColumns lc{};
    if (leaf) {
        lc.info = colinfo;
        lc.bfs = colbfs;
        lc.w3 = w3;
        lc.w3_shift = (u32)__builtin_ctz(w3);
        if (col_list)
            device_scan(ColPairIn{colinfo}, ColListOut{colinfo, clist, nclist, ncol}, ncol, OpSum64{}, nullptr, scan_scratch, s);
        else
            device_scan(ColCountIn{colinfo}, ColStartOut{colinfo}, ncol, OpSum{}, nullptr, scan_scratch, s);
    }

    // --- 1/2. BFS classification -------------------------------------------------------------
    const u64 root = pack_xyz(0, 0, 2 * t);  // create(): root offset (0,0,2*zTileIndex)
    int deepest = R;  // lowest level with nodes (the counts are exact)
    while (deepest > bottom && cnt[deepest - 1] > 0) --deepest;
    // Small levels [Ls, R] go through the one-workgroup kernels (with the hash-table merge).
    int Ls = R + 1;
    if (use_table && leaf)
        for (int L = R; L >= std::max(deepest, 3) && cnt[L] <= kSmallMax; --L) Ls = L;
    SmallArgs sa{};
    sa.R = R;
    sa.Ls = Ls;
    sa.deepest = deepest;
    sa.meta = meta;
    sa.host_meta = reinterpret_cast<u32*>(pinned);
    for (int L = bottom; L <= R; ++L) {
        sa.count[L] = lv[L].count;
        sa.coord[L] = lv[L].coord;
        sa.mask[L] = lv[L].mask;
        sa.cstart[L] = lv[L].cstart;
        sa.rank[L] = lv[L].rank;
        sa.uniq[L] = lv[L].uniq;
        sa.off[L] = lv[L].off;
        sa.lv[L] = p->lv[L];
    }
    const int top_big = std::min(R, Ls - 1);  // highest level of the per-level path
    if (Ls <= R)
        CPVS_LAUNCH("small_topdown", k_small_topdown, dim3(1), dim3(kSmallThreads), 0, s, sa, p->depth, N, T, root, lc);
    else
        CPVS_LAUNCH("start", k_start, dim3(1), dim3(1), 0, s, lv[R].coord, root, meta, (u32)kMetaWords);
    for (int L = top_big; L >= deepest; --L) {
        const u32 c = lv[L].count;
        if (leaf && L == 2) {
            // With a column list, at most min(columns, leaf nodes) columns have nodes.
            const u32 bound = col_list ? std::min(ncol, c) : ncol;
            const u32 units = ceil_div<u32>(bound, kLeafWaves * kLeafCols);
            const u32 unit_chunk = ceil_div<u32>(units, kXcds);
            const u32 slots = use_table ? leaf_table_slots(c, small_leaf) : 0u;
            // The leaf level is the first merged: its table is cleared by this kernel.
            if (col_list)
                CPVS_LAUNCH("leaf_waves", k_leaf_waves<true>, dim3(unit_chunk * kXcds), dim3(64 * kLeafWaves), 0, s, lc,
                            bound, c, p->depth, N, T, p->lv[3], lv[2].mask, leafm, runhead, unit_chunk, tab,
                            slots, clist, nclist);
            else
                CPVS_LAUNCH("leaf_waves", k_leaf_waves<false>, dim3(unit_chunk * kXcds), dim3(64 * kLeafWaves), 0, s, lc,
                            bound, c, p->depth, N, T, p->lv[3], lv[2].mask, leafm, runhead, unit_chunk, tab,
                            slots, nullptr, nullptr);
            break;
        }
        // A level is classified inside its emit scan when that scan loads its outputs' prefetch
        // before the block scan.  Above 1 M nodes the output stage loads after the scan, where
        // classifying (two dependent loads) costs more than the classify launch it saves.
        if (L > bottom && scan_prefetches(c)) {
            const Classify cls{lv[L].coord, p->lv[L], p->depth, N, L, T};
            EmitChildrenOut emit{lv[L].mask, lv[L].coord, lv[L].cstart, lv[L - 1].coord, lv[L - 1].count, LeafColumns{},
                                 lv[L].mask, cls};
            if (leaf && L == 3) emit.cols = lc;
            device_scan(ClassifyCountIn{cls}, emit, c, OpSum{}, meta + kMetaC + L, scan_scratch, s);
        } else {
            CPVS_LAUNCH("classify", k_classify, grid1d(c), dim3(256), 0, s, lv[L].coord, c, p->lv[L], p->depth, N, L, T,
                        lv[L].mask);
            if (L > bottom) {
                EmitChildrenOut emit{lv[L].mask, lv[L].coord, lv[L].cstart, lv[L - 1].coord, lv[L - 1].count, LeafColumns{}};
                if (leaf && L == 3) emit.cols = lc;
                device_scan(ChildCountIn{lv[L].mask}, emit, c, OpSum{}, meta + kMetaC + L, scan_scratch, s);
            }
        }
    }
    CPVS_HIP_CHECK(hipGetLastError());

    // Levels merged below (not the root, not the small levels) whose word count surely fits 32
    // bits get ranks and offsets from one 64-bit scan.
    auto fused_sizes = [&](int L) {
        const u64 per_node = (leaf && L == 2) ? 17ull : 9ull;
        return L >= deepest && L < std::min(R, Ls) && (u64)lv[L].count * per_node < 0xFFFFFFFFull;
    };
    // With every larger level's offsets known from its ranks scan, the small-level merge also
    // computes the small levels' offsets and all level bases (no separate compress launch).
    sa.sizes_in_merge = 1u;
    for (int L = deepest; L < std::min(R, Ls); ++L)
        if (!fused_sizes(L)) sa.sizes_in_merge = 0u;

    // --- 3. merge, bottom-up (the root is never merged) --------------------------------------
    for (int L = deepest; L < std::min(R, Ls); ++L) {
        const u32 c = lv[L].count;
        const bool is_leaf = leaf && L == 2;
        const u32* child_rank = (L > deepest) ? lv[L - 1].rank : nullptr;
        const u32* rh = is_leaf ? runhead : nullptr;
        // Levels ranked by the 64-bit scan take its tile partials from k_resolve (table merge): no
        // reduce pass.  Tiles of more than one round are summed with atomics into partials that the
        // insert kernel clears.
        u64* rpart = (use_table && fused_sizes(L)) ? reinterpret_cast<u64*>(scan_scratch) : nullptr;
        const u32 rrounds = scan_rounds(c);
        const u32 nzero = (rpart && rrounds > 1) ? ceil_div<u32>(c, rrounds * kScanThreads) : 0u;
        if (use_table) {
            const u32 Sl = is_leaf ? leaf_table_slots(c, small_leaf) : table_slots(c);
            // Cleared by the leaf kernel (leaf level) or the previous level's k_rank_copies.
            if (L == deepest && !is_leaf) CPVS_LAUNCH("table_clear", k_table_clear, grid1d(Sl), dim3(256), 0, s, tab, Sl);
            if (is_leaf) {
                CPVS_LAUNCH("insert_leaf", k_insert_leaf, dim3(ceil_div<u32>(c, kLeafWindow)), dim3(kLeafInsertThreads), 0, s,
                            lv[2].mask, leafm, runhead, c, tab, Sl, tslot, meta + kMetaFull, rpart, nzero,
                            ctx->leaf_hash_bits.load());
                CPVS_LAUNCH("resolve", k_resolve<LeafEq>, grid1d(c, kNodeThreads), dim3(kNodeThreads), 0, s, tslot, tab, runhead, c,
                            LeafEq{lv[2].mask, leafm, nullptr, c}, rep, meta + kMetaColl, lv[2].mask, 1u, rpart, rrounds);
            } else {
                CPVS_LAUNCH("insert_inner", k_insert_inner, grid1d(c, kInsertBlock), dim3(kInsertBlock), 0, s, lv[L].mask,
                            lv[L].cstart, child_rank, c,
                            tab, Sl, tslot, meta + kMetaFull, rpart, nzero, ctx->inner_hash_bits.load());
                CPVS_LAUNCH("resolve", k_resolve<InnerEq>, grid1d(c, kNodeThreads), dim3(kNodeThreads), 0, s, tslot, tab, nullptr, c,
                            InnerEq{lv[L].mask, lv[L].cstart, child_rank}, rep, meta + kMetaColl, lv[L].mask, 0u, rpart, rrounds);
            }
        } else {
            if (is_leaf)
                CPVS_LAUNCH("sortkeys", k_sortkeys_leaf, grid1d(c), dim3(256), 0, s, lv[2].mask, leafm, runhead, c, skey,
                            sval);
            else
                CPVS_LAUNCH("sortkeys", k_sortkeys_inner, grid1d(c), dim3(256), 0, s, lv[L].mask, lv[L].cstart, child_rank, c,
                            skey, sval);
            radix_sort_pairs(skey, sval, ktmp, vtmp, c, 32, sort_scratch, s);
            device_scan(SegHeadIn{skey}, SegStartOut{seg}, c, OpMax{}, nullptr, scan_scratch, s);
            if (is_leaf)
                CPVS_LAUNCH("representative", k_representative<LeafEq>, grid1d(c), dim3(256), 0, s, sval, seg, c,
                            LeafEq{lv[2].mask, leafm, runhead, c}, rep, meta + kMetaSortColl);
            else
                CPVS_LAUNCH("representative", k_representative<InnerEq>, grid1d(c), dim3(256), 0, s, sval, seg, c,
                            InnerEq{lv[L].mask, lv[L].cstart, child_rank}, rep, meta + kMetaSortColl);
            rh = nullptr;  // every node got its representative directly
        }
        if (fused_sizes(L) && rpart)  // tile partials came from k_resolve
            device_scan_from_partials(FirstSizeIn{rep, lv[L].mask, is_leaf ? 1u : 0u},
                                      FirstSizeOut{lv[L].rank, lv[L].uniq, lv[L].off, meta + kMetaU + L, meta + kMetaS + L, c},
                                      c, OpSum64{}, nullptr, rpart, s);
        else if (fused_sizes(L))
            device_scan(FirstSizeIn{rep, lv[L].mask, is_leaf ? 1u : 0u},
                        FirstSizeOut{lv[L].rank, lv[L].uniq, lv[L].off, meta + kMetaU + L, meta + kMetaS + L, c}, c, OpSum64{},
                        nullptr, scan_scratch, s);
        else
            device_scan(FirstFlagIn{rep}, FirstRankOut{lv[L].rank, lv[L].uniq}, c, OpSum{}, meta + kMetaU + L, scan_scratch, s);
        const u32 S_next = (use_table && L + 1 < std::min(R, Ls)) ? table_slots(lv[L + 1].count) : 0u;
        CPVS_LAUNCH("rank_copies", k_rank_copies, grid1d(c, kNodeThreads), dim3(kNodeThreads), 0, s, rep, rh, c, lv[L].rank, tab, S_next);
    }
    if (Ls <= R)
        CPVS_LAUNCH("small_merge", k_small_merge, dim3(1), dim3(kSmallThreads), 0, s, sa);
    else
        CPVS_LAUNCH("root_init", k_root_init, dim3(1), dim3(1), 0, s, lv[R].rank, lv[R].uniq, meta + kMetaU + R);
    CPVS_HIP_CHECK(hipGetLastError());

    // --- 4. compress: sizes, bases, one readback, write -------------------------------------
    for (int L = top_big; L >= deepest; --L) {
        if (fused_sizes(L)) continue;  // offsets came with the ranks
        const u32 c = lv[L].count;
        const u32 is_leaf = (leaf && L == 2) ? 1u : 0u;
        device_scan(UniqueSizeIn{lv[L].uniq, lv[L].mask, meta + kMetaU + L, is_leaf},
                    UniqueOffOut{lv[L].off, meta + kMetaU + L, meta + kMetaS + L, c}, c, OpSum64{}, nullptr, scan_scratch, s);
    }
    if (Ls <= R) {
        if (!sa.sizes_in_merge) CPVS_LAUNCH("small_compress", k_small_compress, dim3(1), dim3(kSmallThreads), 0, s, sa);
    } else {
        CPVS_LAUNCH("level_bases", k_level_bases, dim3(1), dim3(128), 0, s, meta, R, deepest, lv[R].mask,
                    reinterpret_cast<u32*>(pinned));
    }
    // The word writer's per-level arrays (the counts and bases come from meta or the readback).
    WriteArgs wa{};
    u32 wmax = 0;  // blocks of the largest possible write (every node unique)
    for (int L = R; L >= deepest; --L) {
        WriteLevel& w = wa.lv[wa.nlev++];
        const bool kids = L > deepest;
        w.uniq = lv[L].uniq;
        w.mask = lv[L].mask;
        w.cstart = lv[L].cstart;
        w.off = lv[L].off;
        w.child_rank = kids ? lv[L - 1].rank : nullptr;
        w.child_off = kids ? lv[L - 1].off : nullptr;
        w.child_n = kids ? std::max<u32>(lv[L - 1].count, 1u) : 1u;
        w.leaf = (leaf && L == 2) ? 1u : 0u;
        w.L = L;
        wmax += ceil_div<u32>(lv[L].count ? lv[L].count : 1, kNodeThreads);
    }
    wa.leafm = leafm;
    wa.nleaf = (u32)c2;
    if (staged)
        CPVS_LAUNCH("write_levels", k_write_levels_dev, dim3(std::min<u32>(wmax, 2048)), dim3(kNodeThreads), 0, s, wa, meta,
                    staging, bound_words + 17);
    CPVS_HIP_CHECK(hipGetLastError());
    CPVS_HIP_CHECK(hipStreamSynchronize(s));  // the bases kernel copied meta into pinned
    u32 hm[kMetaWords];
    std::memcpy(hm, pinned, sizeof(hm));

    // Counts first: a level whose count exceeds its emitted children merged unwritten nodes too.
    for (int L = deepest + 1; L <= R; ++L) {
        if (hm[kMetaC + L] != lv[L - 1].count) {
            if (!exact) return kRetryExactCounts;  // cells under a NaN texel were counted: count exactly
            set_error("build: level %d emitted %u children, expected %u", L, hm[kMetaC + L], lv[L - 1].count);
            return CPVS_EINTERNAL;
        }
    }
    if (use_table && hm[kMetaFull] != 0) return small_leaf ? kRetryFullTable : kRetryWithSort;
    if (use_table && hm[kMetaColl] != 0) return kRetryWithSort;
    if (deepest > bottom && hm[kMetaC + deepest] != 0) {
        set_error("build: level %d emitted children below the deepest level", deepest);
        return CPVS_EINTERNAL;
    }
    u64 total = 0;
    for (int L = deepest; L <= R; ++L) total += hm[kMetaS + L];
    if (total >= 0xFFFFFFFEull) {
        set_error("build: DAG of %llu words exceeds 32-bit offsets", (unsigned long long)total);
        return CPVS_EINVAL;
    }

    if (into && total > into->cap) {
        set_error("build: the DAG has %llu words, the output holds %llu", (unsigned long long)total,
                  (unsigned long long)into->cap);
        return CPVS_EINVAL;
    }
    u32* d_words = into ? into->words : static_cast<u32*>(alloc_words(ctx, admit, total ? total : 1));
    if (!d_words) {
        set_error("build: hipMalloc of %llu DAG words failed", (unsigned long long)total);
        return CPVS_ENOMEM;
    }
    if (staged) {
        if (total) CPVS_LAUNCH("copy_words", k_copy_words, dim3(std::min<u64>(ceil_div<u64>(total, 1024), 4096)),
                               dim3(256), 0, s, staging, d_words, total);
    } else {
        u32 wblocks = 0;
        for (int k = 0; k < wa.nlev; ++k) {
            WriteLevel& w = wa.lv[k];
            w.U = hm[kMetaU + w.L];
            w.base = hm[kMetaB + w.L];
            w.child_base = w.child_rank ? hm[kMetaB + w.L - 1] : 0u;
            wa.first[k] = wblocks;
            wblocks += ceil_div<u32>(w.U, kNodeThreads);
        }
        wa.first[wa.nlev] = wblocks;
        if (wblocks) CPVS_LAUNCH("write_levels", k_write_levels, dim3(wblocks), dim3(kNodeThreads), 0, s, wa, d_words);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        if (!into) ctx->buffers.give(d_words);
        return hip_fail(e, "build write", __FILE__, __LINE__);
    }
    if (into) {
        into->nwords = total;
        into->root_mask = hm[kMetaRoot];
        return CPVS_OK;
    }

    cpvs_dag* d = new cpvs_dag();
    d->ctx = ctx;
    d->n = n;
    d->words = total;
    d->d_words = d_words;
    d->root_mask = hm[kMetaRoot];
    d->N = N;
    d->zTileIndex = t;
    d->zTileNum = T;
    d->merge_mode = merge_mode;
    d->key_collisions = hm[kMetaSortColl];
    d->leafmasks = leaf;
    for (int L = 0; L < n && L < kMaxLevels; ++L) {
        d->nodes[L] = (L >= bottom && L <= R) ? cnt[L] : 0;
        d->uniques[L] = (L >= deepest && L <= R) ? hm[kMetaU + L] : 0;
    }
    *out = d;
    return CPVS_OK;
}

}  // namespace

// A build that finds its merge table too small runs again with a full-size table, one with a key
// collision again with the sort merge, and one whose counts exceed the emitted nodes again with
// exact counts; each switch is made at most once.
int build_dag(cpvs_ctx* ctx, Arena& A, u64* pinned, int merge_mode, const cpvs_pyramid* p, u32 t, u32 T, hipStream_t s,
              cpvs_dag** out, const Admit* admit, DagOut* into) {
    bool small_leaf = true, exact = false;
    int rc = CPVS_OK;
    for (int attempt = 0; attempt < 4; ++attempt) {
        rc = build_once(ctx, A, pinned, p, t, T, s, merge_mode, out, admit, small_leaf, exact, into);
        if (rc == kRetryFullTable && small_leaf) small_leaf = false;
        else if (rc == kRetryWithSort && merge_mode != CPVS_MERGE_SORT) merge_mode = CPVS_MERGE_SORT, small_leaf = false;
        else if (rc == kRetryExactCounts && !exact) exact = true;
        else break;
    }
    if (rc == kRetryFullTable || rc == kRetryWithSort || rc == kRetryExactCounts) {
        set_error("build: no consistent build after every fallback (status %d)", rc);
        return CPVS_EINTERNAL;
    }
    return rc;
}

}  // namespace cpvs
