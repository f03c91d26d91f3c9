// build_parts.hpp -- device code shared by the single-map build (build.hip) and the forest build
// of many maps (forest.hip): classification, leaf hashing, the merge tables and their kernels,
// the rank and size scans' functors, the metadata layout and the word writers.
//
// The subtree merge works on node indices, masks, child ranks and leafmasks only, so a forest of
// many DAGs goes through the same merge kernels: its mask words carry the DAG's index in their
// upper 16 bits (hash and key compares then never merge across DAGs, and num_children() reads
// the low 16 bits only).
#pragma once

#include "internal.hpp"
#include "scan.hpp"

namespace cpvs {

constexpr u64 kHashSeed = 0x6A09E667F3BCC909ull;
constexpr u32 kNoRep = 0xFFFFFFFFu;

__device__ inline u32 fold32(u64 h) { return (u32)(h ^ (h >> 32)); }

__device__ inline u32 fmix32(u32 h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// ---------------------------------------------------------------------------------------
// 1. Inner classification: createChildmask (src/CompressedShadowUtil.cpp:20-54).
// ---------------------------------------------------------------------------------------
// The cell is clamped into the level: a node past a level's emitted children (counts that
// exceed them, which the host then reports) still reads inside the pyramid.
__device__ inline u32 classify_node(u64 c, const float2* __restrict__ lv, const float* __restrict__ depth, u32 N, int L,
                                    u32 T) {
    const u32 w = N >> L;
    const u32 x = min(unpack_x(c), w - 2) & ~1u, y = min(unpack_y(c), w - 2), z = unpack_z(c);
    const float h = (float)((N >> L) * T);
    float mn[4], mx[4];  // texel (x + (q&1), y + (q>>1))
    if (L > 0) {
        const float4 a = *reinterpret_cast<const float4*>(lv + (size_t)y * w + x);
        const float4 b = *reinterpret_cast<const float4*>(lv + (size_t)(y + 1) * w + x);
        mn[0] = a.x; mx[0] = a.y; mn[1] = a.z; mx[1] = a.w;
        mn[2] = b.x; mx[2] = b.y; mn[3] = b.z; mx[3] = b.w;
    } else {
        const float2 a = *reinterpret_cast<const float2*>(depth + (size_t)y * N + x);
        const float2 b = *reinterpret_cast<const float2*>(depth + (size_t)(y + 1) * N + x);
        mn[0] = mx[0] = a.x; mn[1] = mx[1] = a.y;
        mn[2] = mx[2] = b.x; mn[3] = mx[3] = b.y;
    }
    u32 mask = 0;
#pragma unroll
    for (u32 ch = 0; ch < 8; ++ch) {
        const u32 q = ch & 3, cz = z + (ch >> 2);
        u32 bits;
        if (L > 0) bits = classify_cell(cz, mn[q] * h, mx[q] * h);
        else bits = classify_mid(cz, mn[q] * h);
        mask |= bits << (2 * ch);
    }
    return mask;
}


// Leaf columns: per level-3 texel one record {partial cells, first leaf's index in column order,
// first partial z cell, 0}, so a column's scalars come in one load.  The counts (x, z) come from
// the node-count pass, the starts (y) from a scan of the counts.
struct LeafColumns {
    uint4* info;
    u32* bfs;       // (column start + z - zs) -> breadth-first index of the leaf node
    u32 w3;         // N >> 3
    u32 w3_shift;   // log2(w3)
};

struct ColCountIn {
    const uint4* info;
    __device__ u32 operator()(u32 i) const { return info[i].x; }
};
struct ColStartOut {
    uint4* info;
    __device__ void operator()(u32 i, u32 ex, u32) const { info[i].y = ex; }
};
// With several z-tiles most columns have no leaf node in a given tile.  One scan of
// (column has nodes ? 1 : 0) << 32 | nodes gives the column starts and the list of the columns
// with nodes, in column order (the halves cannot carry: node counts stay below 2^32).
struct ColPairIn {
    const uint4* info;
    __device__ u64 operator()(u32 i) const {
        const u32 c = info[i].x;
        return (c ? (1ull << 32) : 0ull) | c;
    }
};
struct ColListOut {
    uint4* info;
    u32* list;   // list[k] = the k-th column with nodes
    u32* nlist;  // device: the number of such columns
    u32 n;
    __device__ void operator()(u32 i, u64 ex, u64 v) const {
        info[i].y = (u32)ex;
        if ((u32)v) list[(u32)(ex >> 32)] = i;
        if (i == n - 1) *nlist = (u32)((ex + v) >> 32);
    }
};

// Scan functors: partial-child count in, (first child index, children coordinates) out.
struct ChildCountIn {
    const u32* mask;
    __device__ u32 operator()(u32 i) const { return num_children(mask[i]); }
};
// A level's classification inside its emit scan: the input classifies node i, and the output
// (which classifies it again from the same cached loads) stores the mask.  One launch per level
// fewer than a separate classify pass.
struct Classify {
    const u64* coord;
    const float2* lv;
    const float* depth;
    u32 N;
    int L;
    u32 T;
    __device__ u32 operator()(u32 i) const { return classify_node(coord[i], lv, depth, N, L, T); }
};
struct ClassifyCountIn {
    Classify cls;
    __device__ u32 operator()(u32 i) const { return num_children(cls(i)); }
};
struct EmitChildrenOut {
    const u32* mask;
    const u64* coord;
    u32* cstart;
    u64* child_coord;   // used when the child level is classified node by node
    u32 child_cap;      // nodes of the child level (exact count); writes beyond it are dropped
    LeafColumns cols;   // used when cols.bfs != null (the children are leaves)
    u32* mask_out = nullptr;  // when set, the node is classified here (cls) and its mask stored
    Classify cls = {};
    // The node's mask and cell, loaded by the scan before any output is written.
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
        const u32 x = unpack_x(c), y = unpack_y(c), z = unpack_z(c);
        u32 j = first;
        // getChildCoordinates (src/CompressedShadowUtil.cpp:101-117): partial children in
        // ascending child index, child cell doubled.
        if (cols.bfs) {
            // (the column records read lie inside the level even for a node past its level's
            // emitted children)
            const u32 xc = min(x, cols.w3 - 2) & ~1u, yc = min(y, cols.w3 - 2);
            // Leaf children: children ch and ch + 4 lie in one column, so the node's (at most
            // four) column records are loaded first, all together, and the stores follow.
            uint4 ci[4];
#pragma unroll
            for (u32 q = 0; q < 4; ++q) {
                ci[q] = make_uint4(0u, 0u, 0u, 0u);
                if (m & ((2u << (2 * q)) | (2u << (2 * (q + 4))))) ci[q] = cols.info[(yc + (q >> 1)) * cols.w3 + xc + (q & 1)];
            }
#pragma unroll
            for (u32 ch = 0; ch < 8; ++ch) {
                if (!(m & (2u << (2 * ch)))) continue;
                const uint4 r = ci[ch & 3];
                const u32 slot = r.y + (z + (ch >> 2) - r.z);
                if (j < child_cap && slot < child_cap) cols.bfs[slot] = j;
                ++j;
            }
            return;
        }
#pragma unroll
        for (u32 ch = 0; ch < 8; ++ch) {
            if (!(m & (2u << (2 * ch)))) continue;
            const u32 cx = x + (ch & 1), cy = y + ((ch >> 1) & 1), cz = z + (ch >> 2);
            if (j < child_cap) child_coord[j] = pack_xyz(2 * cx, 2 * cy, 2 * cz);
            ++j;
        }
    }
};


// Empty slot of the merge hash tables (see table_insert).
constexpr u64 kEmptySlot = ~0ull;

// Grid-stride clear of a merge table, folded into a kernel that runs before the table is used.
__device__ inline void clear_table(u64* tab, u32 slots) {
    const u32 stride = gridDim.x * blockDim.x;
    for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < slots; j += stride) tab[j] = kEmptySlot;
}

// Key hash of a leaf node (wave-uniform): two 32-bit multiply-xorshift streams over the low and
// high halves of the partial leafmasks, seeded by the childmask and coupled at every step, then
// a finaliser that is a bijection on the two streams (so it adds no collisions).
struct LeafHash {
    u32 a, b;
    __device__ explicit LeafHash(u32 mask) : a(mask * 0x9E3779B1u + 0x7F4A7C15u), b((mask ^ 0xA5A5u) * 0x85EBCA77u + 1u) {}
    __device__ void add(u64 m) {
        a = (a ^ (u32)m) * 0x9E3779B1u;
        a ^= a >> 15;
        b = (b ^ (u32)(m >> 32)) * 0x85EBCA77u;
        b ^= b >> 13;
        a += b;
    }
    __device__ u64 value() const {
        u32 x = (a ^ (b >> 16)) * 0x2C1B3C6Du;
        u32 y = (b ^ (x >> 15)) * 0x297A2D39u;
        x ^= y >> 13;
        return ((u64)y << 32) | x;
    }
};


// f = floor(D - 0.5) of a texel's scaled depth D = d * N * T: slice z of its column is lit iff
// z <= f (the midpoint test float(z) + 0.5 <= D, exact for |D| < 2^22).  NaN is never lit.  f is
// non-decreasing in d, so the f of a block's minimum (maximum) depth is the minimum (maximum) of
// its texels' f when the block holds no NaN.
__device__ inline int leaf_f(float D) {
    int f = -(1 << 30);
    if (D == D) f = (int)floorf(fminf(fmaxf(D, -2.0f), 4194304.0f) - 0.5f);
    return f;
}

// Key hash of a leaf node from its mask and its partial slices' leafmasks (slot8: the node's
// leafmask row, partial slices in ascending order).  All loads are issued before the hash.
__device__ inline u64 leaf_key_hash(u32 m, const u64* __restrict__ slot8) {
    const u32 k = num_children(m);
    u64 v[8];
#pragma unroll
    for (u32 j = 0; j < 8; ++j) v[j] = (j < k) ? slot8[j] : 0ull;
    LeafHash hs(m);
#pragma unroll
    for (u32 j = 0; j < 8; ++j)
        if (j < k) hs.add(v[j]);
    return hs.value();
}

// Lane l of a vector value takes the wave-uniform v (v_writelane_b32, an ALU instruction this
// compiler has no builtin for); every other lane keeps `old`.  Works whatever the exec mask.  The
// lane select goes through M0 and the value through a plain SGPR (the instruction takes neither a
// second SGPR nor VCC beside M0).
// v and l must be wave-uniform; readfirstlane says so to the compiler (and costs nothing when
// they already sit in SGPRs).
__device__ inline u32 write_lane(u32 old, u32 v, u32 l) {
    u32 t;
    asm("s_mov_b32 %1, %2\n\ts_mov_b32 m0, %3\n\tv_writelane_b32 %0, %1, m0"
        : "+v"(old), "=&s"(t)
        : "s"(__builtin_amdgcn_readfirstlane(v)), "s"(__builtin_amdgcn_readfirstlane(l))
        : "m0");
    return old;
}

// One column's leaf nodes, the walk both leaf kernels (one map's and a forest's) run per column:
// nodes q = 0 .. cnt-1 of the column at level-2 z cells zs + q, lane = texel of the column's 8x8
// block with f = floor(depth * H - 0.5) (fmin, fmax: the column's extremes), first_bfs = the
// breadth-first index of node lane (the first 64 nodes), bfs_run[q] that of node q.  Writes each
// node's run head (the breadth-first index of its run's first node), and for every run head its
// mask (with gtag, a forest's DAG index, in the upper half) and the leafmasks of its partial slices
// in the row at its own index (8 slots per node).  fl: the wave's 128 bytes of LDS flags.
//
// Node q differs from node q-1 iff some texel's h = f + 8 - 8*zs has h - 8q in [0, 14]: q =
// floor(h/8), or floor(h/8) - 1 when h mod 8 != 7.  Each lane flags its (at most two) nodes in LDS,
// then reads back the flag of node q0 + lane: one ballot gives a 64-node chunk's changed nodes.  LDS
// instructions of one wave run in program order, so the clear, the flags and the read need no wait
// between them; the fences keep the compiler from moving them across one another.
__device__ inline void leaf_column(int f, int fmin, int fmax, int zs, u32 cnt, u32 first_bfs, const u32* bfs_run,
                                   u32 gtag, u32 nleaf, uint8_t* fl, u32 lane, u32* __restrict__ mask_out,
                                   u64* __restrict__ leafm, u32* __restrict__ runhead) {
    const int h = f + 8 - 8 * zs;
    const int qa = h >> 3;  // floor(h / 8)
    const bool below = (h & 7) != 7;
    u32 head_bfs = 0;  // the run head carried into the next chunk
    // Nodes [q0, q0 + 64) of the column; bfs = breadth-first index of node q0 + lane.
    auto chunk = [&](u32 q0, u32 bfs) {
        const u32 qn = min(cnt - q0, 64u);
        const int ra = qa - (int)q0;
        // Node n's flag is byte n + 1 (n = -1 .. 63); every node outside the chunk lands on byte 65.
        // Bytes 0 .. 67 are cleared first (lanes past 16 repeat the last word).
        reinterpret_cast<u32*>(fl)[min(lane, 16u)] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        fl[min((u32)(ra + 1), 65u)] = 1;           // node ra
        fl[below ? min((u32)ra, 65u) : 65u] = 1;   // node ra - 1
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bool hit = fl[lane + 1] != 0;           // read by every lane (no branch around it)
        u64 chg = __ballot(hit) & __ballot(lane < qn);  // wave-uniform
        if (q0 == 0) chg |= 1ull;            // a column's first node starts a run
        // The run heads of this chunk, one after the other (wave-uniform): mask and leafmasks (their
        // key hash is computed lane-parallel by the leaf insert, which reads them anyway).
        u32 my_mask = 0;
        // Node q0 + lane's run head: the last changed node at or before it, set head by head
        // (lanes before the chunk's first head continue the previous chunk's run).
        u32 my_head_bfs = head_bfs;
        // Node q0 + j lies at z8 = zb + 8j; its lit and partial slice counts come from the column
        // extremes relative to it.
        const int zb = 8 * (zs + (int)q0);
        const int fmin1 = fmin + 1 - zb, fmax1 = fmax + 1 - zb, fb = f - zb;
        for (u64 pend = chg; pend;) {
            const u32 j = (u32)__builtin_ctzll(pend);
            pend &= ~(1ull << j);
            const int j8 = 8 * (int)j;
            const int u = fb - j8;
            const u32 cmin = (u32)min(max(fmin1 - j8, 0), 8);
            const u32 cmax = (u32)min(max(fmax1 - j8, 0), 8);
            const u32 below_min = (1u << (2 * cmin)) - 1u, below_max = (1u << (2 * cmax)) - 1u;
            const u32 mask = (0x5555u & below_min) | (0xAAAAu & below_max & ~below_min) | gtag;
            u32 mine_lo = 0, mine_hi = 0;  // lane s gets the leafmask of partial slice cmin + s
            const u32 kp = cmax - cmin;
            for (u32 sl = 0; sl < kp; ++sl) {
                const u64 m = __ballot(u >= (int)(cmin + sl));
                mine_lo = write_lane(mine_lo, (u32)m, sl);
                mine_hi = write_lane(mine_hi, (u32)(m >> 32), sl);
            }
            // The head's leafmask row sits at its own breadth-first index.
            const u32 hrow = __builtin_amdgcn_readlane(bfs, j);
            if (lane < kp && hrow < nleaf)  // the two halves stored as they are (no 64-bit assembly)
                *reinterpret_cast<uint2*>(leafm + (size_t)hrow * 8 + lane) = make_uint2(mine_lo, mine_hi);
            my_mask = write_lane(my_mask, mask, j);
            my_head_bfs = (lane >= j) ? hrow : my_head_bfs;
        }
        head_bfs = __builtin_amdgcn_readlane(my_head_bfs, qn - 1);
        // nleaf bounds every index (only exceeded if the column map were inconsistent).
        if (lane < qn && bfs < nleaf) {
            runhead[bfs] = my_head_bfs;
            if ((chg >> lane) & 1ull) mask_out[bfs] = my_mask;
        }
    };
    // The first chunk's indices are staged; a longer column (box walls) loads the rest here.
    chunk(0u, first_bfs);
    for (u32 q0 = 64; q0 < cnt; q0 += 64) chunk(q0, bfs_run[min(q0 + lane, cnt - 1)]);
}

constexpr u32 kXcds = 8;
constexpr u32 kLeafCols = 8;
constexpr u32 kLeafWaves = 1;

// Children ranks of inner node i (slots j >= k are 0).  All eight loads are issued before any is
// used, so a node costs one memory round trip, not one per child.
__device__ inline void load_child_ranks(const u32* cstart, const u32* child_rank, u32 i, u32 k, u32 r[8]) {
    const u32 s = k ? cstart[i] : 0u;
#pragma unroll
    for (u32 j = 0; j < 8; ++j) r[j] = (j < k) ? child_rank[s + j] : 0u;
}

__device__ inline u64 inner_key_hash(const u32* mask, const u32* cstart, const u32* child_rank, u32 i) {
    const u32 m = mask[i];
    const u32 k = num_children(m);
    u64 h = hash_step(kHashSeed, m);
    if (k) {
        u32 r[8];
        load_child_ranks(cstart, child_rank, i, k, r);
#pragma unroll
        for (u32 j = 0; j < 8; ++j)
            if (j < k) h = hash_step(h, r[j]);
    }
    return h;
}

// A merge hash cut to its low `bits` bits, in both the tag and the slot position, so that distinct
// keys collide (bits = 0: the full hash).  Builds use it only to exercise collision detection.
__device__ inline u64 weaken_hash(u64 h, u32 bits) {
    if (!bits) return h;
    const u32 t = (u32)h & ((1u << bits) - 1u);
    return ((u64)t << 32) | t;
}

// Run heads are bounded by the level's n nodes where they are used as indices:
// past a level's emitted children (counts that exceed them, which the host then reports) the
// entries were never written.
struct LeafEq {
    const u32* mask;
    const u64* leafm;     // a run's first node's partial leafmasks: the row at that node's index
    const u32* runhead;   // when set, a node's key is read from its run's first node
    u32 n;                // nodes of the leaf level
    __device__ bool operator()(u32 a, u32 b) const {
        if (runhead) {
            const u32 ha = runhead[a], hb = runhead[b];
            a = (ha < n) ? ha : a;
            b = (hb < n) ? hb : b;
            if (a == b) return true;
        }
        const u32 m = mask[a], mb = mask[b];
        if (m != mb) return false;
        const u32 k = num_children(m);
        const u64* pa = leafm + (size_t)a * 8;
        const u64* pb = leafm + (size_t)b * 8;
        bool eq = true;
#pragma unroll
        for (u32 j = 0; j < 8; ++j)
            if (j < k) eq = eq && (pa[j] == pb[j]);
        return eq;
    }
};
struct InnerEq {
    const u32* mask;
    const u32* cstart;
    const u32* child_rank;
    __device__ bool operator()(u32 a, u32 b) const {
        // Masks and first-child indices are loaded together (one round trip before the ranks).
        const u32 m = mask[a], mb = mask[b];
        const u32 ca = cstart[a], cb = cstart[b];
        if (m != mb) return false;
        const u32 k = num_children(m);
        if (!k) return true;
        u32 ra[8], rb[8];
#pragma unroll
        for (u32 j = 0; j < 8; ++j) {
            ra[j] = (j < k) ? child_rank[ca + j] : 0u;
            rb[j] = (j < k) ? child_rank[cb + j] : 0u;
        }
        bool eq = true;
#pragma unroll
        for (u32 j = 0; j < 8; ++j) eq = eq && (ra[j] == rb[j]);
        return eq;
    }
};


// Open addressing, linear probing.  Large groups of equal keys hit one slot, so the slot is read
// first, and only an empty slot or a larger index costs an atomic.  A slot only ever goes from
// empty to one tag, and its index only decreases, so a stale read merely leads to an atomic that
// was not needed.  The read is an agent-scope load, which is served coherently across the XCDs:
// each XCD's L2 would otherwise keep showing a slot as empty after another XCD's atomic (atomics
// are performed at memory) had filled it, and every later insert of that key from the XCD would
// pay an atomic round trip.  Measured at 8192^2: leaf insert 77 -> 55 µs, inner inserts 137 ->
// 104 µs against a plain load.
// A probe gives up after kMaxProbes occupied slots (a table too small for its keys: the leaf
// level's table is sized for a quarter of its nodes), counts that in *full and returns a slot in
// range; the host then builds again with a larger table.  Node i's representative is then still a
// node r with rep[r] = r (k_resolve reads whatever the returned slot holds and checks the keys),
// so every index the rest of the build derives stays in range.
constexpr u32 kMaxProbes = 256;
__device__ inline u32 table_insert(u64* tab, u32 S, u64 h, u32 i, u32* full) {
    const u32 tag = (u32)(h >> 32);
    const u64 mine = ((u64)tag << 32) | i;
    u32 slot = (u32)(h ^ (h >> 31)) & (S - 1);
    for (u32 probe = 0; probe < kMaxProbes; ++probe) {
        u64 cur = __hip_atomic_load(tab + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == kEmptySlot) {
            cur = atomicCAS(reinterpret_cast<unsigned long long*>(tab + slot), (unsigned long long)kEmptySlot,
                            (unsigned long long)mine);
            if (cur == kEmptySlot) return slot;
        }
        if ((u32)(cur >> 32) == tag) {
            if ((u32)cur > i) atomicMin(reinterpret_cast<unsigned long long*>(tab + slot), (unsigned long long)mine);
            return slot;
        }
        slot = (slot + 1) & (S - 1);
    }
    atomicAdd(full, 1u);
    return slot;
}

// A table slot hit by many threads at once serialises their atomics, and equal keys are common
// (flat regions repeat the same subtrees all over a level; a level of a few hundred thousand nodes
// runs almost all at once, before any first occurrence has landed).  So each 256-thread workgroup
// first merges its own equal hashes in an LDS table: only the smallest index of each distinct
// hash in the group touches the global table, and the others take its slot.  Every thread of the
// workgroup calls this (it has barriers).
// Threads per insert workgroup (the LDS merge window).  Measured at 8192^2: 512 takes the inner
// inserts 0.142 -> 0.136 ms against 256; 1024 gains nothing more.
constexpr u32 kInsertBlock = 512;
constexpr u32 kGroupSlots = 2 * kInsertBlock;

__device__ inline u32 table_insert_group(u64* tab, u32 S, u64 h, u32 i, bool active, u32* full) {
    __shared__ u64 gh[kGroupSlots];
    __shared__ u32 gi[kGroupSlots];
    __shared__ u32 gs[kGroupSlots];
    for (u32 j = threadIdx.x; j < kGroupSlots; j += blockDim.x) {
        gh[j] = 0ull;
        gi[j] = kNoRep;
    }
    __syncthreads();
    if (h == 0ull) h = 1ull;  // 0 marks an empty LDS slot
    u32 ls = 0;
    if (active) {
        ls = (u32)(h ^ (h >> 29)) & (kGroupSlots - 1);
        for (;;) {
            const u64 cur = atomicCAS(reinterpret_cast<unsigned long long*>(gh + ls), 0ull, (unsigned long long)h);
            if (cur == 0ull || cur == h) break;
            ls = (ls + 1) & (kGroupSlots - 1);
        }
        atomicMin(gi + ls, i);
    }
    __syncthreads();
    if (active && gi[ls] == i) gs[ls] = table_insert(tab, S, h, i, full);
    __syncthreads();
    return active ? gs[ls] : 0u;
}

// zero[0 .. nzero): the ranks scan's tile partials that k_resolve adds into next (cleared here, in a
// launch that runs anyway).
__device__ inline void clear_partials(u64* zero, u32 nzero) {
    const u32 stride = gridDim.x * blockDim.x;
    for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < nzero; j += stride) zero[j] = 0ull;
}

// The leaf level: only a run's first node is inserted (1.3 M of 6.7 M nodes at 8192^2), so each
// thread takes kLeafPer nodes and a workgroup merges a window of kLeafWindow nodes in LDS.  The
// window's distinct hashes are listed as they are found, and the list is spread over the threads
// for the global inserts: a workgroup waits for about one global insert per thread, not per node.
constexpr u32 kLeafInsertThreads = 256;
constexpr u32 kLeafPer = 4;
constexpr u32 kLeafWindow = kLeafInsertThreads * kLeafPer;
constexpr u32 kLeafSlots = 2 * kLeafWindow;  // LDS table at most half full, so every probe ends
static_assert(kLeafSlots <= 65536, "LDS slot numbers are listed as 16-bit values");


// Threads per workgroup of the one-thread-per-node merge and write kernels (at 8192^2, 128 and 256
// tie; 512 and 1024 make resolve and the writes slower by 5-10 µs).
constexpr u32 kNodeThreads = 256;

// With partials set, each workgroup also reduces the ranks scan's input over its nodes (FirstSizeIn:
// 1 << 32 | the node's words, for a first occurrence) into the partial of the scan tile it lies in
// (tiles of `rounds` workgroups: a store for one, an atomic add into a partial the insert kernel
// cleared otherwise), so the scan that follows needs no reduce pass.
static_assert(kNodeThreads == kScanThreads, "resolve workgroups are the ranks scan's rounds");
template <class Eq>
__global__ __launch_bounds__(kNodeThreads) void k_resolve(const u32* __restrict__ tslot, const u64* __restrict__ tab,
                                                 const u32* __restrict__ runhead, u32 n, Eq eq, u32* __restrict__ rep,
                                                 u32* __restrict__ collisions, const u32* __restrict__ mask, u32 leaf,
                                                 u64* __restrict__ partials, u32 rounds) {
    __shared__ u64 s_wave[kNodeThreads / kWave];
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    u64 v = 0;
    if (i < n) {
        if (runhead && runhead[i] != i) {
            rep[i] = kNoRep;  // resolved through its run's first node
        } else {
            u32 r = (u32)tab[tslot[i]];  // the smallest index with this node's tag
            if (r >= n) r = i;
            if (r != i && !eq(i, r)) {
                atomicAdd(collisions, 1u);  // two keys share a 64-bit hash: the build is redone by sorting
                r = i;
            }
            rep[i] = r;
            if (partials && r == i) v = (1ull << 32) | (u64)(1u + (num_children(mask[i]) << leaf));
        }
    }
    if (partials) {  // uniform across the launch
        v = block_reduce(v, s_wave, OpSum64{});
        if (threadIdx.x == 0) {
            if (rounds == 1)
                partials[blockIdx.x] = v;
            else if (v)
                atomicAdd(reinterpret_cast<unsigned long long*>(partials + blockIdx.x / rounds), (unsigned long long)v);
        }
    }
}


// Segment start of each sorted position: max-scan of (key differs from predecessor ? p : 0).
struct SegHeadIn {
    const u32* skey;
    __device__ u32 operator()(u32 p) const { return (p == 0 || skey[p] != skey[p - 1]) ? p : 0u; }
};
struct SegStartOut {
    u32* seg;
    __device__ void operator()(u32 p, u32 ex, u32 v) const { seg[p] = ex > v ? ex : v; }
};

// Representative (first occurrence) of each node.  Within a segment of equal 32-bit keys the
// stable sort keeps node indices ascending, so the head is the smallest index.
template <class Eq>
__global__ __launch_bounds__(256) void k_representative(const u32* __restrict__ sval, const u32* __restrict__ seg, u32 count,
                                                        Eq eq, u32* __restrict__ rep, u32* __restrict__ collisions) {
    const u32 p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const u32 i = sval[p];
    const u32 s = seg[p];
    const u32 head = sval[s];
    u32 r = head;
    if (head != i && !eq(i, head)) {
        atomicAdd(collisions, 1u);
        r = i;
        for (u32 q = s + 1; q < p; ++q) {
            const u32 j = sval[q];
            if (eq(i, j)) {
                r = j;
                break;
            }
        }
    }
    rep[i] = r;
}


struct FirstFlagIn {
    const u32* rep;
    __device__ u32 operator()(u32 i) const { return rep[i] == i ? 1u : 0u; }
};
struct FirstRankOut {
    u32* rank;
    u32* uniq;
    __device__ void operator()(u32 i, u32 ex, u32 first) const {
        if (first) {
            rank[i] = ex;
            uniq[ex] = i;
        }
    }
};

// Ranks and word offsets in one scan: first occurrences in index order are the uniques in rank
// order, so a scan of (first ? 1 : 0, first ? node size : 0) -- packed as a u64 whose halves
// cannot carry (size sums stay below 2^32) -- gives each first occurrence its rank and its offset.
struct FirstSizeIn {
    const u32* rep;
    const u32* mask;
    u32 leaf;  // 1 at the leafmask level (two words per partial child)
    __device__ u64 operator()(u32 i) const {
        const u32 r = rep[i], m = mask[i];  // both loads issued together (a select, not a branch)
        const u64 v = (1ull << 32) | (u64)(1u + (num_children(m) << leaf));
        return (r == i) ? v : 0ull;
    }
};
struct FirstSizeOut {
    u32* rank;
    u32* uniq;
    u32* off;
    u32* n_uniq;  // device: the level's unique count
    u32* words;   // device: the level's word count
    u32 n;
    __device__ void operator()(u32 i, u64 ex, u64 v) const {
        if (v) {
            const u32 r = (u32)(ex >> 32);
            rank[i] = r;
            uniq[r] = i;
            off[r] = (u32)ex;
        }
        if (i == n - 1) {
            const u64 t = ex + v;
            *n_uniq = (u32)(t >> 32);
            *words = (u32)t;
        }
    }
};


// Sizes of the unique nodes of a level whose word count may pass 2^32 (the levels not ranked by
// the fused 64-bit scan hold at least 2^32 / 17 nodes).  The sums run in 64 bits; an offset or a
// level total at or above 2^32 is stored saturated, so the host sees the overflow and rejects the
// DAG before anything is written.
struct UniqueSizeIn {
    const u32* uniq;
    const u32* mask;
    const u32* n_uniq;  // device count of uniques of the level
    u32 leaf;           // 1 at the leafmask level (two words per partial child)
    __device__ u64 operator()(u32 r) const {
        if (r >= *n_uniq) return 0ull;
        return 1ull + (u64)(num_children(mask[uniq[r]]) << leaf);
    }
};

__device__ inline u32 saturate32(u64 v) { return v > 0xFFFFFFFFull ? 0xFFFFFFFFu : (u32)v; }

// Offsets of the unique nodes only (the scan runs over the level's node count, which the host
// knows, while the unique count is on the device), and the level's word count.
struct UniqueOffOut {
    u32* off;
    const u32* n_uniq;
    u32* words;  // device: the level's word count (saturated)
    u32 n;
    __device__ void operator()(u32 r, u64 ex, u64 v) const {
        if (r < *n_uniq) off[r] = saturate32(ex);
        if (r == n - 1) *words = saturate32(ex + v);
    }
};

// Device metadata block (u32 words).
enum : int {
    kMetaU = 0,                          // uniques per level
    kMetaS = kMaxLevels,                 // words per level
    kMetaB = 2 * kMaxLevels,             // level base offsets
    kMetaC = 3 * kMaxLevels,             // children emitted per level (scan totals)
    kMetaW = 4 * kMaxLevels,             // total words
    kMetaColl = 4 * kMaxLevels + 1,      // key collisions in the hash table (table merge)
    kMetaRoot = 4 * kMaxLevels + 2,      // root mask
    kMetaSortColl = 4 * kMaxLevels + 3,  // 32-bit key collisions resolved by a scan (sort merge)
    kMetaFull = 4 * kMaxLevels + 4,      // table inserts that found no free slot (table merge)
    kMetaWords = 4 * kMaxLevels + 5,
};


// Unique node r of an inner level: [mask, absolute offsets of its partial children].
// child_n (the child level's nodes) bounds the child ranks read as indices: ranks past a level's
// emitted children were never written (the host then rejects the build's words).
__device__ inline void write_inner_node(u32 r, const u32* __restrict__ uniq, const u32* __restrict__ mask,
                                        const u32* __restrict__ cstart, const u32* __restrict__ off, u32 base,
                                        const u32* __restrict__ child_rank, const u32* __restrict__ child_off, u32 child_base,
                                        u32 child_n, u32* __restrict__ out) {
    const u32 i = uniq[r];
    const u32 o = base + off[r];
    const u32 m = mask[i];
    out[o] = m;
    const u32 k = num_children(m);
    if (!k) return;
    u32 cr[8];
    load_child_ranks(cstart, child_rank, i, k, cr);
    u32 co[8];
#pragma unroll
    for (u32 j = 0; j < 8; ++j) co[j] = (j < k) ? child_off[min(cr[j], child_n - 1)] : 0u;
#pragma unroll
    for (u32 j = 0; j < 8; ++j)
        if (j < k) out[o + 1 + j] = child_base + co[j];
}

// Unique leaf node r: [mask, (lo, hi) of each partial slice's leafmask].
__device__ inline void write_leaf_node(u32 r, const u32* __restrict__ uniq, const u32* __restrict__ mask,
                                       const u64* __restrict__ leafm, u32 nleaf, const u32* __restrict__ off, u32 base,
                                       u32* __restrict__ out) {
    const u32 i = min(uniq[r], nleaf - 1);  // a first occurrence, which heads its run: its own row
    const u32 o = base + off[r];
    const u32 m = mask[i];
    out[o] = m;
    const u32 k = num_children(m);
    const u64* p = leafm + (size_t)i * 8;
    u64 v[8];
#pragma unroll
    for (u32 j = 0; j < 8; ++j) v[j] = (j < k) ? p[j] : 0ull;
#pragma unroll
    for (u32 j = 0; j < 8; ++j)
        if (j < k) {
            out[o + 1 + 2 * j] = (u32)v[j];
            out[o + 2 + 2 * j] = (u32)(v[j] >> 32);
        }
}


inline u32 table_slots(u32 n) {
    u32 S = 1024;
    while (S < n + n / 2 + 1) S <<= 1;
    return S;
}
// Below 2^20 leaf nodes the table is small anyway, and small maps have more distinct leaves per
// node (measured: 256^2 and 512^2 maps fill 85-96% of a quarter-size table, 1024^2 40%, 8192^2 4%).
constexpr u32 kSmallLeafTableMin = 1u << 20;
inline u32 leaf_table_slots(u32 c, bool small) { return table_slots((small && c >= kSmallLeafTableMin) ? c / 4 : c); }

// Non-template kernels defined in build.hip.
__global__ __launch_bounds__(256) void k_table_clear(u64* __restrict__ tab, u32 S);
__global__ __launch_bounds__(kLeafInsertThreads) void k_insert_leaf(const u32* __restrict__ mask, const u64* __restrict__ leafm,
                                                                    const u32* __restrict__ runhead, u32 n,
                                                                    u64* __restrict__ tab, u32 S, u32* __restrict__ tslot,
                                                                    u32* __restrict__ full, u64* __restrict__ zero,
                                                                    u32 nzero, u32 hbits);
__global__ __launch_bounds__(kInsertBlock) void k_insert_inner(const u32* __restrict__ mask, const u32* __restrict__ cstart,
                                                               const u32* __restrict__ child_rank, u32 n, u64* __restrict__ tab,
                                                               u32 S, u32* __restrict__ tslot, u32* __restrict__ full,
                                                               u64* __restrict__ zero, u32 nzero, u32 hbits);
__global__ __launch_bounds__(256) void k_sortkeys_leaf(const u32* __restrict__ mask, const u64* __restrict__ leafm,
                                                       const u32* __restrict__ runhead,
                                                       u32 n, u32* __restrict__ skey, u32* __restrict__ sval);
__global__ __launch_bounds__(256) void k_sortkeys_inner(const u32* __restrict__ mask, const u32* __restrict__ cstart,
                                                        const u32* __restrict__ child_rank, u32 n, u32* __restrict__ skey,
                                                        u32* __restrict__ sval);
__global__ __launch_bounds__(kNodeThreads) void k_rank_copies(const u32* __restrict__ rep, const u32* __restrict__ runhead, u32 count,
                                                     u32* __restrict__ rank, u64* __restrict__ tab, u32 tab_slots);

}  // namespace cpvs
