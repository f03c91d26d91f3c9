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
//   2. Leaves (level 2): one wave64 per node, lane = x + 8y of the 8x8 texel block, eight
//      __ballot()s give the eight 64-bit leafmasks (src/CompressedShadowUtil.cpp:59-99).
//   3. Merge, bottom-up from the leaves to the root's children: a 64-bit hash of each node's
//      key (mask + partial leafmasks, or mask + merged ranks of its children), a stable LSD
//      radix sort of (hash, node index), segment starts by a max-scan, a full-key compare
//      against the segment head (a mismatch -- a hash collision -- falls back to a scan of
//      the segment), and a scan of first-occurrence flags: rank = order of first occurrence,
//      as mergeLevel's linear search yields (src/CompressedShadowUtil.h:154-182).
//   4. Compress: per level a scan of unique node sizes (1+k words, 1+2k at the leaf level),
//      level bases root first, then one pass writes [mask, absolute child offsets] or
//      [mask, lo, hi, ...] (src/CompressedShadow.cpp:326-392).
#include <cstdio>
#include <cstring>

#include "internal.hpp"
#include "scan.hpp"

namespace cpvs {

constexpr u64 kHashSeed = 0x6A09E667F3BCC909ull;

__device__ inline u32 fold32(u64 h) { return (u32)(h ^ (h >> 32)); }

// ---------------------------------------------------------------------------------------
// 1. Inner classification: createChildmask (src/CompressedShadowUtil.cpp:20-54).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_classify(const u64* __restrict__ coord, u32 count, const float2* __restrict__ lv,
                                                  const float* __restrict__ depth, u32 N, int L, u32 T,
                                                  u32* __restrict__ mask_out) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const u64 c = coord[i];
    const u32 x = unpack_x(c), y = unpack_y(c), z = unpack_z(c);
    const float h = (float)((N >> L) * T);
    float mn[4], mx[4];  // texel (x + (q&1), y + (q>>1))
    if (L > 0) {
        const u32 w = N >> L;
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
    mask_out[i] = mask;
}

// Scan functors: partial-child count in, (first child index, children coordinates) out.
struct ChildCountIn {
    const u32* mask;
    __device__ u32 operator()(u32 i) const { return num_children(mask[i]); }
};
struct EmitChildrenOut {
    const u32* mask;
    const u64* coord;
    u32* cstart;
    u64* child_coord;
    u32 child_cap;  // nodes of the child level (exact count); writes beyond it are dropped
    __device__ void operator()(u32 i, u32 first, u32 k) const {
        cstart[i] = first;
        if (!k) return;
        const u32 m = mask[i];
        const u64 c = coord[i];
        const u32 x = unpack_x(c), y = unpack_y(c), z = unpack_z(c);
        u32 j = first;
        // getChildCoordinates (src/CompressedShadowUtil.cpp:101-117): partial children in
        // ascending child index, child cell doubled.
#pragma unroll
        for (u32 ch = 0; ch < 8; ++ch)
            if (m & (2u << (2 * ch))) {
                if (j < child_cap) child_coord[j] = pack_xyz(2 * (x + (ch & 1)), 2 * (y + ((ch >> 1) & 1)), 2 * (z + (ch >> 2)));
                ++j;
            }
    }
};

// ---------------------------------------------------------------------------------------
// 2. Leaves: createChildmask1x1x8 + createLeafmask (src/CompressedShadowUtil.cpp:59-99).
// One wave per node, lane = x + 8y; slice s is partial unless its ballot is 0 or ~0.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_leaf(const u64* __restrict__ coord, u32 count, const float* __restrict__ depth,
                                              u32 N, u32 T, u32* __restrict__ mask_out, u64* __restrict__ leafm,
                                              u32* __restrict__ skey, u32* __restrict__ sval) {
    const u32 lane = threadIdx.x & 63;
    const u32 node = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (node >= count) return;  // uniform per wave
    const u64 c = coord[node];
    const u32 bx = 4 * unpack_x(c), by = 4 * unpack_y(c), bz = 4 * unpack_z(c);
    const float d = depth[(size_t)(by + (lane >> 3)) * N + bx + (lane & 7)];
    const float dh = d * (float)(N * T);
    u32 mask = 0, k = 0;
    u64 mine = 0;
    u64 h = hash_step(kHashSeed, 0);
#pragma unroll
    for (u32 s = 0; s < 8; ++s) {
        const u64 m = __ballot(classify_mid(bz + s, dh) != 0);
        if (m == ~0ull) {
            mask |= 1u << (2 * s);
        } else if (m != 0ull) {
            mask |= 2u << (2 * s);
            if (lane == k) mine = m;
            h = hash_step(h, m);
            ++k;
        }
    }
    h = hash_step(h, mask);
    if (lane < 8) leafm[(size_t)node * 8 + lane] = (lane < k) ? mine : 0ull;
    if (lane == 0) {
        mask_out[node] = mask;
        skey[node] = fold32(h);
        sval[node] = node;
    }
}

// ---------------------------------------------------------------------------------------
// 3. Merge.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hash_inner(const u32* __restrict__ mask, const u32* __restrict__ cstart,
                                                    const u32* __restrict__ child_rank, u32 count, u32* __restrict__ skey,
                                                    u32* __restrict__ sval) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const u32 m = mask[i];
    const u32 k = num_children(m);
    u64 h = hash_step(kHashSeed, 1);
    if (k) {
        const u32 s = cstart[i];
        for (u32 j = 0; j < k; ++j) h = hash_step(h, child_rank[s + j]);
    }
    h = hash_step(h, m);
    skey[i] = fold32(h);
    sval[i] = i;
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

struct LeafEq {
    const u32* mask;
    const u64* leafm;
    __device__ bool operator()(u32 a, u32 b) const {
        const u32 m = mask[a];
        if (m != mask[b]) return false;
        const u32 k = num_children(m);
        const u64* pa = leafm + (size_t)a * 8;
        const u64* pb = leafm + (size_t)b * 8;
        for (u32 j = 0; j < k; ++j)
            if (pa[j] != pb[j]) return false;
        return true;
    }
};
struct InnerEq {
    const u32* mask;
    const u32* cstart;
    const u32* child_rank;
    __device__ bool operator()(u32 a, u32 b) const {
        const u32 m = mask[a];
        if (m != mask[b]) return false;
        const u32 k = num_children(m);
        if (!k) return true;
        const u32 sa = cstart[a], sb = cstart[b];
        for (u32 j = 0; j < k; ++j)
            if (child_rank[sa + j] != child_rank[sb + j]) return false;
        return true;
    }
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

__global__ __launch_bounds__(256) void k_rank_copies(const u32* __restrict__ rep, u32 count, u32* __restrict__ rank) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const u32 r = rep[i];
    if (r != i) rank[i] = rank[r];
}

__global__ void k_root_init(u32* rank, u32* uniq, u32* n_uniq) {
    rank[0] = 0;
    uniq[0] = 0;
    *n_uniq = 1;
}

__global__ void k_set_u64(u64* dst, u64 v) { *dst = v; }

// ---------------------------------------------------------------------------------------
// 4. Compress.
// ---------------------------------------------------------------------------------------
struct UniqueSizeIn {
    const u32* uniq;
    const u32* mask;
    const u32* n_uniq;  // device count of uniques of the level
    u32 leaf;           // 1 at the leafmask level (two words per partial child)
    __device__ u32 operator()(u32 r) const {
        if (r >= *n_uniq) return 0u;
        return 1u + (num_children(mask[uniq[r]]) << leaf);
    }
};

__global__ void k_level_bases(const u32* __restrict__ S, u32* __restrict__ B, u32* __restrict__ total, int R, int bottom) {
    // B[R] = 0, B[L-1] = B[L] + S[L]; total = B[bottom] + S[bottom] (levels absent have S = 0).
    unsigned long long acc = 0;
    for (int L = R; L >= bottom; --L) {
        B[L] = (u32)acc;
        acc += S[L];
    }
    *total = (acc > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (u32)acc;
}

__global__ __launch_bounds__(256) void k_write_inner(const u32* __restrict__ uniq, u32 n_uniq, const u32* __restrict__ mask,
                                                     const u32* __restrict__ cstart, const u32* __restrict__ off, u32 base,
                                                     const u32* __restrict__ child_rank, const u32* __restrict__ child_off,
                                                     u32 child_base, u32* __restrict__ out) {
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_uniq) return;
    const u32 i = uniq[r];
    const u32 o = base + off[r];
    const u32 m = mask[i];
    out[o] = m;
    const u32 k = num_children(m);
    if (!k) return;
    const u32 s = cstart[i];
    for (u32 j = 0; j < k; ++j) out[o + 1 + j] = child_base + child_off[child_rank[s + j]];
}

__global__ __launch_bounds__(256) void k_write_leaf(const u32* __restrict__ uniq, u32 n_uniq, const u32* __restrict__ mask,
                                                    const u64* __restrict__ leafm, const u32* __restrict__ off, u32 base,
                                                    u32* __restrict__ out) {
    const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_uniq) return;
    const u32 i = uniq[r];
    const u32 o = base + off[r];
    const u32 m = mask[i];
    out[o] = m;
    const u32 k = num_children(m);
    const u64* p = leafm + (size_t)i * 8;
    for (u32 j = 0; j < k; ++j) {
        const u64 v = p[j];
        out[o + 1 + 2 * j] = (u32)v;
        out[o + 2 + 2 * j] = (u32)(v >> 32);
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

// Device metadata block (u32 words).
enum : int {
    kMetaU = 0,                    // uniques per level
    kMetaS = kMaxLevels,           // words per level
    kMetaB = 2 * kMaxLevels,       // level base offsets
    kMetaC = 3 * kMaxLevels,       // children emitted per level (scan totals)
    kMetaW = 4 * kMaxLevels,       // total words
    kMetaColl = 4 * kMaxLevels + 1,  // hash collisions seen
    kMetaRoot = 4 * kMaxLevels + 2,  // root mask
    kMetaWords = 4 * kMaxLevels + 4,
};

inline dim3 grid1d(u32 n, u32 block = 256) { return dim3(ceil_div<u32>(n ? n : 1, block)); }

}  // namespace

int build_dag(cpvs_ctx* ctx, const cpvs_pyramid* p, u32 t, u32 T, hipStream_t s, cpvs_dag** out) {
    const int n = p->n;
    const int R = n - 2;
    const bool leaf = n >= 5;
    const int bottom = leaf ? 2 : 0;
    const u32 N = p->N;

    // --- exact node counts (one readback) ---------------------------------------------
    Arena& A = ctx->arena;
    const size_t meta_bytes = arena_bytes(kMaxLevels, sizeof(unsigned long long)) + arena_bytes(kMetaWords, sizeof(u32));
    // The counts live at the front of the arena until they are read back.
    int rc = A.reserve(meta_bytes, s);
    if (rc) return rc;
    unsigned long long* d_counts = A.take<unsigned long long>(kMaxLevels);
    rc = pyramid_node_counts(p, t, T, d_counts, s);
    if (rc) return rc;
    CPVS_HIP_CHECK(hipMemcpyAsync(ctx->pinned, d_counts, sizeof(u64) * kMaxLevels, hipMemcpyDeviceToHost, s));
    CPVS_HIP_CHECK(hipStreamSynchronize(s));
    u64 cnt[kMaxLevels] = {};
    cnt[R] = 1;
    for (int L = bottom; L < R; ++L) cnt[L] = ctx->pinned[L];
    for (int L = bottom; L <= R; ++L) {
        if (cnt[L] > 0xFFFFFFF0ull) {
            set_error("build: level %d has %llu nodes (limit 2^32)", L, (unsigned long long)cnt[L]);
            return CPVS_EINVAL;
        }
    }

    // --- scratch sizing ----------------------------------------------------------------------
    u32 cmax = 1;
    for (int L = bottom; L <= R; ++L) cmax = cnt[L] > cmax ? (u32)cnt[L] : cmax;
    size_t need = meta_bytes;
    for (int L = bottom; L <= R; ++L) {
        const size_t c = cnt[L] ? cnt[L] : 1;
        need += arena_bytes(c, 8) + 5 * arena_bytes(c, 4);
    }
    if (leaf) need += arena_bytes((size_t)(cnt[2] ? cnt[2] : 1) * 8, 8);
    need += 6 * arena_bytes(cmax, 4);
    need += arena_bytes(sort_scratch_words(cmax), 4) + arena_bytes(scan_scratch_words(cmax), 4);
    need += 1 << 20;
    rc = A.reserve(need, s);  // resets the arena (the counts were already read back)
    if (rc) return rc;
    (void)A.take<unsigned long long>(kMaxLevels);
    u32* meta = A.take<u32>(kMetaWords);
    CPVS_HIP_CHECK(hipMemsetAsync(meta, 0, sizeof(u32) * kMetaWords, s));

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
    u64* leafm = leaf ? A.take<u64>((size_t)(cnt[2] ? cnt[2] : 1) * 8) : nullptr;
    u32* skey = A.take<u32>(cmax);
    u32* sval = A.take<u32>(cmax);
    u32* ktmp = A.take<u32>(cmax);
    u32* vtmp = A.take<u32>(cmax);
    u32* seg = A.take<u32>(cmax);
    u32* rep = A.take<u32>(cmax);
    u32* sort_scratch = A.take<u32>(sort_scratch_words(cmax));
    u32* scan_scratch = A.take<u32>(scan_scratch_words(cmax));

    // --- 1/2. BFS classification -------------------------------------------------------------
    const u64 root = pack_xyz(0, 0, 2 * t);  // create(): root offset (0,0,2*zTileIndex)
    CPVS_LAUNCH("set_u64", k_set_u64, dim3(1), dim3(1), 0, s, lv[R].coord, root);
    int deepest = R;  // lowest level with nodes
    for (int L = R; L >= bottom; --L) {
        const u32 c = lv[L].count;
        if (c == 0) break;
        deepest = L;
        if (leaf && L == 2) {
            CPVS_LAUNCH("leaf", k_leaf, grid1d(c, 4), dim3(256), 0, s, lv[2].coord, c, p->depth, N, T, lv[2].mask, leafm,
                               skey, sval);
            break;
        }
        CPVS_LAUNCH("classify", k_classify, grid1d(c), dim3(256), 0, s, lv[L].coord, c, p->lv[L], p->depth, N, L, T,
                           lv[L].mask);
        if (L > bottom) {
            device_scan(ChildCountIn{lv[L].mask},
                        EmitChildrenOut{lv[L].mask, lv[L].coord, lv[L].cstart, lv[L - 1].coord, lv[L - 1].count}, c,
                        OpSum{}, meta + kMetaC + L, scan_scratch, s);
        }
    }
    CPVS_HIP_CHECK(hipGetLastError());

    // --- 3. merge, bottom-up (the root is never merged) --------------------------------------
    for (int L = deepest; L < R; ++L) {
        const u32 c = lv[L].count;
        if (leaf && L == 2) {
            // keys already written by k_leaf
        } else {
            const u32* child_rank = (L > deepest) ? lv[L - 1].rank : nullptr;
            CPVS_LAUNCH("hash_inner", k_hash_inner, grid1d(c), dim3(256), 0, s, lv[L].mask, lv[L].cstart, child_rank, c, skey, sval);
        }
        radix_sort_pairs(skey, sval, ktmp, vtmp, c, 32, sort_scratch, s);
        device_scan(SegHeadIn{skey}, SegStartOut{seg}, c, OpMax{}, nullptr, scan_scratch, s);
        if (leaf && L == 2) {
            CPVS_LAUNCH("representative", k_representative<LeafEq>, grid1d(c), dim3(256), 0, s, sval, seg, c, LeafEq{lv[2].mask, leafm},
                               rep, meta + kMetaColl);
        } else {
            const u32* child_rank = (L > deepest) ? lv[L - 1].rank : nullptr;
            CPVS_LAUNCH("representative", k_representative<InnerEq>, grid1d(c), dim3(256), 0, s, sval, seg, c,
                               InnerEq{lv[L].mask, lv[L].cstart, child_rank}, rep, meta + kMetaColl);
        }
        device_scan(FirstFlagIn{rep}, FirstRankOut{lv[L].rank, lv[L].uniq}, c, OpSum{}, meta + kMetaU + L, scan_scratch, s);
        CPVS_LAUNCH("rank_copies", k_rank_copies, grid1d(c), dim3(256), 0, s, rep, c, lv[L].rank);
    }
    CPVS_LAUNCH("root_init", k_root_init, dim3(1), dim3(1), 0, s, lv[R].rank, lv[R].uniq, meta + kMetaU + R);
    CPVS_HIP_CHECK(hipGetLastError());

    // --- 4. compress: sizes, bases, one readback, write -------------------------------------
    for (int L = R; L >= deepest; --L) {
        const u32 c = lv[L].count;
        const u32 is_leaf = (leaf && L == 2) ? 1u : 0u;
        device_scan(UniqueSizeIn{lv[L].uniq, lv[L].mask, meta + kMetaU + L, is_leaf}, ArrOutExclusive{lv[L].off}, c, OpSum{},
                    meta + kMetaS + L, scan_scratch, s);
    }
    CPVS_LAUNCH("level_bases", k_level_bases, dim3(1), dim3(1), 0, s, meta + kMetaS, meta + kMetaB, meta + kMetaW, R, deepest);
    CPVS_HIP_CHECK(hipMemcpyAsync(meta + kMetaRoot, lv[R].mask, sizeof(u32), hipMemcpyDeviceToDevice, s));
    CPVS_HIP_CHECK(hipGetLastError());
    CPVS_HIP_CHECK(hipMemcpyAsync(ctx->pinned, meta, sizeof(u32) * kMetaWords, hipMemcpyDeviceToHost, s));
    CPVS_HIP_CHECK(hipStreamSynchronize(s));
    u32 hm[kMetaWords];
    std::memcpy(hm, ctx->pinned, sizeof(hm));

    for (int L = deepest + 1; L <= R; ++L) {
        if (hm[kMetaC + L] != lv[L - 1].count) {
            set_error("build: level %d emitted %u children, expected %u", L, hm[kMetaC + L], lv[L - 1].count);
            return CPVS_EINTERNAL;
        }
    }
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

    u32* d_words = nullptr;
    if (hipMalloc(&d_words, sizeof(u32) * (total ? total : 1)) != hipSuccess) {
        set_error("build: hipMalloc of %llu DAG words failed", (unsigned long long)total);
        return CPVS_ENOMEM;
    }
    for (int L = R; L >= deepest; --L) {
        const u32 U = hm[kMetaU + L];
        const u32 base = hm[kMetaB + L];
        if (leaf && L == 2) {
            CPVS_LAUNCH("write_leaf", k_write_leaf, grid1d(U), dim3(256), 0, s, lv[2].uniq, U, lv[2].mask, leafm, lv[2].off, base,
                               d_words);
        } else {
            const bool has_children = L > deepest;
            CPVS_LAUNCH("write_inner", k_write_inner, grid1d(U), dim3(256), 0, s, lv[L].uniq, U, lv[L].mask, lv[L].cstart, lv[L].off,
                               base, has_children ? lv[L - 1].rank : nullptr, has_children ? lv[L - 1].off : nullptr,
                               has_children ? hm[kMetaB + L - 1] : 0u, d_words);
        }
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        hipFree(d_words);
        return hip_fail(e, "build write", __FILE__, __LINE__);
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
    for (int L = 0; L < n && L < kMaxLevels; ++L) {
        d->nodes[L] = (L >= bottom && L <= R) ? cnt[L] : 0;
        d->uniques[L] = (L >= deepest && L <= R) ? hm[kMetaU + L] : 0;
    }
    *out = d;
    return CPVS_OK;
}

}  // namespace cpvs
