// pyramid.hip -- min/max pyramid (MinMaxHierarchy, src/MinMaxHierarchy.cpp:9-97) and the
// exact per-level octree node counts derived from it.
//
// k_pyr_tile: one 256-thread block reduces a 64x64 tile of its input level to six output
// levels.  Each thread loads a 4x4 patch with 16-byte loads (coalesced rows), reduces it in
// registers to level+1 (2x2) and level+2 (1x1), and the block finishes levels +3..+6 in LDS.
// For the root launch the input is the depth map; later launches read the float2 level the
// previous launch wrote.  Levels 1-2 are written only when requested: with leafmasks the
// build reads depth directly and pyramid levels >= 3 (src/CompressedShadow.cpp:104-106,162-166).
// Every (min, max) is combined as min(min(a,b), min(c,d)) over the 2x2 block [a b; c d]
// exactly as constructRange/setRow/find do (src/MinMaxHierarchy.cpp:29-49).
#include "internal.hpp"

namespace cpvs {

struct PyrOut {
    float2* lv[6];  // outputs of level in+1 .. in+6 (null = not stored)
};

__device__ inline float2 combine4(float2 a, float2 b, float2 c, float2 d) {
    return make_float2(smin(smin(a.x, b.x), smin(c.x, d.x)), smax(smax(a.y, b.y), smax(c.y, d.y)));
}

template <bool ROOT>
__global__ __launch_bounds__(256) void k_pyr_tile(const float* __restrict__ in, u32 W, PyrOut out) {
    __shared__ float2 s_a[16 * 16];
    __shared__ float2 s_b[8 * 8];
    const u32 t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const u32 bx = blockIdx.x * 64, by = blockIdx.y * 64;
    const u32 x0 = bx + 4 * tx, y0 = by + 4 * ty;

    float2 p[4][4];  // p[row][col] of the 4x4 input patch as (min, max)
    if (ROOT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // Streamed past the caches (nontemporal): the map is read once here; measured at
            // 8192^2 the root launch takes 60 -> 52 µs, the leaf kernel's later depth reads 302 ->
            // 306 µs.
            typedef float f4v __attribute__((ext_vector_type(4)));
            const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(in + (size_t)(y0 + r) * W + x0));
            p[r][0] = make_float2(v.x, v.x);
            p[r][1] = make_float2(v.y, v.y);
            p[r][2] = make_float2(v.z, v.z);
            p[r][3] = make_float2(v.w, v.w);
        }
    } else {
        const float2* in2 = reinterpret_cast<const float2*>(in);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float4* row = reinterpret_cast<const float4*>(in2 + (size_t)(y0 + r) * W + x0);
            const float4 v0 = row[0], v1 = row[1];
            p[r][0] = make_float2(v0.x, v0.y);
            p[r][1] = make_float2(v0.z, v0.w);
            p[r][2] = make_float2(v1.x, v1.y);
            p[r][3] = make_float2(v1.z, v1.w);
        }
    }
    float2 q[2][2];
#pragma unroll
    for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int ox = 0; ox < 2; ++ox)
            q[oy][ox] = combine4(p[2 * oy][2 * ox], p[2 * oy][2 * ox + 1], p[2 * oy + 1][2 * ox], p[2 * oy + 1][2 * ox + 1]);
    if (out.lv[0]) {
        const u32 W1 = W / 2;
#pragma unroll
        for (int oy = 0; oy < 2; ++oy) {
            float4* dst = reinterpret_cast<float4*>(out.lv[0] + (size_t)(y0 / 2 + oy) * W1 + x0 / 2);
            *dst = make_float4(q[oy][0].x, q[oy][0].y, q[oy][1].x, q[oy][1].y);
        }
    }
    const float2 r2 = combine4(q[0][0], q[0][1], q[1][0], q[1][1]);
    if (out.lv[1]) out.lv[1][(size_t)(y0 / 4) * (W / 4) + x0 / 4] = r2;
    s_a[ty * 16 + tx] = r2;
    __syncthreads();

    // Levels +3 .. +6 inside the block: 16x16 -> 8x8 -> 4x4 -> 2x2 -> 1x1.
    float2* src = s_a;
    float2* dst = s_b;
    u32 w = 16;
#pragma unroll
    for (int k = 2; k < 6; ++k) {
        const u32 hw = w / 2;
        if (t < hw * hw) {
            const u32 x = t % hw, y = t / hw;
            const float2 v = combine4(src[(2 * y) * w + 2 * x], src[(2 * y) * w + 2 * x + 1], src[(2 * y + 1) * w + 2 * x],
                                      src[(2 * y + 1) * w + 2 * x + 1]);
            dst[y * hw + x] = v;
            if (out.lv[k]) {
                const u32 Wk = W >> (k + 1);
                out.lv[k][(size_t)(by >> (k + 1)) * Wk + (size_t)(y) * Wk + (bx >> (k + 1)) + x] = v;
            }
        }
        __syncthreads();
        float2* tmp = src;
        src = dst;
        dst = tmp;
        w = hw;
    }
}

// One block reduces an input level of width W <= 32 (all of it in LDS) down to 1x1.
template <bool ROOT>
__global__ __launch_bounds__(256) void k_pyr_small(const float* __restrict__ in, u32 W, PyrOut out, int nout) {
    __shared__ float2 s_a[32 * 32];
    __shared__ float2 s_b[16 * 16];
    for (u32 i = threadIdx.x; i < W * W; i += blockDim.x) {
        if (ROOT) {
            const float v = in[i];
            s_a[i] = make_float2(v, v);
        } else {
            s_a[i] = reinterpret_cast<const float2*>(in)[i];
        }
    }
    __syncthreads();
    float2* src = s_a;
    float2* dst = s_b;
    u32 w = W;
    for (int k = 0; k < nout; ++k) {
        const u32 hw = w / 2;
        for (u32 i = threadIdx.x; i < hw * hw; i += blockDim.x) {
            const u32 x = i % hw, y = i / hw;
            const float2 v = combine4(src[(2 * y) * w + 2 * x], src[(2 * y) * w + 2 * x + 1], src[(2 * y + 1) * w + 2 * x],
                                      src[(2 * y + 1) * w + 2 * x + 1]);
            dst[i] = v;
            if (out.lv[k]) out.lv[k][i] = v;
        }
        __syncthreads();
        float2* tmp = src;
        src = dst;
        dst = tmp;
        w = hw;
    }
}

size_t pyramid_levels_bytes(u32 N, bool store_all) {
    int n = 1;
    while ((1u << (n - 1)) < N) ++n;
    const bool keep12 = store_all || n < 5;
    size_t total = 0;
    for (int L = 1; L < n; ++L)
        if (L > 2 || keep12) total += (size_t)(N >> L) * (N >> L);
    return arena_bytes(total ? total : 1, sizeof(float2));
}

int pyramid_build_levels(cpvs_pyramid* p, bool store_all, hipStream_t s, Arena* into, bool append) {
    const u32 N = p->N;
    const int n = p->n;
    // Store levels 1-2 only when asked, or when the octree uses them (no leafmasks, N < 16).
    const bool keep12 = store_all || n < 5;
    size_t total = 0;
    size_t off[kMaxLevels] = {};
    for (int L = 1; L < n; ++L) {
        if (L <= 2 && !keep12) continue;
        off[L] = total;
        const size_t w = N >> L;
        total += w * w;
    }
    if (into) {
        if (!append) {
            const int rc = into->reserve(arena_bytes(total ? total : 1, sizeof(float2)), s);
            if (rc) return rc;
        }
        p->levels = into->take<float2>(total ? total : 1);
        p->levels_owned = false;
        if (into->used > into->cap) {
            set_error("pyramid: the level arena holds %zu bytes, %zu needed", into->cap, into->used);
            return CPVS_EINTERNAL;
        }
    } else if (!(p->levels = static_cast<float2*>(p->ctx->buffers.take(sizeof(float2) * (total ? total : 1))))) {
        set_error("pyramid: hipMalloc of %zu bytes failed", sizeof(float2) * total);
        return CPVS_ENOMEM;
    }
    for (int L = 1; L < n; ++L) p->lv[L] = (L <= 2 && !keep12) ? nullptr : p->levels + off[L];

    int in_level = 0;
    u32 W = N;
    while (W > 1) {
        PyrOut o = {};
        const float* src = (in_level == 0) ? p->depth : reinterpret_cast<const float*>(p->lv[in_level]);
        if (W >= 64) {
            for (int k = 0; k < 6; ++k) o.lv[k] = (in_level + 1 + k < n) ? p->lv[in_level + 1 + k] : nullptr;
            const dim3 grid(W / 64, W / 64);
            if (in_level == 0)
                CPVS_LAUNCH("pyr_tile_root", k_pyr_tile<true>, grid, dim3(256), 0, s, src, W, o);
            else
                CPVS_LAUNCH("pyr_tile", k_pyr_tile<false>, grid, dim3(256), 0, s, src, W, o);
            in_level += 6;
            W /= 64;
        } else {
            int nout = 0;
            for (u32 w = W; w > 1; w /= 2) ++nout;
            for (int k = 0; k < nout && k < 6; ++k) o.lv[k] = p->lv[in_level + 1 + k];
            if (in_level == 0)
                CPVS_LAUNCH("pyr_small", k_pyr_small<true>, dim3(1), dim3(256), 0, s, src, W, o, nout);
            else
                CPVS_LAUNCH("pyr_small", k_pyr_small<false>, dim3(1), dim3(256), 0, s, src, W, o, nout);
            in_level += nout;
            W = 1;
        }
        if (in_level < n - 1 && p->lv[in_level] == nullptr) {
            set_error("pyramid: intermediate level %d not stored", in_level);
            return CPVS_EINTERNAL;
        }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pyramid launch", __FILE__, __LINE__);
    return CPVS_OK;
}

// ------------------------------------------------------------------------------------------
// Exact node counts.  A cell at level L is partial iff float(z+1) > mn*h and float(z) < mx*h
// (src/CompressedShadowUtil.h:35-43 with h = (N>>L)*T, src/CompressedShadowUtil.cpp:16-18),
// i.e. z in [floor(mn*h), ceil(mx*h)).  Without NaN a partial cell implies a partial parent
// cell, so the nodes of level L-1 are exactly the partial level-L cells inside the z-tile's slab
// [t*(N>>L), (t+1)*(N>>L)).  A NaN texel can be partial where its parent's cell is not; such a
// map's counts exceed what the classification emits, the build sees the difference after its
// readback, and counts again with `exact`, which cuts every cell by its ancestors' cells
// (reachable_cells).
// ------------------------------------------------------------------------------------------

struct CountArgs {
    int nlev;
    u32 block_start[kMaxLevels + 1];  // first block of each level in the grid
    u32 width[kMaxLevels];
    int level[kMaxLevels];
    const float2* lv[kMaxLevels];
    u32 N, T, t;
    u32* partials;               // partials[block] = the block's partial cells (host memory; may be null)
    uint4* col_info;             // per level-3 texel (leaf column): x = partial cells, z = first partial z
                                 // (y is left for the column's start); may be null
    const float2* by_level[kMaxLevels];  // exact: level M's texels, M in [level[0], R]
    int R;
    u32 exact;                   // count only cells under partial cells at every level above
};

constexpr u32 kCountPerThread = 8;
constexpr u32 kCountTile = 256 * kCountPerThread;

__global__ __launch_bounds__(256) void k_node_counts(CountArgs a) {
    __shared__ u32 s_wave[4];
    int k = 0;
    while (k + 1 < a.nlev && blockIdx.x >= a.block_start[k + 1]) ++k;
    const u32 w = a.width[k];
    const int L = a.level[k];
    const float h = (float)((a.N >> L) * a.T);
    const int zlo = (int)(a.t * (a.N >> L)), zhi = (int)((a.t + 1) * (a.N >> L));
    uint4* colc = (L == 3) ? a.col_info : nullptr;
    const u32 base = (blockIdx.x - a.block_start[k]) * kCountTile;
    u32 acc = 0;
#pragma unroll
    for (u32 r = 0; r < kCountPerThread; ++r) {
        const u32 j = base + r * 256 + threadIdx.x;
        if (j < w * w) {
            int s;
            const u32 c = a.exact ? reachable_cells(a.by_level, j % w, j / w, L, a.R, a.N, a.T, zlo, zhi, &s)
                                  : partial_cells(a.lv[k][j], h, zlo, zhi, &s);
            if (colc) colc[j] = make_uint4(c, 0u, (u32)max(s, 0), 0u);
            acc += c;
        }
    }
    // Block reduction; every block writes its own partial (a block stays within one level and
    // sums at most kCountTile columns of at most N/8 cells, so 32 bits hold it).  The host adds
    // the partials up: no counter to clear before the launch and no copy after it.
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0 && a.partials) {
        a.partials[blockIdx.x] = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        // The root level (2x2 texels, one block): the root's childmask follows the partials.  With
        // no partial child the DAG is that one word.  Cells z = 2t, 2t+1 of level R, as
        // classify_node reads them for the root (0, 0, 2t).
        if (k == a.nlev - 1) {
            const float2* l = a.lv[k];
            u32 m = 0;
#pragma unroll
            for (u32 ch = 0; ch < 8; ++ch) {
                const float2 v = l[ch & 3];
                m |= classify_cell(2 * a.t + (ch >> 2), v.x * h, v.y * h) << (2 * ch);
            }
            a.partials[a.block_start[a.nlev]] = m;
        }
    }
}

void sum_node_counts(const CountLayout& lay, const u32* partials, u64* cnt) {
    for (int k = 0; k < lay.nlev; ++k) {
        u64 c = 0;
        for (u32 b = lay.block_start[k]; b < lay.block_start[k + 1]; ++b) c += partials[b];
        cnt[lay.level[k] - 1] = c;  // partial cells of level L are the nodes of octree level L-1
    }
}

int pyramid_node_counts(const cpvs_pyramid* p, u32 t, u32 T, bool leaf, u32* partials, uint4* col_info, hipStream_t s,
                        CountLayout* layout, bool exact) {
    CountArgs a = {};
    const int n = p->n;
    const int R = n - 2;
    leaf = leaf && n >= 5;
    const int Lmin = leaf ? 3 : 1;
    a.N = p->N;
    a.T = T;
    a.t = t;
    a.R = R;
    a.exact = exact ? 1u : 0u;
    for (int L = Lmin; L <= R; ++L) a.by_level[L] = p->lv[L];
    a.partials = partials;
    a.col_info = leaf ? col_info : nullptr;
    u32 blocks = 0;
    for (int L = Lmin; L <= R; ++L) {
        const u32 w = p->N >> L;
        a.block_start[a.nlev] = blocks;
        a.width[a.nlev] = w;
        a.level[a.nlev] = L;
        a.lv[a.nlev] = p->lv[L];
        if (!p->lv[L]) {
            set_error("node counts: pyramid level %d not stored (a build without leafmasks needs a pyramid made with "
                      "every level: create it in that mode or with all levels)", L);
            return CPVS_EINTERNAL;
        }
        blocks += ceil_div<u32>(w * w, kCountTile);
        ++a.nlev;
    }
    a.block_start[a.nlev] = blocks;
    if (partials && blocks + 1 > kPinnedPartials) {
        set_error("node counts: %u count blocks exceed the %zu pinned partials", blocks, kPinnedPartials);
        return CPVS_EINTERNAL;
    }
    if (layout) {
        layout->nlev = a.nlev;
        for (int k = 0; k < a.nlev; ++k) layout->level[k] = a.level[k];
        for (int k = 0; k <= a.nlev; ++k) layout->block_start[k] = a.block_start[k];
    }
    if (blocks) CPVS_LAUNCH("node_counts", k_node_counts, dim3(blocks), dim3(256), 0, s, a);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

}  // namespace cpvs
