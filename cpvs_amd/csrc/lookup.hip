// lookup.hip -- shadow lookups: shader/traverse.cs (via CompressedShadowContainer::evaluate,
// src/CompressedShadowContainer.cpp:93-124) and CompressedShadow::traverse
// (src/CompressedShadow.cpp:404-463).
//
// One thread per query.  The walk is a chain of dependent 4-byte loads, one per level; the
// upper levels are shared by every query and stay in L2, so throughput is set by how many
// walks are in flight and by the locality of neighbouring queries.  The image entry point
// gives each wave an 8x8 pixel tile so neighbouring lanes follow the same DAG paths.
// Compiled with -ffp-contract=off: the matrix product, the divide and the quantisation round
// exactly as written, as the reference's float expressions do.
#include "internal.hpp"

namespace cpvs {

struct GridDag {
    const u32* __restrict__ dag;
    const u32* __restrict__ grid;
    int dl, gl;
    u32 words;  // DAG word count: every read is bounds-checked against it
    u32 leaf;   // 1: leafmask nodes at level 2; 0: the DAG goes down to level 0
    const u64* __restrict__ top;  // the walk's state after the top topk levels (null: none)
    u32 topk;
};

// Workgroups are dealt to the XCDs in turn (block b to XCD b mod kXcds), each XCD with its own L2.
// XCD x takes the contiguous workgroups [x*chunk, (x+1)*chunk): neighbouring queries (an image's
// neighbouring pixels) then walk the same DAG nodes through one L2, and each L2 holds the part of
// the DAG under its share of the queries rather than all of it.
constexpr u32 kXcds = 8;
__device__ inline u64 xcd_query(u32 chunk) {
    return (u64)((blockIdx.x % kXcds) * chunk + blockIdx.x / kXcds) * blockDim.x + threadIdx.x;
}

// Threads per workgroup of the flat and PCF lookups (measured on 16.6 M G-buffer queries: 64, 128,
// 256, 512 and 1024 threads took 0.301, 0.297, 0.289, 0.277 and 0.276 ms walking every level; with
// the 7-level top table, 512 and 1024 threads take 0.159 and 0.150 ms).
constexpr u32 kLookupThreads = 1024;

// traverse.cs:43-48: path = int(((ndc + 1) * 0.5) * float(R - 1)), clamped to [0, R-1]
// (out-of-range positions index outside the grid in the reference).
__device__ inline u32 quantise(float v, u32 R) {
    v = v + 1.0f;
    v = v * 0.5f;
    v = v * (float)(R - 1);
    if (!(v > 0.0f)) return 0u;
    if (v >= (float)(R - 1)) return R - 1;
    return (u32)v;
}

// traverse.cs:89-133: the walk below one grid cell whose DAG starts at word `base` of `dag`
// (pointers inside a DAG are relative to its first word, combineDAGs keeps them so).  With leaf
// the inner levels end at 3 and level 2 holds leafmask nodes (MIN_LEVEL 3, traverse.cs:34-38);
// without, the inner levels go down to level 0, whose children are single voxels, lit or shadow
// (absoluteVisible, src/CompressedShadowUtil.cpp:34-40), never partial; a walk that passed level 0
// would end visible (traverse.cs:130-132).
// walk_node continues a walk at node `off` of level level0, whose mask m (= dag[off], off < words)
// the caller has read; walk_from reads it first (walk_dag starts at the root).
__device__ inline u32 walk_node(const u32* __restrict__ dag, u32 words, u32 leaf, u32 base, u32 off, u32 m, int level0,
                                u32 px, u32 py, u32 pz) {
    const int last = leaf ? 3 : 0;
    for (int level = level0; level >= last; --level) {
        const u32 ci = (((px >> level) & 1u) << 1) | (((py >> level) & 1u) << 2) | (((pz >> level) & 1u) << 3);
        const u32 v = (m >> ci) & 3u;
        if (v != 2u) return v;  // 0 shadow, 1 visible (3 never occurs)
        if (level == 0) return 1u;  // a partial voxel does not occur
        const u32 co = (u32)__builtin_popcount(m & 0xAAAAu & ((1u << ci) - 1u));
        if (off + 1 + co >= words) return 0u;
        off = base + dag[off + 1 + co];
        if (off >= words) return 0u;
        m = dag[off];
    }
    if (!leaf) return 1u;
    const u32 ci = (pz & 7u) * 2u;
    const u32 v = (m >> ci) & 3u;
    if (v != 2u) return v;
    const u32 co = (u32)__builtin_popcount(m & 0xAAAAu & ((1u << ci) - 1u));
    const u32 bit = (px & 7u) + 8u * (py & 7u);
    const u32 idx = off + co * 2 + 1 + (bit >> 5);
    if (idx >= words) return 0u;
    const u32 w = dag[idx];
    return (w >> (bit & 31u)) & 1u;
}
__device__ inline u32 walk_from(const u32* __restrict__ dag, u32 words, u32 leaf, u32 base, u32 off, int level0,
                                u32 px, u32 py, u32 pz) {
    if (off >= words) return 0u;
    return walk_node(dag, words, leaf, base, off, dag[off], level0, px, py, pz);
}
__device__ inline u32 walk_dag(const u32* __restrict__ dag, u32 words, int dl, u32 leaf, u32 base, u32 px, u32 py,
                               u32 pz) {
    return walk_from(dag, words, leaf, base, base, dl - 2, px, py, pz);
}

// A container's top table: for every grid cell and every cell of its DAG's level dl-2-K (K = topk),
// the result of walking the K levels above it: the node offset there (low half) with that node's
// mask (high half), or the value a walk that stopped earlier returns (kTopShadow 0, kTopVisible 1,
// kTopThree 3 for a child state 3, which no valid DAG holds; high half 0).  Entry (cell, t) with
// t = z * 4^K + y * 2^K + x over the K path bits below the cell's.  A lookup reads one 8-byte entry
// instead of walking K levels and reading the next node's mask: the same result by construction.
// Valid offsets lie below kTopThree (the container checks its word count).
constexpr u32 kTopShadow = 0xFFFFFFFFu, kTopVisible = 0xFFFFFFFEu, kTopThree = 0xFFFFFFFDu;

// One thread per entry of level K-1 (the parent): the walk over the parent's K-1 path bits, then
// the last level for its 8 children from the one node they share (one mask load, a pointer per
// partial child).  Each child's value is the one a per-entry walk of all K bits gives, checks
// included.
__global__ __launch_bounds__(256) void k_build_top(const u32* __restrict__ dag, u32 words, const u32* __restrict__ grid,
                                                   u64 parents, u32 K, u64* __restrict__ top) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= parents) return;
    const u32 P = K - 1, pm = (1u << P) - 1u;
    const u32 t = (u32)(i & ((1ull << (3 * P)) - 1ull));
    const u32 px = t & pm, py = (t >> P) & pm, pz = t >> (2 * P);
    const u64 cell = i >> (3 * P);
    const u32 base = grid ? grid[cell] : 0u;
    u32 e = base;      // a sentinel cell keeps its sentinel
    bool node = false; // e is the parent's node offset (possibly past the words)
    if (base < kTopVisible) {
        u32 off = base;
        e = kTopShadow;  // a read past the words ends the walk as shadow, as walk_from does
        node = true;
        for (int b = (int)P - 1; b >= 0; --b) {
            const u32 ci = (((px >> b) & 1u) << 1) | (((py >> b) & 1u) << 2) | (((pz >> b) & 1u) << 3);
            if (off >= words) {
                node = false;
                break;
            }
            const u32 m = dag[off];
            const u32 v = (m >> ci) & 3u;
            if (v != 2u) {
                e = (v == 0u) ? kTopShadow : (v == 1u) ? kTopVisible : kTopThree;
                node = false;
                break;
            }
            const u32 co = (u32)__builtin_popcount(m & 0xAAAAu & ((1u << ci) - 1u));
            if (off + 1 + co >= words) {
                node = false;
                break;
            }
            off = base + dag[off + 1 + co];
        }
        if (node) e = off;
    }
    const bool inside = node && e < words;
    const u32 m = inside ? dag[e] : 0u;
    u64* out = top + ((size_t)cell << (3 * K));
    u32 r[8];
#pragma unroll
    for (u32 c = 0; c < 8; ++c) {
        const u32 cx = c & 1u, cy = (c >> 1) & 1u, cz = c >> 2;
        r[c] = node ? kTopShadow : e;  // a parent past the words: shadow, as walk_from's first check
        if (inside) {
            const u32 ci = (cx << 1) | (cy << 2) | (cz << 3);
            const u32 v = (m >> ci) & 3u;
            if (v != 2u) {
                r[c] = (v == 0u) ? kTopShadow : (v == 1u) ? kTopVisible : kTopThree;
            } else {
                const u32 co = (u32)__builtin_popcount(m & 0xAAAAu & ((1u << ci) - 1u));
                if (e + 1 + co < words) {
                    const u32 nx = base + dag[e + 1 + co];
                    r[c] = (nx < words) ? nx : kTopShadow;  // walk_from's first check on it
                }
            }
        }
    }
#pragma unroll
    for (u32 c = 0; c < 8; ++c) {
        const u32 cx = c & 1u, cy = (c >> 1) & 1u, cz = c >> 2;
        const u32 idx = (((2 * pz + cz) << (2 * K)) | ((2 * py + cy) << K)) | (2 * px + cx);
        const u32 cm = (r[c] < kTopThree) ? dag[r[c]] : 0u;  // the child node's mask (r[c] < words)
        out[idx] = ((u64)cm << 32) | r[c];
    }
}

int launch_build_top(const u32* dag, u32 words, const u32* grid, u32 cells, int dl, u32 K, u64* top, hipStream_t s) {
    (void)dl;
    if (K == 0) return CPVS_OK;
    const u64 parents = (u64)cells << (3 * (K - 1));
    CPVS_LAUNCH("build_top", k_build_top, dim3((u32)((parents + 255) / 256)), dim3(256), 0, s, dag, words, grid, parents,
                K, top);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

// traverse.cs:75-88: the grid cell, its sentinels, then the walk (from the top table's entry when
// the container has one).
__device__ inline u32 walk(const GridDag& g, u32 px, u32 py, u32 pz) {
    const u32 G = 1u << g.gl;
    const u32 sh = (u32)(g.dl - 1);
    const u32 cell = ((pz >> sh) * G + (py >> sh)) * G + (px >> sh);
    const u32 base = g.grid[cell];
    if (base == 0xFFFFFFFFu) return 0u;
    if (base == 0xFFFFFFFEu) return 1u;
    if (g.topk) {
        const u32 K = g.topk, s = sh - K, msk = (1u << K) - 1u;
        const u32 t = (((pz >> s) & msk) << (2 * K)) | (((py >> s) & msk) << K) | ((px >> s) & msk);
        const u64 te = g.top[((size_t)cell << (3 * K)) + t];
        const u32 e = (u32)te;
        if (e >= kTopThree) return (e == kTopShadow) ? 0u : (e == kTopVisible) ? 1u : 3u;
        return walk_node(g.dag, g.words, g.leaf, base, e, (u32)(te >> 32), g.dl - 2 - (int)K, px, py, pz);
    }
    return walk_dag(g.dag, g.words, g.dl, g.leaf, base, px, py, pz);
}

// traverse.cs:141-144: p = lightViewProj * (pos, 1); p /= p.w; then the path.
__device__ inline void project(const GridDag& g, const Mat4& M, float x, float y, float z, u32& px, u32& py, u32& pz) {
    float r[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float acc = M.m[0 * 4 + c] * x;
        acc = acc + M.m[1 * 4 + c] * y;
        acc = acc + M.m[2 * 4 + c] * z;
        acc = acc + M.m[3 * 4 + c] * 1.0f;
        r[c] = acc;
    }
    const u32 R = 1u << (g.dl + g.gl - 1);
    // x / 1 is x exactly (infinities and NaN included), so an affine matrix's w = 1 skips the
    // three IEEE divides, about ten instructions each.
    float nx = r[0], ny = r[1], nz = r[2];
    if (r[3] != 1.0f) {
        nx = nx / r[3];
        ny = ny / r[3];
        nz = nz / r[3];
    }
    px = quantise(nx, R);
    py = quantise(ny, R);
    pz = quantise(nz, R);
}

__device__ inline u32 lookup_point(const GridDag& g, const Mat4& M, float x, float y, float z) {
    u32 px, py, pz;
    project(g, M, x, y, z, px, py, pz);
    return walk(g, px, py, pz);
}

// Percentage-closer filtering (the reference declares the kernel width, filterSize, and leaves it
// unused: shader/traverse.cs:16-17): the fraction of lit leaf cells among the f x f paths
// (px + dx, py + dy, pz), dx and dy in [-(f-1)/2, f/2] (integer halves), each clamped to the grid,
// as float(lit) / float(f*f).  f = 1 is the plain lookup.
__global__ __launch_bounds__(kLookupThreads) void k_evaluate_pcf(GridDag g, Mat4 M, const float* __restrict__ pos, u32 stride,
                                                      u64 count, u32 f, float* __restrict__ vis, u32 chunk) {
    const u64 i = xcd_query(chunk);
    if (i >= count) return;
    const float* q = pos + i * stride;
    u32 px, py, pz;
    project(g, M, q[0], q[1], q[2], px, py, pz);
    const int R = 1 << (g.dl + g.gl - 1);
    const int lo = -(int)((f - 1) / 2);
    u32 lit = 0;
    for (u32 a = 0; a < f; ++a) {
        const int y = min(max((int)py + lo + (int)a, 0), R - 1);
        for (u32 b = 0; b < f; ++b) {
            const int x = min(max((int)px + lo + (int)b, 0), R - 1);
            lit += walk(g, (u32)x, (u32)y, pz);
        }
    }
    vis[i] = (float)lit / (float)(f * f);
}

int launch_evaluate_pcf(const u32* dag, u32 words, const u32* grid, int dl, int gl, int leaf, const u64* top, u32 topk,
                        const float* pos, u32 stride, u64 count, const Mat4& m, u32 f, float* vis, hipStream_t s) {
    if (count == 0) return CPVS_OK;
    GridDag g{dag, grid, dl, gl, words, leaf ? 1u : 0u, top, top ? topk : 0u};
    const u64 blocks = (count + kLookupThreads - 1) / kLookupThreads;
    if (blocks > 0x7FFFFFF0ull) {
        set_error("evaluate_pcf: %llu points exceed one launch", (unsigned long long)count);
        return CPVS_EINVAL;
    }
    const u32 chunk = (u32)((blocks + kXcds - 1) / kXcds);
    CPVS_LAUNCH("evaluate_pcf", k_evaluate_pcf, dim3(chunk * kXcds), dim3(kLookupThreads), 0, s, g, m, pos, stride, count, f, vis, chunk);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

__global__ __launch_bounds__(kLookupThreads) void k_evaluate4(GridDag g, Mat4 M, const float4* __restrict__ pos, u64 count,
                                                   uint8_t* __restrict__ vis, u32 chunk) {
    const u64 i = xcd_query(chunk);
    if (i >= count) return;
    // Positions and results stream past the caches (nontemporal), which keeps the L2 for the DAG:
    // 0.276 -> 0.261 ms for the 16.6 M G-buffer lookups at 8192^2.
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v q = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(pos) + i);
    __builtin_nontemporal_store((uint8_t)lookup_point(g, M, q.x, q.y, q.z), vis + i);
}

__global__ __launch_bounds__(kLookupThreads) void k_evaluate3(GridDag g, Mat4 M, const float* __restrict__ pos, u64 count,
                                                   uint8_t* __restrict__ vis, u32 chunk) {
    const u64 i = xcd_query(chunk);
    if (i >= count) return;
    const float* q = pos + 3 * i;
    vis[i] = (uint8_t)lookup_point(g, M, q[0], q[1], q[2]);
}

// Image form: a 256-thread block covers a 32x8 pixel strip as four 8x8 tiles, one per wave.  The
// blocks (gx per 8 rows, nblocks in all) are dealt to the XCDs in contiguous runs of chunk.
__global__ __launch_bounds__(256) void k_evaluate_image(GridDag g, Mat4 M, const float4* __restrict__ pos, u32 width,
                                                        u32 height, uint8_t* __restrict__ vis, u32 gx, u32 nblocks,
                                                        u32 chunk) {
    const u32 b = (blockIdx.x % kXcds) * chunk + blockIdx.x / kXcds;
    if (b >= nblocks) return;
    const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u32 x = (b % gx) * 32 + wave * 8 + (lane & 7);
    const u32 y = (b / gx) * 8 + (lane >> 3);
    if (x >= width || y >= height) return;
    const size_t i = (size_t)y * width + x;
    const float4 q = pos[i];
    vis[i] = (uint8_t)lookup_point(g, M, q.x, q.y, q.z);
}

int launch_evaluate(const u32* dag, u32 words, const u32* grid, int dl, int gl, int leaf, const u64* top, u32 topk,
                    const float* pos, u32 stride, u64 count, const Mat4& m, uint8_t* vis, hipStream_t s) {
    if (count == 0) return CPVS_OK;
    GridDag g{dag, grid, dl, gl, words, leaf ? 1u : 0u, top, top ? topk : 0u};
    const u64 blocks = (count + kLookupThreads - 1) / kLookupThreads;
    if (blocks > 0x7FFFFFF0ull) {
        set_error("evaluate: %llu points exceed one launch", (unsigned long long)count);
        return CPVS_EINVAL;
    }
    const u32 chunk = (u32)((blocks + kXcds - 1) / kXcds);
    if (stride == 4)
        CPVS_LAUNCH("evaluate4", k_evaluate4, dim3(chunk * kXcds), dim3(kLookupThreads), 0, s, g, m, reinterpret_cast<const float4*>(pos),
                           count, vis, chunk);
    else
        CPVS_LAUNCH("evaluate3", k_evaluate3, dim3(chunk * kXcds), dim3(kLookupThreads), 0, s, g, m, pos, count, vis, chunk);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

int launch_evaluate_image(const u32* dag, u32 words, const u32* grid, int dl, int gl, int leaf, const u64* top, u32 topk,
                          const float* pos, u32 width, u32 height, const Mat4& m, uint8_t* vis, hipStream_t s) {
    if (width == 0 || height == 0) return CPVS_OK;
    GridDag g{dag, grid, dl, gl, words, leaf ? 1u : 0u, top, top ? topk : 0u};
    const u64 gx = ceil_div<u32>(width, 32), nblocks = gx * ceil_div<u32>(height, 8);
    if (nblocks > 0x7FFFFFF0ull) {
        set_error("evaluate_image: %ux%u pixels exceed one launch", width, height);
        return CPVS_EINVAL;
    }
    const u32 chunk = (u32)((nblocks + kXcds - 1) / kXcds);
    CPVS_LAUNCH("evaluate_image", k_evaluate_image, dim3(chunk * kXcds), dim3(256), 0, s, g, m, reinterpret_cast<const float4*>(pos),
                width, height, vis, (u32)gx, (u32)nblocks, chunk);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

// CompressedShadow::traverse (src/CompressedShadow.cpp:404-463) with getPathFromNDC
// (src/CompressedShadowUtil.h:70-75): no grid, pointers absolute, PARTIAL when the walk
// passes level 0 without leafmasks.
__global__ __launch_bounds__(256) void k_traverse_single(const u32* __restrict__ w, u32 words, int n,
                                                         const float* __restrict__ ndc, u64 count, int try_leaf,
                                                         int* __restrict__ out) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int res = (1 << (n - 1)) - 1;
    int p[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = ndc[3 * i + c] + 1.0f;
        v = v * 0.5f;
        p[c] = (int)(v * (float)res);
    }
    u32 off = 0;
    int level = n - 2;
    const int minL = try_leaf ? 3 : 0;
    int result = 2;
    bool done = false;
    while (level >= minL) {
        const int b = 1 << level;
        const u32 c = ((p[0] & b) ? 1u : 0u) + ((p[1] & b) ? 2u : 0u) + ((p[2] & b) ? 4u : 0u);
        if (off >= words) { result = 0; done = true; break; }
        const u32 m = w[off];
        if (m & (1u << (2 * c))) { result = 1; done = true; break; }
        if (!(m & (2u << (2 * c)))) { result = 0; done = true; break; }
        const u32 slot = off + 1 + (u32)__builtin_popcount(m & 0xAAAAu & ((1u << (2 * c)) - 1u));
        if (slot >= words) { result = 0; done = true; break; }
        off = w[slot];
        --level;
    }
    if (!done && try_leaf && off < words) {
        const u32 c = (u32)(p[2] & 7);
        const u32 m = w[off];
        if (m & (1u << (2 * c))) result = 1;
        else if (!(m & (2u << (2 * c)))) result = 0;
        else {
            const u32 idx = off + (u32)__builtin_popcount(m & 0xAAAAu & ((1u << (2 * c)) - 1u)) * 2 + 1;
            const int bit = (p[0] & 7) + 8 * (p[1] & 7);
            const u32 widx = idx + (bit < 32 ? 0u : 1u);
            const u32 v = (widx >= words) ? 0u : (bit < 32) ? (w[widx] & (1u << bit)) : (w[widx] & (1u << (bit - 32)));
            result = v ? 1 : 0;
        }
    }
    out[i] = result;
}

int launch_traverse_single(const u32* dag, u32 words, int n, const float* ndc, u64 count, int try_leaf, int* out,
                           hipStream_t s) {
    if (count == 0) return CPVS_OK;
    CPVS_LAUNCH("traverse_single", k_traverse_single, dim3((u32)((count + 255) / 256)), dim3(256), 0, s, dag, words, n, ndc, count,
                       try_leaf, out);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

// ------------------------------------------------------------------------------------------
// DAG sets: many independent DAGs, each walked as a one-cell container
// (CompressedShadowContainer(unique_ptr<CompressedShadow>), src/CompressedShadowContainer.h:27-32,
// with gridLevels 0 in shader/traverse.cs), queried in one launch.  Query i names its DAG by
// ids[i], or by i mod the set size when ids is null.  Queries are NDC (the product with an
// identity matrix and the divide by w = 1 are exact, so this equals evaluate() with identity).
// ------------------------------------------------------------------------------------------

__device__ inline uint8_t dagset_lookup(const DagDesc& d, const float* __restrict__ q) {
    const u32 R = 1u << (d.dl - 1);
    const u32 px = quantise(q[0], R), py = quantise(q[1], R), pz = quantise(q[2], R);
    if (d.base == 0xFFFFFFFFu) return 0;
    if (d.base == 0xFFFFFFFEu) return 1;
    if (d.topk) {  // the DAG's top table: the first topk levels in one read (as a one-cell container's)
        const u32 K = d.topk, s = d.dl - 1 - K, msk = (1u << K) - 1u;
        const u32 t = (((pz >> s) & msk) << (2 * K)) | (((py >> s) & msk) << K) | ((px >> s) & msk);
        const u64 te = d.top[t];
        const u32 e = (u32)te;
        if (e >= kTopThree) return (e == kTopShadow) ? 0 : (e == kTopVisible) ? 1 : 3;
        return (uint8_t)walk_node(d.words, d.nwords, d.leaf, 0u, e, (u32)(te >> 32), (int)d.dl - 2 - (int)K, px, py, pz);
    }
    return (uint8_t)walk_dag(d.words, d.nwords, (int)d.dl, d.leaf, 0u, px, py, pz);
}

// Query i names its DAG by ids[i] (or uses `one` when set is null).
__global__ __launch_bounds__(256) void k_dagset_eval(const DagDesc* __restrict__ set, u32 nset, DagDesc one,
                                                     const u32* __restrict__ ids, const float* __restrict__ ndc, u32 stride,
                                                     u32 count, uint8_t* __restrict__ vis) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    DagDesc d = one;
    if (set) {
        const u32 id = ids[i];
        if (id >= nset) {
            vis[i] = kBadId;
            return;
        }
        d = set[id];
    }
    vis[i] = dagset_lookup(d, ndc + (size_t)i * stride);
}

// Round-robin queries (query i -> DAG i mod nset), walked DAG by DAG.  A workgroup of kRRGroup
// waves takes kRRGroup consecutive DAGs (a group) and 64 consecutive rows of queries (row r holds
// the queries r*nset .. r*nset + nset-1): it loads the tile of positions row by row (each row's
// group of points is contiguous, so every line is fetched once), one wave per DAG walks its 64
// queries from LDS (all lanes on one DAG's upper levels), and the results go back row by row
// through LDS.  Groups are spread over the XCDs in contiguous ranges, so each XCD's L2 holds the
// upper levels of its own DAGs only.
constexpr u32 kRRGroup = 16;

__global__ __launch_bounds__(64 * kRRGroup) void k_dagset_eval_rr(const DagDesc* __restrict__ set, u32 nset, u32 ngroups,
                                                               u32 gper, u32 kblocks, const float* __restrict__ ndc,
                                                               u32 stride, u64 count, uint8_t* __restrict__ vis) {
    __shared__ float s_q[kRRGroup][3][64];
    __shared__ uint8_t s_v[64][kRRGroup];
    // XCD x (blocks x, x + kXcds, ...) takes groups [x*gper, (x+1)*gper), every row block of each.
    const u32 x = blockIdx.x % kXcds, c = blockIdx.x / kXcds;
    const u32 grp = x * gper + c / kblocks, kb = c % kblocks;
    if (grp >= ngroups || c >= gper * kblocks) return;  // whole workgroup: no barrier is left waiting
    const u32 t = threadIdx.x;
    const u32 r = t / kRRGroup, dd = t % kRRGroup;  // tile element of this thread for the load and the store
    const u32 d = grp * kRRGroup + dd;
    const u64 i = ((u64)kb * 64 + r) * nset + d;
    const bool live = d < nset && i < count;
    if (live) {
        const float* q = ndc + i * stride;
        s_q[dd][0][r] = q[0];
        s_q[dd][1][r] = q[1];
        s_q[dd][2][r] = q[2];
    }
    __syncthreads();
    {
        const u32 w = t / 64, lane = t % 64;  // wave w walks DAG grp*kRRGroup + w, lane = row
        const u32 dw = grp * kRRGroup + w;
        const u64 iw = ((u64)kb * 64 + lane) * nset + dw;
        if (dw < nset && iw < count) {
            const float q[3] = {s_q[w][0][lane], s_q[w][1][lane], s_q[w][2][lane]};
            s_v[lane][w] = dagset_lookup(set[dw], q);
        }
    }
    __syncthreads();
    if (live) vis[i] = s_v[r][dd];
}

// One block of 256 queries of one request: the block finds its request by a binary search over
// the (wave-uniform) block starts.
__global__ __launch_bounds__(256) void k_segments_eval(const LookupSeg* __restrict__ segs, u32 nseg) {
    const u32 b = blockIdx.x;
    u32 lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const u32 mid = (lo + hi + 1) >> 1;
        if (segs[mid].block0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const LookupSeg& g = segs[lo];
    const u32 i = (b - g.block0) * 256 + threadIdx.x;
    if (i >= g.count) return;
    g.vis[i] = dagset_lookup(g.d, g.ndc + (size_t)i * g.stride);
}

int launch_segments_eval(const LookupSeg* segs, u32 nseg, u32 blocks, hipStream_t s) {
    if (!nseg || !blocks) return CPVS_OK;
    CPVS_LAUNCH("segments_eval", k_segments_eval, dim3(blocks), dim3(256), 0, s, segs, nseg);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

int launch_dagset_eval(const DagDesc* set, u32 nset, const DagDesc& one, const u32* ids, const float* ndc, u32 stride,
                       u64 count, uint8_t* vis, hipStream_t s) {
    if (count == 0) return CPVS_OK;
    if (count > 0xFFFFFF00ull) {
        set_error("dagset evaluate: %llu queries exceed one launch (2^32 - 256)", (unsigned long long)count);
        return CPVS_EINVAL;
    }
    if (set && !ids) {
        const u64 per = (count + nset - 1) / nset;  // rows of queries (DAG 0 has the most)
        const u64 kblocks = (per + 63) / 64;
        const u64 ngroups = (nset + kRRGroup - 1) / kRRGroup;
        const u64 gper = (ngroups + kXcds - 1) / kXcds;  // groups per XCD
        const u64 blocks = gper * kblocks * kXcds;
        if (blocks > 0x7FFFFFF0ull) {
            set_error("dagset evaluate: %llu queries over %u DAGs exceed one launch", (unsigned long long)count, nset);
            return CPVS_EINVAL;
        }
        CPVS_LAUNCH("dagset_eval_rr", k_dagset_eval_rr, dim3((u32)blocks), dim3(64 * kRRGroup), 0, s, set, nset,
                    (u32)ngroups, (u32)gper, (u32)kblocks, ndc, stride, count, vis);
    } else {
        CPVS_LAUNCH("dagset_eval", k_dagset_eval, dim3((u32)((count + 255) / 256)), dim3(256), 0, s, set, nset, one, ids,
                    ndc, stride, (u32)count, vis);
    }
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

}  // namespace cpvs
