// common.hpp -- shared device/host helpers for the cpvs_amd HIP kernels (gfx950, wave64).
//
// Bit layout of a node mask follows src/CompressedShadowUtil.h:80-132 of the reference:
// two bits per child c = x | y<<1 | z<<2, 01 = VISIBLE, 00 = SHADOW, 10 = PARTIAL.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prof.hpp"

namespace cpvs {

using u32 = uint32_t;
using u64 = uint64_t;

constexpr int kWave = 64;

// Node coordinates (x, y, z) at one octree level, packed 21 bits each.
__host__ __device__ inline u64 pack_xyz(u32 x, u32 y, u32 z) {
    return (u64)x | ((u64)y << 21) | ((u64)z << 42);
}
__host__ __device__ inline u32 unpack_x(u64 c) { return (u32)(c & 0x1FFFFF); }
__host__ __device__ inline u32 unpack_y(u64 c) { return (u32)((c >> 21) & 0x1FFFFF); }
__host__ __device__ inline u32 unpack_z(u64 c) { return (u32)((c >> 42) & 0x1FFFFF); }

// getNumChildren (CompressedShadowUtil.h:80-85): number of partial children.
__host__ __device__ inline u32 num_children(u32 mask) { return (u32)__builtin_popcount(mask & 0xAAAAu); }

// std::min<float>/std::max<float> exactly as the reference's find() applies them
// (MinMaxHierarchy.cpp:29-33): min(a,b) = b<a ? b : a, max(a,b) = a<b ? b : a.
__host__ __device__ inline float smin(float a, float b) { return (b < a) ? b : a; }
__host__ __device__ inline float smax(float a, float b) { return (a < b) ? b : a; }

// visible() (CompressedShadowUtil.h:35-43) on a cell [z, z+1) against scaled min/max.
__device__ inline u32 classify_cell(u32 z, float mn_h, float mx_h) {
    if ((float)(z + 1) <= mn_h) return 1u;
    if ((float)z >= mx_h) return 0u;
    return 2u;
}

// absoluteVisible() (CompressedShadowUtil.h:51-57): midpoint compare.
__device__ inline u32 classify_mid(u32 z, float d_h) {
    const float mid = ((float)z + (float)(z + 1)) * 0.5f;
    return (mid <= d_h) ? 1u : 0u;
}

// Partial cells of one level-L texel (mn, mx) inside a z-tile's slab [zlo, zhi) of cells, h =
// (N>>L)*T: cell z is partial iff !(z + 1 <= mn*h) && !(z >= mx*h) (visible(),
// src/CompressedShadowUtil.h:35-43), i.e. z in [floor(mn*h), ceil(mx*h)).  The bounds are clipped
// to the slab in float before the conversion: a NaN bound fails its compare and leaves that side
// open (as it does in the classification), infinities and huge values clip.  *first = the first
// partial cell (the slab's start when none).
__device__ inline u32 clip_cells(float lo, float hi, int zlo, int zhi, int* first) {
    const float fz0 = (float)zlo, fz1 = (float)zhi;
    const float slo = (lo > fz0) ? ((lo < fz1) ? lo : fz1) : fz0;
    const float shi = (hi < fz1) ? ((hi > fz0) ? hi : fz0) : fz1;
    const int s = (int)slo, e = (int)shi;
    *first = s;
    return (e > s) ? (u32)(e - s) : 0u;
}
__device__ inline u32 partial_cells(float2 v, float h, int zlo, int zhi, int* first) {
    return clip_cells(floorf(v.x * h), ceilf(v.y * h), zlo, zhi, first);
}

// Interval bounds where NaN means open (no bound): the tighter of a and b.
__device__ inline float open_max(float a, float b) { return (b > a || a != a) ? b : a; }
__device__ inline float open_min(float a, float b) { return (b < a || a != a) ? b : a; }

// The cells of texel (x, y) of level L that are nodes of the octree: partial, and under a partial
// cell at every level above, up to the root's children at level R (lvl[M] = level M, M in
// [L, R]).  A cell at level L+d lies over the cells [2^d z, 2^d (z+1)) of level L, so the
// reachable cells stay one interval: the texel's own partial interval cut by each ancestor's,
// scaled.  Without NaN this equals partial_cells (a parent's minimum is at most, and its maximum
// at least, its children's); a NaN, which the pyramid keeps only from a block's first texel
// (std::min/max order, src/MinMaxHierarchy.cpp:29-49), makes a cell partial at every z while its
// parent's range may be narrower.
__device__ inline u32 reachable_cells(const float2* const* lvl, u32 x, u32 y, int L, int R, u32 N, u32 T, int zlo,
                                      int zhi, int* first) {
    const float2 v = lvl[L][(size_t)y * (N >> L) + x];
    const float h = (float)((N >> L) * T);
    float lo = floorf(v.x * h), hi = ceilf(v.y * h);
    for (int M = L + 1; M <= R; ++M) {
        const int d = M - L;
        const float2 a = lvl[M][(size_t)(y >> d) * (N >> M) + (x >> d)];
        const float hm = (float)((N >> M) * T), sc = (float)(1u << d);
        lo = open_max(lo, floorf(a.x * hm) * sc);
        hi = open_min(hi, ceilf(a.y * hm) * sc);
    }
    return clip_cells(lo, hi, zlo, zhi, first);
}

// 64-bit mixers for subtree keys (Stafford variant 13 finaliser of splitmix64).
__host__ __device__ inline u64 mix64(u64 z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline u64 hash_step(u64 h, u64 v) { return mix64(h ^ (v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2))); }

template <class T>
__host__ __device__ inline T ceil_div(T a, T b) { return (a + b - 1) / b; }

}  // namespace cpvs

#define CPVS_HIP_CHECK(expr)                                   \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return ::cpvs::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

namespace cpvs {
// Records a HIP failure in the thread's last-error slot; returns CPVS_EHIP.
int hip_fail(hipError_t e, const char* what, const char* file, int line);
}  // namespace cpvs
