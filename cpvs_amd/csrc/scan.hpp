// scan.hpp -- device-wide scans (u32, or u64 pairs of sums) with fused input/output functors.
//
// Reduce-then-scan in three launches: per-tile reduction, a single-block scan of the tile
// partials, and a per-tile scan that hands (index, exclusive prefix, value) to an output
// functor.  Input is read in striped order (coalesced); every launch is a plain grid with
// no inter-block waiting.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace cpvs {

constexpr int kScanThreads = 256;
constexpr int kScanRounds = 8;
constexpr u32 kScanTile = kScanThreads * kScanRounds;  // 2048 elements per block

// Scan operators; T is the value type (u32, or u64 for a pair of 32-bit sums that cannot carry).
struct OpSum {
    using T = u32;
    static constexpr T identity = 0;
    __device__ T operator()(T a, T b) const { return a + b; }
};
struct OpMax {
    using T = u32;
    static constexpr T identity = 0;
    __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};
struct OpSum64 {
    using T = u64;
    static constexpr T identity = 0;
    __device__ T operator()(T a, T b) const { return a + b; }
};

template <class Op>
__device__ inline typename Op::T wave_inclusive_scan(typename Op::T v, Op op) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const typename Op::T o = __shfl_up(v, d, kWave);
        if (lane >= d) v = op(v, o);
    }
    return v;
}

template <class Op>
__device__ inline typename Op::T wave_reduce(typename Op::T v, Op op) {
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) v = op(v, __shfl_xor(v, d, kWave));
    return v;
}

// Exclusive scan of one value per thread over a 256-thread block; *total gets the block sum.
// s_wave must hold kScanThreads/64 entries of shared memory.
template <class Op>
__device__ inline typename Op::T block_exclusive_scan(typename Op::T v, typename Op::T* s_wave, typename Op::T* total,
                                                      Op op) {
    using T = typename Op::T;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const T inc = wave_inclusive_scan(v, op);
    if (lane == kWave - 1) s_wave[wid] = inc;
    __syncthreads();
    T wpre = Op::identity, tot = Op::identity;
#pragma unroll
    for (int w = 0; w < kScanThreads / kWave; ++w) {
        const T x = s_wave[w];
        if (w < wid) wpre = op(wpre, x);
        tot = op(tot, x);
    }
    __syncthreads();
    T ex = __shfl_up(inc, 1, kWave);
    if (lane == 0) ex = Op::identity;
    *total = tot;
    return op(wpre, ex);
}

template <class Op>
__device__ inline typename Op::T block_reduce(typename Op::T v, typename Op::T* s_wave, Op op) {
    using T = typename Op::T;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    v = wave_reduce(v, op);
    if (lane == 0) s_wave[wid] = v;
    __syncthreads();
    T tot = Op::identity;
#pragma unroll
    for (int w = 0; w < kScanThreads / kWave; ++w) tot = op(tot, s_wave[w]);
    __syncthreads();
    return tot;
}

// Exclusive scans of a tile held striped across a 256-thread block (thread t holds elements
// r*256 + t, r < R), all R rounds at once: R independent wave scans and one LDS exchange, so a
// tile costs two barriers instead of two per round.  ex[r] = the element's prefix within the tile;
// returns the tile total.
template <int R, class Op>
__device__ inline typename Op::T block_exclusive_scan_rounds(const typename Op::T (&v)[R], typename Op::T (&ex)[R], Op op) {
    using T = typename Op::T;
    __shared__ T s_w[R][kScanThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    T inc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) inc[r] = v[r];
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const T o = __shfl_up(inc[r], d, kWave);
            if (lane >= d) inc[r] = op(inc[r], o);
        }
    }
    if (lane == kWave - 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) s_w[r][wid] = inc[r];
    }
    __syncthreads();
    T run = Op::identity;  // total of the earlier rounds
#pragma unroll
    for (int r = 0; r < R; ++r) {
        T wpre = Op::identity, tot = Op::identity;
#pragma unroll
        for (int w = 0; w < kScanThreads / kWave; ++w) {
            const T x = s_w[r][w];
            if (w < wid) wpre = op(wpre, x);
            tot = op(tot, x);
        }
        T e = __shfl_up(inc[r], 1, kWave);
        if (lane == 0) e = Op::identity;
        ex[r] = op(run, op(wpre, e));
        run = op(run, tot);
    }
    __syncthreads();
    return run;
}

// An output functor may declare `Pre` and `Pre load(u32 i) const`: what out(i, ...) reads, loaded
// by the latency path below together with the inputs, and handed over as a fourth argument.
template <class Out, class = void>
struct HasPre : std::false_type {};
template <class Out>
struct HasPre<Out, std::void_t<typename Out::Pre>> : std::true_type {};

// One tile for launches of few blocks, where each block's chain of memory round trips is the cost:
// every input (and prefetch) of the tile's 8 rounds is loaded first, one multi-round block scan
// follows, then the outputs, with no load after the scan.  Returns the tile total.
template <int R, class In, class Out, class Op>
__device__ inline typename Op::T scan_tile_latency(In in, Out out, u32 n, u32 base, typename Op::T carry, Op op) {
    using T = typename Op::T;
    T v[R], ex[R];
    if constexpr (HasPre<Out>::value) {
        typename Out::Pre pre[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const u32 i = base + r * kScanThreads + threadIdx.x;
            v[r] = (i < n) ? in(i) : Op::identity;
            pre[r] = (i < n) ? out.load(i) : typename Out::Pre{};
        }
        const T tot = block_exclusive_scan_rounds<R>(v, ex, op);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const u32 i = base + r * kScanThreads + threadIdx.x;
            if (i < n) out(i, op(carry, ex[r]), v[r], pre[r]);
        }
        return tot;
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const u32 i = base + r * kScanThreads + threadIdx.x;
            v[r] = (i < n) ? in(i) : Op::identity;
        }
        const T tot = block_exclusive_scan_rounds<R>(v, ex, op);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const u32 i = base + r * kScanThreads + threadIdx.x;
            if (i < n) out(i, op(carry, ex[r]), v[r]);
        }
        return tot;
    }
}

// Calls out(i, ex, v), loading the prefetched part itself where the functor has one.
template <class Out, class T>
__device__ inline void call_out(const Out& out, u32 i, T ex, T v) {
    if constexpr (HasPre<Out>::value)
        out(i, ex, v, out.load(i));
    else
        out(i, ex, v);
}

template <class In, class Op, int R = kScanRounds>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(In in, u32 n, typename Op::T* partials, Op op) {
    using T = typename Op::T;
    __shared__ T s_wave[kScanThreads / kWave];
    const u32 base = blockIdx.x * (R * kScanThreads);
    T acc = Op::identity;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const u32 i = base + r * kScanThreads + threadIdx.x;
        if (i < n) acc = op(acc, in(i));
    }
    const T tot = block_reduce(acc, s_wave, op);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Exclusive scan of the tile partials in place (one block); *total = overall reduction.  Each of
// 1024 threads sums a contiguous run of partials, one block scan of those sums follows, and each
// thread writes its run's prefixes: one pass instead of one block scan per 256 partials.
constexpr int kPartialsThreads = 1024;

template <class Op>
__global__ __launch_bounds__(kPartialsThreads) void k_scan_partials(typename Op::T* partials, u32 nparts, typename Op::T* total,
                                                                    Op op) {
    using T = typename Op::T;
    __shared__ T s_w[kPartialsThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const u32 E = ceil_div<u32>(nparts, kPartialsThreads);
    const u32 b = threadIdx.x * E, e = min(b + E, nparts);
    T sum = Op::identity;
#pragma unroll 4
    for (u32 i = b; i < e; ++i) sum = op(sum, partials[i]);
    const T inc = wave_inclusive_scan(sum, op);
    if (lane == kWave - 1) s_w[wid] = inc;
    __syncthreads();
    T wpre = Op::identity, tot = Op::identity;
#pragma unroll
    for (int w = 0; w < kPartialsThreads / kWave; ++w) {
        const T x = s_w[w];
        if (w < wid) wpre = op(wpre, x);
        tot = op(tot, x);
    }
    T ex = __shfl_up(inc, 1, kWave);
    if (lane == 0) ex = Op::identity;
    T run = op(wpre, ex);
    for (u32 i = b; i < e; ++i) {
        const T v = partials[i];
        partials[i] = run;
        run = op(run, v);
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

template <class In, class Out, class Op>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(In in, Out out, u32 n, const typename Op::T* partials, Op op) {
    using T = typename Op::T;
    __shared__ T s_wave[kScanThreads / kWave];
    const u32 base = blockIdx.x * kScanTile;
    T carry = partials[blockIdx.x];
    for (int r = 0; r < kScanRounds; ++r) {
        const u32 r0 = base + r * kScanThreads;
        if (r0 >= n) break;  // uniform across the block
        const u32 i = r0 + threadIdx.x;
        const T v = (i < n) ? in(i) : Op::identity;
        T tot;
        const T ex = block_exclusive_scan(v, s_wave, &tot, op);
        if (i < n) call_out(out, i, op(carry, ex), v);
        carry = op(carry, tot);
    }
}

// Single-block scan for n <= kScanTile: one launch.
template <class In, class Out, class Op>
__global__ __launch_bounds__(kScanThreads) void k_scan_single(In in, Out out, u32 n, typename Op::T* total, Op op) {
    using T = typename Op::T;
    const T all = scan_tile_latency<kScanRounds>(in, out, n, 0u, Op::identity, op);
    if (threadIdx.x == 0 && total) *total = all;
}

// Down-sweep that derives its own carry from the tile partials (nparts <= kScanPartialsInline):
// two launches in total instead of three.
constexpr u32 kScanPartialsInline = 4 * kScanThreads;

// Up to kScanThreads * kScanPartialsInline elements, tiles are one round (256 elements): a level
// of a few thousand to a few hundred thousand nodes then spreads over up to 1024 workgroups rather
// than a handful, and each tile loads its inputs (and the outputs' prefetch) before its scan.
// Up to 2x and 4x that many, tiles of 2 and 4 rounds keep the workgroup count at most 1024.
constexpr u32 kScanSmallTileMax = kScanThreads * kScanPartialsInline;

template <class In, class Out, class Op, int R>
__global__ __launch_bounds__(kScanThreads) void k_scan_down_inline(In in, Out out, u32 n, const typename Op::T* partials,
                                                                   u32 nparts, typename Op::T* total, Op op) {
    using T = typename Op::T;
    __shared__ T s_wave[kScanThreads / kWave];
    T pre = Op::identity, all = Op::identity;
    for (u32 k = threadIdx.x; k < nparts; k += kScanThreads) {
        const T v = partials[k];
        if (k < blockIdx.x) pre = op(pre, v);
        all = op(all, v);
    }
    T carry = block_reduce(pre, s_wave, op);
    const T grand = block_reduce(all, s_wave, op);
    if (blockIdx.x == 0 && threadIdx.x == 0 && total) *total = grand;
    const u32 base = blockIdx.x * (R * kScanThreads);
    if constexpr (R <= 4) {
        scan_tile_latency<R>(in, out, n, base, carry, op);
    } else {
        for (int r = 0; r < R; ++r) {
            const u32 r0 = base + r * kScanThreads;
            if (r0 >= n) break;  // uniform across the block
            const u32 i = r0 + threadIdx.x;
            const T v = (i < n) ? in(i) : Op::identity;
            T tot;
            const T ex = block_exclusive_scan(v, s_wave, &tot, op);
            if (i < n) call_out(out, i, op(carry, ex), v);
            carry = op(carry, tot);
        }
    }
}

// Plain-array functors for device_scan.
struct ArrIn {
    const u32* a;
    __device__ u32 operator()(u32 i) const { return a[i]; }
};
struct ArrOutExclusive {
    u32* a;
    __device__ void operator()(u32 i, u32 ex, u32) const { a[i] = ex; }
};

// Scratch words (u32) needed by device_scan and device_scan_from_partials for n elements: one
// 64-bit partial per one-round tile (the finest tiling any path uses) and one more.  Monotone in
// n, so it covers every scan of at most n elements (callers size scratch for their largest scan).
inline size_t scan_scratch_words(u32 n) { return 2 * ((size_t)ceil_div<u32>(n ? n : 1, kScanThreads) + 1); }

// Reduce and down-sweep with R-round tiles (at most kScanPartialsInline of them).
template <class In, class Out, class Op, int R>
inline void scan_two_launches(In in, Out out, u32 n, Op op, typename Op::T* d_total, typename Op::T* part, hipStream_t s) {
    const u32 nb = ceil_div<u32>(n, R * kScanThreads);
    CPVS_LAUNCH("scan_reduce", (k_scan_reduce<In, Op, R>), dim3(nb), dim3(kScanThreads), 0, s, in, n, part, op);
    CPVS_LAUNCH("scan_down", (k_scan_down_inline<In, Out, Op, R>), dim3(nb), dim3(kScanThreads), 0, s, in, out, n, part, nb,
                d_total, op);
}

// Whether device_scan over n elements loads every input and output prefetch of a tile before its
// block scan (one and two-launch scans up to 4 * kScanSmallTileMax elements).
inline bool scan_prefetches(u32 n) { return n <= 4 * kScanSmallTileMax; }

// out(i, exclusive_prefix_i, value_i) for all i < n; *d_total (device, may be null) gets the
// reduction of all values.  scratch: scan_scratch_words(n) u32 of device memory (8-byte aligned).
// One launch for n <= 2048, two up to 2M elements, three beyond.
template <class In, class Out, class Op>
inline void device_scan(In in, Out out, u32 n, Op op, typename Op::T* d_total, u32* scratch, hipStream_t s) {
    using T = typename Op::T;
    T* part = reinterpret_cast<T*>(scratch);
    if (n <= kScanTile) {
        CPVS_LAUNCH("scan_single", (k_scan_single<In, Out, Op>), dim3(1), dim3(kScanThreads), 0, s, in, out, n, d_total, op);
        return;
    }
    if (n <= kScanSmallTileMax) {
        scan_two_launches<In, Out, Op, 1>(in, out, n, op, d_total, part, s);
        return;
    }
    if (n <= 2 * kScanSmallTileMax) {
        scan_two_launches<In, Out, Op, 2>(in, out, n, op, d_total, part, s);
        return;
    }
    if (n <= 4 * kScanSmallTileMax) {
        scan_two_launches<In, Out, Op, 4>(in, out, n, op, d_total, part, s);
        return;
    }
    const u32 nb = ceil_div<u32>(n, kScanTile);
    CPVS_LAUNCH("scan_reduce", (k_scan_reduce<In, Op>), dim3(nb), dim3(kScanThreads), 0, s, in, n, part, op);
    if (nb <= kScanPartialsInline) {
        CPVS_LAUNCH("scan_down", (k_scan_down_inline<In, Out, Op, kScanRounds>), dim3(nb), dim3(kScanThreads), 0, s, in, out, n,
                    part, nb, d_total, op);
        return;
    }
    CPVS_LAUNCH("scan_partials", k_scan_partials<Op>, dim3(1), dim3(kPartialsThreads), 0, s, part, nb, d_total, op);
    CPVS_LAUNCH("scan_down", (k_scan_down<In, Out, Op>), dim3(nb), dim3(kScanThreads), 0, s, in, out, n, part, op);
}

// Rounds per tile that device_scan uses for n > kScanTile elements (a single-block scan below).
inline u32 scan_rounds(u32 n) {
    return n <= kScanSmallTileMax ? 1u : n <= 2 * kScanSmallTileMax ? 2u : n <= 4 * kScanSmallTileMax ? 4u : (u32)kScanRounds;
}

// device_scan for an input whose producer already stored part[t] = the reduction of in(i) over tile
// t of scan_rounds(n) * kScanThreads elements, t < ceil(n / that): the same down pass as
// device_scan without its reduce pass.  part: scan_scratch_words(n) words.
template <class In, class Out, class Op>
inline void device_scan_from_partials(In in, Out out, u32 n, Op op, typename Op::T* d_total, typename Op::T* part,
                                      hipStream_t s) {
    if (!n) return;
    const u32 R = scan_rounds(n);
    const u32 nt = ceil_div<u32>(n, R * kScanThreads);
    if (nt <= kScanPartialsInline) {
        if (R == 1)
            CPVS_LAUNCH("scan_down", (k_scan_down_inline<In, Out, Op, 1>), dim3(nt), dim3(kScanThreads), 0, s, in, out, n, part, nt,
                        d_total, op);
        else if (R == 2)
            CPVS_LAUNCH("scan_down", (k_scan_down_inline<In, Out, Op, 2>), dim3(nt), dim3(kScanThreads), 0, s, in, out, n, part, nt,
                        d_total, op);
        else if (R == 4)
            CPVS_LAUNCH("scan_down", (k_scan_down_inline<In, Out, Op, 4>), dim3(nt), dim3(kScanThreads), 0, s, in, out, n, part, nt,
                        d_total, op);
        else
            CPVS_LAUNCH("scan_down", (k_scan_down_inline<In, Out, Op, kScanRounds>), dim3(nt), dim3(kScanThreads), 0, s, in, out, n,
                        part, nt, d_total, op);
        return;
    }
    // R == kScanRounds here: more than kScanPartialsInline tiles only above 4 * kScanSmallTileMax.
    CPVS_LAUNCH("scan_partials", k_scan_partials<Op>, dim3(1), dim3(kPartialsThreads), 0, s, part, nt, d_total, op);
    CPVS_LAUNCH("scan_down", (k_scan_down<In, Out, Op>), dim3(nt), dim3(kScanThreads), 0, s, in, out, n, part, op);
}

// ------------------------------------------------------------------------------------------
// LSD radix sort of (u32 key, u32 value) pairs, 8 bits per pass, stable.
// ------------------------------------------------------------------------------------------

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr u32 kSortTile = kSortThreads * kSortRounds;  // 4096 pairs per block

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const u32* __restrict__ keys, u32 n, int shift,
                                                            u32* __restrict__ hist, u32 nblocks);
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const u32* __restrict__ kin, const u32* __restrict__ vin,
                                                               u32* __restrict__ kout, u32* __restrict__ vout, u32 n,
                                                               int shift, const u32* __restrict__ offsets, u32 nblocks);

// Pairs up to this count are sorted by one block in LDS (bitonic network on key<<32 | value).
constexpr u32 kSmallSort = 4096;

// Scratch words for radix_sort_pairs with n pairs.
inline size_t sort_scratch_words(u32 n) {
    const u32 nb = ceil_div<u32>(n ? n : 1, kSortTile);
    return (size_t)256 * nb + scan_scratch_words(256 * nb);
}

// Sorts keys[0..n) with values by the low `bits` bits of the key (a multiple of 8).
// keys/vals hold the result; ktmp/vtmp are ping-pong buffers of n entries.
void radix_sort_pairs(u32* keys, u32* vals, u32* ktmp, u32* vtmp, u32 n, int bits, u32* scratch, hipStream_t s);

}  // namespace cpvs
