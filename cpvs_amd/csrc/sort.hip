// sort.hip -- stable LSD radix sort of (u32 key, u32 value) pairs on gfx950.
//
// Per 8-bit digit: (1) per-tile digit histograms, digit-major so that one exclusive scan
// yields every (digit, tile) output base; (2) scan; (3) per-tile stable scatter.  Stable
// ranking inside a tile uses wave64 ballots to find lanes holding the same digit and a small
// LDS table of per-wave digit counts, processed round by round in input order.
#include "scan.hpp"

namespace cpvs {

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const u32* __restrict__ keys, u32 n, int shift,
                                                            u32* __restrict__ hist, u32 nblocks) {
    __shared__ u32 s_hist[256];
    s_hist[threadIdx.x] = 0;
    __syncthreads();
    const u32 base = blockIdx.x * kSortTile;
#pragma unroll 4
    for (int r = 0; r < kSortRounds; ++r) {
        const u32 i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&s_hist[(keys[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = s_hist[threadIdx.x];
}

// Stable scatter of one tile.  Round by round (256 pairs, input order) every pair gets its rank
// among equal digits: lanes of a wave with the same digit are found with eight __ballot()s, and
// earlier waves of the round add their per-digit counts from LDS.  Pairs are first placed in
// LDS in tile-local digit order, then copied out so that consecutive threads write consecutive
// addresses of each digit's run.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const u32* __restrict__ kin, const u32* __restrict__ vin,
                                                               u32* __restrict__ kout, u32* __restrict__ vout, u32 n,
                                                               int shift, const u32* __restrict__ offsets, u32 nblocks) {
    constexpr int kWaves = kSortThreads / kWave;
    __shared__ u32 s_key[kSortTile];
    __shared__ u32 s_val[kSortTile];
    __shared__ u32 s_global[256];  // global output position of the digit's first pair of this tile
    __shared__ u32 s_local[256];   // tile-local start of the digit
    __shared__ u32 s_run[256];     // pairs of the digit placed so far
    __shared__ u32 s_cnt[kWaves][256];
    __shared__ u32 s_wave[kWaves];
    const u32 tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    const u32 b = blockIdx.x;
    const u32 base = b * kSortTile;
    const u32 tile_n = min(kSortTile, n - base);

    // Tile digit counts from the scanned digit-major offsets: count[d][b] = off[d][b+1] - off[d][b].
    const size_t idx = (size_t)tid * nblocks + b;
    const u32 g = offsets[idx];
    const u32 next = (tid == 255 && b == nblocks - 1) ? n : offsets[idx + 1];
    const u32 cnt = next - g;
    u32 tot;
    const u32 local = block_exclusive_scan(cnt, s_wave, &tot, OpSum{});
    s_global[tid] = g;
    s_local[tid] = local;
    s_run[tid] = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_cnt[w][tid] = 0;
    __syncthreads();

    const u64 lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < kSortRounds; ++r) {
        const u32 r0 = r * kSortThreads;
        if (r0 >= tile_n) break;  // uniform across the block
        const u32 i = r0 + tid;
        const bool valid = i < tile_n;
        const u32 key = valid ? kin[base + i] : 0u;
        const u32 val = valid ? vin[base + i] : 0u;
        const u32 d = (key >> shift) & 0xFF;
        u64 peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1;
            const u64 m = __ballot(on);
            peers &= on ? m : ~m;
        }
        const u32 rank = (u32)__popcll(peers & lt);
        if (valid && rank == 0) s_cnt[wid][d] = (u32)__popcll(peers);
        __syncthreads();
        if (valid) {
            u32 pos = s_local[d] + s_run[d] + rank;
            for (int w = 0; w < wid; ++w) pos += s_cnt[w][d];
            s_key[pos] = key;
            s_val[pos] = val;
        }
        __syncthreads();
        u32 add = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            add += s_cnt[w][tid];
            s_cnt[w][tid] = 0;
        }
        s_run[tid] += add;
        __syncthreads();
    }
    // Copy out in tile-local digit order: runs of equal digits land on consecutive addresses.
    for (u32 i = tid; i < tile_n; i += kSortThreads) {
        const u32 key = s_key[i];
        const u32 d = (key >> shift) & 0xFF;
        const u32 pos = s_global[d] + (i - s_local[d]);
        kout[pos] = key;
        vout[pos] = s_val[i];
    }
}

// One block sorts n <= kSmallSort pairs: bitonic network over key<<32 | value padded to a power
// of two with all-ones (values are distinct node indices, so the order equals the stable one).
__global__ __launch_bounds__(1024) void k_sort_small(u32* __restrict__ keys, u32* __restrict__ vals, u32 n, u32 P) {
    __shared__ u64 s_v[kSmallSort];
    for (u32 i = threadIdx.x; i < P; i += blockDim.x) s_v[i] = (i < n) ? (((u64)keys[i] << 32) | vals[i]) : ~0ull;
    __syncthreads();
    for (u32 k = 2; k <= P; k <<= 1) {
        for (u32 j = k >> 1; j > 0; j >>= 1) {
            for (u32 t = threadIdx.x; t < P / 2; t += blockDim.x) {
                const u32 i = 2 * t - (t & (j - 1));  // lower index of the pair, bit j clear
                const u32 l = i + j;
                const bool up = (i & k) == 0;
                const u64 a = s_v[i], b = s_v[l];
                if ((a > b) == up) {
                    s_v[i] = b;
                    s_v[l] = a;
                }
            }
            __syncthreads();
        }
    }
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
        keys[i] = (u32)(s_v[i] >> 32);
        vals[i] = (u32)s_v[i];
    }
}

void radix_sort_pairs(u32* keys, u32* vals, u32* ktmp, u32* vtmp, u32 n, int bits, u32* scratch, hipStream_t s) {
    if (n <= 1) return;
    if (n <= kSmallSort && bits == 32) {
        u32 P = 2;
        while (P < n) P <<= 1;
        CPVS_LAUNCH("sort_small", k_sort_small, dim3(1), dim3(1024), 0, s, keys, vals, n, P);
        return;
    }
    const u32 nb = ceil_div<u32>(n, kSortTile);
    u32* hist = scratch;
    u32* scan_tmp = scratch + (size_t)256 * nb;
    u32 *ki = keys, *vi = vals, *ko = ktmp, *vo = vtmp;
    for (int shift = 0; shift < bits; shift += 8) {
        CPVS_LAUNCH("radix_hist", k_radix_hist, dim3(nb), dim3(kSortThreads), 0, s, ki, n, shift, hist, nb);
        device_scan(ArrIn{hist}, ArrOutExclusive{hist}, 256 * nb, OpSum{}, nullptr, scan_tmp, s);
        CPVS_LAUNCH("radix_scatter", k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, s, ki, vi, ko, vo, n, shift, hist, nb);
        u32* t;
        t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    if (ki != keys) {
        hipMemcpyAsync(keys, ki, sizeof(u32) * n, hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(vals, vi, sizeof(u32) * n, hipMemcpyDeviceToDevice, s);
    }
}

}  // namespace cpvs
