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

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const u32* __restrict__ kin, const u32* __restrict__ vin,
                                                               u32* __restrict__ kout, u32* __restrict__ vout, u32 n,
                                                               int shift, const u32* __restrict__ offsets, u32 nblocks) {
    constexpr int kWaves = kSortThreads / kWave;
    __shared__ u32 s_run[256];
    __shared__ u32 s_cnt[kWaves][256];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    s_run[threadIdx.x] = offsets[(size_t)threadIdx.x * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_cnt[w][threadIdx.x] = 0;
    __syncthreads();
    const u32 base = blockIdx.x * kSortTile;
    const u64 lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < kSortRounds; ++r) {
        const u32 r0 = base + r * kSortThreads;
        if (r0 >= n) break;  // uniform across the block
        const u32 i = r0 + threadIdx.x;
        const bool valid = i < n;
        const u32 key = valid ? kin[i] : 0u;
        const u32 val = valid ? vin[i] : 0u;
        const u32 d = (key >> shift) & 0xFF;
        u64 peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const u64 m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const u32 rank = (u32)__popcll(peers & lt);
        const bool leader = valid && rank == 0;
        if (leader) s_cnt[wid][d] = (u32)__popcll(peers);
        __syncthreads();
        if (valid) {
            u32 pos = s_run[d] + rank;
            for (int w = 0; w < wid; ++w) pos += s_cnt[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        u32 add = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            add += s_cnt[w][threadIdx.x];
            s_cnt[w][threadIdx.x] = 0;
        }
        s_run[threadIdx.x] += add;
        __syncthreads();
    }
}

void radix_sort_pairs(u32* keys, u32* vals, u32* ktmp, u32* vtmp, u32 n, int bits, u32* scratch, hipStream_t s) {
    if (n <= 1) return;
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
