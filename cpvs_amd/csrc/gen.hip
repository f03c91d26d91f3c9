// gen.hip -- synthetic depth maps and query sets on the device.
//
// The reference renders a scene into a float depth texture (src/ShadowMap.cpp:23-30); the
// workloads here are synthetic light-space depth maps instead:
//   D = 0.6 + 0.1*u  (a tilted plane, u = (x + 0.5)/N)
//     + 0.05*(noise - 0.5) with 4-octave value noise (lattices of 8, 16, 32, 64 cells per side,
//       smoothstep-weighted bilinear interpolation, octave weights 1, 1/2, 1/4, 1/8 over 1.875),
//   clamped to [0, 1], then D = min(D, dz) under each of nboxes axis-aligned squares of side
//   16..128 and depth dz in [0.1, 0.5).
// Box parameters and noise lattices come from splitmix64(seed) on the host (all boxes first,
// then the lattices), so any tile of a large map can be produced independently.
#include <vector>

#include "internal.hpp"

namespace cpvs {

namespace {

// splitmix64 (Steele, Lea & Flood 2014).
inline u64 sm64_next(u64& s) {
    u64 z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline u64 sm64_at(u64 seed, u64 k) {
    u64 z = seed + k * 0x9E3779B97F4A7C15ull + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline float u01(u64 r) { return (float)(r >> 40) * (1.0f / 16777216.0f); }

struct Box {
    int s, x0, y0;
    float dz;
};

constexpr int kOctaves = 4;
constexpr int kLatticeFloats = 9 * 9 + 17 * 17 + 33 * 33 + 65 * 65;

}  // namespace

__global__ __launch_bounds__(256) void k_gen_base(float* __restrict__ out, u32 fullN, u32 x0, u32 y0, u32 tileN,
                                                  const float* __restrict__ lat, int noise) {
    const u32 tx = blockIdx.x * 64 + (threadIdx.x & 63);
    const u32 ty = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (tx >= tileN || ty >= tileN) return;
    const u32 gx = x0 + tx, gy = y0 + ty;
    const float fN = (float)fullN;
    float d = 0.6f + 0.1f * (((float)gx + 0.5f) / fN);
    if (noise) {
        float acc = 0.0f, w = 1.0f;
        int latoff = 0;
        for (int o = 0; o < kOctaves; ++o) {
            const int cells = 8 << o;
            const float sc = (float)cells / fN;
            const float fx = ((float)gx + 0.5f) * sc;
            const float fy = ((float)gy + 0.5f) * sc;
            int ix = (int)fx, iy = (int)fy;
            if (ix > cells - 1) ix = cells - 1;
            if (iy > cells - 1) iy = cells - 1;
            const float rx = fx - (float)ix, ry = fy - (float)iy;
            const float sx = rx * rx * (3.0f - 2.0f * rx);
            const float sy = ry * ry * (3.0f - 2.0f * ry);
            const int L = cells + 1;
            const float* l = lat + latoff;
            const float a = l[iy * L + ix], b = l[iy * L + ix + 1];
            const float c = l[(iy + 1) * L + ix], e = l[(iy + 1) * L + ix + 1];
            const float top = a + (b - a) * sx;
            const float bot = c + (e - c) * sx;
            const float v = top + (bot - top) * sy;
            acc = acc + w * v;
            w = w * 0.5f;
            latoff += L * L;
        }
        d = d + 0.05f * (acc / 1.875f - 0.5f);
    }
    if (d < 0.0f) d = 0.0f;
    if (d > 1.0f) d = 1.0f;
    out[(size_t)ty * tileN + tx] = d;
}

// One block per box; min() on non-negative floats equals min() on their bit patterns, so the
// result does not depend on the order overlapping boxes land in.
__global__ __launch_bounds__(256) void k_gen_boxes(float* __restrict__ out, u32 x0, u32 y0, u32 tileN,
                                                   const Box* __restrict__ boxes) {
    const Box b = boxes[blockIdx.x];
    const int bx0 = max(b.x0, (int)x0), by0 = max(b.y0, (int)y0);
    const int bx1 = min(b.x0 + b.s, (int)(x0 + tileN)), by1 = min(b.y0 + b.s, (int)(y0 + tileN));
    if (bx1 <= bx0 || by1 <= by0) return;
    const int w = bx1 - bx0;
    const int area = w * (by1 - by0);
    const u32 dz = __float_as_uint(b.dz);
    for (int i = threadIdx.x; i < area; i += blockDim.x) {
        const int gx = bx0 + i % w, gy = by0 + i / w;
        u32* p = reinterpret_cast<u32*>(out + (size_t)(gy - (int)y0) * tileN + (gx - (int)x0));
        atomicMin(p, dz);
    }
}

int launch_gen_depth(float* out, u32 fullN, u64 seed, u32 nboxes, int noise, u32 x0, u32 y0, u32 tileN, hipStream_t s) {
    u64 st = seed;
    std::vector<Box> boxes(nboxes ? nboxes : 1);
    for (u32 i = 0; i < nboxes; ++i) {
        int sz = 16 + (int)(sm64_next(st) % 113u);
        if (sz > (int)fullN) sz = (int)fullN;
        boxes[i].s = sz;
        boxes[i].x0 = (int)(sm64_next(st) % (u64)(fullN - sz + 1));
        boxes[i].y0 = (int)(sm64_next(st) % (u64)(fullN - sz + 1));
        boxes[i].dz = 0.1f + 0.4f * u01(sm64_next(st));
    }
    std::vector<float> lat(kLatticeFloats);
    if (noise)
        for (int i = 0; i < kLatticeFloats; ++i) lat[i] = u01(sm64_next(st));
    Box* d_boxes = nullptr;
    float* d_lat = nullptr;
    CPVS_HIP_CHECK(hipMalloc(&d_boxes, sizeof(Box) * boxes.size()));
    CPVS_HIP_CHECK(hipMalloc(&d_lat, sizeof(float) * lat.size()));
    CPVS_HIP_CHECK(hipMemcpyAsync(d_boxes, boxes.data(), sizeof(Box) * boxes.size(), hipMemcpyHostToDevice, s));
    CPVS_HIP_CHECK(hipMemcpyAsync(d_lat, lat.data(), sizeof(float) * lat.size(), hipMemcpyHostToDevice, s));
    CPVS_LAUNCH("gen_base", k_gen_base, dim3(ceil_div<u32>(tileN, 64), ceil_div<u32>(tileN, 4)), dim3(256), 0, s, out, fullN, x0,
                       y0, tileN, d_lat, noise);
    if (nboxes) CPVS_LAUNCH("gen_boxes", k_gen_boxes, dim3(nboxes), dim3(256), 0, s, out, x0, y0, tileN, d_boxes);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d_boxes);
    hipFree(d_lat);
    if (e != hipSuccess) return hip_fail(e, "gen depth", __FILE__, __LINE__);
    return CPVS_OK;
}

// Output point i is point first + i*step of the counter-based stream.
__global__ __launch_bounds__(256) void k_gen_uniform(float* __restrict__ out, u64 seed, u64 first, u64 step, u64 count,
                                                     u32 stride) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const u64 k = first + i * step;
    for (u32 c = 0; c < 3; ++c) {
        const float p = u01(sm64_at(seed, 3 * k + c));
        out[i * stride + c] = 2.0f * p - 1.0f;
    }
    if (stride == 4) out[i * stride + 3] = 1.0f;
}

int launch_gen_queries_uniform(float* out, u64 seed, u64 first, u64 step, u64 count, u32 stride, hipStream_t s) {
    if (!count) return CPVS_OK;
    CPVS_LAUNCH("gen_uniform", k_gen_uniform, dim3((u32)((count + 255) / 256)), dim3(256), 0, s, out, seed, first, step, count,
                stride);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

__global__ __launch_bounds__(256) void k_gen_gbuffer(float4* __restrict__ out, const float* __restrict__ D, u32 N, u64 seed,
                                                     u32 W, u32 H) {
    const u64 pix = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= (u64)W * H) return;
    const u32 i = (u32)(pix % W), j = (u32)(pix / W);
    const float fN = (float)N;
    const float u = ((float)i + 0.5f) / (float)W;
    const float v = ((float)j + 0.5f) / (float)H;
    int tx = (int)(u * fN), ty = (int)(v * fN);
    if (tx > (int)N - 1) tx = (int)N - 1;
    if (ty > (int)N - 1) ty = (int)N - 1;
    const float z0 = D[(size_t)ty * N + tx];
    float z1 = z0 + (2.0f * u01(sm64_at(seed, pix)) - 1.0f) * (2.0f / fN);
    if (z1 < 0.0f) z1 = 0.0f;
    if (z1 > 1.0f) z1 = 1.0f;
    const float nx = 2.0f * u - 1.0f, ny = 2.0f * v - 1.0f;
    out[2 * pix] = make_float4(nx, ny, 2.0f * z0 - 1.0f, 1.0f);
    out[2 * pix + 1] = make_float4(nx, ny, 2.0f * z1 - 1.0f, 1.0f);
}

// Plain device copy, one 16-byte element per thread: the reference kernel for bandwidth and for
// scaling HBM traffic counters (its grid spreads over all XCDs as the path's kernels do).
__global__ __launch_bounds__(256) void k_device_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, u64 n) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

int launch_device_copy(void* dst, const void* src, u64 bytes, hipStream_t s) {
    const u64 n = bytes / 16;
    if (!n) return CPVS_OK;
    const u64 blocks = (n + 255) / 256;
    if (blocks > 0x7FFFFFFFull) {
        set_error("device_copy: %llu bytes exceed one launch", (unsigned long long)bytes);
        return CPVS_EINVAL;
    }
    CPVS_LAUNCH("device_copy", k_device_copy, dim3((u32)blocks), dim3(256), 0, s, static_cast<const uint4*>(src),
                static_cast<uint4*>(dst), n);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

int launch_gen_queries_gbuffer(float* out, const float* depth, u32 N, u64 seed, u32 W, u32 H, hipStream_t s) {
    const u64 n = (u64)W * H;
    if (!n) return CPVS_OK;
    CPVS_LAUNCH("gen_gbuffer", k_gen_gbuffer, dim3((u32)((n + 255) / 256)), dim3(256), 0, s, reinterpret_cast<float4*>(out), depth, N,
                       seed, W, H);
    CPVS_HIP_CHECK(hipGetLastError());
    return CPVS_OK;
}

}  // namespace cpvs
