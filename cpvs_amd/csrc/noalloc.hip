// noalloc.hip -- CompressedShadow::create without any device allocation, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// cpvs_build takes its scratch from the context's arena (grown with hipMalloc when a map needs
// more) and its DAG words from the context's buffer cache (hipMalloc on a miss).  Here the scratch
// is one file-scope device array of fixed size, and the words go to a buffer the caller passes.
// The array's name is a host-side stand-in in host code: its device address is looked up once
// per device (hipGetSymbolAddress, at context creation, with that device current) and only that
// address is handed to kernels.  Each device holds its own copy of the array until the process
// ends.  Every call on a device shares that device's array and leaves its contents behind; the
// build clears what it needs cleared (its metadata words and merge tables) in its own kernels.
// Calls that overlap on one device race on the array: calls through one context are serialised
// by the context's lock; callers with several contexts on one device must serialise themselves.
#include <mutex>

#include "abi.hpp"

namespace cpvs {

// 64 MiB: the scratch of maps up to 1024^2 (about 10 MB for the 1024^2 plane + 32 boxes map) and of
// most 2048^2 maps.  A map that needs more is rejected with CPVS_EINVAL before any kernel of the
// build proper runs (its need is known from the exact node counts).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_mu;
void* g_ptr[kNoAllocMaxDevices] = {};  // device address of `scratch` on each device, once found
}  // namespace

int noalloc_setup(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) {
        set_error("noalloc: device %d beyond the %d supported", device, kNoAllocMaxDevices);
        return CPVS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return CPVS_OK;
    int cur = -1;
    CPVS_HIP_CHECK(hipGetDevice(&cur));
    if (cur != device) CPVS_HIP_CHECK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    if (!p) {
        set_error("noalloc: hipGetSymbolAddress gave no address for the scratch array");
        return CPVS_EHIP;
    }
    g_ptr[device] = p;
    return CPVS_OK;
}

}  // namespace cpvs

using namespace cpvs;

extern "C" {

size_t cpvs_build_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

int cpvs_build_noalloc(cpvs_ctx* ctx, const cpvs_pyramid* p, uint32_t t, uint32_t T, void* stream, uint32_t* d_words,
                       uint64_t cap_words, uint64_t* num_words, uint32_t* root_mask) {
    CHECK_ARG(ctx && p && d_words && num_words, "build_noalloc: NULL argument");
    CHECK_ARG(p->ctx == ctx, "build_noalloc: pyramid belongs to another context");
    CHECK_ARG(p->n >= 4, "build: a %ux%u map has %d levels; the octree needs N >= 8", p->N, p->N, p->n);
    CHECK_ARG(T >= 1 && t < T, "zTileIndex %u must be < zTileNum %u", t, T);
    CHECK_ARG((uint64_t)p->N * T <= (1ull << 21), "N*zTileNum = %llu exceeds 2^21", (unsigned long long)p->N * T);
    GUARD_BEGIN
    std::lock_guard<std::mutex> lk(ctx->mu);
    ProfInstall prof(ctx->prof_on ? &ctx->prof : nullptr);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        base = (ctx->device >= 0 && ctx->device < kNoAllocMaxDevices) ? g_ptr[ctx->device] : nullptr;
    }
    if (!base) {
        set_error("build_noalloc: the fixed scratch of device %d was not set up", ctx->device);
        return CPVS_ESTATE;
    }
    hipStream_t s = pick_stream(ctx, stream);
    CPVS_HIP_CHECK(pyramid_wait(p, s));
    Arena fixed;
    fixed.base = static_cast<char*>(base);
    fixed.cap = kNoAllocScratchBytes;
    fixed.fixed = true;
    DagOut into;
    into.words = d_words;
    into.cap = cap_words;
    const int rc = build_dag(ctx, fixed, ctx->pinned, ctx->merge_mode, p, t, T, s, nullptr, nullptr, &into);
    if (rc != CPVS_OK) return rc;
    *num_words = into.nwords;
    if (root_mask) *root_mask = into.root_mask;
    return CPVS_OK;
    GUARD_END
}

}  // extern "C"
