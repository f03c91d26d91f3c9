// format.hip -- a versioned byte format for DAGs (host memory or files), with a structural check
// on load.
//
// The reference keeps DAGs in memory only (CompressedShadow::getDAG, src/CompressedShadow.h:75-77,
// a vector<uint>); this format is that word array behind a 64-byte header, so a DAG built once can
// be stored and loaded into a server's resident store or a container later.
//
//   offset  size  field
//        0     8  magic "CPVSDAG\0"
//        8     4  version (1)
//       12     4  header bytes (64)
//       16     4  numLevels
//       20     4  N (depth-map side, 2^(numLevels-1))
//       24     4  zTileIndex
//       28     4  zTileNum
//       32     4  root mask (= word 0)
//       36     4  flags: bit 0 = no leafmasks (a DAG of numLevels >= 5 built down to level 0,
//                  the reference without LEAFMASKS, src/CompressedShadow.cpp:17); other bits 0
//       40     8  word count W
//       48     8  checksum of the words (below)
//       56     8  reserved (0)
//       64   4*W  words, little-endian
//
// Checksum: h = 0x243F6A8885A308D3; for each 64-bit little-endian pair of words (the last word
// alone zero-extended when W is odd): h = (h ^ v) * 0x9E3779B97F4A7C15; h ^= h >> 29;
// finally h ^= W.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "abi.hpp"

namespace cpvs {
namespace {

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    uint32_t num_levels;
    uint32_t N;
    uint32_t z_tile_index;
    uint32_t z_tile_num;
    uint32_t root_mask;
    uint32_t flags;
    uint64_t words;
    uint64_t checksum;
    uint64_t reserved;
};
static_assert(sizeof(FileHeader) == 64, "DAG file header is 64 bytes");

constexpr char kMagic[8] = {'C', 'P', 'V', 'S', 'D', 'A', 'G', '\0'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kFlagNoLeafmasks = 1u;

// useLeafmasks (src/CompressedShadow.cpp:20-27) unless the DAG was built without them.
bool leaf_dag(const FileHeader& h) { return h.num_levels >= 5 && !(h.flags & kFlagNoLeafmasks); }

uint64_t words_checksum(const uint32_t* w, uint64_t n) {
    uint64_t h = 0x243F6A8885A308D3ull;
    uint64_t i = 0;
    for (; i + 1 < n; i += 2) {
        const uint64_t v = (uint64_t)w[i] | ((uint64_t)w[i + 1] << 32);
        h = (h ^ v) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 29;
    }
    if (i < n) {
        h = (h ^ (uint64_t)w[i]) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 29;
    }
    return h ^ n;
}

// Decodes the DAG level by level (root at word 0, levels contiguous, every child pointer the
// start of a node of the next level, every node below the root referenced, nothing after the
// last level).  nodes[L] receives the node count of level L.  leaf: level 2 holds leafmask nodes.
int check_structure(const uint32_t* w, uint64_t W, int n, bool leaf, uint64_t* nodes) {
    const int R = n - 2, bottom = leaf ? 2 : 0;
    for (int L = 0; L < kMaxLevels; ++L) nodes[L] = 0;
    if (W == 0) {
        set_error("dag format: no words");
        return CPVS_EINVAL;
    }
    std::vector<uint64_t> targets{0};  // node starts the current level must hold (sorted, unique)
    uint64_t pos = 0;
    for (int L = R; L >= bottom && !targets.empty(); --L) {
        const bool leaf_level = leaf && L == 2;
        std::vector<uint64_t> next;
        size_t t = 0;
        uint64_t count = 0;
        while (t < targets.size()) {
            if (pos >= W) {
                set_error("dag format: level %d runs past the %llu words", L, (unsigned long long)W);
                return CPVS_EINVAL;
            }
            if (pos != targets[t]) {
                set_error("dag format: word %llu of level %d is not referenced (next reference %llu)",
                          (unsigned long long)pos, L, (unsigned long long)targets[t]);
                return CPVS_EINVAL;
            }
            ++t;
            ++count;
            const uint32_t m = w[pos];
            const uint32_t k = (uint32_t)__builtin_popcount(m & 0xAAAAu);
            if (m & ~0xFFFFu) {
                set_error("dag format: mask 0x%x at word %llu has bits above 16", m, (unsigned long long)pos);
                return CPVS_EINVAL;
            }
            for (uint32_t c = 0; c < 8; ++c)
                if (((m >> (2 * c)) & 3u) == 3u) {
                    set_error("dag format: mask 0x%x at word %llu has an invalid child state", m, (unsigned long long)pos);
                    return CPVS_EINVAL;
                }
            const uint64_t size = 1 + (leaf_level ? 2ull * k : (uint64_t)k);
            if (pos + size > W) {
                set_error("dag format: node at word %llu runs past the end", (unsigned long long)pos);
                return CPVS_EINVAL;
            }
            if (!leaf_level && k) {
                if (L == bottom) {
                    set_error("dag format: node at word %llu of the last level %d has children", (unsigned long long)pos,
                              L);
                    return CPVS_EINVAL;
                }
                for (uint32_t j = 0; j < k; ++j) next.push_back(w[pos + 1 + j]);
            }
            pos += size;
        }
        nodes[L] = count;
        std::sort(next.begin(), next.end());
        next.erase(std::unique(next.begin(), next.end()), next.end());
        if (!next.empty() && next.front() < pos) {
            set_error("dag format: a pointer of level %d points backwards (%llu)", L, (unsigned long long)next.front());
            return CPVS_EINVAL;
        }
        targets.swap(next);
    }
    if (!targets.empty()) {
        set_error("dag format: pointers below the last level");
        return CPVS_EINVAL;
    }
    if (pos != W) {
        set_error("dag format: %llu words after the last level", (unsigned long long)(W - pos));
        return CPVS_EINVAL;
    }
    return CPVS_OK;
}

}  // namespace
}  // namespace cpvs

using namespace cpvs;

extern "C" {

int cpvs_dag_serialize(const cpvs_dag* d, void* buf, uint64_t cap, uint64_t* size) {
    CHECK_ARG(d && size, "dag_serialize: NULL argument");
    const uint64_t need = sizeof(FileHeader) + 4 * d->words;
    *size = need;
    if (!buf) return CPVS_OK;
    CHECK_ARG(cap >= need, "dag_serialize: capacity %llu < %llu bytes", (unsigned long long)cap, (unsigned long long)need);
    GUARD_BEGIN
    uint32_t* w = reinterpret_cast<uint32_t*>(static_cast<char*>(buf) + sizeof(FileHeader));
    CPVS_HIP_CHECK(hipSetDevice(d->ctx->device));
    if (d->words) CPVS_HIP_CHECK(hipMemcpy(w, d->d_words, 4 * d->words, hipMemcpyDeviceToHost));
    FileHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.header_bytes = sizeof(FileHeader);
    h.num_levels = (uint32_t)d->n;
    h.N = d->N;
    h.z_tile_index = d->zTileIndex;
    h.z_tile_num = d->zTileNum;
    h.root_mask = d->root_mask;
    h.flags = (d->n >= 5 && !d->leafmasks) ? kFlagNoLeafmasks : 0u;
    h.words = d->words;
    h.checksum = words_checksum(w, d->words);
    std::memcpy(buf, &h, sizeof(h));
    return CPVS_OK;
    GUARD_END
}

int cpvs_dag_check(const void* buf, uint64_t size, uint32_t* num_levels, uint64_t* words, uint64_t* level_nodes,
                   int cap) {
    CHECK_ARG(buf, "dag_check: NULL buffer");
    CHECK_ARG(size >= sizeof(FileHeader), "dag_check: %llu bytes is shorter than the header", (unsigned long long)size);
    FileHeader h;
    std::memcpy(&h, buf, sizeof(h));
    CHECK_ARG(std::memcmp(h.magic, kMagic, 8) == 0, "dag_check: not a DAG (bad magic)");
    CHECK_ARG(h.version == kVersion, "dag_check: version %u (this library reads %u)", h.version, kVersion);
    CHECK_ARG(h.header_bytes == sizeof(FileHeader), "dag_check: header of %u bytes", h.header_bytes);
    CHECK_ARG(h.num_levels >= 4 && h.num_levels <= 22, "dag_check: numLevels %u", h.num_levels);
    CHECK_ARG(h.N == (1u << (h.num_levels - 1)), "dag_check: N %u does not match numLevels %u", h.N, h.num_levels);
    CHECK_ARG(h.z_tile_num >= 1 && h.z_tile_index < h.z_tile_num, "dag_check: z tile %u of %u", h.z_tile_index,
              h.z_tile_num);
    CHECK_ARG(h.words >= 1 && h.words < 0xFFFFFFFEull, "dag_check: %llu words", (unsigned long long)h.words);
    CHECK_ARG((h.flags & ~kFlagNoLeafmasks) == 0, "dag_check: unknown flags 0x%x", h.flags);
    CHECK_ARG(size == sizeof(FileHeader) + 4 * h.words, "dag_check: %llu bytes for %llu words", (unsigned long long)size,
              (unsigned long long)h.words);
    GUARD_BEGIN
    // Copy the words out (the buffer may be unaligned) before reading them.
    std::vector<uint32_t> w(h.words);
    std::memcpy(w.data(), static_cast<const char*>(buf) + sizeof(FileHeader), 4 * h.words);
    CHECK_ARG(words_checksum(w.data(), h.words) == h.checksum, "dag_check: checksum mismatch");
    CHECK_ARG(w[0] == h.root_mask, "dag_check: root mask 0x%x, header says 0x%x", w[0], h.root_mask);
    uint64_t nodes[kMaxLevels];
    const int rc = check_structure(w.data(), h.words, (int)h.num_levels, leaf_dag(h), nodes);
    if (rc) return rc;
    if (num_levels) *num_levels = h.num_levels;
    if (words) *words = h.words;
    if (level_nodes)
        for (int L = 0; L < cap && L < kMaxLevels; ++L) level_nodes[L] = nodes[L];
    return CPVS_OK;
    GUARD_END
}

int cpvs_dag_deserialize(cpvs_ctx* ctx, const void* buf, uint64_t size, cpvs_dag** out) {
    CHECK_ARG(ctx && buf && out, "dag_deserialize: NULL argument");
    uint64_t nodes[kMaxLevels];
    int rc = cpvs_dag_check(buf, size, nullptr, nullptr, nodes, kMaxLevels);
    if (rc) return rc;
    GUARD_BEGIN
    FileHeader h;
    std::memcpy(&h, buf, sizeof(h));
    std::vector<uint32_t> w(h.words);
    std::memcpy(w.data(), static_cast<const char*>(buf) + sizeof(FileHeader), 4 * h.words);
    CPVS_HIP_CHECK(hipSetDevice(ctx->device));
    uint32_t* dw = static_cast<uint32_t*>(ctx->buffers.take(4 * h.words));
    if (!dw) {
        set_error("dag_deserialize: hipMalloc of %llu words failed", (unsigned long long)h.words);
        return CPVS_ENOMEM;
    }
    const hipError_t e = hipMemcpy(dw, w.data(), 4 * h.words, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ctx->buffers.give(dw);
        return hip_fail(e, "dag_deserialize copy", __FILE__, __LINE__);
    }
    cpvs_dag* d = new cpvs_dag();
    d->ctx = ctx;
    d->n = (int)h.num_levels;
    d->words = h.words;
    d->d_words = dw;
    d->root_mask = h.root_mask;
    d->N = h.N;
    d->zTileIndex = h.z_tile_index;
    d->zTileNum = h.z_tile_num;
    d->leafmasks = leaf_dag(h);
    for (int L = 0; L < kMaxLevels; ++L) d->uniques[L] = nodes[L];  // pre-merge counts are not stored
    *out = d;
    return CPVS_OK;
    GUARD_END
}

int cpvs_dag_save(const cpvs_dag* d, const char* path) {
    CHECK_ARG(d && path, "dag_save: NULL argument");
    GUARD_BEGIN
    uint64_t size = 0;
    int rc = cpvs_dag_serialize(d, nullptr, 0, &size);
    if (rc) return rc;
    std::vector<char> buf(size);
    rc = cpvs_dag_serialize(d, buf.data(), size, &size);
    if (rc) return rc;
    FILE* f = std::fopen(path, "wb");
    CHECK_ARG(f, "dag_save: cannot open %s for writing", path);
    const size_t put = std::fwrite(buf.data(), 1, size, f);
    const int closed = std::fclose(f);
    if (put != size || closed != 0) {
        set_error("dag_save: writing %s failed", path);
        return CPVS_EINVAL;
    }
    return CPVS_OK;
    GUARD_END
}

int cpvs_dag_load(cpvs_ctx* ctx, const char* path, cpvs_dag** out) {
    CHECK_ARG(ctx && path && out, "dag_load: NULL argument");
    GUARD_BEGIN
    FILE* f = std::fopen(path, "rb");
    CHECK_ARG(f, "dag_load: cannot open %s", path);
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    CHECK_ARG(!err, "dag_load: reading %s failed", path);
    return cpvs_dag_deserialize(ctx, buf.data(), buf.size(), out);
    GUARD_END
}

}  // extern "C"
