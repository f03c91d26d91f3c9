// cpvs_amd.hpp -- header-only C++ facade over the C ABI (cpvs.h) with the class names, signatures
// and ownership rules of TobiasRp/cpvs, so C++ callers of the reference switch by changing the
// include and the namespace:
//
//   reference                                   here
//   ImageF (src/Image.h:10-69)                   cpvs_amd::ImageF
//   MinMaxHierarchy (src/MinMaxHierarchy.h)      cpvs_amd::MinMaxHierarchy      (device pyramid)
//   CompressedShadow (src/CompressedShadow.h)    cpvs_amd::CompressedShadow     (device DAG)
//   CompressedShadowContainer                    cpvs_amd::CompressedShadowContainer
//     (src/CompressedShadowContainer.h)
//
// Differences a caller sees: failures throw cpvs_amd::Error (the reference asserts);
// evaluate() takes device buffers instead of GL textures; getDAG() copies the device words to a
// host vector on first use.  All work runs on the GPU of the process-wide default context.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cpvs.h"

namespace cpvs_amd {

using uint = unsigned int;

class Error : public std::runtime_error {
public:
    Error(const std::string& what, int code) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const char* what) {
    if (rc != CPVS_OK) throw Error(std::string(what) + ": " + cpvs_last_error(), rc);
}

// The process-wide context on device 0 (the reference runs in one GL context).  The reference
// compiles leafmasks in or out (#define LEAFMASKS, src/CompressedShadow.cpp:17); compiling this
// header with CPVS_AMD_NO_LEAFMASKS defined is the same switch (setLeafmasks at run time).
class Context {
public:
    static Context& instance() {
        static Context c;
        return c;
    }
    cpvs_ctx* get() const { return ctx_; }
    // Builds from now on end in leafmask nodes (true, the default) or go down to level 0.
    void setLeafmasks(bool on) { check(cpvs_ctx_set_leafmasks(ctx_, on ? 1 : 0), "cpvs_ctx_set_leafmasks"); }
    bool leafmasks() const {
        int on = 1;
        check(cpvs_ctx_leafmasks(ctx_, &on), "cpvs_ctx_leafmasks");
        return on != 0;
    }
    ~Context() { cpvs_ctx_destroy(ctx_); }

private:
    Context() {
        check(cpvs_ctx_create(0, &ctx_), "cpvs_ctx_create");
#ifdef CPVS_AMD_NO_LEAFMASKS
        setLeafmasks(false);
#endif
    }
    cpvs_ctx* ctx_ = nullptr;
};

// Image<float> (src/Image.h:10-69): row-major, interleaved channels.
class ImageF {
public:
    ImageF(size_t width, size_t height, size_t numChannels)
        : w_(width), h_(height), c_(numChannels), v_(width * height * numChannels) {}
    void setAll(const float* p) { v_.assign(p, p + w_ * h_ * c_); }
    void setAll(const std::vector<float>& v) { v_.assign(v.begin(), v.end()); }
    float get(size_t x, size_t y, size_t ch) const { return v_[(y * w_ + x) * c_ + ch]; }
    void set(size_t x, size_t y, size_t ch, float val) { v_[(y * w_ + x) * c_ + ch] = val; }
    const float* data() const { return v_.data(); }
    float* data() { return v_.data(); }
    size_t getNumChannels() const { return c_; }
    size_t getWidth() const { return w_; }
    size_t getHeight() const { return h_; }

private:
    size_t w_, h_, c_;
    std::vector<float> v_;
};

// MinMaxHierarchy (src/MinMaxHierarchy.h:14-72), built on the device.
class MinMaxHierarchy {
public:
    // MinMaxHierarchy(const ImageF&) (src/MinMaxHierarchy.h:23): square, power-of-two, 1 channel.
    explicit MinMaxHierarchy(const ImageF& orig, bool allLevels = true) {
        if (orig.getWidth() != orig.getHeight() || orig.getNumChannels() != 1)
            throw Error("MinMaxHierarchy: square one-channel image expected", CPVS_EINVAL);
        check(cpvs_pyramid_build(Context::instance().get(), orig.data(), 0, (uint32_t)orig.getWidth(),
                                 allLevels ? CPVS_PYR_ALL_LEVELS : 0u, nullptr, &p_),
              "cpvs_pyramid_build");
    }
    // From depth already in device memory (borrowed: the caller keeps it alive).
    MinMaxHierarchy(const float* deviceDepth, uint32_t size, void* stream = nullptr) {
        check(cpvs_pyramid_build(Context::instance().get(), deviceDepth, 1, size, CPVS_PYR_BORROW_DEPTH, stream, &p_),
              "cpvs_pyramid_build");
    }
    ~MinMaxHierarchy() { cpvs_pyramid_destroy(p_); }
    MinMaxHierarchy(const MinMaxHierarchy&) = delete;
    MinMaxHierarchy& operator=(const MinMaxHierarchy&) = delete;

    float getMin(size_t level, size_t x, size_t y) const { return get(level, x, y).first; }
    float getMax(size_t level, size_t x, size_t y) const { return get(level, x, y).second; }
    int getNumLevels() const {
        int n = 0;
        check(cpvs_pyramid_num_levels(p_, &n), "cpvs_pyramid_num_levels");
        return n;
    }
    // getLevel (src/MinMaxHierarchy.h:67-72) as a host image: 1 channel at level 0, else 2.
    ImageF getLevel(size_t level) const {
        uint32_t N = 1u << (getNumLevels() - 1);
        const size_t w = N >> level;
        ImageF img(w, w, level == 0 ? 1 : 2);
        check(cpvs_pyramid_copy_level(p_, (uint32_t)level, img.data(), (uint64_t)(w * w * img.getNumChannels())),
              "cpvs_pyramid_copy_level");
        return img;
    }
    const cpvs_pyramid* handle() const { return p_; }

private:
    std::pair<float, float> get(size_t level, size_t x, size_t y) const {
        float mn = 0, mx = 0;
        check(cpvs_pyramid_get(p_, (uint32_t)level, (uint32_t)x, (uint32_t)y, &mn, &mx), "cpvs_pyramid_get");
        return {mn, mx};
    }
    cpvs_pyramid* p_ = nullptr;
};

// CompressedShadow (src/CompressedShadow.h:20-125).
class CompressedShadow {
public:
    enum NodeVisibility { SHADOW = CPVS_SHADOW, VISIBLE = CPVS_VISIBLE, PARTIAL = CPVS_PARTIAL };

    // create(const MinMaxHierarchy&, uint zTileIndex, uint zTileNum) (src/CompressedShadow.h:48-49).
    static std::unique_ptr<CompressedShadow> create(const MinMaxHierarchy& minMax, uint zTileIndex = 0,
                                                    uint zTileNum = 1) {
        cpvs_dag* d = nullptr;
        check(cpvs_build(Context::instance().get(), minMax.handle(), zTileIndex, zTileNum, nullptr, &d), "cpvs_build");
        return std::unique_ptr<CompressedShadow>(new CompressedShadow(d));
    }

    ~CompressedShadow() {
        if (owns_) cpvs_dag_destroy(d_);
    }
    CompressedShadow(const CompressedShadow&) = delete;
    CompressedShadow& operator=(const CompressedShadow&) = delete;

    // traverse(vec3, bool) (src/CompressedShadow.h:64); V is any type with x, y, z (glm::vec3).
    template <class V>
    NodeVisibility traverse(const V& position, bool tryLeafmasks = true) {
        const float ndc[3] = {(float)position.x, (float)position.y, (float)position.z};
        int vis = 0;
        check(cpvs_dag_traverse(d_, ndc, 1, tryLeafmasks ? 1 : 0, &vis), "cpvs_dag_traverse");
        return (NodeVisibility)vis;
    }
    NodeVisibility getTotalVisibility() const {
        int tv = 0;
        check(cpvs_dag_info(d_, nullptr, nullptr, &tv), "cpvs_dag_info");
        return (NodeVisibility)tv;
    }
    uint getNumLevels() const {
        uint32_t n = 0;
        check(cpvs_dag_info(d_, &n, nullptr, nullptr), "cpvs_dag_info");
        return n;
    }
    // getDAG (src/CompressedShadow.h:75-77): the words, copied from the device on first use.
    const std::vector<uint>& getDAG() const {
        if (host_.empty()) {
            uint64_t n = 0;
            check(cpvs_dag_info(d_, nullptr, &n, nullptr), "cpvs_dag_info");
            host_.resize(n);
            check(cpvs_dag_copy_words(d_, host_.data(), n), "cpvs_dag_copy_words");
        }
        return host_;
    }
    cpvs_dag* release() {
        cpvs_dag* d = d_;
        d_ = nullptr;
        return d;
    }
    // The DAG in the library's byte format (64-byte header + words), and back.
    void save(const std::string& path) const { check(cpvs_dag_save(d_, path.c_str()), "cpvs_dag_save"); }
    static std::unique_ptr<CompressedShadow> load(const std::string& path) {
        cpvs_dag* d = nullptr;
        check(cpvs_dag_load(Context::instance().get(), path.c_str(), &d), "cpvs_dag_load");
        return std::unique_ptr<CompressedShadow>(new CompressedShadow(d));
    }

    const cpvs_dag* handle() const { return d_; }

private:
    friend class CompressedShadowContainer;
    // owns = false: a view of a DAG a container owns (what CompressedShadowContainer::get returns).
    explicit CompressedShadow(cpvs_dag* d, bool owns = true) : d_(d), owns_(owns) {}
    cpvs_dag* d_ = nullptr;
    bool owns_ = true;
    mutable std::vector<uint> host_;
};

// CompressedShadowContainer (src/CompressedShadowContainer.h:18-92).
class CompressedShadowContainer {
public:
    explicit CompressedShadowContainer(uint length) : length_(length) {
        check(cpvs_container_create(Context::instance().get(), length, &c_), "cpvs_container_create");
        views_.resize((size_t)length * length * length);
    }
    explicit CompressedShadowContainer(std::unique_ptr<CompressedShadow> shadow) : CompressedShadowContainer(1u) {
        set(std::move(shadow), 0, 0, 0);
    }
    ~CompressedShadowContainer() { cpvs_container_destroy(c_); }
    CompressedShadowContainer(const CompressedShadowContainer&) = delete;
    CompressedShadowContainer& operator=(const CompressedShadowContainer&) = delete;

    void set(std::unique_ptr<CompressedShadow> shadow, uint x, uint y, uint z) {
        cpvs_dag* d = shadow->d_;
        check(cpvs_container_set(c_, d, x, y, z), "cpvs_container_set");
        shadow->release();  // the container owns the DAG now
        views_[slot(x, y, z)].reset(new CompressedShadow(d, false));
    }
    // get (src/CompressedShadowContainer.h:41-46): the shadow in slot (x, y, z), owned by the
    // container; null when the slot is empty or after freeOnCPU / moveToGPU.
    const CompressedShadow* get(uint x, uint y, uint z) const {
        const cpvs_dag* d = nullptr;
        check(cpvs_container_get(c_, x, y, z, &d), "cpvs_container_get");
        return d ? views_[slot(x, y, z)].get() : nullptr;
    }
    // evaluate (src/CompressedShadowContainer.cpp:93-124) over device positions (3 or 4 floats
    // per point) with a column-major light view-projection matrix; vis gets 0/1 per point.
    void evaluate(const float* devicePositions, uint32_t strideFloats, uint64_t count, const float lightViewProj[16],
                  uint8_t* deviceVisibilities, void* stream = nullptr) {
        check(cpvs_container_evaluate(c_, devicePositions, strideFloats, count, lightViewProj, deviceVisibilities, stream),
              "cpvs_container_evaluate");
    }
    // PCF with the filter size of setFilterSize (the reference declares it, shader/traverse.cs:16-17):
    // one float per point, the fraction of lit leaf cells over the f x f neighbourhood.
    void evaluatePCF(const float* devicePositions, uint32_t strideFloats, uint64_t count, const float lightViewProj[16],
                     float* deviceVisibilities, void* stream = nullptr) {
        check(cpvs_container_evaluate_pcf(c_, devicePositions, strideFloats, count, lightViewProj, deviceVisibilities,
                                          stream),
              "cpvs_container_evaluate_pcf");
    }
    void freeOnCPU() {
        check(cpvs_container_free_on_cpu(c_), "cpvs_container_free_on_cpu");
        for (auto& v : views_) v.reset();
    }
    void copyToGPU() { check(cpvs_container_copy_to_gpu(c_), "cpvs_container_copy_to_gpu"); }
    void moveToGPU() {
        check(cpvs_container_move_to_gpu(c_), "cpvs_container_move_to_gpu");
        for (auto& v : views_) v.reset();
    }
    void setFilterSize(uint size) { check(cpvs_container_set_filter_size(c_, size), "cpvs_container_set_filter_size"); }
    // Lookup top table levels for the next copyToGPU (-1 automatic, 0 none, 1..9); no counterpart in
    // the reference, results are the same with any setting.
    void setTopLevels(int levels) { check(cpvs_container_set_top_levels(c_, levels), "cpvs_container_set_top_levels"); }
    int topLevels() const {
        int k = 0;
        check(cpvs_container_top_levels(c_, &k), "cpvs_container_top_levels");
        return k;
    }
    cpvs_container* handle() const { return c_; }

private:
    size_t slot(uint x, uint y, uint z) const { return ((size_t)z * length_ + y) * length_ + x; }
    cpvs_container* c_ = nullptr;
    uint length_ = 1;
    std::vector<std::unique_ptr<CompressedShadow>> views_;  // non-owning, one per slot
};

}  // namespace cpvs_amd
