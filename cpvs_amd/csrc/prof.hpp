// prof.hpp -- optional per-kernel timing with HIP events on the launch stream.
//
// When a context has profiling on, the entry point installs its profiler for the calling
// thread; every kernel launched through CPVS_LAUNCH is bracketed by two events on the stream
// it runs on.  Events are resolved after the entry point's final stream synchronisation and
// accumulated per kernel name (total ms, launches).
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

namespace cpvs {

struct Prof {
    struct Rec {
        const char* name;
        hipEvent_t a, b;
    };
    // Builds (under the context's lock) and lookups (without it) may record at once, and a reader
    // may resolve meanwhile: every access to pending and acc holds mu.
    std::mutex mu;
    std::vector<Rec> pending;
    std::map<std::string, std::pair<double, long>> acc;  // name -> (total ms, launches)
    void add(const char* name, hipEvent_t a, hipEvent_t b) {
        std::lock_guard<std::mutex> lk(mu);
        pending.push_back({name, a, b});
    }
    void resolve();
    void reset();
    std::map<std::string, std::pair<double, long>> snapshot() {
        resolve();
        std::lock_guard<std::mutex> lk(mu);
        return acc;
    }
};

extern thread_local Prof* t_prof;

struct ProfScope {
    const char* name;
    hipStream_t s;
    hipEvent_t a = nullptr;
    ProfScope(const char* n, hipStream_t st) : name(n), s(st) {
        if (t_prof && hipEventCreate(&a) == hipSuccess) (void)hipEventRecord(a, s);
    }
    ~ProfScope() {
        if (!t_prof || !a) return;
        hipEvent_t b = nullptr;
        if (hipEventCreate(&b) == hipSuccess) {
            (void)hipEventRecord(b, s);
            t_prof->add(name, a, b);
        } else {
            (void)hipEventDestroy(a);
        }
    }
};

// Installs a profiler for the current thread for the lifetime of the guard.
struct ProfInstall {
    Prof* prev;
    explicit ProfInstall(Prof* p) : prev(t_prof) { t_prof = p; }
    ~ProfInstall() {
        if (t_prof) t_prof->resolve();
        t_prof = prev;
    }
};

}  // namespace cpvs

#define CPVS_LAUNCH(name, kernel, grid, block, shmem, stream, ...)           \
    do {                                                                     \
        ::cpvs::ProfScope _cpvs_ps(name, stream);                            \
        hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__); \
    } while (0)
