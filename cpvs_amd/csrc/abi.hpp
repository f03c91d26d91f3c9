// abi.hpp -- argument checks and exception guards shared by the extern "C" entry points.
#pragma once

#include <new>

#include "internal.hpp"

#define CHECK_ARG(cond, ...)              \
    do {                                  \
        if (!(cond)) {                    \
            ::cpvs::set_error(__VA_ARGS__); \
            return CPVS_EINVAL;           \
        }                                 \
    } while (0)

#define GUARD_BEGIN try {
#define GUARD_END                                               \
    }                                                           \
    catch (const std::bad_alloc&) {                             \
        ::cpvs::set_error("host allocation failed");            \
        return CPVS_ENOMEM;                                     \
    }                                                           \
    catch (...) {                                               \
        ::cpvs::set_error("internal exception");                \
        return CPVS_EINTERNAL;                                  \
    }

static inline hipStream_t pick_stream(cpvs_ctx* ctx, void* stream) {
    return stream ? reinterpret_cast<hipStream_t>(stream) : ctx->stream;
}

static inline bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }
