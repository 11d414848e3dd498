// Shared plumbing of the C ABI: the context and the error guard every entry point runs in.
#pragma once
#include "../../include/maelstrom_hip.h"
#include "common.hpp"
#include <string>

struct mael_ctx {
    int device = 0;
    std::string err;
};

namespace {

// Run `f` on the context's device; the library's errors become a status code and the
// context's last-error text.
template <class F> int guard(mael_ctx* ctx, F&& f) {
    try {
        if (ctx) MAEL_HIP(hipSetDevice(ctx->device));
        f();
        return MAEL_OK;
    } catch (const mael::Error& e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return MAEL_E_ARG;
    }
}

struct ArrRef { void* p; int64_t n; int32_t es; };

}  // namespace
