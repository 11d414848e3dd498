// Shared plumbing of the C ABI: the context and the error guard every entry point runs in.
#pragma once
#include "../../include/maelstrom_hip.h"
#include "common.hpp"
#include "sparse.hpp"
#include "stir.hpp"
#include <memory>
#include <string>

struct mael_ctx {
    int device = 0;
    std::string err;
};
struct mael_dense {
    mael_ctx* ctx;
    std::unique_ptr<mael::DenseBasin> b;
};
struct mael_sparse {
    mael_ctx* ctx;
    std::unique_ptr<mael::SparseBasin> b;
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

namespace mael {
// The stirring model's parameters from the C ABI's struct.
inline StirParams stir_params(const mael_stir& p) {
    StirParams q;
    q.radius = p.radius; q.band = p.band; q.relax = p.relax; q.core = p.core; q.Gamma = p.gamma_circ;
    q.current[0] = p.current[0]; q.current[1] = p.current[1];
    q.eddy = p.eddy; q.eddyscale = p.eddyscale; q.eddytime = p.eddytime; q.guide = p.guide; q.free = p.free;
    q.turbulence = p.turbulence; q.boil = p.boil; q.reach = p.reach;
    q.origin[0] = p.origin[0]; q.origin[1] = p.origin[1]; q.offset = p.offset;
    q.surface = p.surface; q.hradius = p.hradius; q.ramp = p.ramp;
    return q;
}
}  // namespace mael
