// The C ABI (include/maelstrom_hip.h): handles, array access and the phase entry points.
// Every entry point catches the library's errors and turns them into a status code and
// the context's last-error text.
#include "../../include/maelstrom_hip.h"
#include "capi_common.hpp"
#include "dense.hpp"
#include "scan.hpp"
#include <cmath>
#include <memory>
#include <string>



namespace {


ArrRef dense_array(mael::DenseBasin& b, int which) {
    using namespace mael;
    size_t C = b.cells();
    int c = b.cur;
    switch (which) {
        case MAEL_U: case MAEL_V: case MAEL_W: return {b.u[which - MAEL_U].p, (int64_t)b.faces(which - MAEL_U), 4};
        case MAEL_U0: case MAEL_V0: case MAEL_W0: return {b.u0[which - MAEL_U0].p, (int64_t)b.faces(which - MAEL_U0), 4};
        case MAEL_PHI: return {b.phi.p, (int64_t)C, 4};
        case MAEL_PHI0: return {b.phi0.p, (int64_t)C, 4};
        case MAEL_TYPES: return {b.types.p, (int64_t)C, 1};
        case MAEL_P: return {b.P.p, (int64_t)C, 4};
        case MAEL_FU: case MAEL_FV: case MAEL_FW: return {b.flag[which - MAEL_FU].p, (int64_t)b.faces(which - MAEL_FU), 1};
        case MAEL_RHS: return {b.b.p, (int64_t)C, 4};
        case MAEL_KEY: return {b.key.p, b.capacity, 4};
        case MAEL_RANK: return {b.rank.p, b.capacity, 4};
        case MAEL_COUNTS: return {b.counts.p, (int64_t)C, 4};
        case MAEL_TOTALS: return {b.totals.p, (int64_t)C, 4};
        case MAEL_OFFSETS: return {b.offsets.p, (int64_t)C, 4};
        case MAEL_MOM_U: case MAEL_MOM_V: case MAEL_MOM_W:
            return {b.mom[which - MAEL_MOM_U].p, (int64_t)b.faces(which - MAEL_MOM_U), 8};
        case MAEL_WT_U: case MAEL_WT_V: case MAEL_WT_W:
            return {b.wt[which - MAEL_WT_U].p, (int64_t)b.faces(which - MAEL_WT_U), 8};
        case MAEL_UB_U: case MAEL_UB_V: case MAEL_UB_W:
            b.adj_alloc();
            return {b.ub[which - MAEL_UB_U].p, (int64_t)b.faces(which - MAEL_UB_U), 4};
        case MAEL_UB0_U: case MAEL_UB0_V: case MAEL_UB0_W:
            b.adj_alloc();
            return {b.ub0[which - MAEL_UB0_U].p, (int64_t)b.faces(which - MAEL_UB0_U), 4};
        default:
            if (which >= MAEL_PB_VX && which <= MAEL_PB_C8) {
                b.adj_alloc();
                return {b.pb[which - MAEL_PB_VX].p, b.capacity, 4};
            }
            if (which >= MAEL_PX && which <= MAEL_LAG) return {b.pbuf[c][which - MAEL_PX].p, b.capacity, 4};
    }
    throw mael::Error(mael::E_ARG, "unknown array " + std::to_string(which));
}

}  // namespace

extern "C" {

const char* mael_version(void) { return "maelstrom-hip 0.1 (gfx950)"; }

int mael_create(int device, mael_ctx** out) {
    if (!out) return MAEL_E_ARG;
    auto* c = new mael_ctx;
    c->device = device;
    int st = guard(c, [&] {
        int count = 0;
        MAEL_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count)
            throw mael::Error(mael::E_ARG, "no HIP device " + std::to_string(device));
        MAEL_HIP(hipSetDevice(device));
        mael::noalloc_setup(device);   // the static scratch of mael_dense_resample_noalloc
    });
    if (st != MAEL_OK) {
        delete c;
        *out = nullptr;
        return st;
    }
    *out = c;
    return MAEL_OK;
}
void mael_destroy(mael_ctx* ctx) { delete ctx; }
const char* mael_last_error(const mael_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int mael_flipgrid(float radius, float depth, float top, float dx, int32_t pad, mael_dense_grid* out) {
    if (!out || !(dx > 0.0f)) return MAEL_E_ARG;
    // flip.jl:48-62, in the reference's Float32 arithmetic.
    int nxy = (int)std::ceil(2.0f * (radius + (float)pad * dx) / dx);
    float loz = -depth - dx;
    int nz0 = (int)std::ceil((top - loz) / dx);
    for (int levels = 2; levels < 30; ++levels) {
        int m = 1 << (levels - 1);
        int nx = m * ((nxy + m - 1) / m), nz = m * ((nz0 + m - 1) / m);
        if ((long)(nx / m) * (nx / m) * (nz / m) <= 1024) {
            float half = (float)(-nx) * dx / 2.0f;
            out->lo[0] = half; out->lo[1] = half; out->lo[2] = loz;
            out->dx = dx; out->n[0] = nx; out->n[1] = nx; out->n[2] = nz;
            out->radius = radius; out->floor = -depth; out->levels = levels;
            return MAEL_OK;
        }
    }
    return MAEL_E_GRID;
}

void mael_solve_defaults(mael_solve_opts* o) {
    if (!o) return;
    o->tol = 1e-4; o->maxiter = 60; o->sweeps = 2; o->coarse = 40; o->precision = 0; o->refine = 0; o->warm = 1;
}

int mael_dense_create(mael_ctx* ctx, const mael_dense_grid* grid, int64_t capacity, int32_t band, int32_t nmin,
                      mael_dense** out) {
    if (!ctx || !grid || !out || capacity < 0 || band < 1 || nmin < 1) return MAEL_E_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        for (int a = 0; a < 3; ++a)
            if (grid->n[a] < 2 || (grid->n[a] % (1 << (grid->levels - 1))) != 0)
                throw mael::Error(mael::E_ARG, "grid sides must be multiples of 2^(levels-1)");
        long coarsest = (long)(grid->n[0] >> (grid->levels - 1)) * (grid->n[1] >> (grid->levels - 1)) *
                        (grid->n[2] >> (grid->levels - 1));
        if (coarsest > 1024) throw mael::Error(mael::E_GRID, "the coarsest level has more than 1024 cells");
        mael::DGrid g;
        for (int a = 0; a < 3; ++a) { g.lo[a] = grid->lo[a]; g.n[a] = grid->n[a]; }
        g.dx = grid->dx; g.radius = grid->radius; g.floor = grid->floor;
        auto* d = new mael_dense;
        d->ctx = ctx;
        try {
            d->b.reset(new mael::DenseBasin(g, grid->levels, capacity, band, nmin));
        } catch (...) {
            delete d;
            throw;
        }
        *out = d;
    });
}
void mael_dense_destroy(mael_dense* d) {
    if (!d) return;
    if (d->ctx) (void)hipSetDevice(d->ctx->device);
    delete d;
}

int mael_dense_array(mael_dense* d, int32_t which, void** dptr, int64_t* count, int32_t* elsize) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        ArrRef a = dense_array(*d->b, which);
        if (dptr) *dptr = a.p;
        if (count) *count = a.n;
        if (elsize) *elsize = a.es;
    });
}
int mael_dense_upload(mael_dense* d, int32_t which, const void* host, int64_t count) {
    if (!d || (!host && count)) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        ArrRef a = dense_array(*d->b, which);
        if (count < 0 || count > a.n) throw mael::Error(mael::E_ARG, "upload larger than the array");
        MAEL_HIP(hipMemcpyAsync(a.p, host, (size_t)count * a.es, hipMemcpyHostToDevice, d->b->st.s));
        d->b->st.sync();
    });
}
int mael_dense_download(mael_dense* d, int32_t which, void* host, int64_t count) {
    if (!d || (!host && count)) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        ArrRef a = dense_array(*d->b, which);
        if (count < 0 || count > a.n) throw mael::Error(mael::E_ARG, "download larger than the array");
        MAEL_HIP(hipMemcpyAsync(host, a.p, (size_t)count * a.es, hipMemcpyDeviceToHost, d->b->st.s));
        d->b->st.sync();
    });
}
int mael_dense_count(mael_dense* d, int64_t* n) {
    if (!d || !n) return MAEL_E_ARG;
    *n = d->b->n;
    return MAEL_OK;
}
int mael_dense_set_count(mael_dense* d, int64_t n) {
    if (!d || n < 0 || n > d->b->capacity) return MAEL_E_ARG;
    d->b->n = n;
    return MAEL_OK;
}

int mael_dense_resample(mael_dense* d, int32_t seeding, uint32_t seed, int64_t* total) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        int64_t t = d->b->resample(seeding != 0, seed);
        if (total) *total = t;
    });
}
int mael_dense_resample_noalloc(mael_dense* d, int32_t seeding, uint32_t seed, int64_t* total) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        unsigned long long* tilesums = mael::noalloc_scratch(d->ctx->device);
        if (!tilesums) throw mael::Error(mael::E_ARG, "the static scratch is not set up on this device");
        size_t need = d->b->cells() / mael::SCAN_TILE + 2;
        if (need > mael::NOALLOC_TILE_SUMS)
            throw mael::Error(mael::E_ARG, "resampling without allocation needs " + std::to_string(need) +
                                               " scan tile sums, more than the " +
                                               std::to_string(mael::NOALLOC_TILE_SUMS) + " of its static scratch");
        int64_t t = d->b->resample(seeding != 0, seed, tilesums);
        if (total) *total = t;
    });
}
int mael_dense_p2g(mael_dense* d, float dt) {
    if (!d || !(dt > 0.0f)) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->p2g(dt); });
}
int mael_dense_g2p(mael_dense* d, float dt, float gravity) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->g2p(dt, gravity); });
}
static mael::SolveOpts solve_opts(const mael_solve_opts* opts) {
    mael::SolveOpts o;
    if (opts) {
        o.tol = opts->tol; o.maxiter = opts->maxiter; o.sweeps = opts->sweeps; o.coarse = opts->coarse;
        o.precision = opts->precision; o.refine = opts->refine; o.warm = opts->warm;
    }
    if (o.sweeps < 1 || o.coarse < 1 || o.maxiter < 0 || !(o.tol >= 0.0)) throw mael::Error(mael::E_ARG, "bad solver options");
    return o;
}
int mael_dense_project(mael_dense* d, float dt, float gamma, const mael_solve_opts* opts, mael_solve_stats* stats) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        mael::SolveOpts o = solve_opts(opts);
        mael::SolveStats s;
        d->b->project(dt, gamma, o, s);
        if (stats) {
            stats->iterations = s.iterations; stats->restarts = s.restarts; stats->rel_residual = s.rel_residual;
            stats->rel_residual_rec = s.rel_residual_recursive; stats->ms = s.ms_total;
        }
    });
}
int mael_dense_vcycle(mael_dense* d, const mael_solve_opts* opts, int32_t queued) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->precondition(solve_opts(opts), queued != 0); });
}
int mael_dense_g2p_adjoint(mael_dense* d, float dt) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->g2p_adjoint(dt); });
}
int mael_dense_project_adjoint(mael_dense* d, const mael_solve_opts* opts, mael_solve_stats* stats) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        mael::SolveOpts o = solve_opts(opts);
        mael::SolveStats s;
        d->b->project_adjoint(o, s);
        if (stats) {
            stats->iterations = s.iterations; stats->restarts = s.restarts; stats->rel_residual = s.rel_residual;
            stats->rel_residual_rec = s.rel_residual_recursive; stats->ms = s.ms_total;
        }
    });
}
int mael_dense_p2g_adjoint(mael_dense* d, float dt) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->p2g_adjoint(dt); });
}
int mael_dense_classify(mael_dense* d) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->classify(); });
}
int mael_dense_redistance(mael_dense* d, int32_t iterations) {
    if (!d || iterations < 0) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->redistance(iterations); });
}
int mael_dense_volume(mael_dense* d, double* v) {
    if (!d || !v) return MAEL_E_ARG;
    return guard(d->ctx, [&] { *v = d->b->volume(); });
}
int mael_dense_init_vortex(mael_dense* d, const float* eta, int32_t neta, float rmax, float gamma_circ, float core) {
    if (!d || !eta || neta < 2) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->init_vortex(eta, neta, rmax, gamma_circ, core); });
}
int mael_dense_solidify(mael_dense* d) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->solidify(); });
}
int mael_dense_gravity(mael_dense* d, float dt, float gravity) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->gravity(dt, gravity); });
}
int mael_dense_advect(mael_dense* d, float dt, float gravity, int32_t seed) {
    if (!d || !(dt >= 0.0f)) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->advect(dt, gravity, seed); });
}
int mael_dense_filltrapped(mael_dense* d, int32_t maxrounds) {
    if (!d || maxrounds < 1) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->filltrapped(maxrounds); });
}
int mael_dense_set_stir(mael_dense* d, const mael_stir* p, const float* waves, int32_t nwaves) {
    if (!d || !p || nwaves < 0 || (nwaves && !waves) || p->surface) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->set_stir(mael::stir_params(*p), waves, nwaves); });
}
int mael_dense_forces(mael_dense* d, float dt, float time, float gravity) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->forces(dt, time, gravity); });
}
int mael_dense_extrapolate(mael_dense* d, int32_t layers) {
    if (!d || layers < 0) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->extrapolate(layers); });
}
int mael_dense_maxspeed(mael_dense* d, float* speed) {
    if (!d || !speed) return MAEL_E_ARG;
    return guard(d->ctx, [&] { *speed = d->b->maxspeed(); });
}
int mael_dense_pack_band(mael_dense* d, float keep, float throat, uint16_t* pos, uint16_t* vel, int64_t cap, int64_t* count) {
    if (!d || !pos || !vel || !count) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        if (cap < d->b->n) throw mael::Error(mael::E_ARG, "the packing buffers hold fewer than the particles");
        *count = d->b->pack_band(keep, throat, pos, vel);
    });
}
int mael_dense_quantize_levelset(mael_dense* d, float range, uint8_t* out) {
    if (!d || !out || !(range > 0.0f)) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->quantize_levelset(range, out); });
}
int mael_dense_copy(mael_dense* d, int32_t dst, int32_t src) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        ArrRef a = dense_array(*d->b, dst), b = dense_array(*d->b, src);
        if (a.n != b.n || a.es != b.es) throw mael::Error(mael::E_ARG, "copy between arrays of different shapes");
        MAEL_HIP(hipMemcpyAsync(a.p, b.p, (size_t)a.n * a.es, hipMemcpyDeviceToDevice, d->b->st.s));
    });
}
int mael_dense_ktimer(mael_dense* d, int32_t kernel, int32_t launches) {
    if (!d || kernel < 0 || kernel >= mael::K_NKERNELS || launches < 1) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->kt.enable(kernel, (size_t)launches); });
}
int mael_dense_ktime(mael_dense* d, double* ms, int32_t* launches) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        int n = 0;
        double t = d->b->kt.read(&n);
        if (ms) *ms = t;
        if (launches) *launches = n;
    });
}
int mael_dense_sync(mael_dense* d) {
    if (!d) return MAEL_E_ARG;
    return guard(d->ctx, [&] { d->b->st.sync(); });
}

}  // extern "C"
