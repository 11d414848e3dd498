// The C ABI of the sparse basin (include/maelstrom_hip.h, `mael_sparse_*`).
#include "capi_common.hpp"
#include "sparse.hpp"
#include "scan.hpp"
#include <cmath>
#include <memory>


namespace {

ArrRef sparse_array(mael::SparseBasin& b, int which) {
    using namespace mael;
    int64_t nf = (int64_t)b.blocks * SBCELLS, nfa = (int64_t)b.blocks * SBFACES, Cc = (int64_t)b.ncells();
    auto face = [&](int a, bool fine) { return fine ? nfa : (int64_t)b.nfaces(a); };
    int c = b.cur;
    switch (which) {
        case MAEL_S_PHI_F: return {b.phi.fine.p, nf, 4};
        case MAEL_S_PHI_C: return {b.phi.coarse.p, Cc, 4};
        case MAEL_S_PHI0_F: return {b.phi0.fine.p, nf, 4};
        case MAEL_S_PHI0_C: return {b.phi0.coarse.p, Cc, 4};
        case MAEL_S_P_F: return {b.P.fine.p, nf, 4};
        case MAEL_S_P_C: return {b.P.coarse.p, Cc, 4};
        case MAEL_S_TYPES_F: return {b.types.fine.p, nf, 1};
        case MAEL_S_TYPES_C: return {b.types.coarse.p, Cc, 1};
        case MAEL_S_RHS_F: return {b.b.fine.p, nf, 4};
        case MAEL_S_RHS_C: return {b.b.coarse.p, Cc, 4};
        case MAEL_S_KEY: return {b.key.p, b.capacity, 4};
        case MAEL_S_RANK: return {b.rank.p, b.capacity, 4};
        case MAEL_S_COUNTS: return {b.counts.p, nf, 4};
        case MAEL_S_TOTALS: return {b.totals.p, nf, 4};
        case MAEL_S_OFFSETS: return {b.offsets.p, nf, 4};
        case MAEL_S_SLOT: return {b.slot.p, (int64_t)b.nblocks(), 4};
        case MAEL_S_WHERE: return {b.where.p, (int64_t)b.blocks * 4, 4};
        default: break;
    }
    if (which >= MAEL_S_U_F && which <= MAEL_S_W_C) {
        int a = (which - MAEL_S_U_F) / 2;
        bool fine = ((which - MAEL_S_U_F) & 1) == 0;
        return {fine ? (void*)b.u[a].fine.p : (void*)b.u[a].coarse.p, face(a, fine), 4};
    }
    if (which >= MAEL_S_U0_F && which <= MAEL_S_W0_C) {
        int a = (which - MAEL_S_U0_F) / 2;
        bool fine = ((which - MAEL_S_U0_F) & 1) == 0;
        return {fine ? (void*)b.u0[a].fine.p : (void*)b.u0[a].coarse.p, face(a, fine), 4};
    }
    if (which >= MAEL_S_FU_F && which <= MAEL_S_FW_C) {
        int a = (which - MAEL_S_FU_F) / 2;
        bool fine = ((which - MAEL_S_FU_F) & 1) == 0;
        return {fine ? (void*)b.fl[a].fine.p : (void*)b.fl[a].coarse.p, face(a, fine), 1};
    }
    if (which >= MAEL_S_PX && which <= MAEL_S_LAG) return {b.pbuf[c][which - MAEL_S_PX].p, b.capacity, 4};
    if (which >= MAEL_S_MOM_U && which <= MAEL_S_MOM_W) return {b.mom[which - MAEL_S_MOM_U].p, nfa, 8};
    if (which >= MAEL_S_WT_U && which <= MAEL_S_WT_W) return {b.wt[which - MAEL_S_WT_U].p, nfa, 8};
    if (which >= MAEL_S_UB_U_F && which <= MAEL_S_UB0_W_C) {   // allocated on first use
        b.adj_alloc();
        bool zero = which >= MAEL_S_UB0_U_F;
        int k = which - (zero ? MAEL_S_UB0_U_F : MAEL_S_UB_U_F), a = k / 2;
        bool fine = (k & 1) == 0;
        SField<float>& f = zero ? b.ub0[a] : b.ub[a];
        return {fine ? (void*)f.fine.p : (void*)f.coarse.p, face(a, fine), 4};
    }
    if (which >= MAEL_S_PB_VX && which <= MAEL_S_PB_C8) {
        b.adj_alloc();
        return {b.pb[which - MAEL_S_PB_VX].p, b.capacity, 4};
    }
    throw mael::Error(mael::E_ARG, "unknown sparse array " + std::to_string(which));
}

mael::SparseOpts sparse_opts(const mael_sparse_opts* opts) {
    mael::SparseOpts o;
    if (opts) {
        o.tol = opts->tol; o.maxiter = opts->maxiter; o.sweeps = opts->sweeps; o.coarse = opts->coarse;
        o.restart = opts->restart; o.policy = opts->policy; o.precision = opts->precision; o.boost = opts->boost;
        o.refine = opts->refine;
    }
    if (o.sweeps < 1 || o.coarse < 1 || o.maxiter < 0 || !(o.tol >= 0.0) || o.restart < 1 || o.restart > 40 ||
        o.policy < 0 || o.policy > 2 || o.precision < 0 || o.precision > 2 || o.refine < 0)
        throw mael::Error(mael::E_ARG, "bad solver options");
    return o;
}

}  // namespace

extern "C" {

int mael_sparsegrid(float radius, float depth, float top, float dx, int32_t pad, mael_sparse_grid* out) {
    if (!out || !(dx > 0.0f)) return MAEL_E_ARG;
    // sflip.jl:46-53 in the reference's Float32 arithmetic
    float S = 16.0f * dx;
    int nxy = 4 * (((int)std::ceil(2.0f * (radius + (float)pad * dx) / S) + 3) / 4);
    float loz = -depth - 4.0f * dx;
    int nz = 4 * (((int)std::ceil((top - loz) / S) + 3) / 4);
    float half = (float)(-nxy) * S / 2.0f;
    out->lo[0] = half; out->lo[1] = half; out->lo[2] = loz;
    out->dx = dx;
    out->nb[0] = nxy; out->nb[1] = nxy; out->nb[2] = nz;
    out->radius = radius; out->floor = -depth;
    return MAEL_OK;
}

int mael_sparse_create(mael_ctx* ctx, const mael_sparse_grid* grid, int32_t blocks, int64_t capacity, int32_t band,
                       int32_t nmin, mael_sparse** out) {
    if (!ctx || !grid || !out || blocks < 1 || capacity < 0 || band < 1 || nmin < 1) return MAEL_E_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        for (int a = 0; a < 3; ++a)
            if (grid->nb[a] < 1) throw mael::Error(mael::E_ARG, "a sparse grid needs at least one block per axis");
        if ((int64_t)blocks * mael::SBFACES > INT32_MAX) throw mael::Error(mael::E_ARG, "too many blocks for 32-bit pool indices");
        mael::SGrid g;
        for (int a = 0; a < 3; ++a) { g.lo[a] = grid->lo[a]; g.nb[a] = grid->nb[a]; }
        g.dx = grid->dx; g.radius = grid->radius; g.floor = grid->floor;
        auto* d = new mael_sparse;
        d->ctx = ctx;
        try {
            d->b.reset(new mael::SparseBasin(g, blocks, capacity, band, nmin));
        } catch (...) {
            delete d;
            throw;
        }
        *out = d;
    });
}
void mael_sparse_destroy(mael_sparse* s) {
    if (!s) return;
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    delete s;
}

int mael_sparse_array(mael_sparse* s, int32_t which, void** dptr, int64_t* count, int32_t* elsize) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        ArrRef a = sparse_array(*s->b, which);
        if (dptr) *dptr = a.p;
        if (count) *count = a.n;
        if (elsize) *elsize = a.es;
    });
}
int mael_sparse_upload(mael_sparse* s, int32_t which, const void* host, int64_t count) {
    if (!s || (!host && count)) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        if (which == MAEL_S_SLOT || which == MAEL_S_WHERE)
            throw mael::Error(mael::E_ARG, "the topology is set with mael_sparse_set_topology");
        ArrRef a = sparse_array(*s->b, which);
        if (count < 0 || count > a.n) throw mael::Error(mael::E_ARG, "upload larger than the array");
        MAEL_HIP(hipMemcpyAsync(a.p, host, (size_t)count * a.es, hipMemcpyHostToDevice, s->b->st.s));
        s->b->st.sync();
    });
}
int mael_sparse_download(mael_sparse* s, int32_t which, void* host, int64_t count) {
    if (!s || (!host && count)) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        ArrRef a = sparse_array(*s->b, which);
        if (count < 0 || count > a.n) throw mael::Error(mael::E_ARG, "download larger than the array");
        MAEL_HIP(hipMemcpyAsync(host, a.p, (size_t)count * a.es, hipMemcpyDeviceToHost, s->b->st.s));
        s->b->st.sync();
    });
}
int mael_sparse_count(mael_sparse* s, int64_t* n) {
    if (!s || !n) return MAEL_E_ARG;
    *n = s->b->n;
    return MAEL_OK;
}
int mael_sparse_set_count(mael_sparse* s, int64_t n) {
    if (!s || n < 0 || n > s->b->capacity) return MAEL_E_ARG;
    s->b->n = n;
    return MAEL_OK;
}

int mael_sparse_get_topology(mael_sparse* s, int32_t* slot, int32_t* where, int32_t* nslots, int32_t* freelist,
                             int32_t* nfree) {
    if (!s) return MAEL_E_ARG;
    mael::SparseBasin& b = *s->b;
    if (slot) std::copy(b.hslot.begin(), b.hslot.end(), slot);
    if (where)
        for (int k = 0; k < b.blocks; ++k) {
            where[3 * k] = b.hwhere[k].x; where[3 * k + 1] = b.hwhere[k].y; where[3 * k + 2] = b.hwhere[k].z;
        }
    if (nslots) *nslots = b.nslots;
    if (freelist) std::copy(b.freel.begin(), b.freel.end(), freelist);
    if (nfree) *nfree = (int32_t)b.freel.size();
    return MAEL_OK;
}
int mael_sparse_set_topology(mael_sparse* s, const int32_t* slot, const int32_t* where, int32_t nslots,
                             const int32_t* freelist, int32_t nfree) {
    if (!s || !slot || !where || nslots < 0 || nfree < 0 || (nfree && !freelist)) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        mael::SparseBasin& b = *s->b;
        if (nslots > b.blocks || nfree > nslots) throw mael::Error(mael::E_ARG, "topology larger than the pool");
        // validate before taking it: every slot in range, each live slot's block pointing back at it
        for (size_t k = 0; k < b.nblocks(); ++k)
            if (slot[k] < 0 || slot[k] > nslots) throw mael::Error(mael::E_ARG, "slot out of range");
        for (int k = 0; k < b.blocks; ++k) {
            int x = where[3 * k], y = where[3 * k + 1], z = where[3 * k + 2];
            if (x < 0) continue;
            if (k >= nslots || x >= b.g.nb[0] || y < 0 || y >= b.g.nb[1] || z < 0 || z >= b.g.nb[2] ||
                slot[x + (size_t)b.g.nb[0] * (y + (size_t)b.g.nb[1] * z)] != k + 1)
                throw mael::Error(mael::E_ARG, "inconsistent slot table");
        }
        for (size_t k = 0; k < b.nblocks(); ++k)
            if (slot[k] > 0 && where[3 * (slot[k] - 1)] < 0) throw mael::Error(mael::E_ARG, "inconsistent slot table");
        b.hslot.assign(slot, slot + b.nblocks());
        for (int k = 0; k < b.blocks; ++k) b.hwhere[k] = make_int4(where[3 * k], where[3 * k + 1], where[3 * k + 2], 0);
        b.nslots = nslots;
        b.freel.assign(freelist, freelist + nfree);
        b.upload_topology();
    });
}

int mael_sparse_marks(mael_sparse* s, float vary, float reach, int32_t rings, int32_t particles, uint8_t* marks) {
    if (!s || !marks || rings < 0) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        s->b->marks_compute(vary, reach, rings, particles != 0);
        MAEL_HIP(hipMemcpyAsync(marks, s->b->marks.p, s->b->nblocks(), hipMemcpyDeviceToHost, s->b->st.s));
        s->b->st.sync();
    });
}
int mael_sparse_refine_marks(mael_sparse* s, const uint8_t* marks, int32_t* added, int32_t* removed) {
    if (!s || !marks) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        int na = 0, nr = 0;
        s->b->refine_from(marks, &na, &nr);
        if (added) *added = na;
        if (removed) *removed = nr;
    });
}
int mael_sparse_refine(mael_sparse* s, float vary, float reach, int32_t rings, int32_t* added, int32_t* removed) {
    if (!s || rings < 0) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        int na = 0, nr = 0;
        s->b->refine(vary, reach, rings, &na, &nr);
        if (added) *added = na;
        if (removed) *removed = nr;
    });
}

int mael_sparse_resample(mael_sparse* s, int32_t seeding, uint32_t seed, int64_t* total) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        int64_t t = s->b->resample(seeding != 0, seed);
        if (total) *total = t;
    });
}
int mael_sparse_p2g(mael_sparse* s, float dt) {
    if (!s || !(dt > 0.0f)) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->p2g(dt); });
}
int mael_sparse_g2p(mael_sparse* s, float dt, float gravity) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->g2p(dt, gravity); });
}

void mael_sparse_defaults(mael_sparse_opts* o) {
    if (!o) return;
    o->tol = 5e-4; o->maxiter = 240; o->sweeps = 1; o->coarse = 40; o->restart = 8; o->policy = 2; o->precision = 0;
    o->boost = 1.5f; o->refine = 0;
}
int mael_sparse_project(mael_sparse* s, float gamma, const mael_sparse_opts* opts, mael_sparse_stats* stats) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        mael::SparseOpts o = sparse_opts(opts);
        mael::SparseStats st;
        s->b->project(gamma, o, st);
        if (stats) {
            stats->iterations = st.iterations; stats->fcg_iterations = st.fcg_iterations;
            stats->fgmres_iterations = st.fgmres_iterations; stats->used_fgmres = st.used_fgmres;
            stats->rel_residual = st.rel_residual; stats->ms = st.ms;
        }
    });
}

int mael_sparse_g2p_adjoint(mael_sparse* s, float dt) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->g2p_adjoint(dt); });
}
int mael_sparse_p2g_adjoint(mael_sparse* s, float dt) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->p2g_adjoint(dt); });
}
int mael_sparse_project_adjoint(mael_sparse* s, const mael_sparse_opts* opts, mael_sparse_stats* stats) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        mael::SparseStats st;
        s->b->project_adjoint(sparse_opts(opts), st);
        if (stats) {
            stats->iterations = st.iterations; stats->fcg_iterations = st.fcg_iterations;
            stats->fgmres_iterations = st.fgmres_iterations; stats->used_fgmres = st.used_fgmres;
            stats->rel_residual = st.rel_residual; stats->ms = st.ms;
        }
    });
}

int mael_sparse_vcycle(mael_sparse* s, const mael_sparse_opts* opts, int32_t queued) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->precondition(sparse_opts(opts), queued != 0); });
}

int mael_sparse_init_coarse(mael_sparse* s, const float* eta, int32_t neta, float rmax, float gamma_circ, float core) {
    if (!s || !eta || neta < 2) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->init_coarse(eta, neta, rmax, gamma_circ, core); });
}
int mael_sparse_init_fine(mael_sparse* s, const float* eta, int32_t neta, float rmax, float gamma_circ, float core) {
    if (!s || !eta || neta < 2) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->init_fine(eta, neta, rmax, gamma_circ, core); });
}
int mael_sparse_classify(mael_sparse* s) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->classify(); });
}
int mael_sparse_solidify(mael_sparse* s) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->solidify(); });
}
int mael_sparse_redistance(mael_sparse* s, int32_t iterations) {
    if (!s || iterations < 0) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->redistance(iterations); });
}
int mael_sparse_restrict_velocity(mael_sparse* s) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->restrict_velocity(); });
}
int mael_sparse_keep(mael_sparse* s) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->keep(); });
}
int mael_sparse_fill_flags(mael_sparse* s, uint8_t value) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->fill_flags(value); });
}
int mael_sparse_volume(mael_sparse* s, double* volume) {
    if (!s || !volume) return MAEL_E_ARG;
    return guard(s->ctx, [&] { *volume = s->b->volume(); });
}
int mael_sparse_gravity(mael_sparse* s, float dt, float gravity) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->gravity(dt, gravity); });
}
int mael_sparse_sync(mael_sparse* s) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->st.sync(); });
}
int mael_sparse_ktimer(mael_sparse* s, int32_t kernel, int32_t launches) {
    if (!s || kernel < 0 || kernel >= mael::K_NKERNELS || launches < 1) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->kt.enable(kernel, (size_t)launches); });
}
int mael_sparse_ktime(mael_sparse* s, double* ms, int32_t* launches) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        int n = 0;
        double t = s->b->kt.read(&n);
        if (ms) *ms = t;
        if (launches) *launches = n;
    });
}

int mael_sparse_advect(mael_sparse* s, float dt, int32_t carry_reflected, float gravity, int32_t seed) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->advect(dt, carry_reflected != 0, gravity, seed); });
}
int mael_sparse_filltrapped(mael_sparse* s, int32_t maxrounds) {
    if (!s || maxrounds < 0) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->filltrapped(maxrounds); });
}
int mael_sparse_set_stir(mael_sparse* s, const mael_stir* p, const float* waves, int32_t nwaves, const float* heights,
                         int32_t nheights) {
    if (!s || !p || nwaves < 0 || nheights < 0 || (nwaves && !waves) || (nheights && !heights)) return MAEL_E_ARG;
    if (p->surface && nheights < 2) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        s->b->set_stir(mael::stir_params(*p), waves, nwaves, heights, nheights);
    });
}
int mael_sparse_pack_band(mael_sparse* s, float keep, float throat, uint16_t* pos, uint16_t* vel, int64_t cap,
                          int64_t* count) {
    if (!s || !pos || !vel || !count) return MAEL_E_ARG;
    return guard(s->ctx, [&] {
        if (cap < s->b->n) throw mael::Error(mael::E_ARG, "the packing buffers hold fewer than the particles");
        *count = s->b->pack_band(keep, throat, pos, vel);
    });
}
int mael_sparse_quantize_levelset(mael_sparse* s, float range, uint8_t* out) {
    if (!s || !out || !(range > 0.0f)) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->quantize_levelset(range, out); });
}
int mael_sparse_forces(mael_sparse* s, float dt, float time, float gravity) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->forces(dt, time, gravity); });
}
int mael_sparse_save_reflect(mael_sparse* s) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->save_reflect(); });
}
int mael_sparse_kick(mael_sparse* s) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->kick(); });
}
int mael_sparse_reflected(mael_sparse* s) {
    if (!s) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->reflected(); });
}
int mael_sparse_extrapolate(mael_sparse* s, int32_t layers, int32_t coarselayers, int32_t kick) {
    if (!s || layers < 0 || coarselayers < 0) return MAEL_E_ARG;
    return guard(s->ctx, [&] { s->b->extrapolate(layers, coarselayers, kick != 0); });
}
int mael_sparse_set_flow(mael_sparse* s, int32_t reflected) {
    if (!s) return MAEL_E_ARG;
    s->b->reflect_flow = reflected != 0;
    return MAEL_OK;
}
int mael_sparse_maxspeed(mael_sparse* s, float* speed) {
    if (!s || !speed) return MAEL_E_ARG;
    return guard(s->ctx, [&] { *speed = s->b->maxspeed(); });
}

}  // extern "C"
