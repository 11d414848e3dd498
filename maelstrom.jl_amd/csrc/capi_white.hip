// The C ABI of the whitewater (include/maelstrom_hip.h, `mael_white_*`, and the basins'
// `whitewater!`).
#include "capi_common.hpp"
#include "whitewater.hpp"
#include <memory>

struct mael_white {
    mael_ctx* ctx;
    std::unique_ptr<mael::Whitewater> w;
};

namespace {
ArrRef white_array(mael::Whitewater& w, int which) {
    switch (which) {
        case MAEL_W_X: case MAEL_W_Y: case MAEL_W_Z: return {w.x[which - MAEL_W_X].p, w.capacity, 4};
        case MAEL_W_VX: case MAEL_W_VY: case MAEL_W_VZ: return {w.v[which - MAEL_W_VX].p, w.capacity, 4};
        case MAEL_W_LIFE: return {w.life.p, w.capacity, 4};
        case MAEL_W_KIND: return {w.kind.p, w.capacity, 1};
        default: break;
    }
    throw mael::Error(mael::E_ARG, "unknown whitewater array " + std::to_string(which));
}
}  // namespace

extern "C" {

int mael_white_create(mael_ctx* ctx, int64_t capacity, const mael_white_rules* r, mael_white** out) {
    if (!ctx || !r || !out || capacity < 1) return MAEL_E_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        mael::WhiteParams P;
        for (int k = 0; k < 2; ++k) {
            P.air[k] = r->air[k]; P.crest[k] = r->crest[k]; P.energy[k] = r->energy[k]; P.life[k] = r->life[k];
        }
        P.kair = r->kair; P.kcrest = r->kcrest; P.spray = r->spray; P.bubble = r->bubble;
        P.buoyancy = r->buoyancy; P.drag = r->drag; P.kick = r->kick; P.tossed = r->tossed; P.within = r->within;
        P.floor = r->floor; P.gravity = r->gravity; P.stream = r->stream; P.gamma = r->gamma_circ; P.core = r->core;
        auto* h = new mael_white{ctx, std::unique_ptr<mael::Whitewater>(new mael::Whitewater(capacity, P))};
        *out = h;
    });
}
void mael_white_destroy(mael_white* w) { delete w; }

int mael_white_array(mael_white* w, int32_t which, void** dptr, int64_t* count, int32_t* elsize) {
    if (!w) return MAEL_E_ARG;
    return guard(w->ctx, [&] {
        ArrRef a = white_array(*w->w, which);
        if (dptr) *dptr = a.p;
        if (count) *count = a.n;
        if (elsize) *elsize = a.es;
    });
}
int mael_white_upload(mael_white* w, int32_t which, const void* host, int64_t count) {
    if (!w || (!host && count)) return MAEL_E_ARG;
    return guard(w->ctx, [&] {
        ArrRef a = white_array(*w->w, which);
        if (count < 0 || count > a.n) throw mael::Error(mael::E_ARG, "count exceeds the array");
        MAEL_HIP(hipMemcpyAsync(a.p, host, (size_t)count * a.es, hipMemcpyHostToDevice, w->w->st.s));
        w->w->st.sync();
    });
}
int mael_white_download(mael_white* w, int32_t which, void* host, int64_t count) {
    if (!w || (!host && count)) return MAEL_E_ARG;
    return guard(w->ctx, [&] {
        ArrRef a = white_array(*w->w, which);
        if (count < 0 || count > a.n) throw mael::Error(mael::E_ARG, "count exceeds the array");
        MAEL_HIP(hipMemcpyAsync(host, a.p, (size_t)count * a.es, hipMemcpyDeviceToHost, w->w->st.s));
        w->w->st.sync();
    });
}
int mael_white_state(mael_white* w, int64_t* head, int64_t* frames) {
    if (!w) return MAEL_E_ARG;
    if (head) *head = w->w->head;
    if (frames) *frames = w->w->frames;
    return MAEL_OK;
}
int mael_white_set_state(mael_white* w, int64_t head, int64_t frames) {
    if (!w || head < 0 || frames < 0) return MAEL_E_ARG;
    w->w->head = head;
    w->w->frames = frames;
    return MAEL_OK;
}
int mael_white_pack(mael_white* w, uint16_t* out, uint16_t* vel, int64_t cap, int64_t* count) {
    if (!w || !out || !vel || !count) return MAEL_E_ARG;
    return guard(w->ctx, [&] {
        if (cap < w->w->capacity) throw mael::Error(mael::E_ARG, "the packing buffers hold fewer than the capacity");
        *count = w->w->pack(out, vel);
    });
}
int mael_dense_whitewater(mael_dense* d, mael_white* w, double dt, int32_t substeps, float near, float above, float below) {
    if (!d || !w || substeps < 1 || !(dt >= 0.0)) return MAEL_E_ARG;
    return guard(d->ctx, [&] { mael::dense_whitewater(*d->b, *w->w, dt, substeps, near, above, below); });
}
int mael_sparse_whitewater(mael_sparse* s, mael_white* w, double dt, int32_t substeps, float near, float above, float below) {
    if (!s || !w || substeps < 1 || !(dt >= 0.0)) return MAEL_E_ARG;
    return guard(s->ctx, [&] { mael::sparse_whitewater(*s->b, *w->w, dt, substeps, near, above, below); });
}

}  // extern "C"
