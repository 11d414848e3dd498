// TEST INFRASTRUCTURE ONLY: C entry points of the sparse basin's CPU oracle for the
// Python tests (ctypes). A handle holds one basin's state; arrays are read and written
// through pointers into it. Nothing in the product links this library.
#include "sparse_state.hpp"
#include "whitewater.hpp"
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

using namespace orc;

extern "C" {

struct orc_sgrid { float lo[3]; float dx; int32_t nb[3]; float radius; float floor; };

const char* orc_last_error();
void orc_set_error(const char* msg);

struct OrcSparse {
    SparseState s;
    Field<double> P64;   // orc_sparse_solve64's pressure, kept between calls
};

static void grid_out(const SparseGrid& g, orc_sgrid* o) {
    for (int a = 0; a < 3; ++a) { o->lo[a] = g.lo[a]; o->nb[a] = g.nb[a]; }
    o->dx = g.dx; o->radius = g.radius; o->floor = g.floor;
}

int orc_sparsegrid(float radius, float depth, float top, float dx, int32_t pad, orc_sgrid* out) {
    grid_out(sparsegrid(radius, depth, top, dx, pad), out);
    return 0;
}

void* orc_sparse_create(float radius, float depth, float top, float dx, int32_t pad, int32_t blocks, int64_t capacity,
                        int32_t band, int32_t nmin) {
    auto* h = new OrcSparse;
    h->s.create(sparsegrid(radius, depth, top, dx, pad), blocks, capacity, band, nmin);
    return h;
}
void orc_sparse_destroy(void* p) { delete (OrcSparse*)p; }

void orc_sparse_info(void* p, orc_sgrid* g, int32_t* nslots, int32_t* nfree, int64_t* n, int32_t* blocks) {
    auto& s = ((OrcSparse*)p)->s;
    if (g) grid_out(s.g(), g);
    if (nslots) *nslots = s.t.nslots;
    if (nfree) *nfree = (int32_t)s.t.freel.size();
    if (n) *n = s.ps[s.cur].n;
    if (blocks) *blocks = s.blocks;
}
void orc_sparse_free_list(void* p, int32_t* out) {
    auto& s = ((OrcSparse*)p)->s;
    std::copy(s.t.freel.begin(), s.t.freel.end(), out);
}
void orc_sparse_set_n(void* p, int64_t n) { auto& s = ((OrcSparse*)p)->s; s.ps[s.cur].n = n; }
// The slot count and free list after `slot` and `where` were written through their arrays.
void orc_sparse_set_topology_counts(void* p, int32_t nslots, const int32_t* freelist, int32_t nfree) {
    auto& s = ((OrcSparse*)p)->s;
    s.t.nslots = nslots;
    s.t.freel.assign(freelist, freelist + nfree);
}

// A named array of the state: its address and element count (element type by name).
int orc_sparse_array(void* p, const char* name, void** ptr, int64_t* len) {
    auto& s = ((OrcSparse*)p)->s;
    std::string n(name);
    auto F = [&](auto& v) { *ptr = (void*)v.data(); *len = (int64_t)v.size(); return 0; };
    static const char* ax[3] = {"u", "v", "w"};
    for (int a = 0; a < 3; ++a) {
        std::string b = ax[a];
        if (n == b + "_f") return F(s.u[a].fine);
        if (n == b + "_c") return F(s.u[a].coarse);
        if (n == b + "0_f") return F(s.u0[a].fine);
        if (n == b + "0_c") return F(s.u0[a].coarse);
        if (n == "f" + b + "_f") return F(s.fl[a].fine);
        if (n == "f" + b + "_c") return F(s.fl[a].coarse);
    }
    if (n == "phi_f") return F(s.phi.fine);
    if (n == "phi_c") return F(s.phi.coarse);
    if (n == "phi0_f") return F(s.phi0.fine);
    if (n == "phi0_c") return F(s.phi0.coarse);
    if (n == "P_f") return F(s.P.fine);
    if (n == "P_c") return F(s.P.coarse);
    if (n == "types_f") return F(s.types.fine);
    if (n == "types_c") return F(s.types.coarse);
    if (n == "slot") return F(s.t.slot);
    if (n == "where") { *ptr = (void*)s.t.where.data(); *len = (int64_t)s.t.where.size() * 3; return 0; }
    Particles& ps = s.ps[s.cur];
    static const char* pn[16] = {"px", "py", "pz", "vx", "vy", "vz", "c0", "c1", "c2", "c3", "c4", "c5", "c6", "c7", "c8", "lag"};
    for (int k = 0; k < 16; ++k)
        if (n == pn[k]) {
            if (k < 3) return F(ps.x[k]);
            if (k < 6) return F(ps.v[k - 3]);
            if (k < 15) return F(ps.c[k - 6]);
            return F(ps.lag);
        }
    orc_set_error(("no array " + n).c_str());
    return -1;
}

// sflip.jl:516-526 the marks (surface, flow, particles, `rings` dilations), into `out` (nb).
void orc_sparse_marks(void* p, float vary, float reach, int32_t rings, int32_t particles, uint8_t* out) {
    auto& s = ((OrcSparse*)p)->s;
    std::vector<uint8_t> m(s.g().nblocks(), 0);
    marksurface(s.g(), s.phi.coarse.data(), m.data());
    markflow(s.g(), s.u[0].coarse.data(), s.u[1].coarse.data(), s.u[2].coarse.data(), s.phi.coarse.data(), vary, m.data());
    const Particles& ps = s.ps[s.cur];
    if (particles && ps.n > 0) markparticles(s.g(), ps.x[0].data(), ps.x[1].data(), ps.x[2].data(), ps.n, reach, m.data());
    for (int r = 0; r < rings; ++r) m = dilate(s.g(), m);
    std::copy(m.begin(), m.end(), out);
}

int orc_sparse_refine(void* p, const uint8_t* want, int32_t* added, int32_t* nadd, int32_t* removed, int32_t* nrem) {
    auto& s = ((OrcSparse*)p)->s;
    try {
        std::vector<uint8_t> w(want, want + s.g().nblocks());
        std::vector<int32_t> a, r;
        s.refine(w, a, r);
        if (added) std::copy(a.begin(), a.end(), added);
        if (removed) std::copy(r.begin(), r.end(), removed);
        *nadd = (int32_t)a.size(); *nrem = (int32_t)r.size();
        return 0;
    } catch (const std::exception& e) {
        orc_set_error(e.what());
        return -3;
    }
}

void orc_sparse_init_coarse(void* p, const float* eta, int32_t neta, float rmax, float G, float core) {
    ((OrcSparse*)p)->s.init_coarse(eta, neta, rmax, G, core);
}
void orc_sparse_init_fine(void* p, const float* eta, int32_t neta, float rmax, float G, float core) {
    ((OrcSparse*)p)->s.init_fine(eta, neta, rmax, G, core);
}
// sflip.jl:2427-2432 sparsequantize_kernel!: the level set at the centres of cells `cell`
// wide, clamped to +-range, as bytes (round to nearest even); out is n1 x n2 x n3, x fastest.
void orc_sparse_quantize(void* p, float cell, float range, int32_t n1, int32_t n2, int32_t n3, uint8_t* out) {
    auto& s = ((OrcSparse*)p)->s;
    for (int k = 0; k < n3; ++k)
        for (int j = 0; j < n2; ++j)
            for (int i = 0; i < n1; ++i) {
                float px = s.t.g.lo[0] + ((float)(i + 1) - 0.5f) * cell, py = s.t.g.lo[1] + ((float)(j + 1) - 0.5f) * cell,
                      pz = s.t.g.lo[2] + ((float)(k + 1) - 0.5f) * cell;
                float f = jclamp(cellsample(s.t, s.phi.fine.data(), s.phi.coarse.data(), px, py, pz) / range, -1.0f, 1.0f);
                out[(size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k)] = (uint8_t)(int)std::nearbyint(127.5f * (f + 1.0f));
            }
}

WhiteRules white_rules(const float* q);
WhiteBuf white_buf(float* wx, float* wy, float* wz, float* wvx, float* wvy, float* wvz, float* life, uint8_t* kind, int64_t cap,
                   int64_t* hf);
// sflip.jl:2401-2417 whitewater! on the sparse basin's current state (its particles, φ and
// u, v, w); `offsets`/`totals` of the resample that ordered the particles.
void orc_sparse_whitewater(void* p, const int32_t* offsets, const int32_t* totals, const float* rules, float* wx, float* wy,
                           float* wz, float* wvx, float* wvy, float* wvz, float* life, uint8_t* kind, int64_t cap, int64_t* hf,
                           double dt, int32_t substeps, float near, float above, float below) {
    auto& s = ((OrcSparse*)p)->s;
    WhiteBuf W = white_buf(wx, wy, wz, wvx, wvy, wvz, life, kind, cap, hf);
    const Particles& ps = s.ps[s.cur];
    const float* X[3] = {ps.x[0].data(), ps.x[1].data(), ps.x[2].data()};
    const float* V[3] = {ps.v[0].data(), ps.v[1].data(), ps.v[2].data()};
    sparse_whitewater(W, s.t, X, V, ps.n, offsets, totals, s.phi, s.u, white_rules(rules), dt, substeps, near, above, below);
    hf[0] = W.head;
    hf[1] = W.frames;
}

void orc_sparse_classify(void* p) { ((OrcSparse*)p)->s.classify(); }
void orc_sparse_solidify(void* p) { ((OrcSparse*)p)->s.solidify(); }
void orc_sparse_restrict_velocity(void* p) { ((OrcSparse*)p)->s.restrict_velocity(); }
// sflip.jl:583-594 / 596-612 on one named field: "phi", "P", "types" (cells); "u", "v", "w",
// "fu", "fv", "fw" (faces along their axis). The coarse part under the blocks is rewritten.
int orc_sparse_restrict_field(void* p, const char* name) {
    auto& s = ((OrcSparse*)p)->s;
    std::string n(name);
    if (n == "phi") { restrictcells(s.t, s.phi.coarse.data(), (const float*)s.phi.fine.data()); return 0; }
    if (n == "P") { restrictcells(s.t, s.P.coarse.data(), (const float*)s.P.fine.data()); return 0; }
    if (n == "types") { restrictcells(s.t, s.types.coarse.data(), (const uint8_t*)s.types.fine.data()); return 0; }
    static const char* ax[3] = {"u", "v", "w"};
    for (int a = 0; a < 3; ++a) {
        if (n == ax[a]) { restrictfaces(s.t, s.u[a].coarse.data(), (const float*)s.u[a].fine.data(), a); return 0; }
        if (n == std::string("f") + ax[a]) {
            restrictfaces(s.t, s.fl[a].coarse.data(), (const uint8_t*)s.fl[a].fine.data(), a);
            return 0;
        }
    }
    orc_set_error(("no field " + n).c_str());
    return -1;
}
void orc_sparse_redistance(void* p, int32_t iterations) { ((OrcSparse*)p)->s.redistance(iterations); }
double orc_sparse_volume(void* p) { return ((OrcSparse*)p)->s.volume(); }
void orc_sparse_fill_flags(void* p, uint8_t v) {
    auto& s = ((OrcSparse*)p)->s;
    for (int a = 0; a < 3; ++a) {
        std::fill(s.fl[a].fine.begin(), s.fl[a].fine.end(), v);
        std::fill(s.fl[a].coarse.begin(), s.fl[a].coarse.end(), v);
    }
}
// φ0 = φ, u0 = u (sflip.jl:700-703).
void orc_sparse_keep(void* p) {
    auto& s = ((OrcSparse*)p)->s;
    s.phi0 = s.phi;
    for (int a = 0; a < 3; ++a) s.u0[a] = s.u[a];
}

// sflip.jl:2227-2246 resample!, flow = (u, v, w). Optional outputs sized n / nslots*4096.
int orc_sparse_resample(void* p, int32_t seeding, uint32_t seed, int64_t* total, int32_t* key, int32_t* rank, int32_t* counts,
                        int32_t* totals, int32_t* offsets) {
    auto& s = ((OrcSparse*)p)->s;
    try {
        const Particles& src = s.ps[s.cur];
        Particles& dst = s.ps[1 - s.cur];
        const float* flow[3] = {s.u[0].fine.data(), s.u[1].fine.data(), s.u[2].fine.data()};
        ResampleOut o = sparse_resample(s.t, src, dst, SFc{s.phi.fine.data(), s.phi.coarse.data()}, flow, s.band, s.nmin,
                                        seeding != 0, seed);
        if (key) std::copy(o.key.begin(), o.key.end(), key);
        if (rank) std::copy(o.rank.begin(), o.rank.end(), rank);
        if (counts) std::copy(o.counts.begin(), o.counts.end(), counts);
        if (totals) std::copy(o.totals.begin(), o.totals.end(), totals);
        if (offsets) std::copy(o.offsets.begin(), o.offsets.end(), offsets);
        *total = o.total;
        s.cur = 1 - s.cur;
        return 0;
    } catch (const std::exception& e) {
        orc_set_error(e.what());
        return -2;
    }
}

// sflip.jl:808-831 particles2grid! (accumulation in FP64, or FP32 with fp32 != 0);
// `momwt` (optional) receives mu, mv, mw, wu, wv, ww over nslots*4352 each.
void orc_sparse_p2g(void* p, float dt, int32_t fp32, double* momwt) {
    auto& s = ((OrcSparse*)p)->s;
    SFc u0[3] = {{s.u0[0].fine.data(), s.u0[0].coarse.data()}, {s.u0[1].fine.data(), s.u0[1].coarse.data()},
                 {s.u0[2].fine.data(), s.u0[2].coarse.data()}};
    SF u[3] = {{s.u[0].fine.data(), s.u[0].coarse.data()}, {s.u[1].fine.data(), s.u[1].coarse.data()},
               {s.u[2].fine.data(), s.u[2].coarse.data()}};
    SFc phi0{s.phi0.fine.data(), s.phi0.coarse.data()};
    SF phi{s.phi.fine.data(), s.phi.coarse.data()};
    size_t Fn = (size_t)s.t.nslots * BFACES;
    auto out = [&](auto* mom, auto* wt) {
        if (!momwt) return;
        for (int a = 0; a < 3; ++a)
            for (size_t f = 0; f < Fn; ++f) { momwt[a * Fn + f] = (double)mom[a][f]; momwt[(3 + a) * Fn + f] = (double)wt[a][f]; }
    };
    if (fp32) {
        std::vector<float> mom[3], wt[3];
        sparse_p2g<float>(s.t, s.ps[s.cur], dt, s.band, s.nmin, phi0, u0, phi, u, mom, wt);
        out(mom, wt);
    } else {
        std::vector<double> mom[3], wt[3];
        sparse_p2g<double>(s.t, s.ps[s.cur], dt, s.band, s.nmin, phi0, u0, phi, u, mom, wt);
        out(mom, wt);
    }
}

void orc_sparse_g2p(void* p, float dt, float gravity) {
    auto& s = ((OrcSparse*)p)->s;
    const float* u[3] = {s.u[0].fine.data(), s.u[1].fine.data(), s.u[2].fine.data()};
    const uint8_t* f[3] = {s.fl[0].fine.data(), s.fl[1].fine.data(), s.fl[2].fine.data()};
    sparse_g2p(s.t, s.ps[s.cur], u, f, dt, gravity);
}

// ─── pressure ────────────────────────────────────────────────────────────────

static SparsePressure<double>* setup(SparseState& s) {
    auto* S = new SparsePressure<double>(s.t, s.types.fine.data(), s.types.coarse.data());
    S->coefficients(s.phi.fine.data(), s.phi.coarse.data());
    S->hierarchy();
    return S;
}

static void split(const SparsePressure<double>& S, const double* f, const double* c, Field<double>& x) {
    x.fine.assign(f, f + S.nf());
    x.coarse.assign(c, c + S.g.ncoarse());
}

// sflip.jl:1408-1435 the right-hand side (FP64), fine over nslots*4096.
void orc_sparse_rhs(void* p, float gamma, double* bf, double* bc) {
    auto& s = ((OrcSparse*)p)->s;
    SFc u[3] = {{s.u[0].fine.data(), s.u[0].coarse.data()}, {s.u[1].fine.data(), s.u[1].coarse.data()},
                {s.u[2].fine.data(), s.u[2].coarse.data()}};
    Field<double> b;
    sparse_rhs<double>(s.t, s.types.fine.data(), s.types.coarse.data(), u, gamma, b);
    std::copy(b.fine.begin(), b.fine.end(), bf);
    std::copy(b.coarse.begin(), b.coarse.end(), bc);
}

// The composite operator of level 0 (N = 16) or 1 (N = 8) as COO; call with rows == null for nnz.
int64_t orc_sparse_matrix(void* p, int32_t N, int64_t* rows, int64_t* cols, double* vals) {
    auto& s = ((OrcSparse*)p)->s;
    std::unique_ptr<SparsePressure<double>> S(setup(s));
    std::vector<int64_t> r, c;
    std::vector<double> v;
    S->assemble(N, r, c, v);
    if (rows) {
        std::copy(r.begin(), r.end(), rows);
        std::copy(c.begin(), c.end(), cols);
        std::copy(v.begin(), v.end(), vals);
    }
    return (int64_t)r.size();
}

void orc_sparse_apply(void* p, const double* xf, const double* xc, double* qf, double* qc) {
    auto& s = ((OrcSparse*)p)->s;
    std::unique_ptr<SparsePressure<double>> S(setup(s));
    Field<double> x, q = S->field();
    split(*S, xf, xc, x);
    S->apply(q, x);
    std::copy(q.fine.begin(), q.fine.end(), qf);
    std::copy(q.coarse.begin(), q.coarse.end(), qc);
}

void orc_sparse_vcycle(void* p, const double* rf, const double* rc, double* zf, double* zc, int32_t sweeps, int32_t coarse,
                       double boost) {
    auto& s = ((OrcSparse*)p)->s;
    std::unique_ptr<SparsePressure<double>> S(setup(s));
    Field<double> r, z = S->field(), q = S->field(), tmp = S->field();
    split(*S, rf, rc, r);
    S->vcycle(z, r, q, tmp, sweeps, coarse, boost);
    std::copy(z.fine.begin(), z.fine.end(), zf);
    std::copy(z.coarse.begin(), z.coarse.end(), zc);
}

// The hierarchy's dense levels: count, then per level its dims (3 ints each).
int32_t orc_sparse_levels(void* p, int32_t* dims) {
    auto& s = ((OrcSparse*)p)->s;
    SparsePressure<double> S(s.t, s.types.fine.data(), s.types.coarse.data());
    if (dims)
        for (size_t l = 0; l < S.D.size(); ++l)
            for (int a = 0; a < 3; ++a) dims[3 * l + a] = S.D[l].n[a];
    return (int32_t)S.D.size();
}

}  // extern "C"

template <class T>
static int project_t(SparseState& s, float gamma, double tol, int maxiter, int sweeps, int coarse, double boost, int restart,
                     int policy, int32_t* iters, double* relres, int32_t* used_fgmres) {
    SparsePressure<T> S(s.t, s.types.fine.data(), s.types.coarse.data());
    S.coefficients(s.phi.fine.data(), s.phi.coarse.data());
    SFc u[3] = {{s.u[0].fine.data(), s.u[0].coarse.data()}, {s.u[1].fine.data(), s.u[1].coarse.data()},
                {s.u[2].fine.data(), s.u[2].coarse.data()}};
    Field<T> b, P = S.field();
    sparse_rhs<T>(s.t, s.types.fine.data(), s.types.coarse.data(), u, gamma, b);
    for (size_t n = 0; n < S.nf(); ++n)                                          // sflip.jl:1964-1973 clear
        P.fine[n] = (s.live((int)(n >> 12) + 1) && s.types.fine[n] != LIQUID) ? T(0) : (T)s.P.fine[n];
    for (int K = 0; K < s.nc[2]; ++K)
        for (int J = 0; J < s.nc[1]; ++J)
            for (int I = 0; I < s.nc[0]; ++I) {
                size_t ci = cidx(s.nc, I, J, K);
                P.coarse[ci] = (s.types.coarse[ci] != LIQUID || s.t.covered(I, J, K)) ? T(0) : (T)s.P.coarse[ci];
            }
    S.hierarchy();
    int it = 0;
    bool fg = false;
    double rr = 0;
    if (policy == 0) it = S.fcg(P, b, tol, maxiter, sweeps, (T)boost, coarse, &rr);
    else if (policy == 1) { it = S.fgmres(P, b, tol, maxiter, restart, sweeps, (T)boost, coarse, &rr); fg = true; }
    else it = S.policy(P, b, tol, maxiter, restart, sweeps, (T)boost, coarse, &rr, &fg);
    // the true residual of what is returned
    Field<T> q = S.field(), r = S.field();
    S.apply(q, P);
    r = b;
    S.axpy(r, T(-1), q);
    double bn = std::sqrt((double)S.inner(b, b)), rn = std::sqrt((double)S.inner(r, r));
    rr = bn > 0 ? rn / bn : 0.0;
    for (size_t n = 0; n < S.nf(); ++n) s.P.fine[n] = (float)P.fine[n];
    for (size_t c = 0; c < P.coarse.size(); ++c) s.P.coarse[c] = (float)P.coarse[c];
    SF uf[3] = {{s.u[0].fine.data(), s.u[0].coarse.data()}, {s.u[1].fine.data(), s.u[1].coarse.data()},
                {s.u[2].fine.data(), s.u[2].coarse.data()}};
    uint8_t* vf[3] = {s.fl[0].fine.data(), s.fl[1].fine.data(), s.fl[2].fine.data()};
    uint8_t* vc[3] = {s.fl[0].coarse.data(), s.fl[1].coarse.data(), s.fl[2].coarse.data()};
    sparse_project_faces<T>(S, P, s.phi.fine.data(), s.phi.coarse.data(), uf, vf, vc);
    if (iters) *iters = it;
    if (relres) *relres = rr;
    if (used_fgmres) *used_fgmres = fg ? 1 : 0;
    return 0;
}

extern "C" {

// sflip.jl:1998-2015 project! with the given γ: coefficients, rhs, clear, the solve
// (policy 0: flexible CG; 1: FGMRES; 2: CG then FGMRES, fgmres.jl:96-110), the faces.
// The FP64 pressure solve alone (sflip.jl:1867-1904 pressure! with the rhs and clearing of
// project!, sflip.jl:1998-2013), continuable: with `cont` it starts from the FP64 pressure
// the previous call left (not the FP32 state's), so a long solve can run in chunks of
// `maxiter` iterations with the host reporting between them. The velocities are not
// projected; s.P receives the pressure rounded to FP32. Returns the true relative residual.
int orc_sparse_solve64(void* p, float gamma, double tol, int32_t maxiter, int32_t sweeps, int32_t coarse, double boost,
                       int32_t cont, int32_t* iters, double* relres) {
    auto& h = *(OrcSparse*)p;
    auto& s = h.s;
    try {
        SparsePressure<double> S(s.t, s.types.fine.data(), s.types.coarse.data());
        S.coefficients(s.phi.fine.data(), s.phi.coarse.data());
        SFc u[3] = {{s.u[0].fine.data(), s.u[0].coarse.data()}, {s.u[1].fine.data(), s.u[1].coarse.data()},
                    {s.u[2].fine.data(), s.u[2].coarse.data()}};
        Field<double> b, P = S.field();
        sparse_rhs<double>(s.t, s.types.fine.data(), s.types.coarse.data(), u, gamma, b);
        const bool warm = cont && h.P64.fine.size() == P.fine.size() && h.P64.coarse.size() == P.coarse.size();
        for (size_t n = 0; n < S.nf(); ++n)
            P.fine[n] = (s.live((int)(n >> 12) + 1) && s.types.fine[n] != LIQUID) ? 0.0
                                                                                   : (warm ? h.P64.fine[n] : (double)s.P.fine[n]);
        for (int K = 0; K < s.nc[2]; ++K)
            for (int J = 0; J < s.nc[1]; ++J)
                for (int I = 0; I < s.nc[0]; ++I) {
                    size_t ci = cidx(s.nc, I, J, K);
                    P.coarse[ci] = (s.types.coarse[ci] != LIQUID || s.t.covered(I, J, K))
                                       ? 0.0
                                       : (warm ? h.P64.coarse[ci] : (double)s.P.coarse[ci]);
                }
        S.hierarchy();
        double rr = 0;
        int it = S.fcg(P, b, tol, maxiter, sweeps, boost, coarse, &rr);
        Field<double> q = S.field(), r = S.field();
        S.apply(q, P);
        r = b;
        S.axpy(r, -1.0, q);
        double bn = std::sqrt(S.inner(b, b)), rn = std::sqrt(S.inner(r, r));
        h.P64 = P;
        for (size_t n = 0; n < S.nf(); ++n) s.P.fine[n] = (float)P.fine[n];
        for (size_t c = 0; c < P.coarse.size(); ++c) s.P.coarse[c] = (float)P.coarse[c];
        if (iters) *iters = it;
        if (relres) *relres = bn > 0 ? rn / bn : 0.0;
        return 0;
    } catch (const SolveError& e) {
        orc_set_error(e.what());
        return -5;
    } catch (const std::exception& e) {
        orc_set_error(e.what());
        return -1;
    }
}
// The FP64 pressure of orc_sparse_solve64 (fine pools, then the coarse grid), as doubles.
int orc_sparse_p64(void* p, double* fine, double* coarse) {
    auto& h = *(OrcSparse*)p;
    if (h.P64.fine.empty()) return -1;
    std::copy(h.P64.fine.begin(), h.P64.fine.end(), fine);
    std::copy(h.P64.coarse.begin(), h.P64.coarse.end(), coarse);
    return 0;
}
int orc_sparse_project(void* p, float gamma, double tol, int32_t maxiter, int32_t sweeps, int32_t coarse, double boost,
                       int32_t restart, int32_t policy, int32_t fp32, int32_t* iters, double* relres, int32_t* used_fgmres) {
    auto& s = ((OrcSparse*)p)->s;
    try {
        return fp32 ? project_t<float>(s, gamma, tol, maxiter, sweeps, coarse, boost, restart, policy, iters, relres, used_fgmres)
                    : project_t<double>(s, gamma, tol, maxiter, sweeps, coarse, boost, restart, policy, iters, relres,
                                        used_fgmres);
    } catch (const SolveError& e) {
        orc_set_error(e.what());
        return -5;
    } catch (const std::exception& e) {
        orc_set_error(e.what());
        return -1;
    }
}

}  // extern "C"
