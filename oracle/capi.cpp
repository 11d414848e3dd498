// TEST INFRASTRUCTURE ONLY: C entry points of the CPU oracle for the Python tests
// (ctypes). Nothing in the product links this library.
#include "dense.hpp"
#include "dense_step.hpp"
#include "whitewater.hpp"
#include <cstdio>
#include <string>

using namespace orc;

extern "C" {

struct orc_grid { float lo[3]; float dx; int32_t n[3]; float radius; float floor; int32_t levels; };
struct orc_particles { float* x[3]; float* v[3]; float* c[9]; float* lag; int64_t n; int64_t capacity; };

static thread_local std::string g_err;
const char* orc_last_error() { return g_err.c_str(); }
void orc_set_error(const char* msg) { g_err = msg; }

static DenseGrid to_grid(const orc_grid* g) {
    DenseGrid d;
    for (int a = 0; a < 3; ++a) { d.lo[a] = g->lo[a]; d.n[a] = g->n[a]; }
    d.dx = g->dx; d.radius = g->radius; d.floor = g->floor; d.levels = g->levels;
    return d;
}
static void load(Particles& ps, const orc_particles* p) {
    ps.resize(p->capacity);
    ps.n = p->n;
    for (int a = 0; a < 3; ++a) { std::copy(p->x[a], p->x[a] + p->n, ps.x[a].begin()); std::copy(p->v[a], p->v[a] + p->n, ps.v[a].begin()); }
    for (int a = 0; a < 9; ++a) std::copy(p->c[a], p->c[a] + p->n, ps.c[a].begin());
    std::copy(p->lag, p->lag + p->n, ps.lag.begin());
}
static void store(const Particles& ps, orc_particles* p) {
    p->n = ps.n;
    for (int a = 0; a < 3; ++a) { std::copy(ps.x[a].begin(), ps.x[a].begin() + ps.n, p->x[a]); std::copy(ps.v[a].begin(), ps.v[a].begin() + ps.n, p->v[a]); }
    for (int a = 0; a < 9; ++a) std::copy(ps.c[a].begin(), ps.c[a].begin() + ps.n, p->c[a]);
    std::copy(ps.lag.begin(), ps.lag.begin() + ps.n, p->lag);
}

int orc_flipgrid(float radius, float depth, float top, float dx, int32_t pad, orc_grid* out) {
    DenseGrid g = flipgrid(radius, depth, top, dx, pad);
    for (int a = 0; a < 3; ++a) { out->lo[a] = g.lo[a]; out->n[a] = g.n[a]; }
    out->dx = g.dx; out->radius = g.radius; out->floor = g.floor; out->levels = g.levels;
    return 0;
}

float orc_hash01(int32_t i, int32_t j) { return hash01(i, j); }
void orc_hash01_many(const int32_t* i, const int32_t* j, float* out, int64_t n) {
    for (int64_t k = 0; k < n; ++k) out[k] = hash01(i[k], j[k]);
}
void orc_bspline(float q, int32_t* b, float* w) { BS s = bspline(q); *b = s.b; w[0] = s.w[0]; w[1] = s.w[1]; w[2] = s.w[2]; }
int32_t orc_orderkey(float f) { return orderkey(f); }
float orc_orderfloat(int32_t k) { return orderfloat(k); }
float orc_temporal(float lag, float dt) { return temporal(lag, dt); }
float orc_strength(float s, float dt, float h) { return strength(s, dt, h); }
float orc_ahead(float g) { return ahead(g); }
float orc_oseen(float G, float core, float r) { return oseen(G, core, r); }
float orc_balanced(float G, float core, float g, float r) { return balanced(G, core, g, r); }
double orc_theta64(double a, double b) { return theta<double>(a, b); }
float orc_trilinear(const float* A, int32_t n1, int32_t n2, int32_t n3, float q0, float q1, float q2) {
    Arr3<const float> X{A, n1, n2, n3};
    return trilinear(X, q0, q1, q2);
}
void orc_gather(const float* u, const orc_grid* g, int32_t a, float px, float py, float pz, float* vel, float* aff) {
    DenseGrid G = to_grid(g);
    Arr3<const float> U{u, G.n[0] + (a == 0), G.n[1] + (a == 1), G.n[2] + (a == 2)};
    gather(U, px, py, pz, G, a, *vel, aff);
}

int orc_dense_resample(const orc_grid* g, const orc_particles* src, orc_particles* dst, const float* phi,
                       const float* u, const float* v, const float* w, int32_t band, int32_t nmin, int32_t seeding,
                       uint32_t seed, int32_t* key, int32_t* rank, int32_t* counts, int32_t* totals, int32_t* offsets,
                       int64_t* total) {
    try {
        DenseGrid G = to_grid(g);
        Particles s, d;
        load(s, src);
        d.resize(dst->capacity);
        ResampleOut o = dense_resample(G, s, d, phi, u, v, w, band, nmin, seeding != 0, seed);
        std::copy(o.key.begin(), o.key.end(), key);
        std::copy(o.rank.begin(), o.rank.end(), rank);
        std::copy(o.counts.begin(), o.counts.end(), counts);
        std::copy(o.totals.begin(), o.totals.end(), totals);
        std::copy(o.offsets.begin(), o.offsets.end(), offsets);
        *total = o.total;
        store(d, dst);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

// P2G in double (the FP64 oracle); mom/wt are the per-face accumulators (3 arrays each, double).
int orc_dense_p2g(const orc_grid* g, const orc_particles* ps, float dt, int32_t band, int32_t nmin, const float* phi0,
                  const float* u0, const float* v0, const float* w0, float* phi, float* u, float* v, float* w,
                  double* mu, double* mv, double* mw, double* wu, double* wv, double* ww) {
    DenseGrid G = to_grid(g);
    Particles s;
    load(s, ps);
    const float* U0[3] = {u0, v0, w0};
    float* U[3] = {u, v, w};
    std::vector<double> mom[3], wt[3];
    dense_p2g<double>(G, s, dt, band, nmin, phi0, U0, phi, U, mom, wt);
    double* M[3] = {mu, mv, mw};
    double* W[3] = {wu, wv, ww};
    for (int a = 0; a < 3; ++a) {
        if (M[a]) std::copy(mom[a].begin(), mom[a].end(), M[a]);
        if (W[a]) std::copy(wt[a].begin(), wt[a].end(), W[a]);
    }
    return 0;
}
// P2G accumulation only, in float (the FP32 CPU baseline's arithmetic).
int orc_dense_p2g_f32(const orc_grid* g, const orc_particles* ps, float dt, int32_t band, int32_t nmin, const float* phi0,
                      const float* u0, const float* v0, const float* w0, float* phi, float* u, float* v, float* w) {
    DenseGrid G = to_grid(g);
    Particles s;
    load(s, ps);
    const float* U0[3] = {u0, v0, w0};
    float* U[3] = {u, v, w};
    dense_p2g<float>(G, s, dt, band, nmin, phi0, U0, phi, U);
    return 0;
}

int orc_dense_g2p(const orc_grid* g, orc_particles* ps, const float* u, const float* v, const float* w,
                  const uint8_t* fu, const uint8_t* fv, const uint8_t* fw, float dt, float gravity) {
    DenseGrid G = to_grid(g);
    Particles s;
    load(s, ps);
    const float* U[3] = {u, v, w};
    const uint8_t* F[3] = {fu, fv, fw};
    dense_g2p(G, s, U, F, dt, gravity);
    store(s, ps);
    return 0;
}

// flip.jl:438-464 (the particles) and 356-369 (the fields) of advect!.
int orc_dense_advect_particles(const orc_grid* g, orc_particles* ps, const float* u, const float* v, const float* w,
                               const uint8_t* fu, const uint8_t* fv, const uint8_t* fw, float dt, float gravity, int32_t seed) {
    DenseGrid G = to_grid(g);
    Particles s;
    load(s, ps);
    const float* U[3] = {u, v, w};
    const uint8_t* F[3] = {fu, fv, fw};
    dense_advect_particles(DenseFlow(G, U, F), s, dt, gravity, seed);
    store(s, ps);
    return 0;
}
void orc_dense_advect_grid(const orc_grid* g, const float* u, const float* v, const float* w, const float* src, float* out,
                           int32_t n1, int32_t n2, int32_t n3, const float* off, float dt) {
    DenseGrid G = to_grid(g);
    const float* U[3] = {u, v, w};
    dense_advect_grid(DenseFlow(G, U, nullptr), out, src, n1, n2, n3, off, dt);
}
// flip.jl:1194-1206 extrapolate! of component `a`, in place on (u, f).
// flip.jl:672-679 redistance!(s; iterations), phi in place.
void orc_dense_redistance(const orc_grid* g, float* phi, int32_t iterations) { dense_redistance(to_grid(g), phi, iterations); }
void orc_dense_extrapolate(const orc_grid* g, int32_t a, float* u, uint8_t* f, int32_t layers) {
    dense_extrapolate(to_grid(g), a, u, f, layers);
}
// flip.jl:786-798 forces! with the Stir model (basin.jl:153-181); params: radius, band, relax,
// core, gamma, current x, y, eddy, eddyscale, eddytime, guide, free, turbulence, boil, reach,
// origin x, y, offset.
void orc_dense_forces(const orc_grid* g, const float* params, const float* waves, int32_t nwaves, float* u, float* v,
                      float* w, float dt, float t, float gravity) {
    StirModel S;
    const float* q = params;
    S.radius = q[0]; S.band = q[1]; S.relax = q[2]; S.core = q[3]; S.gamma = q[4];
    S.current[0] = q[5]; S.current[1] = q[6]; S.eddy = q[7]; S.eddyscale = q[8]; S.eddytime = q[9]; S.guide = q[10];
    S.free = q[11]; S.turbulence = q[12]; S.boil = q[13]; S.reach = q[14]; S.origin[0] = q[15]; S.origin[1] = q[16];
    S.offset = q[17];
    S.waves = waves;
    S.nwaves = nwaves;
    float* U[3] = {u, v, w};
    dense_forces(to_grid(g), S, U, dt, t, gravity);
}
// basin.jl:153-181 at one point: acc[3] for velocity vel[3] (params as orc_dense_forces).
void orc_stir(const float* params, const float* waves, int32_t nwaves, float x, float y, float z, const float* vel, float t,
              float* acc) {
    StirModel S;
    const float* q = params;
    S.radius = q[0]; S.band = q[1]; S.relax = q[2]; S.core = q[3]; S.gamma = q[4];
    S.current[0] = q[5]; S.current[1] = q[6]; S.eddy = q[7]; S.eddyscale = q[8]; S.eddytime = q[9]; S.guide = q[10];
    S.free = q[11]; S.turbulence = q[12]; S.boil = q[13]; S.reach = q[14]; S.origin[0] = q[15]; S.origin[1] = q[16];
    S.offset = q[17];
    S.waves = waves;
    S.nwaves = nwaves;
    stir(S, x, y, z, vel, t, acc);
}

// The whitewater's rules from a float array: air 2, crest 2, energy 2, kair, kcrest, life 2,
// buoyancy, drag, kick, tossed, within, floor, gravity, stream, gamma, core.
WhiteRules white_rules(const float* q) {
    WhiteRules R;
    R.air[0] = q[0]; R.air[1] = q[1]; R.crest[0] = q[2]; R.crest[1] = q[3]; R.energy[0] = q[4]; R.energy[1] = q[5];
    R.kair = q[6]; R.kcrest = q[7]; R.life[0] = q[8]; R.life[1] = q[9]; R.buoyancy = q[10]; R.drag = q[11];
    R.kick = q[12]; R.tossed = q[13]; R.within = q[14]; R.floor = q[15]; R.gravity = q[16]; R.stream = (int)q[17];
    R.gamma = q[18]; R.core = q[19];
    return R;
}
WhiteBuf white_buf(float* wx, float* wy, float* wz, float* wvx, float* wvy, float* wvz, float* life, uint8_t* kind, int64_t cap,
                   int64_t* hf) {
    WhiteBuf W{{wx, wy, wz}, {wvx, wvy, wvz}, life, kind, cap, hf[0], hf[1]};
    return W;
}
// flip.jl:1496-1511 whitewater! on the dense basin; hf = (head, frames) in and out.
void orc_dense_whitewater(const orc_grid* g, const orc_particles* ps, const int32_t* offsets, const int32_t* totals,
                          const float* phi, const float* u, const float* v, const float* w, const float* rules, float* wx,
                          float* wy, float* wz, float* wvx, float* wvy, float* wvz, float* life, uint8_t* kind, int64_t cap,
                          int64_t* hf, double dt, int32_t substeps, float near, float above, float below) {
    WhiteBuf W = white_buf(wx, wy, wz, wvx, wvy, wvz, life, kind, cap, hf);
    const float* X[3] = {ps->x[0], ps->x[1], ps->x[2]};
    const float* V[3] = {ps->v[0], ps->v[1], ps->v[2]};
    const float* U[3] = {u, v, w};
    dense_whitewater(W, to_grid(g), X, V, ps->n, offsets, totals, phi, U, white_rules(rules), dt, substeps, near, above, below);
    hf[0] = W.head;
    hf[1] = W.frames;
}

void orc_dense_classify(const orc_grid* g, const float* phi, uint8_t* types) { dense_classify(to_grid(g), phi, types); }
double orc_dense_volume(const orc_grid* g, const float* phi) { return dense_volume(to_grid(g), phi); }

void orc_dense_rhs(const orc_grid* g, const float* u, const float* v, const float* w, const uint8_t* types, float gamma,
                   double* b) {
    dense_rhs<double>(to_grid(g), u, v, w, types, gamma, b);
}

// A x on the fine level (FP64).
void orc_dense_apply(const orc_grid* g, const uint8_t* types, const float* phi, const double* x, double* q) {
    DensePressure<double> S;
    S.g = to_grid(g);
    S.setup(types, phi);
    S.apply(0, x, q);
}
// z = M^-1 r (one V-cycle, FP64).
void orc_dense_vcycle(const orc_grid* g, const uint8_t* types, const float* phi, const double* r, double* z,
                      int32_t sweeps, int32_t coarse) {
    DensePressure<double> S;
    S.g = to_grid(g);
    S.setup(types, phi);
    S.vcycle(z, r, sweeps, coarse);
}
// The same V-cycle in FP32 arithmetic, in the reference's order of operations.
void orc_dense_vcycle_f32(const orc_grid* g, const uint8_t* types, const float* phi, const float* r, float* z,
                          int32_t sweeps, int32_t coarse) {
    DensePressure<float> S;
    S.g = to_grid(g);
    S.setup(types, phi);
    S.vcycle(z, r, sweeps, coarse);
}
// The fine level's operator as COO triplets (rows, cols, vals); returns nnz. Pass null to count.
int64_t orc_dense_matrix(const orc_grid* g, const uint8_t* types, const float* phi, int64_t* rows, int64_t* cols,
                         double* vals) {
    DensePressure<double> S;
    S.g = to_grid(g);
    S.setup(types, phi);
    const int* n = S.g.n;
    int64_t nnz = 0;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                size_t c = (size_t)i + (size_t)n[0] * (j + (size_t)n[1] * k);
                if (types[c] != LIQUID) continue;
                if (rows) { rows[nnz] = c; cols[nnz] = c; vals[nnz] = S.diag[c]; }
                ++nnz;
                for (int m = 0; m < 6; ++m) {
                    int a = i + NB[m][0], b = j + NB[m][1], cc = k + NB[m][2];
                    if (typeat(types, n, a, b, cc) != LIQUID) continue;
                    if (rows) { rows[nnz] = c; cols[nnz] = (size_t)a + (size_t)n[0] * (b + (size_t)n[1] * cc); vals[nnz] = -1.0; }
                    ++nnz;
                }
            }
    return nnz;
}

// pressure! (FP64 MG-PCG): P in/out, b out. Returns iterations.
int32_t orc_dense_pressure(const orc_grid* g, const uint8_t* types, const float* phi, const double* b, double* P,
                           double tol, int32_t maxiter, int32_t sweeps, int32_t coarse, double* relres) {
    DensePressure<double> S;
    S.g = to_grid(g);
    S.setup(types, phi);
    size_t N = S.g.cells();
    std::vector<double> bb(b, b + N), PP(P, P + N);
    int it = dense_pcg(S, bb, PP, tol, maxiter, sweeps, coarse, relres);
    std::copy(PP.begin(), PP.end(), P);
    return it;
}
// FP32 MG-PCG (the CPU port's arithmetic), for the CPU baseline.
int32_t orc_dense_pressure_f32(const orc_grid* g, const uint8_t* types, const float* phi, const float* b, float* P,
                               double tol, int32_t maxiter, int32_t sweeps, int32_t coarse, double* relres) {
    DensePressure<float> S;
    S.g = to_grid(g);
    S.setup(types, phi);
    size_t N = S.g.cells();
    std::vector<float> bb(b, b + N), PP(P, P + N);
    int it = dense_pcg(S, bb, PP, tol, maxiter, sweeps, coarse, relres);
    std::copy(PP.begin(), PP.end(), P);
    return it;
}

void orc_dense_project_faces(const orc_grid* g, const double* P, const uint8_t* types, const float* phi, float* u,
                             float* v, float* w, uint8_t* fu, uint8_t* fv, uint8_t* fw) {
    float* U[3] = {u, v, w};
    uint8_t* F[3] = {fu, fv, fw};
    dense_project_faces<double>(to_grid(g), P, types, phi, U, F);
}

}  // extern "C"
