// TEST INFRASTRUCTURE ONLY. Part of the CPU oracle (see dense.hpp): whitewater from the
// narrow band (flip.jl:1411-1511, sflip.jl:2320-2417, whitewater.jl:11-240), restated so
// the library's can be checked. The fluid particles are read through the cell runs of
// the resampled order; new diffuse particles go into the ring buffer in particle order
// from the head (the reference takes slots with an atomic counter; particle order is
// its result when the atomics resolve in thread order).
#pragma once
#include "dense.hpp"
#include "sparse.hpp"

namespace orc {

// whitewater.jl:30-44 WhitewaterRules and the whitewater's floor, gravity and stream.
struct WhiteRules {
    float air[2], crest[2], energy[2], kair, kcrest, life[2], buoyancy, drag, kick, tossed, within, floor, gravity;
    int stream;
    float gamma, core;
};
struct WhiteBuf {   // the ring buffer (host copies)
    float* x[3];
    float* v[3];
    float* life;
    uint8_t* kind;
    int64_t capacity;
    int64_t head, frames;
};

inline float clampfn(float I, const float r[2]) { return (std::min(I, r[1]) - std::min(I, r[0])) / (r[1] - r[0]); }
inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline float sumsq3(const float a[3]) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

// whitewater.jl:306-310 carried.
inline void carried(const WhiteRules& R, const float x[3], float c[3]) {
    c[0] = c[1] = c[2] = 0.0f;
    if (!R.stream) return;
    float r = std::sqrt(x[0] * x[0] + x[1] * x[1]);
    float turn = r > 1e-4f ? oseen(R.gamma, R.core, r) / r : 0.0f;
    c[0] = -turn * x[1];
    c[1] = turn * x[0];
}

// whitewater.jl:131-200 emit_kernel! over fluid particles (SoA x[3], v[3], normals nrm, n
// of them); `cellof` gives a particle's cell coordinates, `range` a cell's run [s, e).
template <class CellOf, class Range>
inline void emit(WhiteBuf& W, const float* const x[3], const float* const v[3], const std::vector<float>& nrm, int64_t n,
                 float h, const WhiteRules& R, float dt, CellOf cellof, Range range) {
    W.frames += 1;
    const int32_t seed = (int32_t)(uint32_t)W.frames;
    for (int64_t i = 0; i < n; ++i) {
        const float xi[3] = {x[0][i], x[1][i], x[2][i]}, vi[3] = {v[0][i], v[1][i], v[2][i]};
        const float ni[3] = {nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]};
        if (xi[2] < R.floor + 2.0f * h || sumsq3(ni) == 0.0f || xi[0] * xi[0] + xi[1] * xi[1] > R.within * R.within)
            continue;
        float c[3];
        carried(R, xi, c);
        const float vr[3] = {vi[0] - c[0], vi[1] - c[1], vi[2] - c[2]};
        const float rs = std::sqrt(sumsq3(vr));
        const float energy = clampfn(0.5f * rs * rs, R.energy);
        const int32_t i1 = (int32_t)(i + 1);
        const float draw = hash01(i1, seed);
        float upper = energy * R.kair * dt + energy * R.kcrest * dt;
        if (0.0f <= upper && upper < 1.0f) {
            uint32_t bits;
            std::memcpy(&bits, &upper, 4);
            bits += 8u;
            std::memcpy(&upper, &bits, 4);
            if (draw >= upper) continue;
        }
        int cc[3];
        cellof(xi, cc);
        float air = 0.0f, crest = 0.0f;
        for (int dk = -1; dk <= 1; ++dk)
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    int64_t s, e;
                    if (!range(cc[0] + di, cc[1] + dj, cc[2] + dk, s, e)) continue;
                    for (int64_t q = s; q < e; ++q) {
                        const float dx[3] = {xi[0] - x[0][q], xi[1] - x[1][q], xi[2] - x[2][q]};
                        const float d = std::sqrt(sumsq3(dx));
                        if (0.0f < d && d < h) {
                            const float w = 1.0f - d / h;
                            const float vij[3] = {vi[0] - v[0][q], vi[1] - v[1][q], vi[2] - v[2][q]};
                            const float xij[3] = {dx[0] / d, dx[1] / d, dx[2] / d};
                            const float lv = std::sqrt(sumsq3(vij));
                            if (lv > 0.0f) {
                                const float vh[3] = {vij[0] / lv, vij[1] / lv, vij[2] / lv};
                                air += lv * (1.0f - dot3(vh, xij)) * w;
                            }
                            const float nq[3] = {nrm[3 * q], nrm[3 * q + 1], nrm[3 * q + 2]};
                            const float mx[3] = {-xij[0], -xij[1], -xij[2]};
                            if (sumsq3(ni) > 0.0f && sumsq3(nq) > 0.0f && dot3(mx, ni) < 0.0f) crest += (1.0f - dot3(ni, nq)) * w;
                        }
                    }
                }
        bool facing = false;
        if (rs > 0.0f && sumsq3(ni) > 0.0f) {
            const float vh[3] = {vr[0] / rs, vr[1] / rs, vr[2] / rs};
            facing = dot3(vh, ni) >= 0.6f;
        }
        const float aired = energy * R.kair * clampfn(air, R.air) * dt;
        const float crested = energy * R.kcrest * (facing ? clampfn(crest, R.crest) : 0.0f) * dt;
        const float tot = aired + crested;
        int32_t whole = (int32_t)std::floor(tot);
        whole += draw < tot - (float)whole ? 1 : 0;
        const float ln = std::sqrt(sumsq3(ni));
        const float nh[3] = {ni[0] / ln, ni[1] / ln, ni[2] / ln};
        for (int32_t m = 1; m <= whole; ++m) {
            W.head += 1;
            const int64_t slot = W.head % W.capacity;
            const float a = hash01((int32_t)((uint32_t)i1 + (uint32_t)m * 7919u), seed + 31);
            const float bb = hash01((int32_t)((uint32_t)i1 - (uint32_t)m * 104729u), seed + 77);
            const float c3 = hash01(m, (int32_t)((uint32_t)i1 + (uint32_t)seed));
            const float jh = 0.5f * h;
            const float jit[3] = {(a - 0.5f) * jh, (bb - 0.5f) * jh, (c3 - 0.5f) * jh};
            const float u = hash01((int32_t)((uint32_t)i1 + (uint32_t)m * 613u), seed + 5);
            const bool thrown = R.kick > 0.0f && hash01((int32_t)((uint32_t)i1 - (uint32_t)m * 331u), seed + 9) * tot < R.tossed * crested;
            for (int k = 0; k < 3; ++k) {
                if (thrown) {
                    W.x[k][slot] = xi[k] + h * nh[k] + 0.5f * jit[k];
                    W.v[k][slot] = vi[k] + (R.kick * energy * (0.5f + u)) * nh[k];
                } else {
                    W.x[k][slot] = xi[k] + jit[k];
                    W.v[k][slot] = vi[k];
                }
            }
            W.life[slot] = R.life[0] + (R.life[1] - R.life[0]) * hash01(seed - m, i1);
        }
    }
}

// flip.jl:1456-1483 leveldiffuse_kernel! / sflip.jl:2365-2391 sparsediffuse_kernel!.
template <class Phi, class Vel>
inline void diffuse(WhiteBuf& W, Phi phi, Vel vel, float radius, float floorz, float dt, const WhiteRules& R, float above,
                    float below) {
    for (int64_t i = 0; i < W.capacity; ++i) {
        float life = W.life[i];
        if (!(life > 0.0f)) continue;
        float x[3] = {W.x[0][i], W.x[1][i], W.x[2][i]}, v[3] = {W.v[0][i], W.v[1][i], W.v[2][i]};
        const float f = phi(x);
        float vf[3];
        vel(x, vf);
        if (f > above) {
            v[2] = v[2] + (-R.gravity * dt);
            W.kind[i] = 1;
        } else if (f < -below) {
            for (int k = 0; k < 3; ++k) {
                const float up = k == 2 ? dt * (R.buoyancy * R.gravity) : dt * 0.0f;
                v[k] = (v[k] + up) + R.drag * (vf[k] - v[k]);
            }
            life -= dt;
            W.kind[i] = 3;
        } else {
            for (int k = 0; k < 3; ++k) v[k] = vf[k];
            life -= dt;
            W.kind[i] = 2;
        }
        for (int k = 0; k < 3; ++k) x[k] = x[k] + dt * v[k];
        const float reach = jmin(R.within, radius);
        if (x[2] < floorz || x[0] * x[0] + x[1] * x[1] > reach * reach) life = 0.0f;
        for (int k = 0; k < 3; ++k) {
            W.x[k][i] = x[k];
            W.v[k][i] = v[k];
        }
        W.life[i] = life;
    }
}

// flip.jl:1496-1511 whitewater!(w, s::FlipBasin, ...): `offsets`/`totals` of the last
// resample, the particles in its order.
inline void dense_whitewater(WhiteBuf& W, const DenseGrid& g, const float* const x[3], const float* const v[3], int64_t n,
                             const int32_t* offsets, const int32_t* totals, const float* phi, const float* const u[3],
                             const WhiteRules& R, double dt, int substeps, float near, float above, float below) {
    Arr3<const float> P{phi, g.n[0], g.n[1], g.n[2]};
    auto qs = [&](float p, int a) { return (p - g.lo[a]) / g.dx - 0.5f; };
    std::vector<float> nrm(3 * (size_t)std::max<int64_t>(n, 1), 0.0f);
    const float nr = near * g.dx;
    for (int64_t p = 0; p < n; ++p) {   // flip.jl:1433-1447 levelnormals_kernel!
        float q0 = qs(x[0][p], 0), q1 = qs(x[1][p], 1), q2 = qs(x[2][p], 2);
        if (trilinear(P, q0, q1, q2) > -nr) {
            float gx = trilinear(P, q0 + 0.5f, q1, q2) - trilinear(P, q0 - 0.5f, q1, q2);
            float gy = trilinear(P, q0, q1 + 0.5f, q2) - trilinear(P, q0, q1 - 0.5f, q2);
            float gz = trilinear(P, q0, q1, q2 + 0.5f) - trilinear(P, q0, q1, q2 - 0.5f);
            float l = std::sqrt(gx * gx + gy * gy + gz * gz);
            if (l > 0.0f) { nrm[3 * p] = gx / l; nrm[3 * p + 1] = gy / l; nrm[3 * p + 2] = gz / l; }
        }
    }
    auto cellof = [&](const float xp[3], int c[3]) {   // interpolation.jl:20-25 cellcoords
        for (int a = 0; a < 3; ++a) c[a] = std::clamp((int)std::floor((xp[a] - g.lo[a]) / g.dx), 0, g.n[a] - 1);
    };
    auto range = [&](int i, int j, int k, int64_t& s, int64_t& e) {
        if (i < 0 || j < 0 || k < 0 || i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) return false;
        size_t l = (size_t)i + (size_t)g.n[0] * ((size_t)j + (size_t)g.n[1] * k);
        e = offsets[l];
        s = e - totals[l];
        return true;
    };
    emit(W, x, v, nrm, n, g.dx, R, (float)dt, cellof, range);
    Arr3<const float> U[3] = {{u[0], g.n[0] + 1, g.n[1], g.n[2]}, {u[1], g.n[0], g.n[1] + 1, g.n[2]},
                              {u[2], g.n[0], g.n[1], g.n[2] + 1}};
    auto phis = [&](const float xp[3]) { return trilinear(P, qs(xp[0], 0), qs(xp[1], 1), qs(xp[2], 2)); };
    auto vels = [&](const float xp[3], float vf[3]) {
        for (int a = 0; a < 3; ++a) {
            float off[3];
            faceoff(a, off);
            vf[a] = trilinear(U[a], (xp[0] - g.lo[0]) / g.dx - off[0], (xp[1] - g.lo[1]) / g.dx - off[1],
                              (xp[2] - g.lo[2]) / g.dx - off[2]);
        }
    };
    const float hs = (float)(dt / substeps);
    for (int k = 0; k < substeps; ++k) diffuse(W, phis, vels, g.radius, g.floor, hs, R, above * g.dx, below * g.dx);
}

// sflip.jl:2401-2417 whitewater!(w, s::SparseBasin, ...).
inline void sparse_whitewater(WhiteBuf& W, const Topo& t, const float* const x[3], const float* const v[3], int64_t n,
                              const int32_t* offsets, const int32_t* totals, const Field<float>& phi, const Field<float>* u,
                              const WhiteRules& R, double dt, int substeps, float near, float above, float below) {
    const SparseGrid& g = t.g;
    const float* pf = phi.fine.data();
    const float* pc = phi.coarse.data();
    std::vector<float> nrm(3 * (size_t)std::max<int64_t>(n, 1), 0.0f);
    const float nr = near * g.dx, hh = 0.5f * g.dx;
    for (int64_t p = 0; p < n; ++p) {   // sflip.jl:2347-2362 sparsenormals_kernel!
        const float X = x[0][p], Y = x[1][p], Z = x[2][p];
        if (cellsample(t, pf, pc, X, Y, Z) > -nr) {
            float gx = cellsample(t, pf, pc, X + hh, Y, Z) - cellsample(t, pf, pc, X - hh, Y, Z);
            float gy = cellsample(t, pf, pc, X, Y + hh, Z) - cellsample(t, pf, pc, X, Y - hh, Z);
            float gz = cellsample(t, pf, pc, X, Y, Z + hh) - cellsample(t, pf, pc, X, Y, Z - hh);
            float l = std::sqrt(gx * gx + gy * gy + gz * gz);
            if (l > 0.0f) { nrm[3 * p] = gx / l; nrm[3 * p + 1] = gy / l; nrm[3 * p + 2] = gz / l; }
        }
    }
    auto cellof = [&](const float xp[3], int c[3]) {   // sflip.jl:2334-2337
        for (int a = 0; a < 3; ++a) c[a] = (int)std::floor((xp[a] - g.lo[a]) / g.dx);
    };
    auto range = [&](int i, int j, int k, int64_t& s, int64_t& e) {
        int64_t idx = t.fineat(i, j, k);
        if (idx < 0) return false;
        e = offsets[idx];
        s = e - totals[idx];
        return true;
    };
    emit(W, x, v, nrm, n, g.dx, R, (float)dt, cellof, range);
    auto phis = [&](const float xp[3]) { return cellsample(t, pf, pc, xp[0], xp[1], xp[2]); };
    auto vels = [&](const float xp[3], float vf[3]) {
        for (int a = 0; a < 3; ++a) vf[a] = facesample(t, u[a].fine.data(), u[a].coarse.data(), a, xp[0], xp[1], xp[2]);
    };
    const float hs = (float)(dt / substeps);
    for (int k = 0; k < substeps; ++k) diffuse(W, phis, vels, g.radius, g.floor, hs, R, above * g.dx, below * g.dx);
}

}  // namespace orc
