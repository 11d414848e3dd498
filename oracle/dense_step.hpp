// TEST INFRASTRUCTURE ONLY. Part of the CPU oracle (see dense.hpp): the dense basin's
// support phases around the hot path, restated from the reference so the library's can be
// checked: ST-FLIP particle advection and semi-Lagrangian grid advection
// (flip.jl:350-499), extrapolation (flip.jl:1158-1206) and the stirring forces
// (flip.jl:762-798 with basin.jl:105-181 and sea.jl:60-76).
#pragma once
#include "dense.hpp"

namespace orc {

// flip.jl:80-81 velocity and flip.jl:370-372 reached, on the dense grid.
struct DenseFlow {
    DenseGrid g;
    Arr3<const float> u[3];
    Arr3<const uint8_t> f[3];
    DenseFlow(const DenseGrid& g_, const float* const uu[3], const uint8_t* const ff[3]) : g(g_) {
        for (int a = 0; a < 3; ++a) {
            u[a] = {uu[a], g.n[0] + (a == 0), g.n[1] + (a == 1), g.n[2] + (a == 2)};
            f[a] = {ff ? ff[a] : nullptr, g.n[0] + (a == 0), g.n[1] + (a == 1), g.n[2] + (a == 2)};
        }
    }
    // flip.jl:77 indexspace: (p - lo) / dx - off
    void q(int a, float px, float py, float pz, float& q0, float& q1, float& q2) const {
        float off[3];
        faceoff(a, off);
        q0 = (px - g.lo[0]) / g.dx - off[0];
        q1 = (py - g.lo[1]) / g.dx - off[1];
        q2 = (pz - g.lo[2]) / g.dx - off[2];
    }
    void velocity(float px, float py, float pz, float v[3]) const {
        for (int a = 0; a < 3; ++a) {
            float q0, q1, q2;
            q(a, px, py, pz, q0, q1, q2);
            v[a] = trilinear(u[a], q0, q1, q2);
        }
    }
    bool reached(float px, float py, float pz) const {
        float m[3];
        for (int a = 0; a < 3; ++a) {
            float q0, q1, q2;
            q(a, px, py, pz, q0, q1, q2);
            m[a] = trilinear(f[a], q0, q1, q2);
        }
        return jmin(jmin(m[0], m[1]), m[2]) > 0.999f;
    }
};

// flip.jl:367 substeps.
inline int substeps(float cells) {
    float c = std::ceil(cells);
    if (std::isnan(c)) return 1;
    return c >= 32.0f ? 32 : (c <= 1.0f ? 1 : (int)c);
}
inline float norm3(const float v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// flip.jl:374-380 inwalls.
inline void inwalls(const DenseGrid& g, float& x, float& y, float& z) {
    float m = 0.25f * g.dx;
    float r = std::sqrt(x * x + y * y);
    if (r > g.radius - m) {
        x = x * (g.radius - m) / r;
        y = y * (g.radius - m) / r;
    }
    if (z < g.floor + m) z = g.floor + m;
}

// flip.jl:438-464 advectparticles_kernel! (`p` 1-based in the hash, as there).
inline void dense_advect_particles(const DenseFlow& fl, Particles& ps, float dt, float gravity, int32_t seed) {
    for (int64_t p = 0; p < ps.n; ++p) {
        float y[3] = {ps.x[0][p], ps.x[1][p], ps.x[2][p]};
        float vb[3] = {ps.v[0][p], ps.v[1][p], ps.v[2][p]};
        float h0 = fl.g.dx;
        float speed = norm3(vb);
        float gm = strength(speed, dt, h0);
        float xi = hash01((int32_t)(p + 1), seed) - 0.5f;
        float tau = ps.lag[p];
        float act = jclamp(dt + tau + gm * xi * dt, 0.0f, 2.0f * dt);
        ps.lag[p] = dt + tau - act;
        float vf[3];
        fl.velocity(y[0], y[1], y[2], vf);
        int m = substeps(jmax(speed, norm3(vf)) * act / h0);
        float h = act / (float)m;
        for (int s = 0; s < m; ++s) {
            if (fl.reached(y[0], y[1], y[2])) {
                float v1[3];
                fl.velocity(y[0], y[1], y[2], v1);
                float hh = 0.5f * h;
                fl.velocity(y[0] + hh * v1[0], y[1] + hh * v1[1], y[2] + hh * v1[2], vb);
                for (int a = 0; a < 3; ++a) y[a] = y[a] + h * vb[a];
            } else {
                y[0] = y[0] + h * (vb[0] - 0.0f);
                y[1] = y[1] + h * (vb[1] - 0.0f);
                y[2] = y[2] + h * (vb[2] - 0.5f * gravity * h);
                vb[2] = vb[2] - gravity * h;
            }
            inwalls(fl.g, y[0], y[1], y[2]);
        }
        for (int a = 0; a < 3; ++a) ps.x[a][p] = y[a];
    }
}

// flip.jl:356-369 advect_kernel!: out (shape n1 x n2 x n3, samples at `off`) = src traced
// back along the flow.
inline void dense_advect_grid(const DenseFlow& fl, float* out, const float* src, int n1, int n2, int n3, const float off[3],
                              float dt) {
    const DenseGrid& g = fl.g;
    Arr3<const float> S{src, n1, n2, n3};
    for (int k = 0; k < n3; ++k)
        for (int j = 0; j < n2; ++j)
            for (int i = 0; i < n1; ++i) {
                float p[3] = {g.lo[0] + ((float)i + off[0]) * g.dx, g.lo[1] + ((float)j + off[1]) * g.dx,
                              g.lo[2] + ((float)k + off[2]) * g.dx};
                float v[3];
                fl.velocity(p[0], p[1], p[2], v);
                int m = substeps(norm3(v) * dt / g.dx);
                float h = dt / (float)m;
                for (int s = 0; s < m; ++s) {
                    float v1[3], v2[3];
                    fl.velocity(p[0], p[1], p[2], v1);
                    float hh = 0.5f * h;
                    fl.velocity(p[0] - hh * v1[0], p[1] - hh * v1[1], p[2] - hh * v1[2], v2);
                    for (int a = 0; a < 3; ++a) p[a] = p[a] - h * v2[a];
                }
                out[(size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k)] =
                    trilinear(S, (p[0] - g.lo[0]) / g.dx - off[0], (p[1] - g.lo[1]) / g.dx - off[1],
                              (p[2] - g.lo[2]) / g.dx - off[2]);
            }
}

// flip.jl:624-679 reinit_kernel! and redistance!: `iterations` Jacobi sweeps of the level
// set in place. Cells with a neighbour of the other sign keep their value; the others take
// the upwind (Godunov) distance from the smallest neighbour magnitude along each axis, at
// most phimax = (iterations + 1) dx; a solid cell takes the lowest value of the non-solid
// cells next to it (or its own, clamped to +-phimax, when there is none).
inline void dense_redistance(const DenseGrid& g, float* phi, int iterations) {
    const int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2];
    const size_t N = (size_t)n1 * n2 * n3;
    const float h = g.dx, phimax = (float)(iterations + 1) * g.dx;
    auto at = [&](int i, int j, int k) { return (size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k); };
    auto open = [&](int i, int j, int k) {
        return i >= 0 && j >= 0 && k >= 0 && i < n1 && j < n2 && k < n3 && !issolidcell(g, i, j, k);
    };
    std::vector<float> out(N);
    for (int it = 0; it < iterations; ++it) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < n3; ++k)
            for (int j = 0; j < n2; ++j)
                for (int i = 0; i < n1; ++i) {
                    float c = phi[at(i, j, k)];
                    if (issolidcell(g, i, j, k)) {
                        float lowest = c;
                        bool found = false;
                        for (int m = 0; m < 6; ++m) {
                            int a = i + NB[m][0], b = j + NB[m][1], cc = k + NB[m][2];
                            if (!open(a, b, cc)) continue;
                            float f = phi[at(a, b, cc)];
                            lowest = found ? std::min(lowest, f) : f;
                            found = true;
                        }
                        out[at(i, j, k)] = found ? lowest : std::clamp(c, -phimax, phimax);
                        continue;
                    }
                    bool inside = c < 0.0f, frozen = false;
                    float ax = phimax, ay = phimax, az = phimax;
                    for (int m = 0; m < 6; ++m) {
                        int a = i + NB[m][0], b = j + NB[m][1], cc = k + NB[m][2];
                        if (!open(a, b, cc)) continue;
                        float f = phi[at(a, b, cc)];
                        if ((f < 0.0f) != inside) frozen = true;
                        float d = std::fabs(f);
                        if (m < 2) ax = std::min(ax, d); else if (m < 4) ay = std::min(ay, d); else az = std::min(az, d);
                    }
                    if (frozen) { out[at(i, j, k)] = c; continue; }
                    float d = std::min(eikonal(ax, ay, az, h), phimax);
                    out[at(i, j, k)] = inside ? -d : d;
                }
        std::copy(out.begin(), out.end(), phi);
    }
}

// flip.jl:779-784 walled: face (i, j, k) of axis a has a wall cell on a side or is the
// grid's own edge.
inline bool walled(const DenseGrid& g, int i, int j, int k, int a) {
    int b0 = i - (a == 0), b1 = j - (a == 1), b2 = k - (a == 2);
    if (!(b0 >= 0 && b1 >= 0 && b2 >= 0 && i < g.n[0] && j < g.n[1] && k < g.n[2])) return true;
    return issolidcell(g, b0, b1, b2) || issolidcell(g, i, j, k);
}

// flip.jl:1160-1206 extrapolate!: `layers` layers of extrapolation of component `a` in
// place (u, its flags f), then zero on faces not reached, then the walls (solid_kernel!).
inline void dense_extrapolate(const DenseGrid& g, int a, float* u, uint8_t* f, int layers) {
    const int n1 = g.n[0] + (a == 0), n2 = g.n[1] + (a == 1), n3 = g.n[2] + (a == 2);
    const size_t N = (size_t)n1 * n2 * n3;
    std::vector<float> tmp(u, u + N);
    std::vector<uint8_t> f2(N, 0);
    float* src = u;
    float* dst = tmp.data();
    uint8_t* fin = f;
    uint8_t* fout = f2.data();
    for (int layer = 0; layer < layers; ++layer) {
        for (int k = 0; k < n3; ++k)
            for (int j = 0; j < n2; ++j)
                for (int i = 0; i < n1; ++i) {
                    size_t l = (size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k);
                    if (fin[l] != 0) {
                        dst[l] = src[l];
                        fout[l] = 1;
                        continue;
                    }
                    float s = 0.0f;
                    int m = 0;
                    if (!walled(g, i, j, k, a))
                        for (int q = 0; q < 6; ++q) {
                            int x = i + NB[q][0], y = j + NB[q][1], z = k + NB[q][2];
                            if (x < 0 || y < 0 || z < 0 || x >= n1 || y >= n2 || z >= n3) continue;
                            size_t o = (size_t)x + (size_t)n1 * ((size_t)y + (size_t)n2 * z);
                            if (fin[o] != 0) {
                                s += src[o];
                                ++m;
                            }
                        }
                    dst[l] = m > 0 ? s / (float)m : src[l];
                    fout[l] = m > 0 ? 1 : 0;
                }
        std::swap(src, dst);
        std::swap(fin, fout);
    }
    if (src != u) std::copy(src, src + N, u);
    if (fin != f) std::copy(fin, fin + N, f);
    for (size_t l = 0; l < N; ++l)
        if (f[l] == 0) u[l] = 0.0f;
    for (int k = 0; k < n3; ++k)
        for (int j = 0; j < n2; ++j)
            for (int i = 0; i < n1; ++i)
                if (walled(g, i, j, k, a)) u[(size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k)] = 0.0f;
}

// ─── the stirring model (basin.jl:92-181) ─────────────────────────────────────

struct StirModel {
    float radius, band, relax, core, gamma;        // the basin's (basin.jl:13-30)
    float current[2], eddy, eddyscale, eddytime, guide, free, turbulence, boil, reach;   // the race's (:69-80)
    float origin[2], offset;                        // the sea's coordinates of the centre; film time offset
    const float* waves;                             // 7 floats a wave: d0 d1 k omega a q phase (a scaled)
    int nwaves;
};

// sea.jl:68-76 noise.
inline float noise(float x, float y) {
    float xi = std::floor(x), yi = std::floor(y);
    float fx = x - xi, fy = y - yi;
    float sx = fx * fx * (3.0f - 2.0f * fx), sy = fy * fy * (3.0f - 2.0f * fy);
    int32_t i = (int32_t)xi, j = (int32_t)yi;
    float a = hash01(i, j), b = hash01(i + 1, j), c = hash01(i, j + 1), d = hash01(i + 1, j + 1);
    return (a + (b - a) * sx) + ((c + (d - c) * sx) - (a + (b - a) * sx)) * sy;
}
// basin.jl:109-119 eddies.
inline void eddies(const StirModel& S, float x, float y, float t, float e[2]) {
    float s = S.eddyscale;
    float a = (x - S.current[0] * t) / s, b = (y - S.current[1] * t) / s;
    float tau = t / S.eddytime;
    auto psi = [&](float p, float q) { return noise(p + 17.0f, q + tau) + 0.5f * noise(2.1f * p - 5.0f, 2.1f * q - 1.7f * tau); };
    float d = 0.25f;
    float dpdy = (psi(a, b + d) - psi(a, b - d)) / (2.0f * d);
    float dpdx = (psi(a + d, b) - psi(a - d, b)) / (2.0f * d);
    e[0] = S.eddy * dpdy;
    e[1] = -S.eddy * dpdx;
}
// basin.jl:122-125 boils.
inline float boils(const StirModel& S, float x, float y, float t) {
    float n = noise(0.6f * (x - S.current[0] * t) / S.eddyscale + 41.0f,
                    0.6f * (y - S.current[1] * t) / S.eddyscale + 0.2f * t / S.eddytime);
    return S.boil * (std::max(n - 0.6f, 0.0f) / 0.4f - 0.15f);
}
// basin.jl:139-151 orbital.
inline void orbital(const StirModel& S, float x, float y, float z, float t, float u[2], float& w) {
    u[0] = u[1] = 0.0f;
    w = 0.0f;
    for (int k = 0; k < S.nwaves; ++k) {
        const float* W = S.waves + 7 * k;
        float th = W[2] * (W[0] * x + W[1] * y) - W[3] * t + W[6];
        float e = W[4] * W[3] * std::exp(W[2] * std::min(z, 0.0f));
        float s = std::sin(th), c = std::cos(th);
        u[0] += (e * s) * W[0];
        u[1] += (e * s) * W[1];
        w -= e * c;
    }
}
// basin.jl:153-181 (s::Stir)(coords, velocity, density, pressure, t).
inline void stir(const StirModel& S, float x, float y, float z, const float vel[3], float t, float acc[3]) {
    float r = std::sqrt(x * x + y * y);
    float k = jclamp((r - (S.radius - S.band)) / S.band, 0.0f, 1.0f);
    k = k * k * (3.0f - 2.0f * k);
    float v = oseen(S.gamma, S.core, r);
    float core = jclamp(2.0f - r / S.core, 0.0f, 1.0f);
    float hold = (1.0f - k) * jmax(jclamp((-z - S.free) / S.free, 0.0f, 1.0f), core);
    float churn = jclamp((r - 2.0f) / 2.0f, 0.0f, 1.0f);
    float turn = r > 1e-4f ? v / r : 0.0f;
    float e[2];
    eddies(S, x, y, t, e);
    float aim[2] = {-turn * y + S.current[0] + churn * e[0], turn * x + S.current[1] + churn * e[1]};
    float gh = hold / S.guide;
    float guide[2] = {gh * (aim[0] - vel[0]), gh * (aim[1] - vel[1])};
    float top = (1.0f - k) * churn * jclamp(1.0f + z / S.free, 0.0f, 1.0f);
    float push[2] = {0.0f, 0.0f};
    if (S.eddy > 0.0f) {
        float pf = top * S.turbulence / S.eddy;
        push[0] = pf * e[0];
        push[1] = pf * e[1];
    }
    float rising = (1.0f - k) * std::exp(std::min(z, 0.0f) / S.reach);
    acc[0] = guide[0] + push[0];
    acc[1] = guide[1] + push[1];
    acc[2] = rising * churn * boils(S, x, y, t);
    if (k == 0.0f) return;
    float sea[2], up;
    orbital(S, S.origin[0] + x, S.origin[1] + y, z, t - S.offset, sea, up);
    float target[3] = {-v * y / r + S.current[0] + sea[0], v * x / r + S.current[1] + sea[1], up};
    float kr = k / S.relax;
    for (int a = 0; a < 3; ++a) acc[a] = acc[a] + kr * (target[a] - vel[a]);
}

// flip.jl:764-798 forces!: u += dt (stir - g z) at every face out of the walls, the
// velocity read from a copy; then the walls. `u` are the three face arrays.
inline void dense_forces(const DenseGrid& g, const StirModel& S, float* u[3], float dt, float t, float gravity) {
    std::vector<float> old[3];
    for (int a = 0; a < 3; ++a) old[a].assign(u[a], u[a] + g.faces(a));
    const float* oldp[3] = {old[0].data(), old[1].data(), old[2].data()};
    DenseFlow fl(g, oldp, nullptr);
    for (int a = 0; a < 3; ++a) {
        const int n1 = g.n[0] + (a == 0), n2 = g.n[1] + (a == 1), n3 = g.n[2] + (a == 2);
        float off[3];
        faceoff(a, off);
        for (int k = 0; k < n3; ++k)
            for (int j = 0; j < n2; ++j)
                for (int i = 0; i < n1; ++i) {
                    float px = g.lo[0] + ((float)i + off[0]) * g.dx, py = g.lo[1] + ((float)j + off[1]) * g.dx,
                          pz = g.lo[2] + ((float)k + off[2]) * g.dx;
                    if (issolid(g.radius, g.floor, px, py, pz)) continue;
                    float vel[3], acc[3];
                    fl.velocity(px, py, pz, vel);
                    stir(S, px, py, pz, vel, t, acc);
                    float aa = a == 2 ? acc[2] - gravity : acc[a];
                    u[a][(size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k)] += dt * aa;
                }
    }
    for (int a = 0; a < 3; ++a) {
        const int n1 = g.n[0] + (a == 0), n2 = g.n[1] + (a == 1), n3 = g.n[2] + (a == 2);
        for (int k = 0; k < n3; ++k)
            for (int j = 0; j < n2; ++j)
                for (int i = 0; i < n1; ++i)
                    if (walled(g, i, j, k, a)) u[a][(size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k)] = 0.0f;
    }
}

}  // namespace orc
