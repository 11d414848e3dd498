// TEST INFRASTRUCTURE ONLY. This is the CPU oracle of the hot path: a plain C++
// restatement of the reference's algorithm, used by tests/, the smoke check and the
// CPU-baseline leg of bench.py as the checker. The product (the HIP library) never
// links, loads or calls it.
//
// Dense `FlipBasin` path: grid sizing, the per-particle helpers, resampling (bin,
// totals, scan, scatter, spawn), APIC particle-to-grid and grid-to-particle, and the
// multigrid-preconditioned CG pressure projection. Indices are 0-based here; the
// reference is 1-based Julia, and each function cites the lines it restates.
//
// Float operation order follows the reference expression by expression and the file
// is compiled with -ffp-contract=off, so FP32 results of the integer/byte pipeline
// (keys, counts, offsets, spawned positions) are reproducible bit for bit.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <stdexcept>
#include <string>

namespace orc {

constexpr uint8_t SOLID = 0, LIQUID = 1, AIR = 2;   // flip.jl:24-26

// ─── scalar helpers ──────────────────────────────────────────────────────────

// Julia's min/max/clamp on Float32 (NaN-propagating min/max; clamp keeps NaN).
inline float jmin(float x, float y) {
    if (std::isnan(x) || std::isnan(y)) return NAN;
    if (y < x) return y;
    if (x < y) return x;
    return (std::signbit(y) && !std::signbit(x)) ? y : x;
}
inline float jmax(float x, float y) {
    if (std::isnan(x) || std::isnan(y)) return NAN;
    if (y > x) return y;
    if (x > y) return x;
    return (std::signbit(x) && !std::signbit(y)) ? y : x;
}
inline float jclamp(float x, float lo, float hi) { return x > hi ? hi : (x < lo ? lo : x); }

// sea.jl:61-66 hash01: 32-bit mix, Float32(k) / 2^32 (range [0, 1], 1.0 for k >= 0xFFFFFF80).
inline float hash01(int32_t i, int32_t j) {
    uint32_t k = ((uint32_t)i * 0x27d4eb2du) ^ ((uint32_t)j * 0x165667b1u);
    k = (k ^ (k >> 15)) * 0x85ebca6bu;
    k = (k ^ (k >> 13)) * 0xc2b2ae35u;
    return (float)(k ^ (k >> 16)) / 4.2949673e9f;
}

// flip.jl:99-105 bspline: first node floor(q - 1/2) and the quadratic B-spline weights.
struct BS { int b; float w[3]; };
inline BS bspline(float q) {
    float b = std::floor(q - 0.5f);
    float f = q - b;
    BS r;
    r.b = (int)b;
    float a0 = 1.5f - f, a1 = f - 1.0f, a2 = f - 0.5f;
    r.w[0] = 0.5f * (a0 * a0);
    r.w[1] = 0.75f - a1 * a1;
    r.w[2] = 0.5f * (a2 * a2);
    return r;
}

// flip.jl:107-110 orderkey/orderfloat: order-preserving Float32 <-> Int32.
inline int32_t orderkey(float f) {
    int32_t b; std::memcpy(&b, &f, 4);
    return b >= 0 ? b : (b ^ 0x7fffffff);
}
inline float orderfloat(int32_t k) {
    int32_t b = k >= 0 ? k : (k ^ 0x7fffffff);
    float f; std::memcpy(&f, &b, 4);
    return f;
}

// flip.jl:112-113 theta: ghost-fluid fraction.
template <class T> inline T theta(T pi, T pj) {
    T t = pi / (pi - pj);
    return t > T(1) ? T(1) : (t < T(0.01f) ? T(0.01f) : t);
}

// flip.jl:503-511 temporal: ST-FLIP's poly6 weight in time.
inline float temporal(float lag, float dt) {
    float r = jclamp(-lag / dt, -0.5f, 0.5f) - 0.5f;
    float a = 1.0f - r * r;
    return (35.0f / 16.0f) * (a * a * a);
}
// flip.jl:466-470 strength.
inline float strength(float speed, float dt, float h) {
    float c = jclamp(speed * dt / h, 0.0f, 1.0f);
    return c * c * (3.0f - 2.0f * c);
}
// flip.jl:484-487 ahead.
inline float ahead(float g) {
    float h2 = 0.25f * g * g;
    float num = h2 * (9.0f / 16 - h2 * (7.0f / 10 - h2 * 3.0f / 7));
    float den = 27.0f / 64 + h2 * (3.0f / 16 - h2 * (3.0f / 20 + h2 / 7.0f));
    return num / den;
}

// sea.jl:39 oseen (Lamb-Oseen swirl speed) and sea.jl:48-58 balanced (free surface).
inline float oseen(float G, float core, float r) {
    float rr = std::max(r, 1e-4f);
    float q = r / core;
    return G / (2.0f * (float)M_PI * rr) * (1.0f - std::exp(-(q * q)));
}
inline float balanced(float G, float core, float g, float r, int steps = 2000) {
    float a = std::log(std::max(r, 1e-3f));
    float z = std::log(50.0f * core + 50.0f * std::max(r, 1.0f));
    float h = (z - a) / (float)steps;
    float s = 0.0f;
    for (int i = 0; i <= steps; ++i) {
        float rho = std::exp(a + (float)i * h);
        float o = oseen(G, core, rho);
        s += ((i == 0 || i == steps) ? 0.5f : 1.0f) * (o * o) / g;
    }
    return -s * h;
}

// ─── the dense grid ──────────────────────────────────────────────────────────

// flip.jl:30-62 FlipGrid.
struct DenseGrid {
    float lo[3];
    float dx;
    int n[3];
    float radius, floor;
    int levels;
    size_t cells() const { return (size_t)n[0] * n[1] * n[2]; }
    size_t faces(int a) const {
        return (size_t)(n[0] + (a == 0)) * (n[1] + (a == 1)) * (n[2] + (a == 2));
    }
};

inline DenseGrid flipgrid(float radius, float depth, float top, float dx, int pad = 2) {
    int nxy = (int)std::ceil(2.0f * (radius + (float)pad * dx) / dx);
    float loz = -depth - dx;
    int nz0 = (int)std::ceil((top - loz) / dx);
    for (int levels = 2;; ++levels) {
        int m = 1 << (levels - 1);
        int nx = m * ((nxy + m - 1) / m), nz = m * ((nz0 + m - 1) / m);
        if ((long)(nx / m) * (nx / m) * (nz / m) <= 1024) {
            DenseGrid g;
            float half = (float)(-nx) * dx / 2.0f;
            g.lo[0] = half; g.lo[1] = half; g.lo[2] = loz;
            g.dx = dx; g.n[0] = nx; g.n[1] = nx; g.n[2] = nz;
            g.radius = radius; g.floor = -depth; g.levels = levels;
            return g;
        }
    }
}

// flip.jl:611-622 eikonal.
inline float eikonal(float a, float b, float c, float h) {
    float lo = std::min(a, b), hi = std::max(a, b); a = lo; b = hi;
    lo = std::min(b, c); hi = std::max(b, c); b = lo; c = hi;
    lo = std::min(a, b); hi = std::max(a, b); a = lo; b = hi;
    float d = a + h;
    if (d > b) {
        d = 0.5f * (a + b + std::sqrt(std::max(2.0f * h * h - (a - b) * (a - b), 0.0f)));
        if (d > c) {
            float s = a + b + c;
            d = (s + std::sqrt(std::max(s * s - 3.0f * (a * a + b * b + c * c - h * h), 0.0f))) / 3.0f;
        }
    }
    return d;
}

// flip.jl:64-66 issolid / cellcentre.
inline bool issolid(float radius, float fl, float px, float py, float pz) {
    return px * px + py * py > radius * radius || pz < fl;
}
inline bool issolidcell(const DenseGrid& g, int i, int j, int k) {
    float cx = g.lo[0] + ((float)i + 0.5f) * g.dx;
    float cy = g.lo[1] + ((float)j + 0.5f) * g.dx;
    float cz = g.lo[2] + ((float)k + 0.5f) * g.dx;
    return issolid(g.radius, g.floor, cx, cy, cz);
}

// Face-sample offsets (flip.jl:71-74): u on x faces, v on y faces, w on z faces, scalars at centres.
inline void faceoff(int a, float off[3]) {
    off[0] = a == 0 ? 0.0f : 0.5f; off[1] = a == 1 ? 0.0f : 0.5f; off[2] = a == 2 ? 0.0f : 0.5f;
}

// A 3-D array, x fastest (the reference's column-major layout).
template <class V> struct Arr3 {
    V* p; int n1, n2, n3;
    V& operator()(int i, int j, int k) const { return p[(size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k)]; }
    bool in(int i, int j, int k) const { return i >= 0 && i < n1 && j >= 0 && j < n2 && k >= 0 && k < n3; }
};

// interpolation.jl:4-17 trilinear, index-space q, clamped; element (i,j,k) at q = (i,j,k).
template <class V> inline float trilinear(const Arr3<V>& A, float q0, float q1, float q2) {
    float x = jclamp(q0, 0.0f, (float)(A.n1 - 1));
    float y = jclamp(q1, 0.0f, (float)(A.n2 - 1));
    float z = jclamp(q2, 0.0f, (float)(A.n3 - 1));
    int i = std::min((int)x, A.n1 - 2), j = std::min((int)y, A.n2 - 2), k = std::min((int)z, A.n3 - 2);
    float fx = x - (float)i, fy = y - (float)j, fz = z - (float)k;
    float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    float c00 = (float)A(i, j, k) * gx + (float)A(i + 1, j, k) * fx;
    float c10 = (float)A(i, j + 1, k) * gx + (float)A(i + 1, j + 1, k) * fx;
    float c01 = (float)A(i, j, k + 1) * gx + (float)A(i + 1, j, k + 1) * fx;
    float c11 = (float)A(i, j + 1, k + 1) * gx + (float)A(i + 1, j + 1, k + 1) * fx;
    return (c00 * gy + c10 * fy) * gz + (c01 * gy + c11 * fy) * fz;
}

// flip.jl:1210-1224 gather: APIC gather of one face component (clamped nodes).
template <class V> inline void gather(const Arr3<V>& u, float px, float py, float pz, const DenseGrid& g, int a,
                                      float& vel, float aff[3]) {
    float off[3]; faceoff(a, off);
    float q0 = (px - g.lo[0]) / g.dx - off[0];
    float q1 = (py - g.lo[1]) / g.dx - off[1];
    float q2 = (pz - g.lo[2]) / g.dx - off[2];
    BS bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
    vel = 0.0f; aff[0] = aff[1] = aff[2] = 0.0f;
    for (int mz = 0; mz < 3; ++mz)
        for (int my = 0; my < 3; ++my)
            for (int mx = 0; mx < 3; ++mx) {
                int i = std::clamp(bx.b + mx, 0, u.n1 - 1);
                int j = std::clamp(by.b + my, 0, u.n2 - 1);
                int k = std::clamp(bz.b + mz, 0, u.n3 - 1);
                float wgt = bx.w[mx] * by.w[my] * bz.w[mz];
                float val = (float)u(i, j, k);
                vel += wgt * val;
                float t = wgt * val;
                aff[0] += (t * ((float)(bx.b + mx) - q0)) * g.dx;
                aff[1] += (t * ((float)(by.b + my) - q1)) * g.dx;
                aff[2] += (t * ((float)(bz.b + mz) - q2)) * g.dx;
            }
    float c = 4.0f / (g.dx * g.dx);
    aff[0] *= c; aff[1] *= c; aff[2] *= c;
}

// ─── particles ───────────────────────────────────────────────────────────────

// flip.jl:131-142: SoA particles (x, y, z; vx, vy, vz; the rows cx, cy, cz of the APIC
// matrix as 9 arrays, c[3a + b] = component b of row a; lag).
struct Particles {
    std::vector<float> x[3], v[3], c[9], lag;
    int64_t n = 0;
    void resize(size_t cap) {
        for (auto& a : x) a.assign(cap, 0.0f);
        for (auto& a : v) a.assign(cap, 0.0f);
        for (auto& a : c) a.assign(cap, 0.0f);
        lag.assign(cap, 0.0f);
    }
    size_t capacity() const { return lag.size(); }
};

// ─── resampling ──────────────────────────────────────────────────────────────

// flip.jl:1289-1297 cellcounts.
inline void cellcounts(int32_t count, float phic, bool solid, int32_t nmin, float deep, float dx, bool seeding,
                       int32_t& kept, int32_t& add) {
    kept = std::min(count, 2 * nmin);
    bool inband = !solid && -deep <= phic && phic <= (seeding ? 0.5f * dx : -dx);
    add = (inband && kept < nmin) ? nmin - kept : 0;
}

struct ResampleOut {
    std::vector<int32_t> key, rank, counts, totals, offsets;  // offsets inclusive, as AK.accumulate!
    int64_t total = 0;
};

// flip.jl:1258-1287 bin_kernel!: key (0-based cell, -1 dropped) and the stable rank.
// The reference ranks by atomic arrival; the canonical order here is by particle index.
inline void dense_bin(const DenseGrid& g, const Particles& ps, const float* phi, float deep, float above,
                      std::vector<int32_t>& key, std::vector<int32_t>& rank, std::vector<int32_t>& counts) {
    Arr3<const float> P{phi, g.n[0], g.n[1], g.n[2]};
    key.assign(ps.n, -1); rank.assign(ps.n, -1); counts.assign(g.cells(), 0);
    for (int64_t p = 0; p < ps.n; ++p) {
        float xp = ps.x[0][p], yp = ps.x[1][p], zp = ps.x[2][p];
        float c0 = (xp - g.lo[0]) / g.dx, c1 = (yp - g.lo[1]) / g.dx, c2 = (zp - g.lo[2]) / g.dx;
        bool keep = std::isfinite(c0) && std::isfinite(c1) && std::isfinite(c2);
        int i = 0, j = 0, k = 0;
        if (keep) {
            // 0 <= floor(c) < n  <=>  0 <= c < n for finite c.
            keep = c0 >= 0.0f && c0 < (float)g.n[0] && c1 >= 0.0f && c1 < (float)g.n[1] && c2 >= 0.0f &&
                   c2 < (float)g.n[2] && !issolid(g.radius, g.floor, xp, yp, zp);
            if (keep) { i = (int)std::floor(c0); j = (int)std::floor(c1); k = (int)std::floor(c2); }
        }
        if (keep) {
            float f = trilinear(P, c0 - 0.5f, c1 - 0.5f, c2 - 0.5f);
            keep = -deep <= f && f < above;
        }
        if (keep) {
            int32_t l = i + g.n[0] * (j + g.n[1] * k);
            key[p] = l;
            rank[p] = counts[l]++;
        }
    }
}

// flip.jl:1364-1382 resample!: bin, totals, inclusive scan, capacity check, scatter, spawn.
// `src` is read, `dst` receives the resampled particles. `seed` is steps + 1.
inline ResampleOut dense_resample(const DenseGrid& g, const Particles& src, Particles& dst, const float* phi,
                                  const float* u, const float* v, const float* w, int32_t band, int32_t nmin,
                                  bool seeding, uint32_t seed) {
    ResampleOut o;
    float deep = (float)band * g.dx;
    dense_bin(g, src, phi, deep, 3.0f * g.dx, o.key, o.rank, o.counts);
    size_t C = g.cells();
    o.totals.assign(C, 0); o.offsets.assign(C, 0);
    for (int k = 0; k < g.n[2]; ++k)                                           // flip.jl:1299-1305
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                size_t l = (size_t)i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k);
                int32_t kept, add;
                cellcounts(o.counts[l], phi[l], issolidcell(g, i, j, k), nmin, deep, g.dx, seeding, kept, add);
                o.totals[l] = kept + add;
            }
    int64_t run = 0;                                                            // flip.jl:1371
    for (size_t l = 0; l < C; ++l) { run += o.totals[l]; o.offsets[l] = (int32_t)run; }
    o.total = run;
    if ((size_t)o.total > dst.capacity())                                       // flip.jl:1373
        throw std::runtime_error("the narrow band needs " + std::to_string(o.total) + " particles, more than the " +
                                 std::to_string(dst.capacity()) + " allocated");
    for (int64_t p = 0; p < src.n; ++p) {                                       // flip.jl:1307-1319
        int32_t l = o.key[p];
        if (l >= 0 && o.rank[p] < 2 * nmin) {
            int64_t d = (int64_t)o.offsets[l] - o.totals[l] + o.rank[p];
            for (int a = 0; a < 3; ++a) { dst.x[a][d] = src.x[a][p]; dst.v[a][d] = src.v[a][p]; }
            for (int a = 0; a < 9; ++a) dst.c[a][d] = src.c[a][p];
            dst.lag[d] = src.lag[p];
        }
    }
    Arr3<const float> P{phi, g.n[0], g.n[1], g.n[2]};                           // flip.jl:1321-1354
    Arr3<const float> U{u, g.n[0] + 1, g.n[1], g.n[2]}, V{v, g.n[0], g.n[1] + 1, g.n[2]},
        W{w, g.n[0], g.n[1], g.n[2] + 1};
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                size_t l = (size_t)i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k);
                int32_t kept, add;
                cellcounts(o.counts[l], phi[l], issolidcell(g, i, j, k), nmin, deep, g.dx, seeding, kept, add);
                if (add <= 0) continue;
                int64_t base = (int64_t)o.offsets[l] - o.totals[l] + kept;
                float corner[3] = {g.lo[0] + (float)i * g.dx, g.lo[1] + (float)j * g.dx, g.lo[2] + (float)k * g.dx};
                int32_t l1 = (int32_t)(l + 1);   // the reference's 1-based linear index
                for (int32_t m = 0; m < add; ++m) {
                    int32_t h = (int32_t)((uint32_t)l1 * 31u + (uint32_t)m);
                    int32_t s = (int32_t)seed;
                    float jit[3] = {hash01(h, s + 1), hash01(h, s + 2), hash01(h, s + 3)};
                    int32_t oo = m % 8;
                    float sub[3] = {(float)(oo & 1), (float)((oo >> 1) & 1), (float)((oo >> 2) & 1)};
                    float half = 0.5f * g.dx;
                    float xp[3];
                    for (int a = 0; a < 3; ++a) xp[a] = corner[a] + (sub[a] + jit[a]) * half;
                    int64_t d = base + m;
                    dst.lag[d] = 0.0f;
                    bool air = false;
                    if (seeding) {
                        float q0 = (xp[0] - g.lo[0]) / g.dx - 0.5f, q1 = (xp[1] - g.lo[1]) / g.dx - 0.5f,
                              q2 = (xp[2] - g.lo[2]) / g.dx - 0.5f;
                        air = trilinear(P, q0, q1, q2) >= 0.0f;
                    }
                    if (air) {
                        for (int a = 0; a < 3; ++a) { dst.x[a][d] = NAN; dst.v[a][d] = 0.0f; }
                        for (int a = 0; a < 9; ++a) dst.c[a][d] = 0.0f;
                    } else {
                        float vel, aff[3];
                        for (int a = 0; a < 3; ++a) dst.x[a][d] = xp[a];
                        gather(U, xp[0], xp[1], xp[2], g, 0, vel, aff);
                        dst.v[0][d] = vel; for (int b = 0; b < 3; ++b) dst.c[b][d] = aff[b];
                        gather(V, xp[0], xp[1], xp[2], g, 1, vel, aff);
                        dst.v[1][d] = vel; for (int b = 0; b < 3; ++b) dst.c[3 + b][d] = aff[b];
                        gather(W, xp[0], xp[1], xp[2], g, 2, vel, aff);
                        dst.v[2][d] = vel; for (int b = 0; b < 3; ++b) dst.c[6 + b][d] = aff[b];
                    }
                }
            }
    dst.n = o.total;
    return o;
}

// ─── particles to grid ───────────────────────────────────────────────────────

// flip.jl:513-537 p2g_kernel! for component a, accumulated in T (double for the FP64
// oracle): mom += w (v_a + c_a . d), wt += w over the 27 in-range nodes.
template <class T>
inline void dense_p2g_component(const DenseGrid& g, const Particles& ps, int a, float dt, T* mom, T* wt) {
    float off[3]; faceoff(a, off);
    int n1 = g.n[0] + (a == 0), n2 = g.n[1] + (a == 1), n3 = g.n[2] + (a == 2);
    for (int64_t p = 0; p < ps.n; ++p) {
        float wT = temporal(ps.lag[p], dt);
        float q0 = (ps.x[0][p] - g.lo[0]) / g.dx - off[0];
        float q1 = (ps.x[1][p] - g.lo[1]) / g.dx - off[1];
        float q2 = (ps.x[2][p] - g.lo[2]) / g.dx - off[2];
        BS bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
        T vp = ps.v[a][p];
        T c0 = ps.c[3 * a][p], c1 = ps.c[3 * a + 1][p], c2 = ps.c[3 * a + 2][p];
        for (int mz = 0; mz < 3; ++mz)
            for (int my = 0; my < 3; ++my)
                for (int mx = 0; mx < 3; ++mx) {
                    int I = bx.b + mx, J = by.b + my, K = bz.b + mz;
                    if (I < 0 || I >= n1 || J < 0 || J >= n2 || K < 0 || K >= n3) continue;
                    T wgt = (T)wT * (T)bx.w[mx] * (T)by.w[my] * (T)bz.w[mz];
                    T d0 = ((T)I - (T)q0) * (T)g.dx, d1 = ((T)J - (T)q1) * (T)g.dx, d2 = ((T)K - (T)q2) * (T)g.dx;
                    size_t idx = (size_t)I + (size_t)n1 * (J + (size_t)n2 * K);
                    mom[idx] += wgt * (vp + (c0 * d0 + c1 * d1 + c2 * d2));
                    wt[idx] += wgt;
                }
    }
}

// flip.jl:539-560 splat_kernel!: the particles' level set, as order keys (min).
inline void dense_splat(const DenseGrid& g, const Particles& ps, float dt, int32_t* phip) {
    float radius = 0.5f * g.dx;
    for (int64_t p = 0; p < ps.n; ++p) {
        float vx = ps.v[0][p], vy = ps.v[1][p], vz = ps.v[2][p];
        float sp = std::sqrt(vx * vx + vy * vy + vz * vz);
        float s = ps.lag[p] + ahead(strength(sp, dt, g.dx)) * dt;
        float xp[3] = {ps.x[0][p] + s * vx, ps.x[1][p] + s * vy, ps.x[2][p] + s * vz};
        int ci = (int)std::floor((xp[0] - g.lo[0]) / g.dx), cj = (int)std::floor((xp[1] - g.lo[1]) / g.dx),
            ck = (int)std::floor((xp[2] - g.lo[2]) / g.dx);
        for (int dk = -1; dk <= 1; ++dk)
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    int I = ci + di, J = cj + dj, K = ck + dk;
                    if (I < 0 || I >= g.n[0] || J < 0 || J >= g.n[1] || K < 0 || K >= g.n[2]) continue;
                    float cx = g.lo[0] + ((float)I + 0.5f) * g.dx - xp[0];
                    float cy = g.lo[1] + ((float)J + 0.5f) * g.dx - xp[1];
                    float cz = g.lo[2] + ((float)K + 0.5f) * g.dx - xp[2];
                    float d = std::sqrt(cx * cx + cy * cy + cz * cz) - radius;
                    size_t l = (size_t)I + (size_t)g.n[0] * (J + (size_t)g.n[1] * K);
                    phip[l] = std::min(phip[l], orderkey(d));
                }
    }
}

// flip.jl:587-606 particles2grid!: P2G of the three components, splat, union
// (flip.jl:562-565), combine (flip.jl:567-585). Writes phi, u, v, w; returns the
// accumulators when asked (mom[a], wt[a]).
template <class T>
inline void dense_p2g(const DenseGrid& g, const Particles& ps, float dt, int32_t band, int32_t nmin, const float* phi0,
                      const float* u0[3], float* phi, float* u[3], std::vector<T>* mom_out = nullptr,
                      std::vector<T>* wt_out = nullptr) {
    size_t C = g.cells();
    std::vector<int32_t> phip(C, orderkey(1e30f));
    std::vector<T> mom[3], wt[3];
#pragma omp parallel for schedule(static, 1)
    for (int a = 0; a < 3; ++a) {
        mom[a].assign(g.faces(a), T(0)); wt[a].assign(g.faces(a), T(0));
        dense_p2g_component(g, ps, a, dt, mom[a].data(), wt[a].data());
    }
    dense_splat(g, ps, dt, phip.data());
    for (size_t l = 0; l < C; ++l) phi[l] = jmin(phi0[l] + g.dx, orderfloat(phip[l]));
    float r = (float)(band - 1) * g.dx;
    float full = (float)nmin;
    Arr3<const float> P{phi, g.n[0], g.n[1], g.n[2]};
    for (int a = 0; a < 3; ++a) {
        int n1 = g.n[0] + (a == 0), n2 = g.n[1] + (a == 1), n3 = g.n[2] + (a == 2);
        for (int k = 0; k < n3; ++k)
            for (int j = 0; j < n2; ++j)
                for (int i = 0; i < n1; ++i) {
                    int bi = i - (a == 0), bj = j - (a == 1), bk = k - (a == 2);
                    float fa = P.in(bi, bj, bk) ? P(bi, bj, bk)
                                                : P(std::min(i, g.n[0] - 1), std::min(j, g.n[1] - 1), std::min(k, g.n[2] - 1));
                    float fb = P.in(i, j, k) ? P(i, j, k) : fa;
                    float phif = 0.5f * (fa + fb);
                    size_t idx = (size_t)i + (size_t)n1 * (j + (size_t)n2 * k);
                    float wgt = (float)wt[a][idx];
                    float m = (float)mom[a][idx];
                    u[a][idx] = phif >= -r ? (m + jmax(full - wgt, 0.0f) * u0[a][idx]) / jmax(full, wgt) : u0[a][idx];
                }
    }
    if (mom_out) for (int a = 0; a < 3; ++a) mom_out[a] = mom[a];
    if (wt_out) for (int a = 0; a < 3; ++a) wt_out[a] = wt[a];
}

// ─── grid to particles ───────────────────────────────────────────────────────

// flip.jl:1226-1244 g2p_kernel! with the dense Flow (flip.jl:401-405 gathered): APIC where
// every face read was reached (flag gather > 0.999), else ballistic.
inline void dense_g2p(const DenseGrid& g, Particles& ps, const float* u[3], const uint8_t* f[3], float dt, float gravity) {
    Arr3<const float> U[3] = {{u[0], g.n[0] + 1, g.n[1], g.n[2]}, {u[1], g.n[0], g.n[1] + 1, g.n[2]},
                              {u[2], g.n[0], g.n[1], g.n[2] + 1}};
    Arr3<const uint8_t> F[3] = {{f[0], g.n[0] + 1, g.n[1], g.n[2]}, {f[1], g.n[0], g.n[1] + 1, g.n[2]},
                                {f[2], g.n[0], g.n[1], g.n[2] + 1}};
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < ps.n; ++p) {
        float xp = ps.x[0][p], yp = ps.x[1][p], zp = ps.x[2][p];
        float vel[3], aff[3][3], fl, faff[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            gather(U[a], xp, yp, zp, g, a, vel[a], aff[a]);
            gather(F[a], xp, yp, zp, g, a, fl, faff);
            ok = ok && fl > 0.999f;
        }
        if (ok) {
            for (int a = 0; a < 3; ++a) {
                ps.v[a][p] = vel[a];
                for (int b = 0; b < 3; ++b) ps.c[3 * a + b][p] = aff[a][b];
            }
        } else {
            ps.v[2][p] -= gravity * dt;
            for (int b = 0; b < 9; ++b) ps.c[b][p] = 0.0f;
        }
    }
}

// ─── the pressure ────────────────────────────────────────────────────────────

// One multigrid level: types and (level 0 only) the ghost-fluid diagonal.
struct MGLevel { int n[3]; std::vector<uint8_t> types; };

inline uint8_t typeat(const uint8_t* t, const int n[3], int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= n[0] || j >= n[1] || k >= n[2]) return SOLID;
    return t[(size_t)i + (size_t)n[0] * (j + (size_t)n[1] * k)];
}
static const int NB[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};  // flip.jl:84-92

template <class T> struct DensePressure {
    DenseGrid g;
    std::vector<MGLevel> lv;   // lv[0] = the fine grid
    std::vector<T> diag;       // fine level (flip.jl:824-836)

    // flip.jl:879-889 coarsen + flip.jl:824-836 diag.
    void setup(const uint8_t* types, const float* phi) {
        lv.clear();
        MGLevel f; f.n[0] = g.n[0]; f.n[1] = g.n[1]; f.n[2] = g.n[2];
        f.types.assign(types, types + g.cells());
        lv.push_back(f);
        for (int l = 1; l < g.levels; ++l) {
            const MGLevel& F = lv.back();
            MGLevel c; for (int a = 0; a < 3; ++a) c.n[a] = F.n[a] / 2;
            c.types.assign((size_t)c.n[0] * c.n[1] * c.n[2], SOLID);
            for (int K = 0; K < c.n[2]; ++K)
                for (int J = 0; J < c.n[1]; ++J)
                    for (int I = 0; I < c.n[0]; ++I) {
                        bool air = false, liq = false;
                        for (int cc = 0; cc < 2; ++cc)
                            for (int b = 0; b < 2; ++b)
                                for (int a = 0; a < 2; ++a) {
                                    uint8_t t = typeat(F.types.data(), F.n, 2 * I + a, 2 * J + b, 2 * K + cc);
                                    if (t == AIR) air = true;
                                    if (t == LIQUID) liq = true;
                                }
                        c.types[(size_t)I + (size_t)c.n[0] * (J + (size_t)c.n[1] * K)] = air ? AIR : (liq ? LIQUID : SOLID);
                    }
            lv.push_back(c);
        }
        diag.assign(g.cells(), T(0));
        const int* n = g.n;
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i) {
                    size_t l = (size_t)i + (size_t)n[0] * (j + (size_t)n[1] * k);
                    if (types[l] != LIQUID) continue;
                    T d = 0;
                    for (int m = 0; m < 6; ++m) {
                        int a = i + NB[m][0], b = j + NB[m][1], c = k + NB[m][2];
                        uint8_t t = typeat(types, n, a, b, c);
                        if (t == LIQUID) d += T(1);
                        if (t == AIR) d += T(1) / theta<T>((T)phi[l], (T)phi[(size_t)a + (size_t)n[0] * (b + (size_t)n[1] * c)]);
                    }
                    diag[l] = d;
                }
    }
    size_t size(int l) const { return (size_t)lv[l].n[0] * lv[l].n[1] * lv[l].n[2]; }
    // flip.jl:838-848 celldiag.
    T celldiag(int l, int i, int j, int k) const {
        const MGLevel& L = lv[l];
        if (l == 0) return diag[(size_t)i + (size_t)L.n[0] * (j + (size_t)L.n[1] * k)];
        T d = 0;
        for (int m = 0; m < 6; ++m)
            if (typeat(L.types.data(), L.n, i + NB[m][0], j + NB[m][1], k + NB[m][2]) != SOLID) d += T(1);
        return d;
    }
    // flip.jl:850-857 liquidsum.
    T liquidsum(int l, const T* x, int i, int j, int k) const {
        const MGLevel& L = lv[l];
        T s = 0;
        for (int m = 0; m < 6; ++m) {
            int a = i + NB[m][0], b = j + NB[m][1], c = k + NB[m][2];
            if (typeat(L.types.data(), L.n, a, b, c) == LIQUID) s += x[(size_t)a + (size_t)L.n[0] * (b + (size_t)L.n[1] * c)];
        }
        return s;
    }
    // A x at level l (q = diag x - liquidsum on LIQUID cells, else 0).
    void apply(int l, const T* x, T* q) const {
        const MGLevel& L = lv[l];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = 0; i < L.n[0]; ++i) {
                    size_t c = (size_t)i + (size_t)L.n[0] * (j + (size_t)L.n[1] * k);
                    q[c] = L.types[c] == LIQUID ? celldiag(l, i, j, k) * x[c] - liquidsum(l, x, i, j, k) : T(0);
                }
    }
    // flip.jl:859-863 residual_kernel!.
    void residual(int l, T* r, const T* x, const T* b) const {
        const MGLevel& L = lv[l];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = 0; i < L.n[0]; ++i) {
                    size_t c = (size_t)i + (size_t)L.n[0] * (j + (size_t)L.n[1] * k);
                    r[c] = L.types[c] == LIQUID ? b[c] - (celldiag(l, i, j, k) * x[c] - liquidsum(l, x, i, j, k)) : T(0);
                }
    }
    // flip.jl:865-877 rbgs_kernel!: colour in the reference's 1-based parity, i.e.
    // 0-based cells with (i + j + k + 1) % 2 == colour.
    void rbgs(int l, T* x, const T* b, int colour) const {
        const MGLevel& L = lv[l];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = 0; i < L.n[0]; ++i) {
                    if ((i + j + k + 1) % 2 != colour) continue;
                    size_t c = (size_t)i + (size_t)L.n[0] * (j + (size_t)L.n[1] * k);
                    if (L.types[c] != LIQUID) continue;
                    T d = celldiag(l, i, j, k);
                    x[c] = d > T(0) ? (b[c] + liquidsum(l, x, i, j, k)) / d : T(0);
                }
    }
    // flip.jl:964-973 start_kernel!: colour 0 from zero.
    void start(int l, T* x, const T* b) const {
        const MGLevel& L = lv[l];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = 0; i < L.n[0]; ++i) {
                    size_t c = (size_t)i + (size_t)L.n[0] * (j + (size_t)L.n[1] * k);
                    T v = 0;
                    if ((i + j + k + 1) % 2 == 0 && L.types[c] == LIQUID) {
                        T d = celldiag(l, i, j, k);
                        if (d > T(0)) v = b[c] / d;
                    }
                    x[c] = v;
                }
    }
    // flip.jl:891-906 restrict_kernel!.
    void restrict_(int l, T* bc, const T* rf) const {
        const MGLevel &F = lv[l], &Cl = lv[l + 1];
        static const T wB[4] = {T(0.125f), T(0.375f), T(0.375f), T(0.125f)};
#pragma omp parallel for schedule(static)
        for (int K = 0; K < Cl.n[2]; ++K)
            for (int J = 0; J < Cl.n[1]; ++J)
                for (int I = 0; I < Cl.n[0]; ++I) {
                    size_t C = (size_t)I + (size_t)Cl.n[0] * (J + (size_t)Cl.n[1] * K);
                    T s = 0;
                    if (Cl.types[C] == LIQUID)
                        for (int c = 0; c < 4; ++c)
                            for (int b = 0; b < 4; ++b)
                                for (int a = 0; a < 4; ++a) {
                                    int i = 2 * I - 1 + a, j = 2 * J - 1 + b, k = 2 * K - 1 + c;
                                    if (typeat(F.types.data(), F.n, i, j, k) == LIQUID)
                                        s += wB[a] * wB[b] * wB[c] * rf[(size_t)i + (size_t)F.n[0] * (j + (size_t)F.n[1] * k)];
                                }
                    bc[C] = T(4) * s;
                }
    }
    // flip.jl:908-929 prolong_kernel!.
    static void pw(int i, int& a0, T& w0, int& a1, T& w1) {
        if (i % 2 == 0) { a0 = i / 2 - 1; w0 = T(0.25f); a1 = i / 2; w1 = T(0.75f); }
        else { a0 = (i - 1) / 2; w0 = T(0.75f); a1 = (i + 1) / 2; w1 = T(0.25f); }
    }
    void prolong(int l, T* xf, const T* xc) const {
        const MGLevel &F = lv[l], &Cl = lv[l + 1];
        for (int k = 0; k < F.n[2]; ++k)
            for (int j = 0; j < F.n[1]; ++j)
                for (int i = 0; i < F.n[0]; ++i) {
                    size_t c = (size_t)i + (size_t)F.n[0] * (j + (size_t)F.n[1] * k);
                    if (F.types[c] != LIQUID) continue;
                    int A[2], B[2], Cc[2]; T WA[2], WB[2], WC[2];
                    pw(i, A[0], WA[0], A[1], WA[1]); pw(j, B[0], WB[0], B[1], WB[1]); pw(k, Cc[0], WC[0], Cc[1], WC[1]);
                    T s = 0;
                    for (int m = 0; m < 8; ++m) {
                        int a = A[m & 1], b = B[(m >> 1) & 1], cc = Cc[(m >> 2) & 1];
                        if (typeat(Cl.types.data(), Cl.n, a, b, cc) == LIQUID)
                            s += WA[m & 1] * WB[(m >> 1) & 1] * WC[(m >> 2) & 1] *
                                 xc[(size_t)a + (size_t)Cl.n[0] * (b + (size_t)Cl.n[1] * cc)];
                    }
                    xf[c] += s;
                }
    }
    // flip.jl:931-962 coarse_kernel!: `sweeps` RBGS sweeps from zero and as many reversed.
    void coarsest(T* x, const T* b, int sweeps) const {
        int l = (int)lv.size() - 1;
        const MGLevel& L = lv[l];
        size_t N = size(l);
        std::fill(x, x + N, T(0));
        for (int s = 1; s <= 2 * sweeps; ++s)
            for (int half = 0; half < 2; ++half) {
                int colour = s <= sweeps ? half : 1 - half;
                std::vector<T> xs(x, x + N);   // one half-sweep reads the other colour only
                for (int k = 0; k < L.n[2]; ++k)
                    for (int j = 0; j < L.n[1]; ++j)
                        for (int i = 0; i < L.n[0]; ++i) {
                            size_t c = (size_t)i + (size_t)L.n[0] * (j + (size_t)L.n[1] * k);
                            if (L.types[c] != LIQUID || (i + j + k + 1) % 2 != colour) continue;
                            T d = 0, acc = b[c];
                            for (int m = 0; m < 6; ++m) {
                                int a = i + NB[m][0], bb = j + NB[m][1], cc = k + NB[m][2];
                                uint8_t tn = typeat(L.types.data(), L.n, a, bb, cc);
                                if (tn != SOLID) d += T(1);
                                if (tn == LIQUID) acc += xs[(size_t)a + (size_t)L.n[0] * (bb + (size_t)L.n[1] * cc)];
                            }
                            x[c] = d > T(0) ? acc / d : T(0);
                        }
            }
    }
    // flip.jl:986-1014 vcycle!: z = M^-1 r.
    std::vector<std::vector<T>> xs, bs, rs;
    void vcycle(T* z, const T* r, int sweeps, int coarse) {
        int L = (int)lv.size();
        xs.resize(L); bs.resize(L); rs.resize(L);
        for (int l = 1; l < L; ++l) { xs[l].assign(size(l), T(0)); bs[l].assign(size(l), T(0)); rs[l].assign(size(l), T(0)); }
        rs[0].assign(size(0), T(0));
        std::vector<int> down, up;
        for (int s = 0; s < sweeps; ++s) { down.push_back(0); down.push_back(1); up.push_back(1); up.push_back(0); }
        down.erase(down.begin());
        auto X = [&](int l) { return l == 0 ? z : xs[l].data(); };
        auto Bv = [&](int l) { return l == 0 ? r : (const T*)bs[l].data(); };
        for (int l = 0; l < L - 1; ++l) {
            start(l, X(l), Bv(l));
            for (int c : down) rbgs(l, X(l), Bv(l), c);
            residual(l, rs[l].data(), X(l), Bv(l));
            restrict_(l, bs[l + 1].data(), rs[l].data());
        }
        coarsest(X(L - 1), Bv(L - 1), coarse);
        for (int l = L - 2; l >= 0; --l) {
            prolong(l, X(l), X(l + 1));
            for (int c : up) rbgs(l, X(l), Bv(l), c);
        }
    }
};

template <class T> inline T dotv(const std::vector<T>& a, const std::vector<T>& b) {
    T s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// flip.jl:811-822 rhs_kernel!.
template <class T>
inline void dense_rhs(const DenseGrid& g, const float* u, const float* v, const float* w, const uint8_t* types, float gamma,
                      T* b) {
    const int* n = g.n;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                size_t l = (size_t)i + (size_t)n[0] * (j + (size_t)n[1] * k);
                if (types[l] != LIQUID) { b[l] = T(0); continue; }
                size_t u0 = (size_t)i + (size_t)(n[0] + 1) * (j + (size_t)n[1] * k);
                size_t v0 = (size_t)i + (size_t)n[0] * (j + (size_t)(n[1] + 1) * k);
                size_t w0 = l;
                T div = ((((T)u[u0 + 1] - (T)u[u0]) + (T)v[v0 + n[0]]) - (T)v[v0]) + (T)w[w0 + (size_t)n[0] * n[1]] - (T)w[w0];
                b[l] = -(T)g.dx * div + (T)g.dx * (T)g.dx * (T)gamma;
            }
}

// flip.jl:1075-1111 pressure!: MG-preconditioned CG from the warm start P. Returns
// iterations; `relres` receives ||r||/||b|| of the recursive residual.
template <class T>
inline int dense_pcg(DensePressure<T>& S, const std::vector<T>& b, std::vector<T>& P, double tol, int maxiter, int sweeps,
                     int coarse, double* relres) {
    size_t N = b.size();
    std::vector<T> r(N), z(N), d(N, T(0)), q(N);
    S.residual(0, r.data(), P.data(), b.data());
    T bnorm = std::sqrt(dotv(b, b));
    if (relres) *relres = 0;
    if (bnorm == T(0)) return 0;
    T target = (T)tol * bnorm;
    T rnorm = std::sqrt(dotv(r, r));
    if (relres) *relres = (double)(rnorm / bnorm);
    if (rnorm <= target) return 0;
    S.vcycle(z.data(), r.data(), sweeps, coarse);
    T rho = dotv(z, r), beta = 0;
    for (int it = 1; it <= maxiter; ++it) {
        for (size_t i = 0; i < N; ++i) d[i] = z[i] + beta * d[i];   // p = z + beta d (flip.jl:1038-1054)
        S.apply(0, d.data(), q.data());
        T alpha = rho / dotv(d, q);
        for (size_t i = 0; i < N; ++i) { P[i] += alpha * d[i]; r[i] -= alpha * q[i]; }   // flip.jl:1059-1073
        rnorm = std::sqrt(dotv(r, r));
        if (relres) *relres = (double)(rnorm / bnorm);
        if (rnorm <= target) return it;
        S.vcycle(z.data(), r.data(), sweeps, coarse);
        T rhonew = dotv(z, r);
        beta = rhonew / rho;
        rho = rhonew;
    }
    return maxiter;
}

// flip.jl:1113-1130 project_kernel!: u -= grad P / dx on faces with water and no wall; flags.
template <class T>
inline void dense_project_faces(const DenseGrid& g, const T* P, const uint8_t* types, const float* phi, float* u[3],
                                uint8_t* valid[3]) {
    const int* n = g.n;
    for (int a = 0; a < 3; ++a) {
        int n1 = n[0] + (a == 0), n2 = n[1] + (a == 1), n3 = n[2] + (a == 2);
        for (int k = 0; k < n3; ++k)
            for (int j = 0; j < n2; ++j)
                for (int i = 0; i < n1; ++i) {
                    int ai = i - (a == 0), bj = j - (a == 1), ck = k - (a == 2);
                    uint8_t t0 = typeat(types, n, ai, bj, ck), t1 = typeat(types, n, i, j, k);
                    bool ok = t0 != SOLID && t1 != SOLID && (t0 == LIQUID || t1 == LIQUID);
                    size_t f = (size_t)i + (size_t)n1 * (j + (size_t)n2 * k);
                    if (ok) {
                        size_t c0 = (size_t)ai + (size_t)n[0] * (bj + (size_t)n[1] * ck);
                        size_t c1 = (size_t)i + (size_t)n[0] * (j + (size_t)n[1] * k);
                        T D;
                        if (t0 == LIQUID && t1 == LIQUID) D = P[c1] - P[c0];
                        else if (t0 == LIQUID) D = -P[c0] / theta<T>((T)phi[c0], (T)phi[c1]);
                        else D = P[c1] / theta<T>((T)phi[c1], (T)phi[c0]);
                        u[a][f] = (float)((T)u[a][f] - D / (T)g.dx);
                    }
                    valid[a][f] = ok ? 1 : 0;
                }
    }
}

// flip.jl:293 volume (flip.jl:262-279): sum over non-solid cells of clamp(1/2 - phi/dx, 0, 1), times dx^3.
inline double dense_volume(const DenseGrid& g, const float* phi) {
    double s = 0;
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                if (issolidcell(g, i, j, k)) continue;
                size_t l = (size_t)i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k);
                s += jclamp(0.5f - phi[l] / g.dx, 0.0f, 1.0f);
            }
    return s * (double)g.dx * g.dx * g.dx;
}

// flip.jl:674-677 classify_kernel!.
inline void dense_classify(const DenseGrid& g, const float* phi, uint8_t* types) {
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                size_t l = (size_t)i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k);
                types[l] = issolidcell(g, i, j, k) ? SOLID : (phi[l] < 0.0f ? LIQUID : AIR);
            }
}

}  // namespace orc
