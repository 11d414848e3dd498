// TEST INFRASTRUCTURE ONLY: the sparse basin's particle phases in the CPU oracle:
// resampling (sflip.jl:2146-2246), particles to grid (sflip.jl:788-831) and grid to
// particles (sflip.jl:2141-2144), through the topology.
#pragma once
#include "sparse.hpp"

namespace orc {

// The fine-pool and coarse arrays of a sparse field, as raw pointers.
struct SF { float* fine; float* coarse; };
struct SFc { const float* fine; const float* coarse; };

// sflip.jl:2148-2171 sparsebin: key = fine pool index; stable rank (particle order).
inline void sparse_bin(const Topo& t, const Particles& ps, SFc phi, float deep, float above, std::vector<int32_t>& key,
                       std::vector<int32_t>& rank, std::vector<int32_t>& counts) {
    const SparseGrid& g = t.g;
    key.assign(ps.n, -1); rank.assign(ps.n, -1);
    counts.assign((size_t)t.nslots * BCELLS, 0);
    for (int64_t p = 0; p < ps.n; ++p) {
        float xp = ps.x[0][p], yp = ps.x[1][p], zp = ps.x[2][p];
        float c0 = (xp - g.lo[0]) / g.dx, c1 = (yp - g.lo[1]) / g.dx, c2 = (zp - g.lo[2]) / g.dx;
        int64_t idx = -1;
        if (std::isfinite(c0) && std::isfinite(c1) && std::isfinite(c2) && !sparse_issolid(g, xp, yp, zp)) {
            // outside [0, 16 nb) no block holds the cell
            if (c0 >= 0.0f && c1 >= 0.0f && c2 >= 0.0f && c0 < (float)(16 * g.nb[0]) && c1 < (float)(16 * g.nb[1]) &&
                c2 < (float)(16 * g.nb[2]))
                idx = t.fineat((int)std::floor(c0), (int)std::floor(c1), (int)std::floor(c2));
            if (idx >= 0) {
                float f = cellsample(t, phi.fine, phi.coarse, xp, yp, zp);
                if (!(-deep <= f && f < above)) idx = -1;
            }
        }
        if (idx >= 0) {
            key[p] = (int32_t)idx;
            rank[p] = counts[idx]++;
        }
    }
}

// sflip.jl:2227-2246 resample!(::SparseBasin): `flow` the face fields new particles gather from.
inline ResampleOut sparse_resample(const Topo& t, const Particles& src, Particles& dst, SFc phi, const float* flow[3],
                                   int32_t band, int32_t nmin, bool seeding, uint32_t seed) {
    const SparseGrid& g = t.g;
    ResampleOut o;
    float deep = (float)band * g.dx;
    sparse_bin(t, src, phi, deep, 3.0f * g.dx, o.key, o.rank, o.counts);
    size_t Cn = (size_t)t.nslots * BCELLS;
    o.totals.assign(Cn, 0); o.offsets.assign(Cn, 0);
    for (size_t n = 0; n < Cn; ++n) {                                           // sflip.jl:2173-2182
        auto w = t.where[n >> 12];
        if (w[0] < 0) continue;
        int c = (int)(n & 4095), i = c & 15, j = (c >> 4) & 15, k = c >> 8;
        float p[3]; finecentre(g, w[0] * 16 + i, w[1] * 16 + j, w[2] * 16 + k, p);
        int32_t kept, add;
        cellcounts(o.counts[n], phi.fine[n], sparse_issolid(g, p[0], p[1], p[2]), nmin, deep, g.dx, seeding, kept, add);
        o.totals[n] = kept + add;
    }
    int64_t run = 0;
    for (size_t n = 0; n < Cn; ++n) { run += o.totals[n]; o.offsets[n] = (int32_t)run; }
    o.total = run;
    if ((size_t)o.total > dst.capacity())
        throw std::runtime_error("the narrow band needs " + std::to_string(o.total) + " particles, more than the " +
                                 std::to_string(dst.capacity()) + " allocated");
    for (int64_t p = 0; p < src.n; ++p) {
        int32_t l = o.key[p];
        if (l >= 0 && o.rank[p] < 2 * nmin) {
            int64_t d = (int64_t)o.offsets[l] - o.totals[l] + o.rank[p];
            for (int a = 0; a < 3; ++a) { dst.x[a][d] = src.x[a][p]; dst.v[a][d] = src.v[a][p]; }
            for (int a = 0; a < 9; ++a) dst.c[a][d] = src.c[a][p];
            dst.lag[d] = src.lag[p];
        }
    }
    for (size_t n = 0; n < Cn; ++n) {                                           // sflip.jl:2184-2218
        auto w = t.where[n >> 12];
        if (w[0] < 0) continue;
        int c = (int)(n & 4095), i = c & 15, j = (c >> 4) & 15, k = c >> 8;
        int I = w[0] * 16 + i, J = w[1] * 16 + j, K = w[2] * 16 + k;
        float pc[3]; finecentre(g, I, J, K, pc);
        int32_t kept, add;
        cellcounts(o.counts[n], phi.fine[n], sparse_issolid(g, pc[0], pc[1], pc[2]), nmin, deep, g.dx, seeding, kept, add);
        if (add <= 0) continue;
        int64_t base = (int64_t)o.offsets[n] - o.totals[n] + kept;
        float corner[3] = {g.lo[0] + (float)I * g.dx, g.lo[1] + (float)J * g.dx, g.lo[2] + (float)K * g.dx};
        int32_t n1 = (int32_t)(n + 1);
        for (int32_t m = 0; m < add; ++m) {
            int32_t h = (int32_t)((uint32_t)n1 * 31u + (uint32_t)m);
            int32_t s = (int32_t)seed;
            float jit[3] = {hash01(h, s + 1), hash01(h, s + 2), hash01(h, s + 3)};
            int32_t oo = m % 8;
            float sub[3] = {(float)(oo & 1), (float)((oo >> 1) & 1), (float)((oo >> 2) & 1)};
            float half = 0.5f * g.dx;
            float xp[3];
            for (int a = 0; a < 3; ++a) xp[a] = corner[a] + (sub[a] + jit[a]) * half;
            int64_t d = base + m;
            dst.lag[d] = 0.0f;
            bool air = seeding && cellsample(t, phi.fine, phi.coarse, xp[0], xp[1], xp[2]) >= 0.0f;
            if (air) {
                for (int a = 0; a < 3; ++a) { dst.x[a][d] = NAN; dst.v[a][d] = 0.0f; }
                for (int a = 0; a < 9; ++a) dst.c[a][d] = 0.0f;
            } else {
                for (int a = 0; a < 3; ++a) dst.x[a][d] = xp[a];
                for (int a = 0; a < 3; ++a) {
                    float vel, aff[3], found;
                    finegather(t, flow[a], a, xp[0], xp[1], xp[2], vel, aff, found);
                    dst.v[a][d] = vel;
                    for (int b = 0; b < 3; ++b) dst.c[3 * a + b][d] = aff[b];
                }
            }
        }
    }
    dst.n = o.total;
    return o;
}

// sflip.jl:808-831 particles2grid!(::SparseBasin), accumulation in T.
template <class T>
inline void sparse_p2g(const Topo& t, const Particles& ps, float dt, int32_t band, int32_t nmin, SFc phi0, const SFc u0[3],
                       SF phi, const SF u[3], std::vector<T>* mom_out = nullptr, std::vector<T>* wt_out = nullptr) {
    const SparseGrid& g = t.g;
    size_t Fn = (size_t)t.nslots * BFACES, Cn = (size_t)t.nslots * BCELLS;
    std::vector<T> mom[3], wt[3];
#pragma omp parallel for schedule(static, 1)
    for (int a = 0; a < 3; ++a) {
        mom[a].assign(Fn, T(0)); wt[a].assign(Fn, T(0));
        float off[3]; faceoff(a, off);
        for (int64_t p = 0; p < ps.n; ++p) {                                   // flip.jl:513-537 over the topology
            float wT = temporal(ps.lag[p], dt);
            float q0 = (ps.x[0][p] - g.lo[0]) / g.dx - off[0];
            float q1 = (ps.x[1][p] - g.lo[1]) / g.dx - off[1];
            float q2 = (ps.x[2][p] - g.lo[2]) / g.dx - off[2];
            BS bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
            T vp = ps.v[a][p], c0 = ps.c[3 * a][p], c1 = ps.c[3 * a + 1][p], c2 = ps.c[3 * a + 2][p];
            for (int mz = 0; mz < 3; ++mz)
                for (int my = 0; my < 3; ++my)
                    for (int mx = 0; mx < 3; ++mx) {
                        int64_t idx = t.finefaceat(a, bx.b + mx, by.b + my, bz.b + mz);
                        if (idx < 0) continue;
                        T wgt = (T)wT * (T)bx.w[mx] * (T)by.w[my] * (T)bz.w[mz];
                        T d0 = ((T)(bx.b + mx) - (T)q0) * (T)g.dx, d1 = ((T)(by.b + my) - (T)q1) * (T)g.dx,
                          d2 = ((T)(bz.b + mz) - (T)q2) * (T)g.dx;
                        mom[a][idx] += wgt * (vp + (c0 * d0 + c1 * d1 + c2 * d2));
                        wt[a][idx] += wgt;
                    }
        }
        syncplanes(t, mom[a].data(), a);
        syncplanes(t, wt[a].data(), a);
    }
    std::vector<int32_t> phip(Cn, orderkey(1e30f));
    float radius = 0.5f * g.dx;
    for (int64_t p = 0; p < ps.n; ++p) {                                       // flip.jl:539-560 over the topology
        float vx = ps.v[0][p], vy = ps.v[1][p], vz = ps.v[2][p];
        float sp = std::sqrt(vx * vx + vy * vy + vz * vz);
        float s = ps.lag[p] + ahead(strength(sp, dt, g.dx)) * dt;
        float xp[3] = {ps.x[0][p] + s * vx, ps.x[1][p] + s * vy, ps.x[2][p] + s * vz};
        int ci = (int)std::floor((xp[0] - g.lo[0]) / g.dx), cj = (int)std::floor((xp[1] - g.lo[1]) / g.dx),
            ck = (int)std::floor((xp[2] - g.lo[2]) / g.dx);
        for (int dk = -1; dk <= 1; ++dk)
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    int64_t idx = t.fineat(ci + di, cj + dj, ck + dk);
                    if (idx < 0) continue;
                    float c[3]; finecentre(g, ci + di, cj + dj, ck + dk, c);
                    float ex = c[0] - xp[0], ey = c[1] - xp[1], ez = c[2] - xp[2];
                    float d = std::sqrt(ex * ex + ey * ey + ez * ez) - radius;
                    phip[idx] = std::min(phip[idx], orderkey(d));
                }
    }
    for (size_t n = 0; n < Cn; ++n)                                            // sflip.jl:790-794
        if (t.where[n >> 12][0] >= 0) phi.fine[n] = jmin(phi0.fine[n] + g.dx, orderfloat(phip[n]));
    std::copy(phi0.coarse, phi0.coarse + g.ncoarse(), phi.coarse);
    restrictcells(t, phi.coarse, (const float*)phi.fine);
    float r = (float)(band - 1) * g.dx, full = (float)nmin;
    for (int a = 0; a < 3; ++a) {                                              // sflip.jl:796-806
        for (size_t f = 0; f < Fn; ++f) {
            int s = (int)(f / BFACES) + 1;
            auto w = t.where[s - 1];
            if (w[0] < 0) continue;
            int fl = (int)(f % BFACES), i, j, k;
            if (a == 0) { i = fl % 17; int rr = fl / 17; j = rr & 15; k = rr >> 4; }
            else if (a == 1) { i = fl & 15; int rr = fl >> 4; j = rr % 17; k = rr / 17; }
            else { i = fl & 15; j = (fl >> 4) & 15; k = fl >> 8; }
            float p[3]; facepos(g, a, w[0] * 16 + i, w[1] * 16 + j, w[2] * 16 + k, g.dx, p);
            float phif = cellsample(t, (const float*)phi.fine, (const float*)phi.coarse, p[0], p[1], p[2]);
            float wg = (float)wt[a][f], m = (float)mom[a][f], v0 = u0[a].fine[f];
            u[a].fine[f] = phif >= -r ? (m + jmax(full - wg, 0.0f) * v0) / jmax(full, wg) : v0;
        }
        std::copy(u0[a].coarse, u0[a].coarse + g.ncoarse_faces(a), u[a].coarse);
        restrictfaces(t, u[a].coarse, (const float*)u[a].fine, a);
    }
    if (mom_out) for (int a = 0; a < 3; ++a) mom_out[a] = mom[a];
    if (wt_out) for (int a = 0; a < 3; ++a) wt_out[a] = wt[a];
}

// sflip.jl:2141-2144 grid2particles! over Flow{Topology} (sflip.jl:226-230 gathered).
inline void sparse_g2p(const Topo& t, Particles& ps, const float* u[3], const uint8_t* f[3], float dt, float gravity) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < ps.n; ++p) {
        float xp = ps.x[0][p], yp = ps.x[1][p], zp = ps.x[2][p];
        float vel[3], aff[3][3], found, fl, faff[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            finegather(t, u[a], a, xp, yp, zp, vel[a], aff[a], found);
            finegather(t, f[a], a, xp, yp, zp, fl, faff, found);
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

}  // namespace orc
