// TEST INFRASTRUCTURE ONLY: the sparse basin's state in the CPU oracle and the support
// phases that make its inputs: the start (sflip.jl:644-713), classification
// (sflip.jl:921-948), walls (sflip.jl:1019-1041), redistancing (sflip.jl:833-919), the
// volume (sflip.jl:1975-1996) and `refine!`'s device side (sflip.jl:457-576).
#pragma once
#include "sparse_pressure.hpp"

namespace orc {

// sflip.jl:644-650 restφ: height above the balanced surface table η (0..rmax).
inline float restphi(const float* eta, int neta, float rmax, const float p[3]) {
    float r = std::sqrt(p[0] * p[0] + p[1] * p[1]);
    float tt = std::clamp(r / rmax, 0.0f, 1.0f) * (float)(neta - 1);
    int a = std::min((int)tt, neta - 2);
    float f = tt - (float)a;
    return p[2] - (eta[a] * (1.0f - f) + eta[a + 1] * f);
}
// sflip.jl:651-655 swirlvelocity.
inline float swirl(float G, float core, const float p[3], int a) {
    float r = std::sqrt(p[0] * p[0] + p[1] * p[1]);
    float turn = r > 1e-4f ? oseen(G, core, r) / r : 0.0f;
    return a == 0 ? -turn * p[1] : (a == 1 ? turn * p[0] : 0.0f);
}

inline void faceijk(int a, int f, int& i, int& j, int& k) {   // sflip.jl:364-377 fineface
    if (a == 0) { i = f % 17; int r = f / 17; j = r & 15; k = r >> 4; }
    else if (a == 1) { i = f & 15; int r = f >> 4; j = r % 17; k = r / 17; }
    else { i = f & 15; j = (f >> 4) & 15; k = f >> 8; }
}

// sflip.jl:1030-1031 walledface.
inline bool walledface(const SparseGrid& g, const float p[3], int a, float h) {
    float e = 0.5f * h;
    float lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
    lo[a] = p[a] - e; hi[a] = p[a] + e;
    return sparse_issolid(g, lo[0], lo[1], lo[2]) || sparse_issolid(g, hi[0], hi[1], hi[2]);
}

struct SparseState {
    Topo t;
    int blocks = 0;
    int32_t band = 4, nmin = 8;
    Field<float> phi, phi0, P;
    Field<uint8_t> types;
    Field<float> u[3], u0[3];
    Field<uint8_t> fl[3];
    Particles ps[2];
    int cur = 0;
    int nc[3];

    void create(const SparseGrid& g, int nblocks_pool, int64_t capacity, int32_t band_, int32_t nmin_) {
        t.init(g, nblocks_pool);
        blocks = nblocks_pool;
        band = band_; nmin = nmin_;
        for (int a = 0; a < 3; ++a) nc[a] = g.nc(a);
        size_t nf = (size_t)blocks * BCELLS, nfa = (size_t)blocks * BFACES;
        for (auto* f : {&phi, &phi0, &P}) { f->fine.assign(nf, 0.0f); f->coarse.assign(g.ncoarse(), 0.0f); }
        types.fine.assign(nf, 0); types.coarse.assign(g.ncoarse(), 0);
        for (int a = 0; a < 3; ++a) {
            for (auto* f : {&u[a], &u0[a]}) { f->fine.assign(nfa, 0.0f); f->coarse.assign(g.ncoarse_faces(a), 0.0f); }
            fl[a].fine.assign(nfa, 0); fl[a].coarse.assign(g.ncoarse_faces(a), 0);
        }
        ps[0].resize(capacity); ps[1].resize(capacity);
        ps[0].n = ps[1].n = 0;
        cur = 0;
    }
    const SparseGrid& g() const { return t.g; }
    bool live(int s) const { return t.where[s - 1][0] >= 0; }

    // sflip.jl:674-689 the coarse grid's start.
    void init_coarse(const float* eta, int neta, float rmax, float G, float core) {
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    float p[3]; coarsecentre(g(), I, J, K, p);
                    phi.coarse[cidx(nc, I, J, K)] = restphi(eta, neta, rmax, p);
                }
        for (int a = 0; a < 3; ++a) {
            int n[3] = {nc[0] + (a == 0), nc[1] + (a == 1), nc[2] + (a == 2)};
            for (int K = 0; K < n[2]; ++K)
                for (int J = 0; J < n[1]; ++J)
                    for (int I = 0; I < n[0]; ++I) {
                        float p[3]; facepos(g(), a, I, J, K, g().cdx(), p);
                        u[a].coarse[cidx(n, I, J, K)] = swirl(G, core, p, a);
                    }
        }
    }
    // sflip.jl:690-694 the fine blocks' start.
    void init_fine(const float* eta, int neta, float rmax, float G, float core) {
        for (int s = 1; s <= t.nslots; ++s) {
            if (!live(s)) continue;
            auto w = t.where[s - 1];
            for (int c = 0; c < BCELLS; ++c) {
                float p[3]; finecentre(g(), w[0] * 16 + (c & 15), w[1] * 16 + ((c >> 4) & 15), w[2] * 16 + (c >> 8), p);
                phi.fine[(size_t)(s - 1) * BCELLS + c] = restphi(eta, neta, rmax, p);
            }
            for (int a = 0; a < 3; ++a)
                for (int f = 0; f < BFACES; ++f) {
                    int i, j, k; faceijk(a, f, i, j, k);
                    float p[3]; facepos(g(), a, w[0] * 16 + i, w[1] * 16 + j, w[2] * 16 + k, g().dx, p);
                    u[a].fine[(size_t)(s - 1) * BFACES + f] = swirl(G, core, p, a);
                }
        }
    }

    // sflip.jl:516-576 refine! after the marks: the allocator, then zeroed and prolonged slots.
    void refine(const std::vector<uint8_t>& want, std::vector<int32_t>& added, std::vector<int32_t>& removed) {
        t.refine(want, added, removed);
        for (int32_t s : removed) {
            size_t c0 = (size_t)(s - 1) * BCELLS, f0 = (size_t)(s - 1) * BFACES;
            for (auto* f : {&phi, &phi0, &P}) std::fill(f->fine.begin() + c0, f->fine.begin() + c0 + BCELLS, 0.0f);
            std::fill(types.fine.begin() + c0, types.fine.begin() + c0 + BCELLS, 0);
            for (int a = 0; a < 3; ++a) {
                for (auto* f : {&u[a], &u0[a]}) std::fill(f->fine.begin() + f0, f->fine.begin() + f0 + BFACES, 0.0f);
                std::fill(fl[a].fine.begin() + f0, fl[a].fine.begin() + f0 + BFACES, 0);
            }
        }
        for (int32_t s : added) {
            prolongcells(t, phi.fine.data(), phi.coarse.data(), s);
            prolongcells(t, P.fine.data(), P.coarse.data(), s);
            for (int a = 0; a < 3; ++a) {
                prolongfaces(t, u[a].fine.data(), u[a].coarse.data(), s, a);
                prolongfaces(t, fl[a].fine.data(), fl[a].coarse.data(), s, a);
            }
        }
    }

    // sflip.jl:921-948 classify!.
    void classify() {
        for (int s = 1; s <= t.nslots; ++s) {
            if (!live(s)) continue;
            auto w = t.where[s - 1];
            for (int c = 0; c < BCELLS; ++c) {
                size_t n = (size_t)(s - 1) * BCELLS + c;
                float p[3]; finecentre(g(), w[0] * 16 + (c & 15), w[1] * 16 + ((c >> 4) & 15), w[2] * 16 + (c >> 8), p);
                types.fine[n] = sparse_issolid(g(), p[0], p[1], p[2]) ? SOLID : (phi.fine[n] < 0.0f ? LIQUID : AIR);
            }
        }
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (t.covered(I, J, K)) {
                        bool air = false, liquid = false;
                        for (int c = 0; c < 4; ++c)
                            for (int b = 0; b < 4; ++b)
                                for (int a = 0; a < 4; ++a) {
                                    uint8_t ty = types.fine[t.fineat(4 * I + a, 4 * J + b, 4 * K + c)];
                                    air |= ty == AIR; liquid |= ty == LIQUID;
                                }
                        types.coarse[ci] = air ? AIR : (liquid ? LIQUID : SOLID);
                    } else {
                        float p[3]; coarsecentre(g(), I, J, K, p);
                        types.coarse[ci] = sparse_issolid(g(), p[0], p[1], p[2]) ? SOLID : (phi.coarse[ci] < 0.0f ? LIQUID : AIR);
                    }
                }
    }

    // sflip.jl:1019-1041 solidify!.
    void solidify() {
        for (int a = 0; a < 3; ++a) {
            for (int s = 1; s <= t.nslots; ++s) {
                if (!live(s)) continue;
                auto w = t.where[s - 1];
                for (int f = 0; f < BFACES; ++f) {
                    int i, j, k; faceijk(a, f, i, j, k);
                    float p[3]; facepos(g(), a, w[0] * 16 + i, w[1] * 16 + j, w[2] * 16 + k, g().dx, p);
                    if (walledface(g(), p, a, g().dx)) u[a].fine[(size_t)(s - 1) * BFACES + f] = 0.0f;
                }
            }
            int n[3] = {nc[0] + (a == 0), nc[1] + (a == 1), nc[2] + (a == 2)};
            float h = g().cdx();
            for (int K = 0; K < n[2]; ++K)
                for (int J = 0; J < n[1]; ++J)
                    for (int I = 0; I < n[0]; ++I) {
                        bool outside = (a == 0 && (I == 0 || I == nc[0])) || (a == 1 && (J == 0 || J == nc[1])) ||
                                       (a == 2 && (K == 0 || K == nc[2]));
                        float p[3]; facepos(g(), a, I, J, K, h, p);
                        if (outside || walledface(g(), p, a, h)) u[a].coarse[cidx(n, I, J, K)] = 0.0f;
                    }
        }
    }

    // sflip.jl:614-618 restrictvelocity!: the coarse faces under fine blocks from theirs.
    void restrict_velocity() {
        for (int a = 0; a < 3; ++a) restrictfaces(t, u[a].coarse.data(), (const float*)u[a].fine.data(), a);
    }

    // sflip.jl:1975-1996 volume (summed in double).
    double volume() const {
        double acc = 0;
        for (int s = 1; s <= t.nslots; ++s) {
            if (!live(s)) continue;
            auto w = t.where[s - 1];
            for (int c = 0; c < BCELLS; ++c) {
                float p[3]; finecentre(g(), w[0] * 16 + (c & 15), w[1] * 16 + ((c >> 4) & 15), w[2] * 16 + (c >> 8), p);
                if (!sparse_issolid(g(), p[0], p[1], p[2]))
                    acc += std::clamp(0.5f - phi.fine[(size_t)(s - 1) * BCELLS + c] / g().dx, 0.0f, 1.0f);
            }
        }
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    float p[3]; coarsecentre(g(), I, J, K, p);
                    if (t.covered(I, J, K) || sparse_issolid(g(), p[0], p[1], p[2])) continue;
                    acc += 64.0f * std::clamp(0.5f - phi.coarse[cidx(nc, I, J, K)] / g().cdx(), 0.0f, 1.0f);
                }
        double dx = g().dx;
        return acc * dx * dx * dx;
    }

    // sflip.jl:833-919 redistance!: fine Jacobi sweeps (the coarse grid read where no block
    // is), then wet coarse walls, then the covered coarse cells' means.
    void redistance(int iterations) {
        float phimax = (float)(iterations + 1) * g().dx, h = g().dx;
        for (int it = 0; it < iterations; ++it) {
            std::vector<float> out = phi.fine;
            for (int s = 1; s <= t.nslots; ++s) {
                if (!live(s)) continue;
                auto w = t.where[s - 1];
                for (int c = 0; c < BCELLS; ++c) {
                    size_t n = (size_t)(s - 1) * BCELLS + c;
                    int I = w[0] * 16 + (c & 15), J = w[1] * 16 + ((c >> 4) & 15), K = w[2] * 16 + (c >> 8);
                    float cv = phi.fine[n], p[3];
                    finecentre(g(), I, J, K, p);
                    auto nphi = [&](int a, int b, int cc) {
                        int64_t idx = t.fineat(a, b, cc);
                        if (idx >= 0) return phi.fine[idx];
                        return phi.coarse[cidx(nc, std::clamp(a >> 2, 0, nc[0] - 1), std::clamp(b >> 2, 0, nc[1] - 1),
                                               std::clamp(cc >> 2, 0, nc[2] - 1))];
                    };
                    if (sparse_issolid(g(), p[0], p[1], p[2])) {
                        float lowest = cv;
                        bool found = false;
                        for (int m = 0; m < 6; ++m) {
                            int a = I + NB[m][0], b = J + NB[m][1], cc = K + NB[m][2];
                            float q[3]; finecentre(g(), a, b, cc, q);
                            if (!sparse_issolid(g(), q[0], q[1], q[2])) {
                                float f = nphi(a, b, cc);
                                lowest = found ? std::min(lowest, f) : f;
                                found = true;
                            }
                        }
                        out[n] = found ? lowest : std::clamp(cv, -phimax, phimax);
                    } else {
                        bool inside = cv < 0.0f, frozen = false;
                        float ax = phimax, ay = phimax, az = phimax;
                        for (int m = 0; m < 6; ++m) {
                            int a = I + NB[m][0], b = J + NB[m][1], cc = K + NB[m][2];
                            float q[3]; finecentre(g(), a, b, cc, q);
                            if (!sparse_issolid(g(), q[0], q[1], q[2])) {
                                float f = nphi(a, b, cc);
                                if ((f < 0.0f) != inside) frozen = true;
                                float d = std::fabs(f);
                                if (m < 2) ax = std::min(ax, d); else if (m < 4) ay = std::min(ay, d); else az = std::min(az, d);
                            }
                        }
                        if (frozen) out[n] = cv;
                        else {
                            float d = std::min(eikonal(ax, ay, az, h), phimax);
                            out[n] = inside ? -d : d;
                        }
                    }
                }
            }
            phi.fine.swap(out);
        }
        std::vector<float> out = phi.coarse;                                      // sflip.jl:890-908 coarsewall
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    float p[3]; coarsecentre(g(), I, J, K, p);
                    if (t.covered(I, J, K) || !sparse_issolid(g(), p[0], p[1], p[2])) continue;
                    float lowest = phi.coarse[cidx(nc, I, J, K)];
                    bool found = false;
                    for (int m = 0; m < 6; ++m) {
                        int a = I + NB[m][0], b = J + NB[m][1], cc = K + NB[m][2];
                        if (a < 0 || b < 0 || cc < 0 || a >= nc[0] || b >= nc[1] || cc >= nc[2]) continue;
                        float q[3]; coarsecentre(g(), a, b, cc, q);
                        if (sparse_issolid(g(), q[0], q[1], q[2])) continue;
                        float f = phi.coarse[cidx(nc, a, b, cc)];
                        lowest = found ? std::min(lowest, f) : f;
                        found = true;
                    }
                    out[cidx(nc, I, J, K)] = lowest;
                }
        phi.coarse.swap(out);
        restrictcells(t, phi.coarse.data(), (const float*)phi.fine.data());
    }
};

}  // namespace orc
