// TEST INFRASTRUCTURE ONLY: the sparse basin's pressure projection in the CPU oracle:
// level-0 coefficients and right-hand side (sflip.jl:1321-1448), the Galerkin hierarchy
// by aggregation (sflip.jl:1450-1570), the composite operator and its damped-Jacobi
// V-cycle (sflip.jl:1139-1311, 1572-1762), flexible Polak-Ribiere CG (sflip.jl:1858-1904),
// restarted FGMRES and the CG-then-FGMRES policy (fgmres.jl), and the velocity update
// (sflip.jl:1906-2015). Templated on the arithmetic type: double is the FP64 oracle.
#pragma once
#include "sparse_particles.hpp"
#include <cfloat>

namespace orc {

struct SolveError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <class T> struct SparsePressure {
    const Topo& t;
    const SparseGrid& g;
    int ns;                      // slots in use
    int nc[3];
    const uint8_t* tf;           // fine types (pool)
    const uint8_t* tc;           // coarse types
    // level 0
    std::vector<T> cf[3], cc[3], dDf, dDc;
    // level 1 (8^3 per block)
    std::vector<T> c1[3], dD1;
    std::vector<uint8_t> act1;
    // dense levels: D[0] is the coarse grid as level 2
    struct DL {
        int n[3];
        std::vector<T> x, b, r, tmp, c[3], dD;
        std::vector<uint8_t> act;
        size_t cells() const { return (size_t)n[0] * n[1] * n[2]; }
    };
    std::vector<DL> D;
    // level-1 vectors (fine pool ns*512 + a coarse array)
    Field<T> X1x, X1b, X1r, X1t;

    SparsePressure(const Topo& topo, const uint8_t* typesf, const uint8_t* typesc)
        : t(topo), g(topo.g), ns(topo.nslots), tf(typesf), tc(typesc) {
        for (int a = 0; a < 3; ++a) nc[a] = g.nc(a);
        int m[3] = {nc[0], nc[1], nc[2]};
        while (true) {                                                          // sflip.jl:320-326
            DL l;
            for (int a = 0; a < 3; ++a) l.n[a] = m[a];
            size_t C = l.cells();
            l.x.assign(C, T(0)); l.b.assign(C, T(0)); l.r.assign(C, T(0)); l.tmp.assign(C, T(0)); l.dD.assign(C, T(0));
            l.act.assign(C, 0);
            for (int a = 0; a < 3; ++a) l.c[a].assign((size_t)(m[0] + (a == 0)) * (m[1] + (a == 1)) * (m[2] + (a == 2)), T(0));
            D.push_back(std::move(l));
            if ((long)m[0] * m[1] * m[2] <= 1024 || (m[0] & 1) || (m[1] & 1) || (m[2] & 1)) break;
            for (int a = 0; a < 3; ++a) m[a] /= 2;
        }
        if ((long)m[0] * m[1] * m[2] > 1024) throw std::runtime_error("the coarse grid does not halve to at most 1024 cells");
        size_t C = g.ncoarse();
        for (auto* f : {&X1x, &X1b, &X1r, &X1t}) { f->fine.assign((size_t)ns * BCELLS1, T(0)); f->coarse.assign(C, T(0)); }
    }
    size_t nf() const { return (size_t)ns * BCELLS; }
    bool live(int s) const { return t.where[s - 1][0] >= 0; }

    // sflip.jl:1334-1342 celltype: a fine cell's type and whether it is fine.
    uint8_t celltype(int I, int J, int K, bool& isfine) const {
        int64_t idx = t.fineat(I, J, K);
        if (idx >= 0) { isfine = true; return tf[idx]; }
        isfine = false;
        int ci = I >> 2, cj = J >> 2, ck = K >> 2;
        if (ci < 0 || cj < 0 || ck < 0 || ci >= nc[0] || cj >= nc[1] || ck >= nc[2]) return SOLID;
        return tc[cidx(nc, ci, cj, ck)];
    }
    // sflip.jl:883-888 neighbourφ / 1908-1914 pressureat: fine value, else the coarse cell's (clamped).
    template <class V> V fineorcoarse(const V* fine, const V* coarse, int I, int J, int K) const {
        int64_t idx = t.fineat(I, J, K);
        if (idx >= 0) return fine[idx];
        return coarse[cidx(nc, std::clamp(I >> 2, 0, nc[0] - 1), std::clamp(J >> 2, 0, nc[1] - 1), std::clamp(K >> 2, 0, nc[2] - 1))];
    }

    // ── level-0 coefficients (sflip.jl:1323-1448) ──
    void coefficients(const float* phif, const float* phic) {
        for (int a = 0; a < 3; ++a) {
            cf[a].assign((size_t)ns * BFACES, T(0));
            for (int s = 1; s <= ns; ++s) {                                     // sflip.jl:1323-1332 finecoef
                if (!live(s)) continue;
                auto w = t.where[s - 1];
                for (int k = 0; k < 16 + (a == 2); ++k)
                    for (int j = 0; j < 16 + (a == 1); ++j)
                        for (int i = 0; i < 16 + (a == 0); ++i) {
                            int I = w[0] * 16 + i, J = w[1] * 16 + j, K = w[2] * 16 + k;
                            bool fp, fm;
                            uint8_t tp = celltype(I, J, K, fp), tm = celltype(I - (a == 0), J - (a == 1), K - (a == 2), fm);
                            cf[a][(size_t)(s - 1) * BFACES + fface(16, a, i, j, k)] =
                                (tp == LIQUID && tm == LIQUID) ? (T)(fp && fm ? FINE_FINE : FINE_COARSE) : T(0);
                        }
            }
            int n[3] = {nc[0] + (a == 0), nc[1] + (a == 1), nc[2] + (a == 2)};   // sflip.jl:1343-1353 coarsecoef
            cc[a].assign((size_t)n[0] * n[1] * n[2], T(0));
            for (int K = 0; K < n[2]; ++K)
                for (int J = 0; J < n[1]; ++J)
                    for (int I = 0; I < n[0]; ++I) {
                        int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
                        bool ok = Im >= 0 && Jm >= 0 && Km >= 0 && I < nc[0] && J < nc[1] && K < nc[2];
                        cc[a][cidx(n, I, J, K)] = (ok && !t.covered(I, J, K) && !t.covered(Im, Jm, Km) &&
                                                   tc[cidx(nc, I, J, K)] == LIQUID && tc[cidx(nc, Im, Jm, Km)] == LIQUID)
                                                      ? (T)COARSE_COARSE : T(0);
                    }
        }
        dDf.assign(nf(), T(0));                                                 // sflip.jl:1355-1374 finedirichlet
        for (size_t nn = 0; nn < nf(); ++nn) {
            auto w = t.where[nn >> 12];
            if (w[0] < 0 || tf[nn] != LIQUID) continue;
            int c = (int)(nn & 4095), I = w[0] * 16 + (c & 15), J = w[1] * 16 + ((c >> 4) & 15), K = w[2] * 16 + (c >> 8);
            T d = 0, pc = phif[nn];
            for (int m = 0; m < 6; ++m) {
                int a = I + NB[m][0], b = J + NB[m][1], cc2 = K + NB[m][2];
                bool isf;
                if (celltype(a, b, cc2, isf) == AIR)
                    d += (T)(isf ? FINE_FINE : FINE_COARSE) / theta<T>(pc, (T)fineorcoarse(phif, phic, a, b, cc2));
            }
            dDf[nn] = d;
        }
        dDc.assign(g.ncoarse(), T(0));                                          // sflip.jl:1375-1406 coarsedirichlet
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (tc[ci] != LIQUID || t.covered(I, J, K)) continue;
                    T d = 0, pc = phic[ci];
                    for (int a = 0; a < 3; ++a)
                        for (int sg = -1; sg <= 1; sg += 2) {
                            int In = I + sg * (a == 0), Jn = J + sg * (a == 1), Kn = K + sg * (a == 2);
                            if (In < 0 || Jn < 0 || Kn < 0 || In >= nc[0] || Jn >= nc[1] || Kn >= nc[2]) continue;
                            if (t.covered(In, Jn, Kn)) {
                                for (int q = 0; q < 4; ++q)
                                    for (int p = 0; p < 4; ++p) {
                                        int FI = a == 0 ? (sg > 0 ? 4 * In : 4 * I - 1) : 4 * I + p;
                                        int FJ = a == 1 ? (sg > 0 ? 4 * Jn : 4 * J - 1) : 4 * J + (a == 0 ? p : q);
                                        int FK = a == 2 ? (sg > 0 ? 4 * Kn : 4 * K - 1) : 4 * K + q;
                                        int64_t idx = t.fineat(FI, FJ, FK);
                                        if (idx >= 0 && tf[idx] == AIR) d += (T)FINE_COARSE / theta<T>(pc, (T)phif[idx]);
                                    }
                            } else if (tc[cidx(nc, In, Jn, Kn)] == AIR) {
                                d += (T)COARSE_COARSE / theta<T>(pc, (T)phic[cidx(nc, In, Jn, Kn)]);
                            }
                        }
                    dDc[ci] = d;
                }
        DL& d0 = D[0];                                                          // sflip.jl:1313-1319 activecoarse
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I)
                    d0.act[cidx(nc, I, J, K)] = (tc[cidx(nc, I, J, K)] == LIQUID && !t.covered(I, J, K)) ? 1 : 0;
    }

    // ── the hierarchy (sflip.jl:1450-1570) ──
    int64_t cellidx8(int s, int i, int j, int k) const { return (int64_t)(s - 1) * 512 + i + 8 * (j + 8 * k); }
    int64_t faceidx8(int s, int a, int i, int j, int k) const { return (int64_t)(s - 1) * 576 + fface(8, a, i, j, k); }
    int64_t cellidx16(int s, int i, int j, int k) const { return (int64_t)(s - 1) * 4096 + i + 16 * (j + 16 * k); }
    int64_t faceidx16(int s, int a, int i, int j, int k) const { return (int64_t)(s - 1) * 4352 + fface(16, a, i, j, k); }

    void hierarchy() {
        for (int a = 0; a < 3; ++a) {                                           // coef1
            c1[a].assign((size_t)ns * BFACES1, T(0));
            for (int s = 1; s <= ns; ++s) {
                if (!live(s)) continue;
                for (int k = 0; k < 8 + (a == 2); ++k)
                    for (int j = 0; j < 8 + (a == 1); ++j)
                        for (int i = 0; i < 8 + (a == 0); ++i) {
                            T acc = 0;
                            for (int q = 0; q < 2; ++q)
                                for (int p = 0; p < 2; ++p) {
                                    int ii = a == 0 ? 2 * i : 2 * i + p;
                                    int jj = a == 1 ? 2 * j : 2 * j + (a == 0 ? p : q);
                                    int kk = a == 2 ? 2 * k : 2 * k + q;
                                    acc += cf[a][faceidx16(s, a, ii, jj, kk)];
                                }
                            c1[a][faceidx8(s, a, i, j, k)] = acc;
                        }
            }
        }
        dD1.assign((size_t)ns * BCELLS1, T(0)); act1.assign((size_t)ns * BCELLS1, 0);   // cells1
        for (int s = 1; s <= ns; ++s) {
            if (!live(s)) continue;
            for (int k = 0; k < 8; ++k)
                for (int j = 0; j < 8; ++j)
                    for (int i = 0; i < 8; ++i) {
                        T d = 0;
                        uint8_t act = 0;
                        for (int c = 0; c < 2; ++c)
                            for (int b = 0; b < 2; ++b)
                                for (int a = 0; a < 2; ++a) {
                                    int64_t m = cellidx16(s, 2 * i + a, 2 * j + b, 2 * k + c);
                                    d += dDf[m];
                                    if (tf[m] == LIQUID) act = 1;
                                }
                        dD1[cellidx8(s, i, j, k)] = d;
                        act1[cellidx8(s, i, j, k)] = act;
                    }
        }
        DL& d2 = D[0];
        for (int a = 0; a < 3; ++a) {                                           // coef2
            int n[3] = {nc[0] + (a == 0), nc[1] + (a == 1), nc[2] + (a == 2)};
            for (int K = 0; K < n[2]; ++K)
                for (int J = 0; J < n[1]; ++J)
                    for (int I = 0; I < n[0]; ++I) {
                        size_t fi = cidx(n, I, J, K);
                        if (t.covered(I, J, K) || t.covered(I - (a == 0), J - (a == 1), K - (a == 2))) {
                            T acc = 0;
                            for (int q = 0; q < 2; ++q)
                                for (int p = 0; p < 2; ++p) {
                                    int FI = a == 0 ? 2 * I : 2 * I + p;
                                    int FJ = a == 1 ? 2 * J : 2 * J + (a == 0 ? p : q);
                                    int FK = a == 2 ? 2 * K : 2 * K + q;
                                    int64_t idx = t.faceat(8, a, FI, FJ, FK);
                                    if (idx >= 0) acc += c1[a][idx];
                                }
                            d2.c[a][fi] = acc;
                        } else {
                            d2.c[a][fi] = cc[a][fi];
                        }
                    }
        }
        for (int K = 0; K < nc[2]; ++K)                                         // cells2
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (t.covered(I, J, K)) {
                        T d = 0;
                        uint8_t act = 0;
                        for (int c = 0; c < 2; ++c)
                            for (int b = 0; b < 2; ++b)
                                for (int a = 0; a < 2; ++a) {
                                    int64_t m = t.cellat(8, 2 * I + a, 2 * J + b, 2 * K + c);
                                    d += dD1[m];
                                    if (act1[m]) act = 1;
                                }
                        d2.dD[ci] = d;
                        d2.act[ci] = act;
                    } else {
                        d2.dD[ci] = dDc[ci];
                    }
                }
        for (size_t l = 1; l < D.size(); ++l) {                                 // coefdense / cellsdense
            DL &F = D[l - 1], &Cl = D[l];
            for (int a = 0; a < 3; ++a) {
                int n[3] = {Cl.n[0] + (a == 0), Cl.n[1] + (a == 1), Cl.n[2] + (a == 2)};
                int nfd[3] = {F.n[0] + (a == 0), F.n[1] + (a == 1), F.n[2] + (a == 2)};
                for (int K = 0; K < n[2]; ++K)
                    for (int J = 0; J < n[1]; ++J)
                        for (int I = 0; I < n[0]; ++I) {
                            T acc = 0;
                            for (int q = 0; q < 2; ++q)
                                for (int p = 0; p < 2; ++p) {
                                    int FI = a == 0 ? 2 * I : 2 * I + p;
                                    int FJ = a == 1 ? 2 * J : 2 * J + (a == 0 ? p : q);
                                    int FK = a == 2 ? 2 * K : 2 * K + q;
                                    acc += F.c[a][cidx(nfd, FI, FJ, FK)];
                                }
                            Cl.c[a][cidx(n, I, J, K)] = acc;
                        }
            }
            for (int K = 0; K < Cl.n[2]; ++K)
                for (int J = 0; J < Cl.n[1]; ++J)
                    for (int I = 0; I < Cl.n[0]; ++I) {
                        T d = 0;
                        uint8_t act = 0;
                        for (int c = 0; c < 2; ++c)
                            for (int b = 0; b < 2; ++b)
                                for (int a = 0; a < 2; ++a) {
                                    size_t m = cidx(F.n, 2 * I + a, 2 * J + b, 2 * K + c);
                                    d += F.dD[m];
                                    if (F.act[m]) act = 1;
                                }
                        Cl.dD[cidx(Cl.n, I, J, K)] = d;
                        Cl.act[cidx(Cl.n, I, J, K)] = act;
                    }
        }
    }

    // ── the composite rows (sflip.jl:1139-1216) for level N in {16, 8} ──
    const std::vector<T>& fcoef(int N, int a) const { return N == 16 ? cf[a] : c1[a]; }
    const std::vector<T>& fdD(int N) const { return N == 16 ? dDf : dD1; }
    // Value accessors: raw arrays, or a probe that reads 1 at one unknown (to assemble rows).
    struct Ptrs {
        const T* f; const T* c;
        T fine(int64_t i) const { return f[i]; }
        T coarse(size_t i) const { return c[i]; }
    };
    struct Probe {
        int64_t nf, target;
        std::vector<int64_t>* seen;
        T fine(int64_t i) const { if (seen) seen->push_back(i); return i == target ? T(1) : T(0); }
        T coarse(size_t i) const { if (seen) seen->push_back(nf + (int64_t)i); return nf + (int64_t)i == target ? T(1) : T(0); }
    };
    template <class X> void finerow(int N, const X& x, int s, int i, int j, int k, T& diag, T& off) const {
        auto w = t.where[s - 1];
        int I = w[0] * N + i, J = w[1] * N + j, K = w[2] * N + k;
        int sh = N == 16 ? 2 : 1;
        int64_t per = N == 16 ? 4096 : 512;
        diag = fdD(N)[(int64_t)(s - 1) * per + i + N * (j + N * k)];
        off = 0;
        auto nv = [&](int In, int Jn, int Kn) -> T {
            int64_t idx = t.cellat(N, In, Jn, Kn);
            if (idx >= 0) return x.fine(idx);
            return x.coarse(cidx(nc, std::clamp(In >> sh, 0, nc[0] - 1), std::clamp(Jn >> sh, 0, nc[1] - 1),
                                 std::clamp(Kn >> sh, 0, nc[2] - 1)));
        };
        int64_t fper = (int64_t)(N + 1) * N * N;
        auto cof = [&](int a, int ii, int jj, int kk) { return fcoef(N, a)[(int64_t)(s - 1) * fper + fface(N, a, ii, jj, kk)]; };
        T c;
        c = cof(0, i, j, k);     if (c > 0) { diag += c; off += c * nv(I - 1, J, K); }
        c = cof(0, i + 1, j, k); if (c > 0) { diag += c; off += c * nv(I + 1, J, K); }
        c = cof(1, i, j, k);     if (c > 0) { diag += c; off += c * nv(I, J - 1, K); }
        c = cof(1, i, j + 1, k); if (c > 0) { diag += c; off += c * nv(I, J + 1, K); }
        c = cof(2, i, j, k);     if (c > 0) { diag += c; off += c * nv(I, J, K - 1); }
        c = cof(2, i, j, k + 1); if (c > 0) { diag += c; off += c * nv(I, J, K + 1); }
    }
    template <class X> void coarserow(int N, const X& x, int I, int J, int K, T& diag, T& off) const {
        int R = N / 4;
        diag = dDc[cidx(nc, I, J, K)];
        off = 0;
        for (int a = 0; a < 3; ++a) {
            int nf3[3] = {nc[0] + (a == 0), nc[1] + (a == 1), nc[2] + (a == 2)};
            for (int sg = -1; sg <= 1; sg += 2) {
                int In = I + sg * (a == 0), Jn = J + sg * (a == 1), Kn = K + sg * (a == 2);
                if (In < 0 || Jn < 0 || Kn < 0 || In >= nc[0] || Jn >= nc[1] || Kn >= nc[2]) continue;
                if (t.covered(In, Jn, Kn)) {
                    for (int q = 0; q < R; ++q)
                        for (int p = 0; p < R; ++p) {
                            int FI = a == 0 ? (sg > 0 ? R * In : R * I - 1) : R * I + p;
                            int FJ = a == 1 ? (sg > 0 ? R * Jn : R * J - 1) : R * J + (a == 0 ? p : q);
                            int FK = a == 2 ? (sg > 0 ? R * Kn : R * K - 1) : R * K + q;
                            int PI = a == 0 ? (sg > 0 ? FI : FI + 1) : FI;
                            int PJ = a == 1 ? (sg > 0 ? FJ : FJ + 1) : FJ;
                            int PK = a == 2 ? (sg > 0 ? FK : FK + 1) : FK;
                            int64_t fidx = t.faceat(N, a, PI, PJ, PK), cid = t.cellat(N, FI, FJ, FK);
                            if (fidx >= 0 && cid >= 0) {
                                T c = fcoef(N, a)[fidx];
                                diag += c;
                                off += c * x.fine(cid);
                            }
                        }
                } else {
                    T c = sg < 0 ? cc[a][cidx(nf3, I, J, K)] : cc[a][cidx(nf3, I + (a == 0), J + (a == 1), K + (a == 2))];
                    if (c > 0) { diag += c; off += c * x.coarse(cidx(nc, In, Jn, Kn)); }
                }
            }
        }
    }

    // The level-N composite operator as COO triplets over its unknowns (fine pool index, then
    // nf + coarse index): rows of the active fine cells (level 0: LIQUID) and of the active
    // uncovered coarse cells, read back through the same row functions the solver uses.
    void assemble(int N, std::vector<int64_t>& rows, std::vector<int64_t>& cols, std::vector<T>& vals) const {
        int64_t per = (int64_t)N * N * N, nfl = (int64_t)ns * per;
        const uint8_t* act = N == 16 ? nullptr : act1.data();
        auto emit = [&](int64_t row, auto&& rowfn) {
            std::vector<int64_t> seen;
            T diag, off;
            rowfn(Probe{nfl, -1, &seen}, diag, off);
            rows.push_back(row); cols.push_back(row); vals.push_back(diag);
            std::sort(seen.begin(), seen.end());
            seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
            for (int64_t col : seen) {
                rowfn(Probe{nfl, col, nullptr}, diag, off);
                if (off != 0) { rows.push_back(row); cols.push_back(col); vals.push_back(-off); }
            }
        };
        for (int s = 1; s <= ns; ++s) {
            if (!live(s)) continue;
            for (int k = 0; k < N; ++k)
                for (int j = 0; j < N; ++j)
                    for (int i = 0; i < N; ++i) {
                        int64_t n = (int64_t)(s - 1) * per + i + N * (j + N * k);
                        if (act ? !act[n] : tf[n] != LIQUID) continue;
                        emit(n, [&](const Probe& x, T& d, T& o) { finerow(N, x, s, i, j, k, d, o); });
                    }
        }
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (!D[0].act[ci] || t.covered(I, J, K)) continue;
                    emit(nfl + (int64_t)ci, [&](const Probe& x, T& d, T& o) { coarserow(N, x, I, J, K, d, o); });
                }
    }

    const uint8_t* factive(int N) const { return N == 16 ? tf : act1.data(); }

    // Jacobi (sflip.jl:1218-1243) and residual (1244-1269) on a composite level.
    void jacobi(int N, Field<T>& dst, const Field<T>& src, const Field<T>& b) const {
        int64_t per = (int64_t)N * N * N;
        const uint8_t* act = factive(N);
#pragma omp parallel for schedule(dynamic, 16)
        for (int s = 1; s <= ns; ++s) {
            if (!live(s)) continue;
            for (int k = 0; k < N; ++k)
                for (int j = 0; j < N; ++j)
                    for (int i = 0; i < N; ++i) {
                        int64_t n = (int64_t)(s - 1) * per + i + N * (j + N * k);
                        if (act[n]) {
                            T diag, off;
                            finerow(N, Ptrs{src.fine.data(), src.coarse.data()}, s, i, j, k, diag, off);
                            dst.fine[n] = diag > 0 ? src.fine[n] + (T)OMEGA * (b.fine[n] - (diag * src.fine[n] - off)) / diag : T(0);
                        } else {
                            dst.fine[n] = 0;
                        }
                    }
        }
#pragma omp parallel for schedule(static)
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (D[0].act[ci] && !t.covered(I, J, K)) {
                        T diag, off;
                        coarserow(N, Ptrs{src.fine.data(), src.coarse.data()}, I, J, K, diag, off);
                        dst.coarse[ci] = diag > 0 ? src.coarse[ci] + (T)OMEGA * (b.coarse[ci] - (diag * src.coarse[ci] - off)) / diag : T(0);
                    } else {
                        dst.coarse[ci] = 0;
                    }
                }
    }
    void residual(int N, Field<T>& r, const Field<T>& x, const Field<T>& b) const {
        int64_t per = (int64_t)N * N * N;
        const uint8_t* act = factive(N);
#pragma omp parallel for schedule(dynamic, 16)
        for (int s = 1; s <= ns; ++s) {
            if (!live(s)) continue;
            for (int k = 0; k < N; ++k)
                for (int j = 0; j < N; ++j)
                    for (int i = 0; i < N; ++i) {
                        int64_t n = (int64_t)(s - 1) * per + i + N * (j + N * k);
                        if (act[n]) {
                            T diag, off;
                            finerow(N, Ptrs{x.fine.data(), x.coarse.data()}, s, i, j, k, diag, off);
                            r.fine[n] = b.fine[n] - (diag * x.fine[n] - off);
                        } else {
                            r.fine[n] = 0;
                        }
                    }
        }
#pragma omp parallel for schedule(static)
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (D[0].act[ci] && !t.covered(I, J, K)) {
                        T diag, off;
                        coarserow(N, Ptrs{x.fine.data(), x.coarse.data()}, I, J, K, diag, off);
                        r.coarse[ci] = b.coarse[ci] - (diag * x.coarse[ci] - off);
                    } else {
                        r.coarse[ci] = 0;
                    }
                }
    }
    // q = A x at level 0 (sflip.jl:1824-1856): LIQUID fine cells, active uncovered coarse cells.
    void apply(Field<T>& q, const Field<T>& x) const {
#pragma omp parallel for schedule(dynamic, 16)
        for (int s = 1; s <= ns; ++s) {
            if (!live(s)) continue;
            for (int c = 0; c < BCELLS; ++c) {
                int64_t n = (int64_t)(s - 1) * BCELLS + c;
                if (tf[n] == LIQUID) {
                    T diag, off;
                    finerow(16, Ptrs{x.fine.data(), x.coarse.data()}, s, c & 15, (c >> 4) & 15, c >> 8, diag, off);
                    q.fine[n] = diag * x.fine[n] - off;
                } else {
                    q.fine[n] = 0;
                }
            }
        }
#pragma omp parallel for schedule(static)
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (D[0].act[ci] && !t.covered(I, J, K)) {
                        T diag, off;
                        coarserow(16, Ptrs{x.fine.data(), x.coarse.data()}, I, J, K, diag, off);
                        q.coarse[ci] = diag * x.coarse[ci] - off;
                    } else {
                        q.coarse[ci] = 0;
                    }
                }
    }

    // ── dense levels (sflip.jl:1574-1657) ──
    void denserow(const DL& L, const std::vector<T>& x, int i, int j, int k, T& diag, T& off) const {
        const int* n = L.n;
        int nu[3] = {n[0] + 1, n[1], n[2]}, nv[3] = {n[0], n[1] + 1, n[2]}, nw[3] = {n[0], n[1], n[2] + 1};
        diag = L.dD[cidx(n, i, j, k)];
        off = 0;
        T c;
        c = L.c[0][cidx(nu, i, j, k)];     if (c > 0 && i > 0) { diag += c; off += c * x[cidx(n, i - 1, j, k)]; }
        c = L.c[0][cidx(nu, i + 1, j, k)]; if (c > 0 && i < n[0] - 1) { diag += c; off += c * x[cidx(n, i + 1, j, k)]; }
        c = L.c[1][cidx(nv, i, j, k)];     if (c > 0 && j > 0) { diag += c; off += c * x[cidx(n, i, j - 1, k)]; }
        c = L.c[1][cidx(nv, i, j + 1, k)]; if (c > 0 && j < n[1] - 1) { diag += c; off += c * x[cidx(n, i, j + 1, k)]; }
        c = L.c[2][cidx(nw, i, j, k)];     if (c > 0 && k > 0) { diag += c; off += c * x[cidx(n, i, j, k - 1)]; }
        c = L.c[2][cidx(nw, i, j, k + 1)]; if (c > 0 && k < n[2] - 1) { diag += c; off += c * x[cidx(n, i, j, k + 1)]; }
    }
    void jacobidense(DL& L, std::vector<T>& dst, const std::vector<T>& src) const {
#pragma omp parallel for schedule(static)
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = 0; i < L.n[0]; ++i) {
                    size_t c = cidx(L.n, i, j, k);
                    if (!L.act[c]) { dst[c] = 0; continue; }
                    T diag, off;
                    denserow(L, src, i, j, k, diag, off);
                    dst[c] = diag > 0 ? src[c] + (T)OMEGA * (L.b[c] - (diag * src[c] - off)) / diag : T(0);
                }
    }
    void residualdense(DL& L) const {
#pragma omp parallel for schedule(static)
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = 0; i < L.n[0]; ++i) {
                    size_t c = cidx(L.n, i, j, k);
                    if (!L.act[c]) { L.r[c] = 0; continue; }
                    T diag, off;
                    denserow(L, L.x, i, j, k, diag, off);
                    L.r[c] = L.b[c] - (diag * L.x[c] - off);
                }
    }
    // sflip.jl:1623-1657 coarsest: RBGS with face coefficients, `sweeps` then as many reversed.
    void coarsest(DL& L, int sweeps) const {
        std::fill(L.x.begin(), L.x.end(), T(0));
        const int* n = L.n;
        int nu[3] = {n[0] + 1, n[1], n[2]}, nv[3] = {n[0], n[1] + 1, n[2]}, nw[3] = {n[0], n[1], n[2] + 1};
        for (int s = 1; s <= 2 * sweeps; ++s)
            for (int half = 0; half < 2; ++half) {
                int colour = s <= sweeps ? half : 1 - half;
                std::vector<T> xs = L.x;
                for (int k = 0; k < n[2]; ++k)
                    for (int j = 0; j < n[1]; ++j)
                        for (int i = 0; i < n[0]; ++i) {
                            size_t c = cidx(n, i, j, k);
                            if (!L.act[c] || (i + j + k + 1) % 2 != colour) continue;
                            T diag = L.dD[c], acc = L.b[c], cf2;
                            cf2 = L.c[0][cidx(nu, i, j, k)];     if (cf2 > 0 && i > 0) { diag += cf2; acc += cf2 * xs[c - 1]; }
                            cf2 = L.c[0][cidx(nu, i + 1, j, k)]; if (cf2 > 0 && i < n[0] - 1) { diag += cf2; acc += cf2 * xs[c + 1]; }
                            cf2 = L.c[1][cidx(nv, i, j, k)];     if (cf2 > 0 && j > 0) { diag += cf2; acc += cf2 * xs[c - n[0]]; }
                            cf2 = L.c[1][cidx(nv, i, j + 1, k)]; if (cf2 > 0 && j < n[1] - 1) { diag += cf2; acc += cf2 * xs[c + n[0]]; }
                            cf2 = L.c[2][cidx(nw, i, j, k)];     if (cf2 > 0 && k > 0) { diag += cf2; acc += cf2 * xs[c - (size_t)n[0] * n[1]]; }
                            cf2 = L.c[2][cidx(nw, i, j, k + 1)]; if (cf2 > 0 && k < n[2] - 1) { diag += cf2; acc += cf2 * xs[c + (size_t)n[0] * n[1]]; }
                            L.x[c] = diag > 0 ? acc / diag : T(0);
                        }
            }
    }

    void smooth(int N, Field<T>& x, const Field<T>& b, Field<T>& tmp, int sweeps) const {
        for (int s = 0; s < sweeps; ++s) {
            jacobi(N, tmp, x, b);
            jacobi(N, x, tmp, b);
        }
    }
    static void zero(Field<T>& f) {
        std::fill(f.fine.begin(), f.fine.end(), T(0));
        std::fill(f.coarse.begin(), f.coarse.end(), T(0));
    }

    // sflip.jl:1722-1762 vcycle!: z = M^-1 r. `q` and `tmp` are level-0 scratch fields.
    void vcycle(Field<T>& z, const Field<T>& r, Field<T>& q, Field<T>& tmp, int sweeps, int coarse, T boost) {
        zero(z);
        smooth(16, z, r, tmp, sweeps);
        residual(16, q, z, r);
        for (int s = 1; s <= ns; ++s) {                                         // restrict01
            if (!live(s)) continue;
            for (int k = 0; k < 8; ++k)
                for (int j = 0; j < 8; ++j)
                    for (int i = 0; i < 8; ++i) {
                        T acc = 0;
                        for (int c = 0; c < 2; ++c)
                            for (int b = 0; b < 2; ++b)
                                for (int a = 0; a < 2; ++a) acc += q.fine[cellidx16(s, 2 * i + a, 2 * j + b, 2 * k + c)];
                        X1b.fine[cellidx8(s, i, j, k)] = acc;
                    }
        }
        X1b.coarse = q.coarse;
        zero(X1x);
        smooth(8, X1x, X1b, X1t, sweeps);
        residual(8, X1r, X1x, X1b);
        DL& d2 = D[0];
        for (int K = 0; K < nc[2]; ++K)                                         // restrict12
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (t.covered(I, J, K)) {
                        T acc = 0;
                        for (int c = 0; c < 2; ++c)
                            for (int b = 0; b < 2; ++b)
                                for (int a = 0; a < 2; ++a) acc += X1r.fine[t.cellat(8, 2 * I + a, 2 * J + b, 2 * K + c)];
                        d2.b[ci] = acc;
                    } else {
                        d2.b[ci] = X1r.coarse[ci];
                    }
                }
        for (size_t l = 0; l + 1 < D.size(); ++l) {
            DL& L = D[l];
            std::fill(L.x.begin(), L.x.end(), T(0));
            for (int s = 0; s < sweeps; ++s) { jacobidense(L, L.tmp, L.x); jacobidense(L, L.x, L.tmp); }
            residualdense(L);
            DL& C = D[l + 1];
            for (int K = 0; K < C.n[2]; ++K)                                    // restrictdense
                for (int J = 0; J < C.n[1]; ++J)
                    for (int I = 0; I < C.n[0]; ++I) {
                        T acc = 0;
                        for (int c = 0; c < 2; ++c)
                            for (int b = 0; b < 2; ++b)
                                for (int a = 0; a < 2; ++a) acc += L.r[cidx(L.n, 2 * I + a, 2 * J + b, 2 * K + c)];
                        C.b[cidx(C.n, I, J, K)] = acc;
                    }
        }
        coarsest(D.back(), coarse);
        for (int l = (int)D.size() - 2; l >= 0; --l) {
            DL &L = D[l], &C = D[l + 1];
            for (int k = 0; k < L.n[2]; ++k)                                    // prolongdense
                for (int j = 0; j < L.n[1]; ++j)
                    for (int i = 0; i < L.n[0]; ++i) {
                        size_t c = cidx(L.n, i, j, k);
                        if (L.act[c]) L.x[c] += boost * C.x[cidx(C.n, i >> 1, j >> 1, k >> 1)];
                    }
            for (int s = 0; s < sweeps; ++s) { jacobidense(L, L.tmp, L.x); jacobidense(L, L.x, L.tmp); }
        }
        for (int s = 1; s <= ns; ++s) {                                         // prolong21fine
            if (!live(s)) continue;
            auto w = t.where[s - 1];
            for (int k = 0; k < 8; ++k)
                for (int j = 0; j < 8; ++j)
                    for (int i = 0; i < 8; ++i) {
                        int64_t n = cellidx8(s, i, j, k);
                        if (act1[n]) X1x.fine[n] += boost * d2.x[cidx(nc, (w[0] * 8 + i) >> 1, (w[1] * 8 + j) >> 1, (w[2] * 8 + k) >> 1)];
                    }
        }
        prolongcoarse(X1x.coarse, d2.x, boost);
        smooth(8, X1x, X1b, X1t, sweeps);
        for (int s = 1; s <= ns; ++s) {                                         // prolong10fine
            if (!live(s)) continue;
            for (int c = 0; c < BCELLS; ++c) {
                int64_t n = (int64_t)(s - 1) * BCELLS + c;
                int i = c & 15, j = (c >> 4) & 15, k = c >> 8;
                if (tf[n] == LIQUID) z.fine[n] += boost * X1x.fine[cellidx8(s, i >> 1, j >> 1, k >> 1)];
            }
        }
        prolongcoarse(z.coarse, X1x.coarse, boost);
        smooth(16, z, r, tmp, sweeps);
    }
    void prolongcoarse(std::vector<T>& xc, const std::vector<T>& from, T boost) const {   // sflip.jl:1694-1698
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (D[0].act[ci] && !t.covered(I, J, K)) xc[ci] += boost * from[ci];
                }
    }

    // sflip.jl:1772-1805 inner: LIQUID fine cells of live slots, LIQUID uncovered coarse cells.
    T inner(const Field<T>& a, const Field<T>& b) const {
        T s = 0;
        for (size_t n = 0; n < nf(); ++n)
            if (tf[n] == LIQUID && t.where[n >> 12][0] >= 0) s += a.fine[n] * b.fine[n];
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I) {
                    size_t ci = cidx(nc, I, J, K);
                    if (tc[ci] == LIQUID && !t.covered(I, J, K)) s += a.coarse[ci] * b.coarse[ci];
                }
        return s;
    }
    void axpy(Field<T>& y, T a, const Field<T>& x) const {   // y += a x over the used fine cells and all coarse
#pragma omp parallel for schedule(static)
        for (size_t n = 0; n < nf(); ++n) y.fine[n] += a * x.fine[n];
        for (size_t c = 0; c < y.coarse.size(); ++c) y.coarse[c] += a * x.coarse[c];
    }
    void xpby(Field<T>& y, const Field<T>& x, T b) const {   // y = x + b y
#pragma omp parallel for schedule(static)
        for (size_t n = 0; n < nf(); ++n) y.fine[n] = x.fine[n] + b * y.fine[n];
        for (size_t c = 0; c < y.coarse.size(); ++c) y.coarse[c] = x.coarse[c] + b * y.coarse[c];
    }
    Field<T> field() const {
        Field<T> f;
        f.fine.assign(nf(), T(0));
        f.coarse.assign(g.ncoarse(), T(0));
        return f;
    }

    // sflip.jl:1867-1904 pressure!: flexible (Polak-Ribiere) CG. Returns iterations; throws
    // the reference's errors.
    int fcg(Field<T>& P, const Field<T>& b, double tol, int maxiter, int sweeps, T boost, int coarse, double* relres) {
        Field<T> r = field(), z = field(), d = field(), q = field(), rold = field(), tmp = field();
        apply(r, P);
        xpby(r, b, T(-1));
        T bnorm = std::sqrt(inner(b, b));
        if (!std::isfinite((double)bnorm)) throw SolveError("Nonfinite pressure right-hand side");
        if (relres) *relres = 0;
        if (bnorm == 0) return 0;
        T target = (T)tol * bnorm;
        T rnorm = std::sqrt(inner(r, r));
        if (relres) *relres = (double)(rnorm / bnorm);
        if (rnorm <= target) return 0;
        vcycle(z, r, q, tmp, sweeps, coarse, boost);
        d = z;
        T rho = inner(z, r);
        if (!std::isfinite((double)rho)) throw SolveError("Nonfinite pressure preconditioner");
        if (!(rho > 0)) throw SolveError("Pressure CG requires a positive preconditioner");
        for (int it = 1; it <= maxiter; ++it) {
            apply(q, d);
            T den = inner(d, q);
            if (!(den > 0)) throw SolveError("Pressure CG requires a positive operator");
            T alpha = rho / den;
            if (!std::isfinite((double)alpha)) throw SolveError("Nonfinite pressure update");
            axpy(P, alpha, d);
            rold = r;
            axpy(r, -alpha, q);
            rnorm = std::sqrt(inner(r, r));
            if (relres) *relres = (double)(rnorm / bnorm);
            if (rnorm <= target) return it;
            vcycle(z, r, q, tmp, sweeps, coarse, boost);
            T rhonew = inner(z, r);
            if (!(rhonew > 0)) throw SolveError("Pressure CG requires a positive preconditioner");
            T beta = std::max(T(0), (rhonew - inner(z, rold)) / rho);
            if (!std::isfinite((double)beta)) throw SolveError("Nonfinite pressure direction");
            rho = rhonew;
            xpby(d, z, beta);
        }
        return maxiter;
    }

    // fgmres.jl:15-23 the diagonal (damped Jacobi from zero) preconditioner.
    void diagonal(Field<T>& z, const Field<T>& r) const {
        Field<T> zero0 = field();
        jacobi(16, z, zero0, r);
    }

    // fgmres.jl:26-92 pressure_fgmres!: right-preconditioned restarted FGMRES.
    int fgmres(Field<T>& P, const Field<T>& b, double tol, int maxiter, int restart, int sweeps, T boost, int coarse,
               double* relres) {
        std::vector<Field<T>> V(restart + 1, field()), Z(restart, field());
        Field<T> r = field(), q = field(), z = field(), tmp = field(), qs = field();
        T bnorm = std::sqrt(inner(b, b));
        if (!std::isfinite((double)bnorm)) throw SolveError("Nonfinite pressure RHS");
        if (bnorm == 0) { zero(P); if (relres) *relres = 0; return 0; }
        double target = tol * (double)bnorm;
        auto actual = [&]() {
            apply(q, P);
            r = b;
            axpy(r, T(-1), q);
            return std::sqrt(inner(r, r));
        };
        T residual = actual();
        int iterations = 0;
        while ((double)residual > target && iterations < maxiter) {
            if (!std::isfinite((double)residual)) throw SolveError("Nonfinite pressure residual");
            std::vector<double> H((restart + 1) * restart, 0.0), gv(restart + 1, 0.0), cs(restart, 0.0), sn(restart, 0.0);
            auto Hs = [&](int i, int j) -> double& { return H[(size_t)i + (size_t)(restart + 1) * j]; };
            gv[0] = (double)residual;
            zero(V[0]);
            axpy(V[0], T(1) / residual, r);
            int used = 0;
            for (int j = 0; j < std::min(restart, maxiter - iterations); ++j) {
                vcycle(z, V[j], qs, tmp, sweeps, coarse, boost);
                diagonal(Z[j], V[j]);
                axpy(Z[j], T(1), z);
                apply(q, Z[j]);
                for (int pass = 0; pass < 2; ++pass)
                    for (int i = 0; i <= j; ++i) {
                        T h = inner(q, V[i]);
                        Hs(i, j) += (double)h;
                        axpy(q, -h, V[i]);
                    }
                T hnext = std::sqrt(inner(q, q));
                if (!std::isfinite((double)hnext)) throw SolveError("Nonfinite pressure Arnoldi vector");
                Hs(j + 1, j) = (double)hnext;
                if (hnext > 0) { zero(V[j + 1]); axpy(V[j + 1], T(1) / hnext, q); }
                for (int i = 0; i < j; ++i) {
                    double a = Hs(i, j), bb = Hs(i + 1, j);
                    Hs(i, j) = cs[i] * a + sn[i] * bb;
                    Hs(i + 1, j) = -sn[i] * a + cs[i] * bb;
                }
                double a = Hs(j, j), bb = Hs(j + 1, j), hyp = std::hypot(a, bb);
                if (!(hyp > 0)) throw SolveError("Pressure Arnoldi breakdown");
                cs[j] = a / hyp; sn[j] = bb / hyp;
                Hs(j, j) = hyp; Hs(j + 1, j) = 0;
                gv[j + 1] = -sn[j] * gv[j]; gv[j] = cs[j] * gv[j];
                used = j + 1;
                ++iterations;
                if (std::fabs(gv[j + 1]) <= target || hnext == 0) break;
            }
            std::vector<double> y(used);
            for (int i = used - 1; i >= 0; --i) {   // UpperTriangular \ g
                double s = gv[i];
                for (int k = i + 1; k < used; ++k) s -= Hs(i, k) * y[k];
                y[i] = s / Hs(i, i);
            }
            for (int j = 0; j < used; ++j) axpy(P, (T)y[j], Z[j]);
            residual = actual();
        }
        if (relres) *relres = (double)(residual / bnorm);
        if (!((double)residual <= target))
            throw SolveError("Pressure did not converge: " + std::to_string((double)(residual / bnorm)) + " after " +
                             std::to_string(iterations) + " iterations");
        return iterations;
    }

    // fgmres.jl:94-110: flexible CG first (at most 80 iterations), FGMRES when it fails or misses.
    int policy(Field<T>& P, const Field<T>& b, double tol, int maxiter, int restart, int sweeps, T boost, int coarse,
               double* relres, bool* used_fgmres) {
        int its = 0;
        if (used_fgmres) *used_fgmres = false;
        try {
            its = fcg(P, b, tol, std::min(80, maxiter), sweeps, boost, coarse, nullptr);
        } catch (const SolveError& e) {
            std::string m = e.what();
            if (m.rfind("Nonfinite pressure", 0) != 0 && m.rfind("Pressure CG ", 0) != 0) throw;
        }
        Field<T> q = field(), r = field();
        apply(q, P);
        r = b;
        axpy(r, T(-1), q);
        T bnorm = std::sqrt(inner(b, b)), rnorm = std::sqrt(inner(r, r));
        if (rnorm <= (T)tol * bnorm) { if (relres) *relres = (double)(rnorm / bnorm); return its; }
        if (!std::isfinite((double)rnorm) || rnorm > T(2) * bnorm) zero(P);
        if (used_fgmres) *used_fgmres = true;
        return its + fgmres(P, b, tol, maxiter, restart, sweeps, boost, coarse, relres);
    }
};

// sflip.jl:1408-1435 the right-hand side; sflip.jl:1964-1973 clear.
template <class T>
inline void sparse_rhs(const Topo& t, const uint8_t* tf, const uint8_t* tc, const SFc u[3], float gamma, Field<T>& b) {
    const SparseGrid& g = t.g;
    int nc[3] = {g.nc(0), g.nc(1), g.nc(2)};
    b.fine.assign((size_t)t.nslots * BCELLS, T(0));
    b.coarse.assign(g.ncoarse(), T(0));
    T h = g.dx;
    for (int s = 1; s <= t.nslots; ++s) {
        if (t.where[s - 1][0] < 0) continue;
        for (int c = 0; c < BCELLS; ++c) {
            int64_t n = (int64_t)(s - 1) * BCELLS + c;
            if (tf[n] != LIQUID) continue;
            int i = c & 15, j = (c >> 4) & 15, k = c >> 8;
            size_t base = (size_t)(s - 1) * BFACES;
            T div = (T)u[0].fine[base + fface(16, 0, i + 1, j, k)] - (T)u[0].fine[base + fface(16, 0, i, j, k)] +
                    (T)u[1].fine[base + fface(16, 1, i, j + 1, k)] - (T)u[1].fine[base + fface(16, 1, i, j, k)] +
                    (T)u[2].fine[base + fface(16, 2, i, j, k + 1)] - (T)u[2].fine[base + fface(16, 2, i, j, k)];
            b.fine[n] = -h * div + (T)gamma * h * h;
        }
    }
    int nu[3] = {nc[0] + 1, nc[1], nc[2]}, nv[3] = {nc[0], nc[1] + 1, nc[2]}, nw[3] = {nc[0], nc[1], nc[2] + 1};
    for (int K = 0; K < nc[2]; ++K)
        for (int J = 0; J < nc[1]; ++J)
            for (int I = 0; I < nc[0]; ++I) {
                size_t ci = cidx(nc, I, J, K);
                if (tc[ci] != LIQUID || t.covered(I, J, K)) continue;
                T div = (T)u[0].coarse[cidx(nu, I + 1, J, K)] - (T)u[0].coarse[cidx(nu, I, J, K)] +
                        (T)u[1].coarse[cidx(nv, I, J + 1, K)] - (T)u[1].coarse[cidx(nv, I, J, K)] +
                        (T)u[2].coarse[cidx(nw, I, J, K + 1)] - (T)u[2].coarse[cidx(nw, I, J, K)];
                b.coarse[ci] = T(-16) * h * div + T(64) * (T)gamma * h * h;
            }
}

// sflip.jl:1916-1962 projectfine / projectcoarse, then sflip.jl:2011-2012 restrict!.
template <class T>
inline void sparse_project_faces(const SparsePressure<T>& S, const Field<T>& P, const float* phif, const float* phic,
                                 const SF u[3], uint8_t* vf[3], uint8_t* vc[3]) {
    const Topo& t = S.t;
    const SparseGrid& g = t.g;
    const int* nc = S.nc;
    for (int a = 0; a < 3; ++a) {
        for (int s = 1; s <= t.nslots; ++s) {
            auto w = t.where[s - 1];
            if (w[0] < 0) continue;
            for (int k = 0; k < 16 + (a == 2); ++k)
                for (int j = 0; j < 16 + (a == 1); ++j)
                    for (int i = 0; i < 16 + (a == 0); ++i) {
                        size_t idx = (size_t)(s - 1) * BFACES + fface(16, a, i, j, k);
                        int I = w[0] * 16 + i, J = w[1] * 16 + j, K = w[2] * 16 + k;
                        int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
                        bool fp, fm;
                        uint8_t tp = S.celltype(I, J, K, fp), tm = S.celltype(Im, Jm, Km, fm);
                        bool ok = tp != SOLID && tm != SOLID && (tp == LIQUID || tm == LIQUID);
                        if (ok) {
                            T d = (T)((fp && fm ? 1.0f : 2.5f) * g.dx);
                            T Pp = S.fineorcoarse(P.fine.data(), P.coarse.data(), I, J, K);
                            T Pm = S.fineorcoarse(P.fine.data(), P.coarse.data(), Im, Jm, Km);
                            T ph = S.fineorcoarse(phif, phic, I, J, K), pm = S.fineorcoarse(phif, phic, Im, Jm, Km);
                            T Dl = (tp == LIQUID && tm == LIQUID) ? Pp - Pm
                                   : (tm == LIQUID ? -Pm / theta<T>(pm, ph) : Pp / theta<T>(ph, pm));
                            u[a].fine[idx] = (float)((T)u[a].fine[idx] - Dl / d);
                        }
                        vf[a][idx] = ok ? 1 : 0;
                    }
        }
        int n[3] = {nc[0] + (a == 0), nc[1] + (a == 1), nc[2] + (a == 2)};
        for (int K = 0; K < n[2]; ++K)
            for (int J = 0; J < n[1]; ++J)
                for (int I = 0; I < n[0]; ++I) {
                    size_t fi = cidx(n, I, J, K);
                    int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
                    bool inside = Im >= 0 && Jm >= 0 && Km >= 0 && I < nc[0] && J < nc[1] && K < nc[2];
                    if (inside && !t.covered(I, J, K) && !t.covered(Im, Jm, Km)) {
                        uint8_t tp = S.tc[cidx(nc, I, J, K)], tm = S.tc[cidx(nc, Im, Jm, Km)];
                        bool ok = tp != SOLID && tm != SOLID && (tp == LIQUID || tm == LIQUID);
                        if (ok) {
                            T Pp = P.coarse[cidx(nc, I, J, K)], Pm = P.coarse[cidx(nc, Im, Jm, Km)];
                            T ph = phic[cidx(nc, I, J, K)], pm = phic[cidx(nc, Im, Jm, Km)];
                            T Dl = (tp == LIQUID && tm == LIQUID) ? Pp - Pm
                                   : (tm == LIQUID ? -Pm / theta<T>(pm, ph) : Pp / theta<T>(ph, pm));
                            u[a].coarse[fi] = (float)((T)u[a].coarse[fi] - Dl / (T)g.cdx());
                        }
                        vc[a][fi] = ok ? 1 : 0;
                    } else if (!inside) {
                        vc[a][fi] = 0;
                    }
                }
        restrictfaces(t, u[a].coarse, (const float*)u[a].fine, a);
        restrictfaces(t, vc[a], (const uint8_t*)vf[a], a);
    }
}

}  // namespace orc
