// TEST INFRASTRUCTURE ONLY: the CPU oracle of the sparse two-resolution basin
// (`SparseBasin`, sflip.jl): grid and topology, the host block allocator of `refine!`,
// sampling, particle transfers and resampling through the topology, and the composite
// multigrid-preconditioned flexible CG / FGMRES pressure projection (fgmres.jl).
// Indices are 0-based; slots are 1-based as in the reference (slot 0 = no fine block),
// pool indices 0-based (-1 = none).
#pragma once
#include "dense.hpp"
#include <array>
#include <functional>

namespace orc {

constexpr int BLOCK = 16, RATIO = 4, CPB = 4, BCELLS = 4096, BFACES = 4352;   // sflip.jl:23-27
constexpr int BCELLS1 = 512, BFACES1 = 576;                                     // sflip.jl:294-295
constexpr float FINE_FINE = 1.0f, FINE_COARSE = 0.4f, COARSE_COARSE = 4.0f;     // sflip.jl:1071-1073
constexpr float OMEGA = 6.0f / 7.0f;                                            // sflip.jl:1074

// sflip.jl:31-53 SparseGrid.
struct SparseGrid {
    float lo[3];
    float dx;
    int nb[3];
    float radius, floor;
    int nc(int a) const { return nb[a] * CPB; }
    size_t nblocks() const { return (size_t)nb[0] * nb[1] * nb[2]; }
    size_t ncoarse() const { return (size_t)nc(0) * nc(1) * nc(2); }
    size_t ncoarse_faces(int a) const { return (size_t)(nc(0) + (a == 0)) * (nc(1) + (a == 1)) * (nc(2) + (a == 2)); }
    float cdx() const { return (float)RATIO * dx; }
};

inline SparseGrid sparsegrid(float radius, float depth, float top, float dx, int pad = 2) {
    SparseGrid g;
    float S = (float)BLOCK * dx;
    int nxy = 4 * (((int)std::ceil(2.0f * (radius + (float)pad * dx) / S) + 3) / 4);
    float loz = -depth - (float)RATIO * dx;
    int nz = 4 * (((int)std::ceil((top - loz) / S) + 3) / 4);
    float half = (float)(-nxy) * S / 2.0f;
    g.lo[0] = half; g.lo[1] = half; g.lo[2] = loz;
    g.dx = dx; g.nb[0] = nxy; g.nb[1] = nxy; g.nb[2] = nz;
    g.radius = radius; g.floor = -depth;
    return g;
}

inline int fface(int N, int a, int i, int j, int k) {   // a level-N block's face along a (sflip.jl:1081-1085)
    int n = N, m = N + 1;
    return a == 0 ? i + m * (j + n * k) : (a == 1 ? i + n * (j + m * k) : i + n * (j + n * k));
}

// sflip.jl:73-128 Topology and its index maps.
struct Topo {
    SparseGrid g;
    std::vector<int32_t> slot;                   // per block: 0, or the 1-based slot
    std::vector<std::array<int32_t, 3>> where;   // per slot: block coordinates, -1 free
    std::vector<int32_t> freel;                  // the free list (LIFO)
    int nslots = 0;                              // slots in use, free ones among them

    void init(const SparseGrid& gr, int capacity) {
        g = gr;
        slot.assign(g.nblocks(), 0);
        where.assign(capacity, {-1, -1, -1});
        freel.clear();
        nslots = 0;
    }
    bool inblocks(int bi, int bj, int bk) const {
        return bi >= 0 && bi < g.nb[0] && bj >= 0 && bj < g.nb[1] && bk >= 0 && bk < g.nb[2];
    }
    size_t blockid(int bi, int bj, int bk) const { return (size_t)bi + (size_t)g.nb[0] * (bj + (size_t)g.nb[1] * bk); }
    int32_t slotat(int sh, int I, int J, int K) const {
        int bi = I >> sh, bj = J >> sh, bk = K >> sh;
        return inblocks(bi, bj, bk) ? slot[blockid(bi, bj, bk)] : 0;
    }
    int32_t cellslot(int I, int J, int K) const { return slotat(4, I, J, K); }
    // level-N cell (N = 16 the fine cells, 8 level 1), sflip.jl:1087-1094
    int64_t cellat(int N, int I, int J, int K) const {
        int sh = N == 16 ? 4 : 3, m = N - 1;
        int32_t s = slotat(sh, I, J, K);
        return s == 0 ? -1 : (int64_t)(s - 1) * N * N * N + (I & m) + N * ((J & m) + N * (K & m));
    }
    int64_t fineat(int I, int J, int K) const { return cellat(16, I, J, K); }
    // level-N face on the low side of cell (I,J,K), in either block that holds it (sflip.jl:1096-1109, 110-126)
    int64_t faceat(int N, int a, int I, int J, int K) const {
        int sh = N == 16 ? 4 : 3, m = N - 1;
        int64_t per = (int64_t)(N + 1) * N * N;
        int bi = I >> sh, bj = J >> sh, bk = K >> sh;
        if (inblocks(bi, bj, bk)) {
            int32_t s = slot[blockid(bi, bj, bk)];
            if (s != 0) return (int64_t)(s - 1) * per + fface(N, a, I & m, J & m, K & m);
        }
        int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
        if ((Im >> sh) == bi && (Jm >> sh) == bj && (Km >> sh) == bk) return -1;
        if (!inblocks(Im >> sh, Jm >> sh, Km >> sh)) return -1;
        int32_t s = slot[blockid(Im >> sh, Jm >> sh, Km >> sh)];
        if (s == 0) return -1;
        return (int64_t)(s - 1) * per + fface(N, a, a == 0 ? N : (Im & m), a == 1 ? N : (Jm & m), a == 2 ? N : (Km & m));
    }
    int64_t finefaceat(int a, int I, int J, int K) const { return faceat(16, a, I, J, K); }
    bool covered(int I, int J, int K) const { return cellslot(I * RATIO, J * RATIO, K * RATIO) != 0; }   // sflip.jl:128

    // sflip.jl:528-551: the host allocator. `want` per block; returns (added, removed)
    // slots in order, or throws the reference's pool error.
    void refine(const std::vector<uint8_t>& want, std::vector<int32_t>& added, std::vector<int32_t>& removed) {
        size_t required = 0;
        for (uint8_t m : want) required += m != 0;
        if (required > where.size())
            throw std::runtime_error("the sparse grid needs " + std::to_string(required) + " fine blocks, more than its " +
                                     std::to_string(where.size()) + " allocated");
        added.clear(); removed.clear();
        for (size_t blk = 0; blk < want.size(); ++blk)
            if (want[blk] == 0 && slot[blk] != 0) {
                int32_t s = slot[blk];
                where[s - 1] = {-1, -1, -1};
                slot[blk] = 0;
                freel.push_back(s);
                removed.push_back(s);
            }
        for (size_t blk = 0; blk < want.size(); ++blk)
            if (want[blk] != 0 && slot[blk] == 0) {
                int32_t s;
                if (freel.empty()) s = ++nslots;
                else { s = freel.back(); freel.pop_back(); }
                where[s - 1] = {(int32_t)(blk % g.nb[0]), (int32_t)((blk / g.nb[0]) % g.nb[1]),
                                (int32_t)(blk / ((size_t)g.nb[0] * g.nb[1]))};
                slot[blk] = s;
                added.push_back(s);
            }
    }
};

// A field: its fine pool and the dense coarse array (sflip.jl:86-90).
template <class T> struct Field {
    std::vector<T> fine, coarse;
};

// Coarse-array helpers (dims n of cells, or of faces along a).
inline size_t cidx(const int n[3], int i, int j, int k) { return (size_t)i + (size_t)n[0] * (j + (size_t)n[1] * k); }

// sflip.jl:132-134 lerp8.
inline float lerp8(float c000, float c100, float c010, float c110, float c001, float c101, float c011, float c111, float fx,
                   float fy, float fz) {
    float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    return ((c000 * gx + c100 * fx) * gy + (c010 * gx + c110 * fx) * fy) * gz +
           ((c001 * gx + c101 * fx) * gy + (c011 * gx + c111 * fx) * fy) * fz;
}

// sflip.jl:136-155 cellsample: fine trilinear where all 8 taps are fine, else coarse.
template <class V>
inline float cellsample(const Topo& t, const V* fine, const V* coarse, float px, float py, float pz) {
    const SparseGrid& g = t.g;
    float q0 = (px - g.lo[0]) / g.dx - 0.5f, q1 = (py - g.lo[1]) / g.dx - 0.5f, q2 = (pz - g.lo[2]) / g.dx - 0.5f;
    float x = std::floor(q0), y = std::floor(q1), z = std::floor(q2);
    int I = (int)x, J = (int)y, K = (int)z;
    int64_t a[8];
    for (int m = 0; m < 8; ++m) a[m] = t.fineat(I + (m & 1), J + ((m >> 1) & 1), K + ((m >> 2) & 1));
    bool all = true;
    for (int m = 0; m < 8; ++m) all = all && a[m] >= 0;
    if (all)
        return lerp8((float)fine[a[0]], (float)fine[a[1]], (float)fine[a[2]], (float)fine[a[3]], (float)fine[a[4]],
                     (float)fine[a[5]], (float)fine[a[6]], (float)fine[a[7]], q0 - x, q1 - y, q2 - z);
    Arr3<const V> C{coarse, g.nc(0), g.nc(1), g.nc(2)};
    float cd = g.cdx();
    return trilinear(C, (px - g.lo[0]) / cd - 0.5f, (py - g.lo[1]) / cd - 0.5f, (pz - g.lo[2]) / cd - 0.5f);
}

// sflip.jl:157-175 facesample.
template <class V>
inline float facesample(const Topo& t, const V* fine, const V* coarse, int a, float px, float py, float pz) {
    const SparseGrid& g = t.g;
    float off[3]; faceoff(a, off);
    float q0 = (px - g.lo[0]) / g.dx - off[0], q1 = (py - g.lo[1]) / g.dx - off[1], q2 = (pz - g.lo[2]) / g.dx - off[2];
    float x = std::floor(q0), y = std::floor(q1), z = std::floor(q2);
    int I = (int)x, J = (int)y, K = (int)z;
    int64_t ix[8];
    for (int m = 0; m < 8; ++m) ix[m] = t.finefaceat(a, I + (m & 1), J + ((m >> 1) & 1), K + ((m >> 2) & 1));
    bool all = true;
    for (int m = 0; m < 8; ++m) all = all && ix[m] >= 0;
    if (all)
        return lerp8((float)fine[ix[0]], (float)fine[ix[1]], (float)fine[ix[2]], (float)fine[ix[3]], (float)fine[ix[4]],
                     (float)fine[ix[5]], (float)fine[ix[6]], (float)fine[ix[7]], q0 - x, q1 - y, q2 - z);
    Arr3<const V> C{coarse, g.nc(0) + (a == 0), g.nc(1) + (a == 1), g.nc(2) + (a == 2)};
    float cd = g.cdx();
    return trilinear(C, (px - g.lo[0]) / cd - off[0], (py - g.lo[1]) / cd - off[1], (pz - g.lo[2]) / cd - off[2]);
}

// sflip.jl:193-215 finegather: APIC gather over fine faces only; `found` the weight found.
template <class V>
inline void finegather(const Topo& t, const V* u, int a, float px, float py, float pz, float& vel, float aff[3],
                       float& found) {
    const SparseGrid& g = t.g;
    float off[3]; faceoff(a, off);
    float q0 = (px - g.lo[0]) / g.dx - off[0], q1 = (py - g.lo[1]) / g.dx - off[1], q2 = (pz - g.lo[2]) / g.dx - off[2];
    BS bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
    vel = 0.0f; aff[0] = aff[1] = aff[2] = 0.0f; found = 0.0f;
    for (int mz = 0; mz < 3; ++mz)
        for (int my = 0; my < 3; ++my)
            for (int mx = 0; mx < 3; ++mx) {
                int64_t idx = t.finefaceat(a, bx.b + mx, by.b + my, bz.b + mz);
                if (idx < 0) continue;
                float wgt = bx.w[mx] * by.w[my] * bz.w[mz];
                float val = (float)u[idx];
                vel += wgt * val;
                float tt = wgt * val;
                aff[0] += (tt * ((float)(bx.b + mx) - q0)) * g.dx;
                aff[1] += (tt * ((float)(by.b + my) - q1)) * g.dx;
                aff[2] += (tt * ((float)(bz.b + mz) - q2)) * g.dx;
                found += wgt;
            }
    float c = 4.0f / (g.dx * g.dx);
    aff[0] *= c; aff[1] *= c; aff[2] *= c;
}

inline bool sparse_issolid(const SparseGrid& g, float px, float py, float pz) { return issolid(g.radius, g.floor, px, py, pz); }
inline void finecentre(const SparseGrid& g, int I, int J, int K, float p[3]) {
    p[0] = g.lo[0] + ((float)I + 0.5f) * g.dx; p[1] = g.lo[1] + ((float)J + 0.5f) * g.dx; p[2] = g.lo[2] + ((float)K + 0.5f) * g.dx;
}
inline void coarsecentre(const SparseGrid& g, int I, int J, int K, float p[3]) {
    float cd = g.cdx();
    p[0] = g.lo[0] + ((float)I + 0.5f) * cd; p[1] = g.lo[1] + ((float)J + 0.5f) * cd; p[2] = g.lo[2] + ((float)K + 0.5f) * cd;
}
inline void facepos(const SparseGrid& g, int a, int I, int J, int K, float h, float p[3]) {
    float off[3]; faceoff(a, off);
    p[0] = g.lo[0] + ((float)I + off[0]) * h; p[1] = g.lo[1] + ((float)J + off[1]) * h; p[2] = g.lo[2] + ((float)K + off[2]) * h;
}

// ─── refinement marks (sflip.jl:384-455) ─────────────────────────────────────

// sflip.jl:384-399 marksurface.
inline void marksurface(const SparseGrid& g, const float* phic, uint8_t* marks) {
    int n[3] = {g.nc(0), g.nc(1), g.nc(2)};
    for (int K = 0; K < n[2]; ++K)
        for (int J = 0; J < n[1]; ++J)
            for (int I = 0; I < n[0]; ++I) {
                float p[3]; coarsecentre(g, I, J, K, p);
                if (sparse_issolid(g, p[0], p[1], p[2])) continue;
                float c = phic[cidx(n, I, J, K)];
                bool surf = std::fabs(c) < 1.5f * g.cdx();
                for (int m = 0; m < 6; ++m) {
                    int a = I + NB[m][0], b = J + NB[m][1], cc = K + NB[m][2];
                    if (a < 0 || b < 0 || cc < 0 || a >= n[0] || b >= n[1] || cc >= n[2]) continue;
                    float q[3]; coarsecentre(g, a, b, cc, q);
                    if (sparse_issolid(g, q[0], q[1], q[2])) continue;
                    if ((phic[cidx(n, a, b, cc)] < 0.0f) != (c < 0.0f)) surf = true;
                }
                if (surf) marks[(size_t)(I >> 2) + (size_t)g.nb[0] * ((J >> 2) + (size_t)g.nb[1] * (K >> 2))] = 1;
            }
}

// sflip.jl:401-424 markflow.
inline void markflow(const SparseGrid& g, const float* uc, const float* vc, const float* wc, const float* phic, float vary,
                     uint8_t* marks) {
    int n[3] = {g.nc(0), g.nc(1), g.nc(2)};
    int nu[3] = {n[0] + 1, n[1], n[2]}, nv[3] = {n[0], n[1] + 1, n[2]}, nw[3] = {n[0], n[1], n[2] + 1};
    auto cellvel = [&](int i, int j, int k, float v[3]) {
        v[0] = 0.5f * (uc[cidx(nu, i, j, k)] + uc[cidx(nu, i + 1, j, k)]);
        v[1] = 0.5f * (vc[cidx(nv, i, j, k)] + vc[cidx(nv, i, j + 1, k)]);
        v[2] = 0.5f * (wc[cidx(nw, i, j, k)] + wc[cidx(nw, i, j, k + 1)]);
    };
    for (int K = 0; K < n[2]; ++K)
        for (int J = 0; J < n[1]; ++J)
            for (int I = 0; I < n[0]; ++I) {
                float p[3]; coarsecentre(g, I, J, K, p);
                if (!(phic[cidx(n, I, J, K)] < 0.0f) || sparse_issolid(g, p[0], p[1], p[2])) continue;
                float c[3]; cellvel(I, J, K, c);
                float jump = 0.0f;
                for (int m = 0; m < 6; ++m) {
                    int a = I + NB[m][0], b = J + NB[m][1], cc = K + NB[m][2];
                    if (a < 0 || b < 0 || cc < 0 || a >= n[0] || b >= n[1] || cc >= n[2]) continue;
                    float q[3]; coarsecentre(g, a, b, cc, q);
                    if (!(phic[cidx(n, a, b, cc)] < 0.0f) || sparse_issolid(g, q[0], q[1], q[2])) continue;
                    float o[3]; cellvel(a, b, cc, o);
                    float dx = o[0] - c[0], dy = o[1] - c[1], dz = o[2] - c[2];
                    jump = jmax(jump, std::sqrt(dx * dx + dy * dy + dz * dz));
                }
                if (jump > vary) marks[(size_t)(I >> 2) + (size_t)g.nb[0] * ((J >> 2) + (size_t)g.nb[1] * (K >> 2))] = 1;
            }
}

// sflip.jl:426-445 markparticles.
inline void markparticles(const SparseGrid& g, const float* x, const float* y, const float* z, int64_t n, float reach,
                          uint8_t* marks) {
    float S = (float)BLOCK * g.dx;
    for (int64_t p = 0; p < n; ++p) {
        if (!(std::isfinite(x[p]) && std::isfinite(y[p]) && std::isfinite(z[p]))) continue;
        for (int c = 0; c < 2; ++c)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    float q0 = (x[p] - g.lo[0] + (float)(2 * a - 1) * reach) / S;
                    float q1 = (y[p] - g.lo[1] + (float)(2 * b - 1) * reach) / S;
                    float q2 = (z[p] - g.lo[2] + (float)(2 * c - 1) * reach) / S;
                    int bi = std::clamp((int)std::floor(q0), 0, g.nb[0] - 1);
                    int bj = std::clamp((int)std::floor(q1), 0, g.nb[1] - 1);
                    int bk = std::clamp((int)std::floor(q2), 0, g.nb[2] - 1);
                    marks[(size_t)bi + (size_t)g.nb[0] * (bj + (size_t)g.nb[1] * bk)] = 1;
                }
    }
}

// sflip.jl:447-455 dilate.
inline std::vector<uint8_t> dilate(const SparseGrid& g, const std::vector<uint8_t>& marks) {
    std::vector<uint8_t> out(marks.size(), 0);
    for (int k = 0; k < g.nb[2]; ++k)
        for (int j = 0; j < g.nb[1]; ++j)
            for (int i = 0; i < g.nb[0]; ++i) {
                uint8_t m = 0;
                for (int dk = -1; dk <= 1; ++dk)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di) {
                            int a = i + di, b = j + dj, c = k + dk;
                            if (a < 0 || b < 0 || c < 0 || a >= g.nb[0] || b >= g.nb[1] || c >= g.nb[2]) continue;
                            if (marks[(size_t)a + (size_t)g.nb[0] * (b + (size_t)g.nb[1] * c)]) m = 1;
                        }
                out[(size_t)i + (size_t)g.nb[0] * (j + (size_t)g.nb[1] * k)] = m;
            }
    return out;
}

// ─── fine <-> coarse (sflip.jl:457-640) ──────────────────────────────────────

// sflip.jl:468-478 prolongcells: a new block's fine cells from the coarse grid (trilinear).
template <class V> inline void prolongcells(const Topo& t, V* fine, const V* coarse, int32_t s) {
    const SparseGrid& g = t.g;
    auto w = t.where[s - 1];
    Arr3<const V> C{coarse, g.nc(0), g.nc(1), g.nc(2)};
    float cd = g.cdx();
    for (int c = 0; c < BCELLS; ++c) {
        int i = c & 15, j = (c >> 4) & 15, k = c >> 8;
        float p[3]; finecentre(g, w[0] * 16 + i, w[1] * 16 + j, w[2] * 16 + k, p);
        float v = trilinear(C, (p[0] - g.lo[0]) / cd - 0.5f, (p[1] - g.lo[1]) / cd - 0.5f, (p[2] - g.lo[2]) / cd - 0.5f);
        fine[(size_t)(s - 1) * BCELLS + c] = std::is_same<V, uint8_t>::value ? (V)(v > 0.999f ? 1 : 0) : (V)v;
    }
}
// sflip.jl:480-496 prolongfaces.
template <class V> inline void prolongfaces(const Topo& t, V* fine, const V* coarse, int32_t s, int a) {
    const SparseGrid& g = t.g;
    auto w = t.where[s - 1];
    Arr3<const V> C{coarse, g.nc(0) + (a == 0), g.nc(1) + (a == 1), g.nc(2) + (a == 2)};
    float cd = g.cdx(), off[3];
    faceoff(a, off);
    for (int f = 0; f < BFACES; ++f) {
        int i, j, k;
        if (a == 0) { i = f % 17; int r = f / 17; j = r & 15; k = r >> 4; }
        else if (a == 1) { i = f & 15; int r = f >> 4; j = r % 17; k = r / 17; }
        else { i = f & 15; j = (f >> 4) & 15; k = f >> 8; }
        float p[3]; facepos(g, a, w[0] * 16 + i, w[1] * 16 + j, w[2] * 16 + k, g.dx, p);
        float v = trilinear(C, (p[0] - g.lo[0]) / cd - off[0], (p[1] - g.lo[1]) / cd - off[1], (p[2] - g.lo[2]) / cd - off[2]);
        fine[(size_t)(s - 1) * BFACES + fface(16, a, i, j, k)] = std::is_same<V, uint8_t>::value ? (V)(v > 0.999f ? 1 : 0) : (V)v;
    }
}

// sflip.jl:583-594 restrictcells: covered coarse cells = mean of 64 fine (flags: all).
template <class V> inline void restrictcells(const Topo& t, V* coarse, const V* fine) {
    const SparseGrid& g = t.g;
    int n[3] = {g.nc(0), g.nc(1), g.nc(2)};
    for (int K = 0; K < n[2]; ++K)
        for (int J = 0; J < n[1]; ++J)
            for (int I = 0; I < n[0]; ++I) {
                if (!t.covered(I, J, K)) continue;
                float acc = 0.0f;
                for (int c = 0; c < 4; ++c)
                    for (int b = 0; b < 4; ++b)
                        for (int a = 0; a < 4; ++a) acc += (float)fine[t.fineat(4 * I + a, 4 * J + b, 4 * K + c)];
                coarse[cidx(n, I, J, K)] = std::is_same<V, uint8_t>::value ? (V)(acc >= 63.5f ? 1 : 0) : (V)(acc / 64.0f);
            }
}
// sflip.jl:596-612 restrictfaces: a coarse face = mean of its 16 fine faces when all exist.
template <class V> inline void restrictfaces(const Topo& t, V* coarse, const V* fine, int a) {
    const SparseGrid& g = t.g;
    int n[3] = {g.nc(0) + (a == 0), g.nc(1) + (a == 1), g.nc(2) + (a == 2)};
    for (int K = 0; K < n[2]; ++K)
        for (int J = 0; J < n[1]; ++J)
            for (int I = 0; I < n[0]; ++I) {
                float acc = 0.0f;
                int found = 0;
                for (int q = 0; q < 4; ++q)
                    for (int p = 0; p < 4; ++p) {
                        int FI = 4 * I + (a == 0 ? 0 : p);
                        int FJ = 4 * J + (a == 0 ? p : (a == 1 ? 0 : q));
                        int FK = 4 * K + (a == 2 ? 0 : q);
                        int64_t idx = t.finefaceat(a, FI, FJ, FK);
                        if (idx >= 0) { acc += (float)fine[idx]; ++found; }
                    }
                if (found == 16)
                    coarse[cidx(n, I, J, K)] = std::is_same<V, uint8_t>::value ? (V)(acc >= 15.5f ? 1 : 0) : (V)(acc / 16.0f);
            }
}

// sflip.jl:620-640 syncplanes: a block's high-plane faces take the low-plane faces of the
// fine block above it.
template <class V> inline void syncplanes(const Topo& t, V* A, int a) {
    for (int s = 1; s <= t.nslots; ++s) {
        auto w = t.where[s - 1];
        if (w[0] < 0) continue;
        int b0 = w[0] + (a == 0), b1 = w[1] + (a == 1), b2 = w[2] + (a == 2);
        if (!t.inblocks(b0, b1, b2)) continue;
        int32_t s2 = t.slot[t.blockid(b0, b1, b2)];
        if (s2 == 0) continue;
        for (int f = 0; f < 256; ++f) {
            int p = f & 15, q = f >> 4;
            int hi = a == 0 ? fface(16, 0, 16, p, q) : (a == 1 ? fface(16, 1, p, 16, q) : fface(16, 2, p, q, 16));
            int lo = a == 0 ? fface(16, 0, 0, p, q) : (a == 1 ? fface(16, 1, p, 0, q) : fface(16, 2, p, q, 0));
            A[(size_t)(s - 1) * BFACES + hi] = A[(size_t)(s2 - 1) * BFACES + lo];
        }
    }
}

}  // namespace orc
