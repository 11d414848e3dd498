// Device state of one sparse two-resolution basin (the reference's `SparseBasin`,
// sflip.jl:242-347) and its phases.
//
// Layout in HBM (0-based throughout):
//  * the block table `slot` (nb0*nb1*nb2 int32: 0, or the 1-based pool slot) and per
//    slot its block coordinates `where` (int4, x = -1 for a free slot);
//  * fine pools, slot-major: a slot's 4096 cells x fastest (i + 16 (j + 16 k)), its
//    faces along axis a in (16+ea) x (16+eb) x (16+ec) order, 4352 per slot and axis;
//    a face on the plane between two fine blocks is held by both, the upper block's
//    copy canonical (sflip.jl:110-126);
//  * the coarse grid (cells 4 fine cells wide) as dense arrays, x fastest;
//  * level 1 of the multigrid at 8^3 a block (512 cells, 576 faces per axis), and the
//    dense halvings of the coarse grid.
#pragma once
#include "dense.hpp"
#include "stir.hpp"
#include <array>
#include <type_traits>

namespace mael {

constexpr int SBLOCK = 16, SRATIO = 4, SBCELLS = 4096, SBFACES = 4352, SBCELLS1 = 512, SBFACES1 = 576;
constexpr float FINE_FINE = 1.0f, FINE_COARSE = 0.4f, COARSE_COARSE = 4.0f;   // sflip.jl:1071-1073
constexpr float OMEGA = 6.0f / 7.0f;                                            // sflip.jl:1074

struct SGrid {   // sflip.jl:38-44
    float lo[3];
    float dx;
    int nb[3];
    float radius, floor;
    __host__ __device__ int nc(int a) const { return nb[a] * 4; }
    __host__ __device__ float cdx() const { return 4.0f * dx; }
};

// The topology as kernels see it (sflip.jl:78-84).
struct STopo {
    SGrid g;
    const int32_t* slot;
    const int4* where;
};

__host__ __device__ inline int ffn(int N, int a, int i, int j, int k) {   // a level-N block's face (sflip.jl:1081-1085)
    return a == 0 ? i + (N + 1) * (j + N * k) : (a == 1 ? i + N * (j + (N + 1) * k) : i + N * (j + N * k));
}
__host__ __device__ inline size_t cl(int i, int j, int k, int n0, int n1) {
    return (size_t)i + (size_t)n0 * ((size_t)j + (size_t)n1 * k);
}

__device__ inline bool inblocks(const SGrid& g, int bi, int bj, int bk) {
    return bi >= 0 && bi < g.nb[0] && bj >= 0 && bj < g.nb[1] && bk >= 0 && bk < g.nb[2];
}
__device__ inline int blockid(const SGrid& g, int bi, int bj, int bk) { return bi + g.nb[0] * (bj + g.nb[1] * bk); }

// Level-N cell (N = 16: the fine cells; 8: level 1) at global (I, J, K), or -1 (sflip.jl:1087-1094).
template <int N> __device__ inline int32_t cellat(const STopo& t, int I, int J, int K) {
    constexpr int sh = N == 16 ? 4 : 3, m = N - 1;
    int bi = I >> sh, bj = J >> sh, bk = K >> sh;
    if (!inblocks(t.g, bi, bj, bk)) return -1;
    int32_t s = t.slot[blockid(t.g, bi, bj, bk)];
    return s == 0 ? -1 : (s - 1) * (N * N * N) + (I & m) + N * ((J & m) + N * (K & m));
}
__device__ inline int32_t fineat(const STopo& t, int I, int J, int K) { return cellat<16>(t, I, J, K); }

// Level-N face along a on the low side of cell (I, J, K), in either block that holds it;
// the upper block's copy first (sflip.jl:1096-1109, 110-126).
template <int N> __device__ inline int32_t faceat(const STopo& t, int a, int I, int J, int K) {
    constexpr int sh = N == 16 ? 4 : 3, m = N - 1, per = (N + 1) * N * N;
    int bi = I >> sh, bj = J >> sh, bk = K >> sh;
    if (inblocks(t.g, bi, bj, bk)) {
        int32_t s = t.slot[blockid(t.g, bi, bj, bk)];
        if (s != 0) return (s - 1) * per + ffn(N, a, I & m, J & m, K & m);
    }
    int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
    if ((Im >> sh) == bi && (Jm >> sh) == bj && (Km >> sh) == bk) return -1;
    if (!inblocks(t.g, Im >> sh, Jm >> sh, Km >> sh)) return -1;
    int32_t s = t.slot[blockid(t.g, Im >> sh, Jm >> sh, Km >> sh)];
    if (s == 0) return -1;
    return (s - 1) * per + ffn(N, a, a == 0 ? N : (Im & m), a == 1 ? N : (Jm & m), a == 2 ? N : (Km & m));
}
__device__ inline int32_t finefaceat(const STopo& t, int a, int I, int J, int K) { return faceat<16>(t, a, I, J, K); }
// flip.jl:513-537 over the topology: the deposit's nodes are fine faces (the upper block's
// copy on a plane between blocks), as the tiled fixed-point reduction indexes them.
struct SparseNodes {
    STopo t;
    __device__ int32_t index(int a, int I, int J, int K) const { return finefaceat(t, a, I, J, K); }
};
// Whether coarse cell (I, J, K) lies in a fine block (sflip.jl:127-128).
__device__ inline bool covered(const STopo& t, int I, int J, int K) {
    int bi = I >> 2, bj = J >> 2, bk = K >> 2;
    return inblocks(t.g, bi, bj, bk) && t.slot[blockid(t.g, bi, bj, bk)] != 0;
}

// A thread's fine cell n of the pool: slot (1-based), local i, j, k and the block coordinates.
struct SCell { int s, i, j, k; int4 w; };
__device__ inline SCell finecell(const STopo& t, int32_t n) {
    SCell c;
    c.s = (n >> 12) + 1;
    int l = n & 4095;
    c.i = l & 15; c.j = (l >> 4) & 15; c.k = l >> 8;
    c.w = t.where[c.s - 1];
    return c;
}
// A thread's fine face n along a (sflip.jl:364-377).
__device__ inline SCell fineface(const STopo& t, int a, int32_t n) {
    SCell c;
    c.s = n / SBFACES + 1;
    int f = n - (c.s - 1) * SBFACES;
    if (a == 0) { c.i = f % 17; int r = f / 17; c.j = r & 15; c.k = r >> 4; }
    else if (a == 1) { c.i = f & 15; int r = f >> 4; c.j = r % 17; c.k = r / 17; }
    else { c.i = f & 15; c.j = (f >> 4) & 15; c.k = f >> 8; }
    c.w = t.where[c.s - 1];
    return c;
}

__device__ inline bool sissolid(const SGrid& g, float px, float py, float pz) { return issolid(g.radius, g.floor, px, py, pz); }
__device__ inline void finecentre(const SGrid& g, int I, int J, int K, float p[3]) {
    p[0] = g.lo[0] + ((float)I + 0.5f) * g.dx; p[1] = g.lo[1] + ((float)J + 0.5f) * g.dx; p[2] = g.lo[2] + ((float)K + 0.5f) * g.dx;
}
__device__ inline void coarsecentre(const SGrid& g, int I, int J, int K, float p[3]) {
    float cd = g.cdx();
    p[0] = g.lo[0] + ((float)I + 0.5f) * cd; p[1] = g.lo[1] + ((float)J + 0.5f) * cd; p[2] = g.lo[2] + ((float)K + 0.5f) * cd;
}
__device__ inline void facepos(const SGrid& g, int a, int I, int J, int K, float h, float p[3]) {
    p[0] = g.lo[0] + ((float)I + (a == 0 ? 0.0f : 0.5f)) * h;
    p[1] = g.lo[1] + ((float)J + (a == 1 ? 0.0f : 0.5f)) * h;
    p[2] = g.lo[2] + ((float)K + (a == 2 ? 0.0f : 0.5f)) * h;
}

// Block slot lookups for the samplers: the global table, or a workgroup's LDS table of the
// 27 blocks around a home block w (as stage27 fills it) with the global table beyond it.
// Both return the table's entry (0 off the grid), so samples are the same either way.
struct SlotsGlobal {
    __device__ int32_t operator()(const STopo& t, int bx, int by, int bz) const {
        return inblocks(t.g, bx, by, bz) ? t.slot[blockid(t.g, bx, by, bz)] : 0;
    }
};
struct Slots27 {
    const int32_t* s27;
    int4 w;
    __device__ int32_t operator()(const STopo& t, int bx, int by, int bz) const {
        const int di = bx - w.x + 1, dj = by - w.y + 1, dk = bz - w.z + 1;
        if ((unsigned)di < 3u && (unsigned)dj < 3u && (unsigned)dk < 3u) return s27[di + 3 * dj + 9 * dk];
        return inblocks(t.g, bx, by, bz) ? t.slot[blockid(t.g, bx, by, bz)] : 0;
    }
};

// The slots (1-based, 0 for none) of the fine blocks that hold the eight trilinear taps
// I..I+1, J..J+1, K..K+1 (tap m at +(m & 1, m >> 1 & 1, m >> 2)): the lower corner's block,
// and across a block side for the taps past it. When all eight are allocated, fineat and
// finefaceat return each tap's entry in its own block (the upper block's copy of a face
// comes first), so the taps index from these slots with no further table reads.
template <class L = SlotsGlobal>
__device__ inline bool tapslots(const STopo& t, int I, int J, int K, int32_t s[8], const L& lk = L()) {
    const SGrid& g = t.g;
    const int bx = I >> 4, by = J >> 4, bz = K >> 4;
    const int sx = (I & 15) == 15, sy = (J & 15) == 15, sz = (K & 15) == 15;
    const bool inx[2] = {bx >= 0 && bx < g.nb[0], bx + sx >= 0 && bx + sx < g.nb[0]};
    const bool iny[2] = {by >= 0 && by < g.nb[1], by + sy >= 0 && by + sy < g.nb[1]};
    const bool inz[2] = {bz >= 0 && bz < g.nb[2], bz + sz >= 0 && bz + sz < g.nb[2]};
    const int nx = g.nb[0], nxy = g.nb[0] * g.nb[1];
    const int id0 = bx + nx * by + nxy * bz;
    // one read per distinct block: a tap past no block side shares the slot of the tap
    // with its offsets along the unstraddled axes dropped
    auto look = [&](int dx, int dy, int dz) -> int32_t {
        if constexpr (std::is_same<L, SlotsGlobal>::value)
            return inx[dx] && iny[dy] && inz[dz] ? t.slot[id0 + dx + nx * dy + nxy * dz] : 0;
        else
            return lk(t, bx + dx, by + dy, bz + dz);
    };
    s[0] = look(0, 0, 0);
    s[1] = sx ? look(1, 0, 0) : s[0];
    s[2] = sy ? look(0, 1, 0) : s[0];
    s[3] = sx && sy ? look(1, 1, 0) : (sx ? s[1] : s[2]);
    s[4] = sz ? look(0, 0, 1) : s[0];
    s[5] = sz ? (sx ? look(1, 0, 1) : s[4]) : s[1];
    s[6] = sz ? (sy ? look(0, 1, 1) : s[4]) : s[2];
    s[7] = sz ? (sx && sy ? look(1, 1, 1) : (sx ? s[5] : s[6])) : s[3];
    bool all = true;
#pragma unroll
    for (int m = 0; m < 8; ++m) all = all && s[m] != 0;
    return all;
}

// sflip.jl:132-134 lerp8.
__device__ inline float lerp8(const float c[8], float fx, float fy, float fz) {
    float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    return ((c[0] * gx + c[1] * fx) * gy + (c[2] * gx + c[3] * fx) * fy) * gz +
           ((c[4] * gx + c[5] * fx) * gy + (c[6] * gx + c[7] * fx) * fy) * fz;
}

// sflip.jl:136-155 cellsample: fine trilinear where all 8 taps are fine, else the coarse grid's.
template <class V, class L = SlotsGlobal>
__device__ inline float cellsample(const STopo& t, const V* __restrict__ fine, const V* __restrict__ coarse, float px, float py,
                                   float pz, const L& lk = L()) {
    const SGrid& g = t.g;
    float q0 = __fdiv_rn(px - g.lo[0], g.dx) - 0.5f, q1 = __fdiv_rn(py - g.lo[1], g.dx) - 0.5f,
          q2 = __fdiv_rn(pz - g.lo[2], g.dx) - 0.5f;
    float x = floorf(q0), y = floorf(q1), z = floorf(q2);
    int I = (int)x, J = (int)y, K = (int)z;
    int32_t s8[8];
    if (tapslots(t, I, J, K, s8, lk)) {   // the 8 taps in allocated blocks: fineat's indices
        float c[8];
#pragma unroll
        for (int m = 0; m < 8; ++m)
            c[m] = (float)fine[(s8[m] - 1) * SBCELLS + ((I + (m & 1)) & 15) +
                               16 * (((J + ((m >> 1) & 1)) & 15) + 16 * ((K + (m >> 2)) & 15))];
        return lerp8(c, q0 - x, q1 - y, q2 - z);
    }
    float cd = g.cdx();
    return trilinear(coarse, g.nc(0), g.nc(1), g.nc(2), __fdiv_rn(px - g.lo[0], cd) - 0.5f, __fdiv_rn(py - g.lo[1], cd) - 0.5f,
                     __fdiv_rn(pz - g.lo[2], cd) - 0.5f);
}

// sflip.jl:157-175 facesample, for faces along A. A tap whose own block is not allocated
// can still be fine: the top face of the block below along A, held there at coordinate 16
// (finefaceat's second choice). That block is the one of the tap's neighbour below along
// A when the taps straddle the plane, else one table read.
template <int A, class V, class L = SlotsGlobal>
__device__ inline float facesample_a(const STopo& t, const V* __restrict__ fine, const V* __restrict__ coarse, float px,
                                     float py, float pz, const L& lk = L()) {
    const SGrid& g = t.g;
    constexpr float o0 = A == 0 ? 0.0f : 0.5f, o1 = A == 1 ? 0.0f : 0.5f, o2 = A == 2 ? 0.0f : 0.5f;
    float q0 = __fdiv_rn(px - g.lo[0], g.dx) - o0, q1 = __fdiv_rn(py - g.lo[1], g.dx) - o1, q2 = __fdiv_rn(pz - g.lo[2], g.dx) - o2;
    float x = floorf(q0), y = floorf(q1), z = floorf(q2);
    int I = (int)x, J = (int)y, K = (int)z;
    int32_t s8[8];
    bool all = tapslots(t, I, J, K, s8, lk);
    int32_t ix[8];
#pragma unroll
    for (int m = 0; m < 8; ++m)
        ix[m] = (s8[m] - 1) * SBFACES + ffn(16, A, (I + (m & 1)) & 15, (J + ((m >> 1) & 1)) & 15, (K + (m >> 2)) & 15);
    if (!all) {
        all = true;
        const int Ic = A == 0 ? I : (A == 1 ? J : K);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if (s8[m] != 0) continue;
            const int dc = (m >> A) & 1;
            int32_t lower = 0;
            if (((Ic + dc) & 15) == 0) {
                if (dc == 1) {
                    lower = s8[m ^ (1 << A)];
                } else {
                    int bx = (I + (m & 1)) >> 4, by = (J + ((m >> 1) & 1)) >> 4, bz = (K + (m >> 2)) >> 4;
                    bx -= A == 0; by -= A == 1; bz -= A == 2;
                    lower = lk(t, bx, by, bz);
                }
            }
            const int li = (I + (m & 1)) & 15, lj = (J + ((m >> 1) & 1)) & 15, lk = (K + (m >> 2)) & 15;
            ix[m] = (lower - 1) * SBFACES + ffn(16, A, A == 0 ? 16 : li, A == 1 ? 16 : lj, A == 2 ? 16 : lk);
            all = all && lower != 0;
        }
    }
    if (all) {
        float c[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) c[m] = (float)fine[ix[m]];
        return lerp8(c, q0 - x, q1 - y, q2 - z);
    }
    float cd = g.cdx();
    return trilinear(coarse, g.nc(0) + (A == 0), g.nc(1) + (A == 1), g.nc(2) + (A == 2), __fdiv_rn(px - g.lo[0], cd) - o0,
                     __fdiv_rn(py - g.lo[1], cd) - o1, __fdiv_rn(pz - g.lo[2], cd) - o2);
}
template <class V, class L = SlotsGlobal>
__device__ inline float facesample(const STopo& t, const V* __restrict__ fine, const V* __restrict__ coarse, int a, float px,
                                   float py, float pz, const L& lk = L()) {
    return a == 0 ? facesample_a<0>(t, fine, coarse, px, py, pz, lk)
                  : (a == 1 ? facesample_a<1>(t, fine, coarse, px, py, pz, lk) : facesample_a<2>(t, fine, coarse, px, py, pz, lk));
}

// sflip.jl:193-215 finegather: APIC gather over the fine faces only, with the weight found.
template <class V>
__device__ inline void finegather(const STopo& t, const V* __restrict__ u, int a, float px, float py, float pz, float& vel,
                                  float& a0, float& a1, float& a2) {
    const SGrid& g = t.g;
    float q0 = ispace(px, g.lo[0], g.dx, a == 0 ? 0.0f : 0.5f);
    float q1 = ispace(py, g.lo[1], g.dx, a == 1 ? 0.0f : 0.5f);
    float q2 = ispace(pz, g.lo[2], g.dx, a == 2 ? 0.0f : 0.5f);
    BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
    vel = 0.0f; a0 = 0.0f; a1 = 0.0f; a2 = 0.0f;
#pragma unroll
    for (int mz = 0; mz < 3; ++mz) {
        float wz = pick(bz, mz), dz = (float)(bz.b + mz) - q2;
#pragma unroll
        for (int my = 0; my < 3; ++my) {
            float wy = pick(by, my), dy = (float)(by.b + my) - q1;
#pragma unroll
            for (int mx = 0; mx < 3; ++mx) {
                int32_t idx = finefaceat(t, a, bx.b + mx, by.b + my, bz.b + mz);
                if (idx >= 0) {
                    float wgt = pick(bx, mx) * wy * wz;
                    float tt = wgt * (float)u[idx];
                    vel += tt;
                    a0 += (tt * ((float)(bx.b + mx) - q0)) * g.dx;
                    a1 += (tt * dy) * g.dx;
                    a2 += (tt * dz) * g.dx;
                }
            }
        }
    }
    float c = __fdiv_rn(4.0f, g.dx * g.dx);
    a0 *= c; a1 *= c; a2 *= c;
}
template <class V>
__device__ inline float finegather_value(const STopo& t, const V* __restrict__ u, int a, float px, float py, float pz) {
    const SGrid& g = t.g;
    float q0 = ispace(px, g.lo[0], g.dx, a == 0 ? 0.0f : 0.5f);
    float q1 = ispace(py, g.lo[1], g.dx, a == 1 ? 0.0f : 0.5f);
    float q2 = ispace(pz, g.lo[2], g.dx, a == 2 ? 0.0f : 0.5f);
    BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
    float vel = 0.0f;
#pragma unroll
    for (int mz = 0; mz < 3; ++mz)
#pragma unroll
        for (int my = 0; my < 3; ++my)
#pragma unroll
            for (int mx = 0; mx < 3; ++mx) {
                int32_t idx = finefaceat(t, a, bx.b + mx, by.b + my, bz.b + mz);
                float wgt = pick(bx, mx) * pick(by, my) * pick(bz, mz);
                if (idx >= 0) vel += wgt * (float)u[idx];
            }
    return vel;
}

// Whether no point of the box of block w's cells widened by two cells is solid (the fluid
// region, inside the wall's cylinder and above the floor, is convex: its eight corners
// decide, and the margin covers the rounding of any cell centre or face position in it).
__device__ inline bool blockclear(const SGrid& g, int4 w) {
    bool clear = true;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        float px = g.lo[0] + (float)(w.x * 16 + (m & 1 ? 18 : -2)) * g.dx;
        float py = g.lo[1] + (float)(w.y * 16 + (m & 2 ? 18 : -2)) * g.dx;
        float pz = g.lo[2] + (float)(w.z * 16 + (m & 4 ? 18 : -2)) * g.dx;
        clear = clear && !sissolid(g, px, py, pz);
    }
    return clear;
}

// The slots of the 27 blocks around block w (0 off the grid or not allocated), into LDS.
__device__ inline void stage27(const STopo& t, int4 w, int32_t* s27) {
    if (threadIdx.x < 27) {
        int bi = w.x + (int)(threadIdx.x % 3) - 1, bj = w.y + (int)((threadIdx.x / 3) % 3) - 1, bk = w.z + (int)(threadIdx.x / 9) - 1;
        s27[threadIdx.x] = inblocks(t.g, bi, bj, bk) ? t.slot[blockid(t.g, bi, bj, bk)] : 0;
    }
}

// sflip.jl:1030-1031 walledface: a wall on either side of the face at p (cells h wide).
__device__ inline bool walledface_dev(const SGrid& g, const float p[3], int a, float h) {
    float e = 0.5f * h;
    float lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
    lo[a] = p[a] - e;
    hi[a] = p[a] + e;
    return sissolid(g, lo[0], lo[1], lo[2]) || sissolid(g, hi[0], hi[1], hi[2]);
}

struct SFlow;

// The mixed-precision hierarchy's storage (16-bit: FP16 or BF16 bit patterns), allocated
// on the first solve that asks for it: the V-cycle's level-0 vectors, level 1's, each
// dense level's, and rounded copies of every coupling.
struct MPStore {
    bool ready = false;
    DBuf<uint16_t> z[2], r[2], q[2], tmp[2];        // fine pool, coarse
    DBuf<uint16_t> X1[4][2];                        // level-1 x, b, r, t
    std::vector<std::array<DBuf<uint16_t>, 3>> Dv;  // dense levels: x, b, t
    DBuf<uint16_t> cc[3], c1[3], dDf, dDc, dD1;   // level 0's couplings follow from the types
    DBuf<uint16_t> rwf;                             // level 0's omega / diag (the Jacobi weights)
    std::vector<std::array<DBuf<uint16_t>, 4>> Dc;  // dense levels: couplings x, y, z, diagonal
};

// A field: its fine pool and the dense coarse array (sflip.jl:86-90).
template <class T> struct SField {
    DBuf<T> fine, coarse;
    void alloc(size_t nf, size_t nc) { fine.alloc(nf); coarse.alloc(nc); }
};

struct SparseOpts {          // the sparse solve's options (sflip.jl:1867, fgmres.jl:26,96)
    double tol = 5e-4;
    int maxiter = 240, sweeps = 1, coarse = 40, restart = 8;
    float boost = 1.5f;
    int policy = 2;          // 0 flexible CG, 1 FGMRES, 2 CG (at most 80) then FGMRES
    int precision = 0;       // 0 FP32 hierarchy
    int refine = 0;          // > 0: FP64 solution, FP64 residual every `refine` iterations
};
struct SparseStats {
    int iterations = 0, fcg_iterations = 0, fgmres_iterations = 0, used_fgmres = 0;
    double rel_residual = 0;
    double ms = 0;
};

// Dense levels of the sparse hierarchy (sflip.jl:234-240).
struct SLevel {
    int n[3];
    DBuf<float> x, b, t, c[3], dD;
    DBuf<uint8_t> act;
    size_t cells() const { return (size_t)n[0] * n[1] * n[2]; }
    size_t faces(int a) const { return (size_t)(n[0] + (a == 0)) * (n[1] + (a == 1)) * (n[2] + (a == 2)); }
};

struct SparseBasin {
    SGrid g;
    int blocks;               // the pool's capacity in blocks
    int32_t band, nmin;
    Stream st;
    KTimer kt;

    // host side of the topology (sflip.jl:252-290: slots, wheres, free, nslots)
    std::vector<int32_t> hslot;
    std::vector<int4> hwhere;
    std::vector<int32_t> freel;
    int nslots = 0;
    DBuf<int32_t> slot;
    DBuf<int4> where;
    DBuf<int32_t> list;       // added / removed slots of a refinement
    DBuf<int32_t> sorder;     // the live slots in block order (x fastest), for per-slot launches
    int32_t nlive = 0;        // their count
    DBuf<uint8_t> marks, marks2;

    SField<float> phi, phi0, P;
    SField<uint8_t> types;
    SField<float> u[3], u0[3];
    SField<uint8_t> fl[3];
    SField<float> ur[3], ur0[3];              // the reflected velocity and its scratch (sflip.jl:279-283)
    SField<uint8_t> fl2[3];                   // the extrapolation's scratch flags
    DBuf<uint8_t> reach;                      // trapped air: the coarse cells the atmosphere reaches
    DBuf<int> changed;
    StirParams stirp;
    DBuf<float> stir_waves, stir_heights;
    DBuf<int32_t> phip;
    DBuf<unsigned long long> mom[3], wt[3];

    // pressure (sflip.jl:1059-1904)
    SField<float> b, r, z, d, q, rold, tmp, dD;
    DBuf<float> dgf;                          // level 0's full fine diagonal (dD and couplings), per solve
    SField<float> d2;                         // flexible CG's second direction buffer, on first use
    DBuf<float> cf[3], cc[3];                 // level-0 face couplings, fine pools and coarse faces
    DBuf<float> c1[3], dD1;                   // level 1 (8^3 a block)
    DBuf<uint8_t> act1;
    DBuf<uint8_t> cmask;                      // coarse cells' cover masks (k_covermask), per solve
    SField<float> X1x, X1b, X1r, X1t;         // level-1 vectors (fine 512 a slot, coarse part)
    std::vector<SLevel> D;                    // D[0] = the coarse grid as level 2, then halvings
    std::vector<SField<float>> V, Z;          // FGMRES bases, allocated on first use
    MPStore mp;
    void mp_alloc();
    SField<double> Pd;                        // the FP64 pressure of iterative refinement, on first use
    void ir_alloc() {
        if (!Pd.fine.p) Pd.alloc((size_t)blocks * SBCELLS, ncells());
    }
    DBuf<double> partials, scal;

    // particles
    DBuf<float> pbuf[2][PFIELDS];
    int cur = 0;
    int64_t n = 0, capacity = 0;
    DBuf<int32_t> key, rank, seg, counts, totals, offsets, cstart;
    DBuf<unsigned long long> bsums;

    SparseBasin(const SGrid& grid, int blocks, int64_t capacity, int32_t band, int32_t nmin);

    size_t ncells() const { return (size_t)g.nc(0) * g.nc(1) * g.nc(2); }
    size_t nfaces(int a) const { return (size_t)(g.nc(0) + (a == 0)) * (g.nc(1) + (a == 1)) * (g.nc(2) + (a == 2)); }
    size_t finecells() const { return (size_t)nslots * SBCELLS; }
    size_t finefaces() const { return (size_t)nslots * SBFACES; }
    size_t nblocks() const { return (size_t)g.nb[0] * g.nb[1] * g.nb[2]; }
    STopo topo() const { return STopo{g, slot.p, where.p}; }
    PView view(int which) const {
        PView v;
        for (int f = 0; f < PFIELDS; ++f) v.f[f] = pbuf[which][f].p;
        return v;
    }

    // topology (sflip.jl:516-576)
    void marks_compute(float vary, float reach, int rings, bool particles);
    void refine_host(const uint8_t* want, std::vector<int32_t>& added, std::vector<int32_t>& removed);
    void refine(float vary, float reach, int rings, int* nadd, int* nrem);
    void refine_from(const uint8_t* want, int* nadd, int* nrem);
    void upload_topology();
    // phases
    int64_t resample(bool seeding, uint32_t seed);                    // sflip.jl:2227-2246
    void p2g(float dt);                                               // sflip.jl:808-831
    void g2p(float dt, float gravity);                                // sflip.jl:2141-2144
    void project(float gamma, const SparseOpts& o, SparseStats& s);   // sflip.jl:1998-2015
    // the bake writers' device side (bake.hip), as the dense basin's
    DBuf<__half> bake_pos, bake_vel;
    DBuf<unsigned long long> bake_counts;
    DBuf<uint8_t> bake_q;
    int64_t pack_band(float keep, float throat, uint16_t* pos, uint16_t* vel);   // flip.jl:1880-1904
    void quantize_levelset(float range, uint8_t* out);                          // sflip.jl:2427-2440
    // z = M b: the flexible CG's V-cycle preconditioner applied once to the right-hand
    // side array, into P (with `queued`, through the path the Krylov loop takes when the
    // V-cycle's first launches were queued ahead of a host read)
    void precondition(const SparseOpts& o, bool queued);
    // support phases
    void init_coarse(const float* eta, int neta, float rmax, float G, float core);
    void init_fine(const float* eta, int neta, float rmax, float G, float core);
    void classify();
    void solidify();
    void redistance(int iterations);
    void restrict_velocity();
    void keep();
    void fill_flags(uint8_t v);
    double volume();
    void gravity(float dt, float gacc);
    // the rest of a step (sflip.jl:2295-2318)
    SFlow flow(bool reflected) const;
    void advect(float dt, bool carry_reflected, float gacc, int32_t seed);
    void filltrapped(int maxrounds);
    void set_stir(const StirParams& p, const float* waves, int nwaves, const float* heights, int nheights);
    void forces(float dt, float time, float gacc);
    void save_reflect();
    void kick();
    void extrapolate(int layers, int coarselayers, bool withkick);
    void reflected();
    float maxspeed();
    bool reflect_flow = false;                // G2P and resampling read ur (the reflected velocity)
    // adjoints of the transfers (sparse_adjoint.hip) and the projection (sparse_pressure.hip),
    // at the state the forward phase left: cotangents of the faces (ub: fine pools and
    // coarse grid), of u0 (ub0), of the particles' v and APIC rows (pb)
    SField<float> ub[3], ub0[3];
    DBuf<float> pb[12];
    DBuf<unsigned long long> adjacc[3];        // fixed-point deposits of the G2P adjoint
    DBuf<unsigned long long> adjcell[2];       // fixed-point sums of the gradient's transpose (fine, coarse)
    DBuf<uint8_t> reached;
    void adj_alloc();
    void g2p_adjoint(float dt);
    void p2g_adjoint(float dt);
    void project_adjoint(const SparseOpts& o, SparseStats& stats);
    // the pressure solve on the right-hand side in b (P the start), shared by the projection
    // and its adjoint; the hierarchy and couplings must be set
    void solve_rhs(const SparseOpts& o, SparseStats& stats);
};

// restriction kernels shared by several phases (sflip.jl:583-612)
void launch_restrictcells(const SparseBasin& s, float* coarse, const float* fine);
void launch_restrictfaces_f(const SparseBasin& s, float* coarse, const float* fine, int a);
void launch_restrictfaces_u8(const SparseBasin& s, uint8_t* coarse, const uint8_t* fine, int a);
// the transpose of the face restriction on cotangents (sparse_adjoint.hip): fine += the
// restricted coarse faces' shares, cout = the coarse faces' own (cout may be cin)
void launch_restrictfaces_T(const SparseBasin& s, float* fine, float* cout, const float* cin, int a);

}  // namespace mael
