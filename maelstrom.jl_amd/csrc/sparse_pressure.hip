// The sparse basin's pressure projection (sflip.jl:1059-2015, fgmres.jl): level-0
// couplings and right-hand side, the Galerkin hierarchy by aggregation, the composite
// damped-Jacobi V-cycle (fine blocks at 16^3 and 8^3 over the coarse grid, then the
// coarse grid and its halvings, the coarsest in one workgroup), flexible
// Polak-Ribiere CG, restarted FGMRES, the CG-then-FGMRES policy, and the velocity update.
//
// Reductions are FP64 partial sums per workgroup summed in a fixed order, so a solve is
// reproducible run to run. Krylov scalars stay on the device; the host reads one value
// per CG iteration (the residual norm, with the error flags) and one per Arnoldi step.
#include "sparse.hpp"
#include "scan.hpp"
#include "precision.hpp"
#include <chrono>
#include <cmath>

namespace mael {

// ─── level-0 couplings (sflip.jl:1321-1406) ──────────────────────────────────

// sflip.jl:1334-1342 celltype: the fine cell's type and whether it is fine.
__device__ inline uint8_t celltype(const STopo& t, const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc, int I, int J,
                                   int K, bool& isfine) {
    int32_t idx = fineat(t, I, J, K);
    if (idx >= 0) { isfine = true; return tf[idx]; }
    isfine = false;
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    int ci = I >> 2, cj = J >> 2, ck = K >> 2;
    if (ci < 0 || cj < 0 || ck < 0 || ci >= n0 || cj >= n1 || ck >= n2) return SOLID;
    return tc[cl(ci, cj, ck, n0, n1)];
}
// The value at fine cell (I, J, K): a fine block's, else the coarse cell's (clamped).
template <class V>
__device__ inline V fineorcoarse(const STopo& t, const V* __restrict__ fine, const V* __restrict__ coarse, int I, int J, int K) {
    int32_t idx = fineat(t, I, J, K);
    if (idx >= 0) return fine[idx];
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    return coarse[cl(min(max(I >> 2, 0), n0 - 1), min(max(J >> 2, 0), n1 - 1), min(max(K >> 2, 0), n2 - 1), n0, n1)];
}

// The faces along a of a live slot s (1-based), for kernels that take a workgroup per
// slot and a thread per 17 faces: a face's two cells lie in the slot's block or in the
// blocks below and above it along a, whose slots are looked up once.
struct SlotFaces {
    int s, a, n0, n1, n2;
    int4 w;
    int32_t slo, shi;
    __device__ SlotFaces(const STopo& t, int slot, int axis) : s(slot), a(axis) {
        w = t.where[s - 1];
        const SlotsGlobal lk;
        slo = lk(t, w.x - (a == 0), w.y - (a == 1), w.z - (a == 2));
        shi = lk(t, w.x + (a == 0), w.y + (a == 1), w.z + (a == 2));
        n0 = t.g.nc(0); n1 = t.g.nc(1); n2 = t.g.nc(2);
    }
    // face f of the slot: its local (i, j, k) (the layout of fineface)
    __device__ void face(int f, int& i, int& j, int& k) const {
        if (a == 0) { i = f % 17; int q = f / 17; j = q & 15; k = q >> 4; }
        else if (a == 1) { i = f & 15; int q = f >> 4; j = q % 17; k = q / 17; }
        else { i = f & 15; j = (f >> 4) & 15; k = f >> 8; }
    }
    // fineat's index of the cell at local (li, lj, lk), its coordinate along a in -1..16
    __device__ int32_t cell(int li, int lj, int lk) const {
        const int la = a == 0 ? li : (a == 1 ? lj : lk);
        const int32_t sl = la < 0 ? slo : (la > 15 ? shi : s);
        return sl == 0 ? -1 : (sl - 1) * 4096 + (li & 15) + 16 * ((lj & 15) + 16 * (lk & 15));
    }
    // celltype from that index (I, J, K global)
    __device__ uint8_t type(const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc, int32_t id, int I, int J, int K,
                            bool& isfine) const {
        if (id >= 0) { isfine = true; return tf[id]; }
        isfine = false;
        const int ci = I >> 2, cj = J >> 2, ck = K >> 2;
        if (ci < 0 || cj < 0 || ck < 0 || ci >= n0 || cj >= n1 || ck >= n2) return SOLID;
        return tc[cl(ci, cj, ck, n0, n1)];
    }
    // fineorcoarse's coarse cell (clamped)
    __device__ size_t coarse(int I, int J, int K) const {
        return cl(min(max(I >> 2, 0), n0 - 1), min(max(J >> 2, 0), n1 - 1), min(max(K >> 2, 0), n2 - 1), n0, n1);
    }
};

// A workgroup per live slot (block order), a thread per 17 of its faces along a.
__global__ __launch_bounds__(256) void k_finecoef(float* __restrict__ c, const uint8_t* __restrict__ tf,
                                                  const uint8_t* __restrict__ tc, STopo t, int a,
                                                  const int32_t* __restrict__ order) {
    const SlotFaces F(t, order[blockIdx.x], a);
    for (int r = 0; r < SBFACES / 256; ++r) {
        const int f = r * 256 + (int)threadIdx.x;
        int i, j, k;
        F.face(f, i, j, k);
        const int I = F.w.x * 16 + i, J = F.w.y * 16 + j, K = F.w.z * 16 + k;
        bool fp, fm;
        const uint8_t tp = F.type(tf, tc, F.cell(i, j, k), I, J, K, fp);
        const uint8_t tm = F.type(tf, tc, F.cell(i - (a == 0), j - (a == 1), k - (a == 2)), I - (a == 0), J - (a == 1),
                                  K - (a == 2), fm);
        c[(F.s - 1) * SBFACES + f] = (tp == LIQUID && tm == LIQUID) ? (fp && fm ? FINE_FINE : FINE_COARSE) : 0.0f;
    }
}
__global__ void k_coarsecoef(float* __restrict__ c, const uint8_t* __restrict__ tc, STopo t, int a) {
    int nc0 = t.g.nc(0), nc1 = t.g.nc(1), nc2 = t.g.nc(2);
    int n0 = nc0 + (a == 0), n1 = nc1 + (a == 1), n2 = nc2 + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
    bool ok = Im >= 0 && Jm >= 0 && Km >= 0 && I < nc0 && J < nc1 && K < nc2;
    c[l] = (ok && !covered(t, I, J, K) && !covered(t, Im, Jm, Km) && tc[cl(I, J, K, nc0, nc1)] == LIQUID &&
            tc[cl(Im, Jm, Km, nc0, nc1)] == LIQUID)
               ? COARSE_COARSE
               : 0.0f;
}
__global__ void k_finedirichlet(float* __restrict__ dD, const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc,
                                const float* __restrict__ phi, const float* __restrict__ phic, STopo t, int32_t nf) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    SCell s = finecell(t, n);
    if (s.w.x < 0) return;
    float d = 0.0f;
    if (tf[n] == LIQUID) {
        int I = s.w.x * 16 + s.i, J = s.w.y * 16 + s.j, K = s.w.z * 16 + s.k;
        float pc = phi[n];
        // neighbour order of flip.jl:84-92
        const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        for (int m = 0; m < 6; ++m) {
            int a = I + nb[m][0], b = J + nb[m][1], cc = K + nb[m][2];
            bool isf;
            if (celltype(t, tf, tc, a, b, cc, isf) == AIR)
                d += __fdiv_rn(isf ? FINE_FINE : FINE_COARSE, theta(pc, fineorcoarse(t, phi, phic, a, b, cc)));
        }
    }
    dD[n] = d;
}
__global__ void k_coarsedirichlet(float* __restrict__ dD, const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc,
                                  const float* __restrict__ phi, const float* __restrict__ phic, STopo t) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    float d = 0.0f;
    if (tc[l] == LIQUID && !covered(t, I, J, K)) {
        float pc = phic[l];
        for (int a = 0; a < 3; ++a)
            for (int sg = -1; sg <= 1; sg += 2) {
                int In = I + sg * (a == 0), Jn = J + sg * (a == 1), Kn = K + sg * (a == 2);
                if (In < 0 || Jn < 0 || Kn < 0 || In >= n0 || Jn >= n1 || Kn >= n2) continue;
                if (covered(t, In, Jn, Kn)) {
                    for (int q = 0; q < 4; ++q)
                        for (int p = 0; p < 4; ++p) {
                            int FI = a == 0 ? (sg > 0 ? 4 * In : 4 * I - 1) : 4 * I + p;
                            int FJ = a == 1 ? (sg > 0 ? 4 * Jn : 4 * J - 1) : 4 * J + (a == 0 ? p : q);
                            int FK = a == 2 ? (sg > 0 ? 4 * Kn : 4 * K - 1) : 4 * K + q;
                            int32_t idx = fineat(t, FI, FJ, FK);
                            if (idx >= 0 && tf[idx] == AIR) d += __fdiv_rn(FINE_COARSE, theta(pc, phi[idx]));
                        }
                } else if (tc[cl(In, Jn, Kn, n0, n1)] == AIR) {
                    d += __fdiv_rn(COARSE_COARSE, theta(pc, phic[cl(In, Jn, Kn, n0, n1)]));
                }
            }
    }
    dD[l] = d;
}
__global__ void k_activecoarse(uint8_t* __restrict__ act, const uint8_t* __restrict__ tc, STopo t) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    act[l] = (tc[l] == LIQUID && !covered(t, I, J, K)) ? 1 : 0;
}

// sflip.jl:1408-1435 the right-hand side; sflip.jl:1964-1973 P cleared off the unknowns.
__global__ void k_finerhs(float* __restrict__ b, float* __restrict__ P, const uint8_t* __restrict__ tf,
                          const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w, STopo t,
                          float gamma, int32_t nf) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    SCell s = finecell(t, n);
    if (s.w.x < 0) return;
    float h = t.g.dx;
    if (tf[n] == LIQUID) {
        size_t base = (size_t)(s.s - 1) * SBFACES;
        float div = u[base + ffn(16, 0, s.i + 1, s.j, s.k)] - u[base + ffn(16, 0, s.i, s.j, s.k)] +
                    v[base + ffn(16, 1, s.i, s.j + 1, s.k)] - v[base + ffn(16, 1, s.i, s.j, s.k)] +
                    w[base + ffn(16, 2, s.i, s.j, s.k + 1)] - w[base + ffn(16, 2, s.i, s.j, s.k)];
        b[n] = -h * div + gamma * h * h;
    } else {
        b[n] = 0.0f;
        P[n] = 0.0f;
    }
}
__global__ void k_coarserhs(float* __restrict__ b, float* __restrict__ P, const uint8_t* __restrict__ tc,
                            const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w, STopo t,
                            float gamma) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    if (tc[l] == LIQUID && !covered(t, I, J, K)) {
        float h = t.g.dx;
        float div = u[cl(I + 1, J, K, n0 + 1, n1)] - u[cl(I, J, K, n0 + 1, n1)] + v[cl(I, J + 1, K, n0, n1 + 1)] -
                    v[cl(I, J, K, n0, n1 + 1)] + w[cl(I, J, K + 1, n0, n1)] - w[cl(I, J, K, n0, n1)];
        b[l] = -16.0f * h * div + 64.0f * gamma * h * h;
    } else {
        b[l] = 0.0f;
        P[l] = 0.0f;
    }
}

// ─── the hierarchy (sflip.jl:1450-1570) ──────────────────────────────────────

__global__ void k_coef1(float* __restrict__ c1, const float* __restrict__ c0, STopo t, int a, int nslots) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nslots * SBFACES1) return;
    int s = n / SBFACES1 + 1, f = n - (s - 1) * SBFACES1, i, j, k;
    if (a == 0) { i = f % 9; int r = f / 9; j = r & 7; k = r >> 3; }
    else if (a == 1) { i = f & 7; int r = f >> 3; j = r % 9; k = r / 9; }
    else { i = f & 7; j = (f >> 3) & 7; k = f >> 6; }
    if (t.where[s - 1].x < 0) return;
    float acc = 0.0f;
    for (int q = 0; q < 2; ++q)
        for (int p = 0; p < 2; ++p) {
            int ii = a == 0 ? 2 * i : 2 * i + p;
            int jj = a == 1 ? 2 * j : 2 * j + (a == 0 ? p : q);
            int kk = a == 2 ? 2 * k : 2 * k + q;
            acc += c0[(size_t)(s - 1) * SBFACES + ffn(16, a, ii, jj, kk)];
        }
    c1[n] = acc;
}
__global__ void k_cells1(float* __restrict__ dD1, uint8_t* __restrict__ act1, const float* __restrict__ dD0,
                         const uint8_t* __restrict__ tf, STopo t, int nslots) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nslots * SBCELLS1) return;
    int s = (n >> 9) + 1, c = n & 511, i = c & 7, j = (c >> 3) & 7, k = c >> 6;
    if (t.where[s - 1].x < 0) return;
    float d = 0.0f;
    uint8_t act = 0;
    for (int cz = 0; cz < 2; ++cz)
        for (int by = 0; by < 2; ++by)
            for (int ax = 0; ax < 2; ++ax) {
                int32_t m = (s - 1) * SBCELLS + (2 * i + ax) + 16 * ((2 * j + by) + 16 * (2 * k + cz));
                d += dD0[m];
                if (tf[m] == LIQUID) act = 1;
            }
    dD1[n] = d;
    act1[n] = act;
}
__global__ void k_coef2(float* __restrict__ c2, const float* __restrict__ c0c, const float* __restrict__ c1, STopo t, int a) {
    int n0 = t.g.nc(0) + (a == 0), n1 = t.g.nc(1) + (a == 1), n2 = t.g.nc(2) + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    if (covered(t, I, J, K) || covered(t, I - (a == 0), J - (a == 1), K - (a == 2))) {
        float acc = 0.0f;
        for (int q = 0; q < 2; ++q)
            for (int p = 0; p < 2; ++p) {
                int FI = a == 0 ? 2 * I : 2 * I + p;
                int FJ = a == 1 ? 2 * J : 2 * J + (a == 0 ? p : q);
                int FK = a == 2 ? 2 * K : 2 * K + q;
                int32_t idx = faceat<8>(t, a, FI, FJ, FK);
                if (idx >= 0) acc += c1[idx];
            }
        c2[l] = acc;
    } else {
        c2[l] = c0c[l];
    }
}
__global__ void k_cells2(float* __restrict__ dD2, uint8_t* __restrict__ act, const float* __restrict__ dD0c,
                         const float* __restrict__ dD1, const uint8_t* __restrict__ act1, STopo t) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    if (covered(t, I, J, K)) {
        float d = 0.0f;
        uint8_t ac = 0;
        for (int c = 0; c < 2; ++c)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    int32_t m = cellat<8>(t, 2 * I + a, 2 * J + b, 2 * K + c);
                    d += dD1[m];
                    if (act1[m]) ac = 1;
                }
        dD2[l] = d;
        act[l] = ac;
    } else {
        dD2[l] = dD0c[l];
    }
}
__global__ void k_coefdense(float* __restrict__ cc, const float* __restrict__ cf, int3 nc, int3 nf, int a) {
    int n0 = nc.x + (a == 0), n1 = nc.y + (a == 1), n2 = nc.z + (a == 2);
    int m0 = nf.x + (a == 0), m1 = nf.y + (a == 1);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    float acc = 0.0f;
    for (int q = 0; q < 2; ++q)
        for (int p = 0; p < 2; ++p) {
            int FI = a == 0 ? 2 * I : 2 * I + p;
            int FJ = a == 1 ? 2 * J : 2 * J + (a == 0 ? p : q);
            int FK = a == 2 ? 2 * K : 2 * K + q;
            acc += cf[cl(FI, FJ, FK, m0, m1)];
        }
    cc[l] = acc;
}
__global__ void k_cellsdense(float* __restrict__ dDc, uint8_t* __restrict__ actc, const float* __restrict__ dDf,
                             const uint8_t* __restrict__ actf, int3 nc, int3 nf) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)nc.x * nc.y * nc.z) return;
    int I, J, K;
    unlin(l, nc.x, nc.y, I, J, K);
    float d = 0.0f;
    uint8_t ac = 0;
    for (int c = 0; c < 2; ++c)
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) {
                size_t m = cl(2 * I + a, 2 * J + b, 2 * K + c, nf.x, nf.y);
                d += dDf[m];
                if (actf[m]) ac = 1;
            }
    dDc[l] = d;
    actc[l] = ac;
}

// ─── the composite rows (sflip.jl:1139-1216) ─────────────────────────────────
// The V-cycle's kernels are templates on the storage of its vectors (T) and of its
// couplings (C): FP32 for the reference's solve, FP16 or BF16 for the mixed-precision
// one. Arithmetic is FP32 either way (the coarsest level's in A).

template <class C> struct FineCoefT { const C* c[3]; };
template <class C> struct CoarseCoefT { const C* c[3]; };

// The slot of block w's face neighbour d (-x, +x, -y, +y, -z, +z), 0 for none.
__device__ inline int32_t nbslot(const STopo& t, int4 w, int d) {
    int ax = d >> 1, sg = (d & 1) ? 1 : -1;
    int bi = w.x + (ax == 0 ? sg : 0), bj = w.y + (ax == 1 ? sg : 0), bk = w.z + (ax == 2 ? sg : 0);
    return inblocks(t.g, bi, bj, bk) ? t.slot[blockid(t.g, bi, bj, bk)] : 0;
}

// Level 0's couplings follow from the cell types (k_finecoef, k_coarsecoef,
// sflip.jl:1323-1353): for a LIQUID uncovered cell, 4 to an uncovered LIQUID coarse
// neighbour and 0.4 to a LIQUID fine cell across a sub-face, else 0. With `ty` set
// (level 0), a coarse row takes them from the types (ty->tf fine, ty->tc coarse) instead of
// the coupling arrays: the same values as the FP32 arrays hold.
struct LevelTypes { const uint8_t* tf; const uint8_t* tc; };
// A level's fine cell (N = 16 or 8) of slot s: diag and the neighbours' sum, A x = diag x
// - off, summed in A (FP32, or FP64 for iterative refinement's residual). Neighbours in
// the block index directly; across a face they are the neighbour block's (slots in nb)
// or, without one, the coarse cell's. Every index is computed and all thirteen loads are
// issued before any value is used (each index is in bounds whatever the couplings); a
// coupling that is not positive adds nothing, in the reference's order.
template <int N, class T, class C, class A = float, bool Z = false>
__device__ inline void finerow(const STopo& t, const int32_t* nb, const T* __restrict__ x, const T* __restrict__ xc,
                               const FineCoefT<C>& cf, const C* __restrict__ dD, int s, int i, int j, int k, int4 w, A& diag,
                               A& off) {
    constexpr int sh = N == 16 ? 2 : 1, V = N * N * N;
    constexpr uint32_t per = (N + 1) * N * N;
    int I = w.x * N + i, J = w.y * N + j, K = w.z * N + k;
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    uint32_t own = (uint32_t)(s - 1) * V + (uint32_t)(i + N * (j + N * k));
    auto nbr = [&](bool inside, int delta, int d, int li, int lj, int lk, int In, int Jn, int Kn) -> A {
        if constexpr (Z) return (A)0;   // x = 0: nothing to read
        int32_t sn = nb[d];
        uint32_t in_block = own + (uint32_t)delta;
        uint32_t in_nb = (uint32_t)(sn - 1) * V + (uint32_t)(li + N * (lj + N * lk));
        int ci = min(max(In >> sh, 0), n0 - 1), cj = min(max(Jn >> sh, 0), n1 - 1), ck = min(max(Kn >> sh, 0), n2 - 1);
        uint32_t in_coarse = (uint32_t)ci + (uint32_t)n0 * ((uint32_t)cj + (uint32_t)n1 * (uint32_t)ck);
        bool fine = inside || sn != 0;
        uint32_t idx = inside ? in_block : (sn != 0 ? in_nb : in_coarse);
        const T* p = fine ? x : xc;
        return (A)lf(p[idx]);
    };
    uint32_t base = (uint32_t)(s - 1) * per;
    A c0 = (A)lf(cf.c[0][base + ffn(N, 0, i, j, k)]), c1 = (A)lf(cf.c[0][base + ffn(N, 0, i + 1, j, k)]);
    A c2 = (A)lf(cf.c[1][base + ffn(N, 1, i, j, k)]), c3 = (A)lf(cf.c[1][base + ffn(N, 1, i, j + 1, k)]);
    A c4 = (A)lf(cf.c[2][base + ffn(N, 2, i, j, k)]), c5 = (A)lf(cf.c[2][base + ffn(N, 2, i, j, k + 1)]);
    A x0 = nbr(i > 0, -1, 0, N - 1, j, k, I - 1, J, K), x1 = nbr(i < N - 1, 1, 1, 0, j, k, I + 1, J, K);
    A x2 = nbr(j > 0, -N, 2, i, N - 1, k, I, J - 1, K), x3 = nbr(j < N - 1, N, 3, i, 0, k, I, J + 1, K);
    A x4 = nbr(k > 0, -N * N, 4, i, j, N - 1, I, J, K - 1), x5 = nbr(k < N - 1, N * N, 5, i, j, 0, I, J, K + 1);
    diag = (A)lf(dD[own]);
    off = (A)0;
    if (c0 > (A)0) { diag += c0; off += c0 * x0; }
    if (c1 > (A)0) { diag += c1; off += c1 * x1; }
    if (c2 > (A)0) { diag += c2; off += c2 * x2; }
    if (c3 > (A)0) { diag += c3; off += c3 * x3; }
    if (c4 > (A)0) { diag += c4; off += c4 * x4; }
    if (c5 > (A)0) { diag += c5; off += c5 * x5; }
}

// Per coarse cell, which of it and its six neighbours lie in fine blocks: bit 0 the cell,
// bit 1 + d neighbour d (-x, +x, -y, +y, -z, +z; 0 off the grid). Fixed for a solve.
__global__ void k_covermask(uint8_t* __restrict__ m, STopo t) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    unsigned v = covered(t, I, J, K) ? 1u : 0u;
    for (int d = 0; d < 6; ++d) {
        int ax = d >> 1, sg = (d & 1) ? 1 : -1;
        int In = I + (ax == 0 ? sg : 0), Jn = J + (ax == 1 ? sg : 0), Kn = K + (ax == 2 ? sg : 0);
        bool in = In >= 0 && Jn >= 0 && Kn >= 0 && In < n0 && Jn < n1 && Kn < n2;
        if (in && covered(t, In, Jn, Kn)) v |= 2u << d;
    }
    m[l] = (uint8_t)v;
}

// An uncovered coarse cell (cover mask msk): coarse faces, or towards a covered
// neighbour its (N/4)^2 sub-faces. The uncovered neighbours' couplings and values are
// loaded first (an index off the grid reads the cell itself); a covered neighbour is
// always at the edge of its block, so its sub-faces and cells index from one slot
// lookup (the upper block's copy of a face, or the lower block's high plane). Sums in
// the reference's order. coarserow_pre issues the loads of the uncovered part (the six
// neighbours' couplings and values, the Dirichlet diagonal), coarserow_post sums them
// with the covered sides', so a caller can issue the first batch before it knows whether
// the row is needed.
template <class A> struct CoarsePre {
    A cc[6], xv[6], dg;
    bool in[6];
};
template <class C, class A, class CX>
__device__ inline CoarsePre<A> coarserow_pre(const STopo& t, CX cx, const CoarseCoefT<C>& ccf, const C* __restrict__ dDc, int I,
                                             int J, int K, const LevelTypes* ty) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t own = cl(I, J, K, n0, n1);
    CoarsePre<A> P;
    A* cc = P.cc;
    A* xv = P.xv;
    bool* in = P.in;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        int a = d >> 1, sg = (d & 1) ? 1 : -1;
        int In = I + sg * (a == 0), Jn = J + sg * (a == 1), Kn = K + sg * (a == 2);
        in[d] = In >= 0 && Jn >= 0 && Kn >= 0 && In < n0 && Jn < n1 && Kn < n2;
        size_t nidx = in[d] ? cl(In, Jn, Kn, n0, n1) : own;
        if (ty) {
            cc[d] = (A)(ty->tc[nidx] == LIQUID ? COARSE_COARSE : 0.0f);
        } else {
            int f0 = n0 + (a == 0), f1 = n1 + (a == 1);
            size_t fi = sg < 0 ? cl(I, J, K, f0, f1) : cl(I + (a == 0), J + (a == 1), K + (a == 2), f0, f1);
            cc[d] = (A)lf(ccf.c[a][fi]);
        }
        xv[d] = cx(nidx);
    }
    P.dg = (A)lf(dDc[own]);
    return P;
}
// The row of a cell none of whose neighbours is covered, branch-free: a zero coupling
// adds exactly nothing (the same sums as coarserow_post's in that case).
template <class A> __device__ inline void coarserow_flat(const CoarsePre<A>& P, A& diag, A& off) {
    diag = P.dg;
    off = (A)0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const A c = (P.in[d] && P.cc[d] > (A)0) ? P.cc[d] : (A)0;
        diag += c;
        off += c * P.xv[d];
    }
}
template <int N, class C, class A, class FX>
__device__ inline void coarserow_post(const STopo& t, unsigned msk, FX fx, const FineCoefT<C>& cf, const CoarsePre<A>& P, int I,
                                      int J, int K, A& diag, A& off, const LevelTypes* ty) {
    constexpr int R = N / 4, V = N * N * N, m = N - 1;
    constexpr uint32_t per = (N + 1) * N * N;
    const A* cc = P.cc;
    const A* xv = P.xv;
    const bool* in = P.in;
    diag = P.dg;
    off = (A)0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (!in[d]) continue;
        int a = d >> 1, sg = (d & 1) ? 1 : -1;
        if (msk & (2u << d)) {
            int In = I + sg * (a == 0), Jn = J + sg * (a == 1), Kn = K + sg * (a == 2);
            int32_t sn = t.slot[blockid(t.g, In >> 2, Jn >> 2, Kn >> 2)];
            uint32_t cbase = (uint32_t)(sn - 1) * V, fbase = (uint32_t)(sn - 1) * per;
            // the R x R patch of the neighbour block's cells on the shared side: along a its
            // near plane (local 0 above, N - 1 below), across it R (coordinate mod 4) + p, q
            const int la = sg > 0 ? 0 : m;
            const int bi = a == 0 ? la : R * (I & 3), bj = a == 1 ? la : R * (J & 3), bk = a == 2 ? la : R * (K & 3);
            const uint32_t c0 = cbase + (uint32_t)(bi + N * (bj + N * bk));
            constexpr uint32_t sp0 = N, sq2 = N;   // cell strides of p and q (a = 0: j, k; 1: i, k; 2: i, j)
            const uint32_t sp = a == 0 ? sp0 : 1u, sq = a == 2 ? sq2 : (uint32_t)(N * N);
            // the faces between: the neighbour's own at its plane 0 above, the lower block's
            // top plane (coordinate N) below
            const int fa = sg > 0 ? 0 : N;
            const uint32_t f0 = fbase + (uint32_t)ffn(N, a, a == 0 ? fa : bi, a == 1 ? fa : bj, a == 2 ? fa : bk);
            const uint32_t fp = a == 0 ? (uint32_t)(N + 1) : 1u,
                           fq = a == 0 ? (uint32_t)((N + 1) * N) : (a == 1 ? (uint32_t)(N * (N + 1)) : (uint32_t)N);
#pragma unroll
            for (int q = 0; q < R; ++q)
#pragma unroll
                for (int p = 0; p < R; ++p) {
                    const uint32_t cidx = c0 + (uint32_t)p * sp + (uint32_t)q * sq;
                    A c;
                    if (ty) c = (A)(ty->tf[cidx] == LIQUID ? FINE_COARSE : 0.0f);
                    else c = (A)lf(cf.c[a][f0 + (uint32_t)p * fp + (uint32_t)q * fq]);
                    diag += c;
                    off += c * fx(cidx);
                }
        } else if (cc[d] > (A)0) {
            diag += cc[d];
            off += cc[d] * xv[d];
        }
    }
}

template <int N, class C, class A, class FX, class CX>
__device__ inline void coarserow_acc(const STopo& t, unsigned msk, FX fx, const FineCoefT<C>& cf, CX cx,
                                     const CoarseCoefT<C>& ccf, const C* __restrict__ dDc, int I, int J, int K, A& diag,
                                     A& off, const LevelTypes* ty = nullptr) {
    const CoarsePre<A> P = coarserow_pre<C, A>(t, cx, ccf, dDc, I, J, K, ty);
    coarserow_post<N, C, A>(t, msk, fx, cf, P, I, J, K, diag, off, ty);
}

// The same reading x from arrays (Z: x = 0, nothing read).
template <int N, class T, class C, class A = float, bool Z = false>
__device__ inline void coarserow(const STopo& t, unsigned msk, const T* __restrict__ xf, const FineCoefT<C>& cf,
                                 const T* __restrict__ xc, const CoarseCoefT<C>& ccf, const C* __restrict__ dDc, int I, int J,
                                 int K, A& diag, A& off, const LevelTypes* ty = nullptr) {
    auto fx = [&](uint32_t i) -> A { if constexpr (Z) return (A)0; else return (A)lf(xf[i]); };
    auto cx = [&](size_t i) -> A { if constexpr (Z) return (A)0; else return (A)lf(xc[i]); };
    coarserow_acc<N, C, A>(t, msk, fx, cf, cx, ccf, dDc, I, J, K, diag, off, ty);
}

// The device scalars of the Krylov loops (slots of SparseBasin::scal).
enum { SS_BB = 0, SS_RR, SS_RHO, SS_DEN, SS_ALPHA, SS_RHONEW, SS_ZROLD, SS_BETA, SS_RNORM, SS_FLAG, SS_TMP, SS_H0 = 16 };

// A workgroup's sum (thread 0 holds it); every thread must call it.
__device__ inline double block_sum_red(double v, double* lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += lds[i];
    return s;
}

// Level views: fine couplings and diagonal at N, the coarse grid's level-0 couplings.
template <class C> struct CompT {
    FineCoefT<C> cf;
    const C* dDf;
    const uint8_t* actf;
    CoarseCoefT<C> cc;
    const C* dDc;
    const uint8_t* actc;
    const uint8_t* cmask;   // k_covermask
    const uint8_t* tc;      // the coarse cell types (level 0's matrix-free couplings)
    const C* rwf = nullptr; // reduced precision, level 0: omega / diag per fine cell
    const C* dgf = nullptr; // FP32, level 0: the full diagonal per fine cell (mode 6)
};

// Flexible CG's fused step (mode 4 of the level-0 rows, sflip.jl:1884-1886 with
// 1903): the new direction d = z + beta d_old (x is z) is applied and written, and
// d.q is summed per workgroup over the unknowns.
struct SearchArgs {
    const float* dof;   // d_old, fine pool and coarse grid
    const float* doc;
    float* dnf;         // d_new
    float* dnc;
    double* partials;   // [4 nslots fine quarter tiles | coarse workgroups]
    int32_t coarse0;    // where the coarse workgroups' sums start
    const double* scal;
};

// sflip.jl:1218-1269 on the fine cells of a level: mode 0 Jacobi (out = x + w (b - A x) /
// diag), 1 residual (out = b - A x), 2 apply (out = A x), 3 Jacobi from x = 0 (x not
// read); inactive cells and the cells of free slots get 0. Mode 5 writes w / diag (0 where
// diag is not positive): in reduced precision the level-0 Jacobi sweep takes that weight
// r per cell, out = (1 - w) x + r (b + off), and needs neither the diagonal nor the
// neighbours' couplings, only their products with x (the weights round to the storage
// type once per solve; the FP32 sweep keeps the reference's division). That sweep relaxes
// the operator with diagonal w / r, and the reduced-precision residual uses the same.
// Level 0, matrix-free and staged in LDS. A workgroup takes a quarter of a slot (4 z
// planes of its 16^3 block), a thread 4 consecutive cells along x of one row. Level 0's
// coupling of a LIQUID fine cell to a neighbour is 1 when the neighbour is a LIQUID fine
// cell, 0.4 when it is a LIQUID coarse cell, else 0 (k_finecoef, sflip.jl:1323-1342), so
// it depends on the neighbour alone: the workgroup stages, for its cells and the layers
// around them (the block's other planes, the face neighbours' slots, or else the coarse
// cells: the value clamped to the grid, the type SOLID off it), each cell's coupling w and
// the product w x (0 where w is 0), once per cell. A LIQUID cell's row then adds its six
// neighbours' w into the diagonal and their w x into the off-diagonal sum in finerow's
// order; a zero coupling adds exactly nothing, so the sums equal those over the coupling
// arrays bit for bit. Loads and stores of a thread's own cells are 4-wide.
template <class T> struct alignas(4 * sizeof(T)) Four { T v[4]; };
template <class T> __device__ inline Four<T> load4(const T* __restrict__ p, uint32_t i) {
    return *reinterpret_cast<const Four<T>*>(p + i);
}
// A 4-cell store marked non-temporal: a sweep's output is read by the next launch, not by
// this one, so it need not stay in L2 (where the end of the launch writes back what is
// dirty before the next can start).
template <class T> __device__ inline void store4_nt(T* __restrict__ p, uint32_t i, const Four<T>& v) {
    if constexpr (sizeof(T) == 2) {
        uint64_t w;
        __builtin_memcpy(&w, &v, 8);
        __builtin_nontemporal_store(w, reinterpret_cast<uint64_t*>(p + i));
    } else {
        typedef uint32_t U4 __attribute__((ext_vector_type(4)));
        U4 w;
        __builtin_memcpy(&w, &v, 16);
        __builtin_nontemporal_store(w, reinterpret_cast<U4*>(p + i));
    }
}
constexpr int RS = 24, PS = 18 * RS;   // LDS rows of 24: cell li (-1..16) at 4 + li; planes of 18 rows
// NQ: quads a thread (a workgroup takes 4 NQ planes of its slot, its 16 / (4 NQ) tiles
// cover the slot); more cells a workgroup keep more loads in flight.
template <int MODE, class T, class C, int NQ = 1>
__device__ __forceinline__ void rowfine16_wg(const uint32_t tile, T* __restrict__ out, const T* __restrict__ x,
                                             const T* __restrict__ b, const T* __restrict__ xc, const CompT<C>& L,
                                             const STopo& t, const SearchArgs& S) {
    constexpr bool Z = MODE == 3 || MODE == 5 || MODE == 6, SEARCH = MODE == 4;
    // reduced precision: the sweep and the residual take the weight r = omega / diag
    constexpr bool RW = (MODE == 0 || MODE == 1) && !std::is_same<T, float>::value;
    // FP32: the rows take the full diagonal computed once a solve (mode 6), the same sum
    constexpr bool FD = MODE <= 4 && std::is_same<T, float>::value;
    constexpr bool NW = !RW && !FD;                                        // the couplings staged
    constexpr int PL = 4 * NQ, HV = (PL + 2) * PS, TPS = 4 / NQ;          // planes, LDS, tiles a slot
    static_assert(NQ == 1 || NQ == 2, "a tile is 4 or 8 planes");
    __shared__ float4 sw4[NW ? HV / 4 : 1];    // the couplings
    __shared__ float4 sp4[Z ? 1 : HV / 4];     // coupling times x
    __shared__ double lds[4];
    float* sw = reinterpret_cast<float*>(sw4);
    float* sp = reinterpret_cast<float*>(sp4);
    const int s = (int)(tile / TPS) + 1, k0 = (int)(tile % TPS) * PL;
    const uint32_t cb = (uint32_t)(s - 1) * 4096u;
    const uint8_t* __restrict__ tf = L.actf;   // level 0: the fine cell types
    const float beta = SEARCH ? (float)S.scal[SS_BETA] : 0.0f;
    // x as the row reads it: in the fused search the direction z + beta d_old, rounded to T
    auto xat = [&](const T* __restrict__ xs, const float* __restrict__ ds, size_t i) -> float {
        if constexpr (SEARCH) return lf(sf<T>(lf(xs[i]) + beta * ds[i]));
        else return lf(xs[i]);
    };
    const int tid = threadIdx.x, i4 = tid & 3, j = (tid >> 2) & 15, kk = tid >> 6;
    // quad q of the thread: plane kk + 4 q of the tile
    auto cell = [&](int q) { return cb + 256u * (uint32_t)(k0 + kk + 4 * q) + 16u * (uint32_t)j + 4u * (uint32_t)i4; };
    int4 w = t.where[s - 1];
    if (w.x < 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const uint32_t n = cell(q);
            Four<T> z4;
#pragma unroll
            for (int e = 0; e < 4; ++e) z4.v[e] = sf<T>(0.0f);
            *reinterpret_cast<Four<T>*>(out + n) = z4;
            if constexpr (SEARCH) {
                float4 d4;
                d4.x = xat(x, S.dof, n); d4.y = xat(x, S.dof, n + 1); d4.z = xat(x, S.dof, n + 2); d4.w = xat(x, S.dof, n + 3);
                *reinterpret_cast<float4*>(S.dnf + n) = d4;
            }
        }
        if constexpr (SEARCH)
            if (threadIdx.x == 0) S.partials[tile] = 0.0;
        return;
    }
    // the thread's own cells: types, diagonal (or weight), right-hand side, x
    uint32_t tq[NQ];
    Four<C> dq[NQ];
    Four<T> bq[NQ];
    float x0[NQ][4], w0[NQ][4], p0[NQ][4];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const uint32_t n = cell(q);
        tq[q] = *reinterpret_cast<const uint32_t*>(tf + n);
        dq[q] = load4(RW ? L.rwf : (FD ? L.dgf : L.dDf), n);   // the Dirichlet part, the weight or the diagonal
        if constexpr (MODE == 2 || MODE == 4 || MODE == 5 || MODE == 6) {
#pragma unroll
            for (int e = 0; e < 4; ++e) bq[q].v[e] = sf<T>(0.0f);
        } else {
            bq[q] = load4(b, n);
        }
        if constexpr (!Z) {
            const Four<T> xq = load4(x, n);
            if constexpr (SEARCH) {
                const float4 dold = *reinterpret_cast<const float4*>(S.dof + n);
                const float dv[4] = {dold.x, dold.y, dold.z, dold.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) x0[q][e] = lf(sf<T>(lf(xq.v[e]) + beta * dv[e]));
                *reinterpret_cast<float4*>(S.dnf + n) = make_float4(x0[q][0], x0[q][1], x0[q][2], x0[q][3]);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) x0[q][e] = lf(xq.v[e]);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) x0[q][e] = 0.0f;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            w0[q][e] = ((tq[q] >> (8 * e)) & 0xffu) == LIQUID ? FINE_FINE : 0.0f;
            p0[q][e] = w0[q][e] > 0.0f ? w0[q][e] * x0[q][e] : 0.0f;
        }
        const int hq = 4 + 4 * i4 + RS * (j + 1) + PS * (kk + 4 * q + 1);   // LDS place of its first cell
        if constexpr (NW) sw4[hq >> 2] = make_float4(w0[q][0], w0[q][1], w0[q][2], w0[q][3]);
        if constexpr (!Z) sp4[hq >> 2] = make_float4(p0[q][0], p0[q][1], p0[q][2], p0[q][3]);
    }

    int32_t nb[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) nb[d] = nbslot(t, w, d);
    const int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    // a staged cell outside the slot with no fine block there: the coarse cell, clamped
    auto coarse = [&](int li, int lj, int lk, float& wc, float& pc) {
        int ci = (w.x * 16 + li) >> 2, cj = (w.y * 16 + lj) >> 2, ck = (w.z * 16 + lk) >> 2;
        bool in = ci >= 0 && cj >= 0 && ck >= 0 && ci < n0 && cj < n1 && ck < n2;
        size_t cidx = cl(min(max(ci, 0), n0 - 1), min(max(cj, 0), n1 - 1), min(max(ck, 0), n2 - 1), n0, n1);
        wc = (in ? L.tc[cidx] : SOLID) == LIQUID ? FINE_COARSE : 0.0f;
        float xv = Z ? 0.0f : xat(xc, S.doc, cidx);
        pc = wc > 0.0f ? wc * xv : 0.0f;
    };
    // 4 cells along x from li0 (a multiple of 4) at row lj, plane lk, in this slot or else
    // in the face neighbour's slot `across` (0: none), staged at LDS place h
    auto stage4 = [&](int li0, int lj, int lk, int32_t across, int h) {
        bool inside = lj >= 0 && lj < 16 && lk >= 0 && lk < 16;
        int32_t sn = inside ? s : across;
        float wv[4], pv[4];
        if (sn != 0) {
            uint32_t idx = (uint32_t)(sn - 1) * 4096u + (uint32_t)(li0 + 16 * ((lj & 15) + 16 * (lk & 15)));
            uint32_t ty = *reinterpret_cast<const uint32_t*>(tf + idx);
            float xv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if constexpr (!Z) {
                const Four<T> x4 = load4(x, idx);
                if constexpr (SEARCH) {
                    const float4 d4 = *reinterpret_cast<const float4*>(S.dof + idx);
                    const float dv[4] = {d4.x, d4.y, d4.z, d4.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) xv[e] = lf(sf<T>(lf(x4.v[e]) + beta * dv[e]));
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) xv[e] = lf(x4.v[e]);
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                wv[e] = ((ty >> (8 * e)) & 0xffu) == LIQUID ? FINE_FINE : 0.0f;
                pv[e] = wv[e] > 0.0f ? wv[e] * xv[e] : 0.0f;
            }
        } else {
            float wc, pc;
            coarse(li0, lj, lk, wc, pc);   // the 4 cells share one coarse cell
#pragma unroll
            for (int e = 0; e < 4; ++e) { wv[e] = wc; pv[e] = pc; }
        }
        if constexpr (NW) sw4[h >> 2] = make_float4(wv[0], wv[1], wv[2], wv[3]);
        if constexpr (!Z) sp4[h >> 2] = make_float4(pv[0], pv[1], pv[2], pv[3]);
    };
    if (tid < 128) {   // the z layers below and above the tile: 2 x 16 rows x 4 quads
        int top = tid >> 6, lj = (tid >> 2) & 15, q = tid & 3;
        int lk = top ? k0 + PL : k0 - 1;
        stage4(4 * q, lj, lk, top ? nb[5] : nb[4], 4 + 4 * q + RS * (lj + 1) + PS * (top ? PL + 1 : 0));
    } else {           // the x sides: 2 x 16 rows x PL planes, one cell each
#pragma unroll
        for (int r = 0; r < NQ; ++r) {
            int e = tid - 128 + 128 * r, side = e & 1, lj = (e >> 1) & 15, p = e >> 5;
            int li = side ? 16 : -1, lk = k0 + p;
            int h = 4 + li + RS * (lj + 1) + PS * (p + 1);
            int32_t sn = side ? nb[1] : nb[0];
            float wv, pv;
            if (sn != 0) {
                uint32_t idx = (uint32_t)(sn - 1) * 4096u + (uint32_t)((li & 15) + 16 * (lj + 16 * lk));
                wv = tf[idx] == LIQUID ? FINE_FINE : 0.0f;
                float xv = Z ? 0.0f : xat(x, S.dof, idx);
                pv = wv > 0.0f ? wv * xv : 0.0f;
            } else {
                coarse(li, lj, lk, wv, pv);
            }
            if constexpr (NW) sw[h] = wv;
            if constexpr (!Z) sp[h] = pv;
        }
    }
    if (tid < 32 * NQ) {   // the y sides: 2 x PL planes x 4 quads
        int side = tid & 1, q = (tid >> 1) & 3, p = tid >> 3;
        int lj = side ? 16 : -1;
        stage4(4 * q, lj, k0 + p, side ? nb[3] : nb[2], 4 + 4 * q + RS * (lj + 1) + PS * (p + 1));
    }
    __syncthreads();

    double pq = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int hq = 4 + 4 * i4 + RS * (j + 1) + PS * (kk + 4 * q + 1);
        // the neighbours' couplings and products: x within the quad from registers, the
        // quad's x ends, y and z from LDS; order -x, +x, -y, +y, -z, +z
        float4 wym = make_float4(0, 0, 0, 0), wyp = wym, wzm = wym, wzp = wym;
        float wl = 0.0f, wr = 0.0f;
        if constexpr (NW) {
            wym = sw4[(hq - RS) >> 2]; wyp = sw4[(hq + RS) >> 2]; wzm = sw4[(hq - PS) >> 2]; wzp = sw4[(hq + PS) >> 2];
            wl = sw[hq - 1]; wr = sw[hq + 4];
        }
        float4 pym = make_float4(0, 0, 0, 0), pyp = pym, pzm = pym, pzp = pym;
        float pl = 0.0f, pr = 0.0f;
        if constexpr (!Z) {
            pym = sp4[(hq - RS) >> 2]; pyp = sp4[(hq + RS) >> 2]; pzm = sp4[(hq - PS) >> 2]; pzp = sp4[(hq + PS) >> 2];
            pl = sp[hq - 1]; pr = sp[hq + 4];
        }
        const float wy0[4] = {wym.x, wym.y, wym.z, wym.w}, wy1[4] = {wyp.x, wyp.y, wyp.z, wyp.w};
        const float wz0[4] = {wzm.x, wzm.y, wzm.z, wzm.w}, wz1[4] = {wzp.x, wzp.y, wzp.z, wzp.w};
        const float py0[4] = {pym.x, pym.y, pym.z, pym.w}, py1[4] = {pyp.x, pyp.y, pyp.z, pyp.w};
        const float pz0[4] = {pzm.x, pzm.y, pzm.z, pzm.w}, pz1[4] = {pzp.x, pzp.y, pzp.z, pzp.w};
        Four<T> oq;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t t0 = (tq[q] >> (8 * e)) & 0xffu;
            const float dd = lf(dq[q].v[e]), bb = lf(bq[q].v[e]);
            const float wn[6] = {e == 0 ? wl : w0[q][e - 1], e == 3 ? wr : w0[q][e + 1], wy0[e], wy1[e], wz0[e], wz1[e]};
            const float pn[6] = {e == 0 ? pl : p0[q][e - 1], e == 3 ? pr : p0[q][e + 1], py0[e], py1[e], pz0[e], pz1[e]};
            float diag = dd, off = 0.0f;
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                if constexpr (NW) diag = diag + wn[d];   // (with FD, dd is the whole diagonal)
                off = off + pn[d];
            }
            if (t0 != LIQUID) { diag = dd; off = 0.0f; }   // only a LIQUID cell couples
            const bool act = MODE == 6 ? true : ((MODE == 2 || SEARCH) ? t0 == LIQUID : t0 != 0);
            float res = 0.0f;
            if (act) {
                if constexpr (RW && MODE == 0) {
                    const float rw = dd;   // omega / diag
                    res = rw > 0.0f ? (1.0f - OMEGA) * x0[q][e] + rw * (bb + off) : 0.0f;
                } else if constexpr (RW) {   // the residual of the operator the weighted sweep relaxes
                    const float rw = dd, dg = rw > 0.0f ? OMEGA / rw : 0.0f;
                    res = bb - (dg * x0[q][e] - off);
                } else if constexpr (MODE == 5) {
                    res = diag > 0.0f ? __fdiv_rn(OMEGA, diag) : 0.0f;
                } else if constexpr (MODE == 6) {
                    res = diag;
                } else {
                    float ax = diag * x0[q][e] - off;
                    if (MODE == 0 || MODE == 3) res = diag > 0.0f ? x0[q][e] + __fdiv_rn(OMEGA * (bb - ax), diag) : 0.0f;
                    else if (MODE == 1) res = bb - ax;
                    else res = ax;
                }
                if constexpr (SEARCH) pq += (double)x0[q][e] * (double)res;
            }
            oq.v[e] = sf<T>(res);
        }
        store4_nt(out, cell(q), oq);
    }
    if constexpr (SEARCH) {
        double sum = block_sum_red(pq, lds);
        if (threadIdx.x == 0) S.partials[tile] = sum;
    }
}
template <int MODE, class T, class C, int NQ = 1>
__global__ __launch_bounds__(256) void k_rowfine16(T* __restrict__ out, const T* __restrict__ x, const T* __restrict__ b,
                                                   const T* __restrict__ xc, CompT<C> L, STopo t, SearchArgs S) {
    rowfine16_wg<MODE, T, C, NQ>(xcd_tile(blockIdx.x, gridDim.x), out, x, b, xc, L, t, S);
}

// The reduced-precision level-0 Jacobi sweep from x = 0: out = r b with r = omega / diag
// per cell (mode 5 of k_rowfine16), on non-solid cells; four cells a thread.
template <class T>
__global__ __launch_bounds__(256) void k_jacobi0w(T* __restrict__ out, const T* __restrict__ b, const T* __restrict__ rw,
                                                  const uint8_t* __restrict__ tf, const int4* __restrict__ where, int32_t nf) {
    const uint32_t n = 4u * (uint32_t)(xcd_tile(blockIdx.x, gridDim.x) * 256u + threadIdx.x);
    if (n >= (uint32_t)nf) return;
    Four<T> o;
    if (where[n >> 12].x < 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) o.v[e] = sf<T>(0.0f);
    } else {
        const uint32_t ty = *reinterpret_cast<const uint32_t*>(tf + n);
        const Four<T> bq = load4(b, n), rq = load4(rw, n);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float r = lf(rq.v[e]);
            o.v[e] = sf<T>(((ty >> (8 * e)) & 0xffu) != 0 && r > 0.0f ? r * lf(bq.v[e]) : 0.0f);
        }
    }
    *reinterpret_cast<Four<T>*>(out + n) = o;
}

// The reduced-precision V-cycle's entry on level 0's fine part in one pass: b = r c in T
// with k_convert's scale (c = 1 / scal[slot], 1 when slot < 0) and rounding, stored for
// the rest of the V-cycle, and k_jacobi0w's sweep from zero on it.
template <class T>
__device__ __forceinline__ void jacobi0w_cvt_wg(const uint32_t tile, T* __restrict__ out, T* __restrict__ bh,
                                                const float* __restrict__ r, const T* __restrict__ rw,
                                                const uint8_t* __restrict__ tf, const int4* __restrict__ where,
                                                int32_t nf, const double* __restrict__ scal, int slot) {
    float c = 1.0f;
    if (slot >= 0) c = (float)(1.0 / scal[slot]);
    if (!(c == c) || c == 0.0f || !(fabsf(c) <= 3.0e38f)) c = 1.0f;
    const uint32_t n = 4u * (tile * 256u + threadIdx.x);
    if (n >= (uint32_t)nf) return;
    const float4 rv = *reinterpret_cast<const float4*>(r + n);
    const float ra[4] = {rv.x, rv.y, rv.z, rv.w};
    Four<T> bq, o;
#pragma unroll
    for (int e = 0; e < 4; ++e) bq.v[e] = sf<T>(ra[e] * c);
    *reinterpret_cast<Four<T>*>(bh + n) = bq;
    if (where[n >> 12].x < 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) o.v[e] = sf<T>(0.0f);
    } else {
        const uint32_t ty = *reinterpret_cast<const uint32_t*>(tf + n);
        const Four<T> rq = load4(rw, n);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float w = lf(rq.v[e]);
            o.v[e] = sf<T>(((ty >> (8 * e)) & 0xffu) != 0 && w > 0.0f ? w * lf(bq.v[e]) : 0.0f);
        }
    }
    store4_nt(out, n, o);
}

// Level 1 (8^3 a slot), staged in LDS: a workgroup per slot stages x of the block and its
// face layers (face neighbours' slots, else the coarse cells, clamped), then each thread
// computes 2 cells with finerow's couplings and order.
template <int MODE, class T, class C>
__device__ __forceinline__ void rowfine8_wg(const uint32_t tile, T* __restrict__ out, const T* __restrict__ x,
                                            const T* __restrict__ b, const T* __restrict__ xc, const CompT<C>& L,
                                            const STopo& t) {
    constexpr bool Z = MODE == 3;
    constexpr int H8 = 10;
    __shared__ T sx[Z ? 1 : H8 * H8 * H8];
    const int s = (int)tile + 1;
    const uint32_t cb = (uint32_t)(s - 1) * 512u, fb = (uint32_t)(s - 1) * 576u;
    int4 w = t.where[s - 1];
    if (w.x < 0) {
        for (int c = threadIdx.x; c < 512; c += 256) out[cb + c] = sf<T>(0.0f);
        return;
    }
    if constexpr (!Z) {
        int32_t nb[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) nb[d] = nbslot(t, w, d);
        const int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
        for (int c = threadIdx.x; c < 512; c += 256) {
            int i = c & 7, j = (c >> 3) & 7, k = c >> 6;
            sx[(i + 1) + H8 * ((j + 1) + H8 * (k + 1))] = x[cb + c];
        }
        for (int e = threadIdx.x; e < 6 * 64; e += 256) {
            int d = e >> 6, p = e & 7, q = (e >> 3) & 7, ax = d >> 1, hi = d & 1;
            int li = ax == 0 ? (hi ? 8 : -1) : p, lj = ax == 1 ? (hi ? 8 : -1) : (ax == 0 ? p : q), lk = ax == 2 ? (hi ? 8 : -1) : q;
            int h = (li + 1) + H8 * ((lj + 1) + H8 * (lk + 1));
            int32_t sn = nb[d];
            if (sn != 0) {
                sx[h] = x[(uint32_t)(sn - 1) * 512u + (uint32_t)((li & 7) + 8 * ((lj & 7) + 8 * (lk & 7)))];
            } else {
                int ci = (w.x * 8 + li) >> 1, cj = (w.y * 8 + lj) >> 1, ck = (w.z * 8 + lk) >> 1;
                sx[h] = xc[cl(min(max(ci, 0), n0 - 1), min(max(cj, 0), n1 - 1), min(max(ck, 0), n2 - 1), n0, n1)];
            }
        }
        __syncthreads();
    }
    for (int c = threadIdx.x; c < 512; c += 256) {
        int i = c & 7, j = (c >> 3) & 7, k = c >> 6;
        int h = (i + 1) + H8 * ((j + 1) + H8 * (k + 1));
        uint32_t n = cb + c;
        bool act = MODE == 2 ? L.actf[n] == LIQUID : L.actf[n] != 0;
        float res = 0.0f;
        if (act) {
            float cf[6] = {lf(L.cf.c[0][fb + ffn(8, 0, i, j, k)]), lf(L.cf.c[0][fb + ffn(8, 0, i + 1, j, k)]),
                           lf(L.cf.c[1][fb + ffn(8, 1, i, j, k)]), lf(L.cf.c[1][fb + ffn(8, 1, i, j + 1, k)]),
                           lf(L.cf.c[2][fb + ffn(8, 2, i, j, k)]), lf(L.cf.c[2][fb + ffn(8, 2, i, j, k + 1)])};
            const int off6[6] = {-1, 1, -H8, H8, -H8 * H8, H8 * H8};
            float diag = lf(L.dDf[n]), off = 0.0f;
#pragma unroll
            for (int d = 0; d < 6; ++d)
                if (cf[d] > 0.0f) {
                    diag += cf[d];
                    off += cf[d] * (Z ? 0.0f : lf(sx[Z ? 0 : h + off6[d]]));
                }
            float x0 = Z ? 0.0f : lf(sx[Z ? 0 : h]);
            float ax = diag * x0 - off;
            if (MODE == 0 || MODE == 3) res = diag > 0.0f ? x0 + __fdiv_rn(OMEGA * (lf(b[n]) - ax), diag) : 0.0f;
            else if (MODE == 1) res = lf(b[n]) - ax;
            else res = ax;
        }
        out[n] = sf<T>(res);
    }
}
template <int MODE, class T, class C>
__global__ __launch_bounds__(256) void k_rowfine8(T* __restrict__ out, const T* __restrict__ x, const T* __restrict__ b,
                                                  const T* __restrict__ xc, CompT<C> L, STopo t) {
    rowfine8_wg<MODE, T, C>(xcd_tile(blockIdx.x, gridDim.x), out, x, b, xc, L, t);
}

// Coarse cells 256 cb .. 256 cb + 255; the search's partial sum goes to slot `ps`.
template <int N, int MODE, class T, class C>
__device__ __forceinline__ void rowcoarse_wg(const uint32_t cb, const uint32_t ps, T* __restrict__ out,
                                             const T* __restrict__ x, const T* __restrict__ b, const T* __restrict__ xf,
                                             const CompT<C>& L, const STopo& t, const SearchArgs& S) {
    constexpr bool SEARCH = MODE == 4;
    __shared__ double lds[4];
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)cb * 256u + threadIdx.x;
    bool live = l < (size_t)n0 * n1 * n2;
    if (!SEARCH && !live) return;
    const float beta = SEARCH ? (float)S.scal[SS_BETA] : 0.0f;
    double pq = 0.0;
    if (live) {
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    // every load of the common row (no covered neighbour) issued in one batch, ahead of
    // the mask and type tests that decide which sums it takes
    const unsigned msk = L.cmask[l];
    const bool row = L.actc[l] != 0 && !(msk & 1u);
    const LevelTypes ty{L.actf, L.tc};   // level 0 (N = 16): the couplings from the types
    const LevelTypes* typ = N == 16 ? &ty : nullptr;
    auto fx = [&](uint32_t i) -> float {
        if constexpr (SEARCH) return lf(xf[i]) + beta * S.dof[i];
        else if constexpr (MODE == 3) return 0.0f;
        else return lf(xf[i]);
    };
    auto cx = [&](size_t i) -> float {
        if constexpr (SEARCH) return lf(x[i]) + beta * S.doc[i];
        else if constexpr (MODE == 3) return 0.0f;
        else return lf(x[i]);
    };
    const CoarsePre<float> P = coarserow_pre<C, float>(t, cx, L.cc, L.dDc, I, J, K, typ);
    const float xl = cx(l);
    const float bl = (MODE == 0 || MODE == 1 || MODE == 3) ? lf(b[l]) : 0.0f;
    if constexpr (SEARCH) S.dnc[l] = xl;
    float diag, off;
    if (msk & 0x7eu) {
        diag = 0.0f;
        off = 0.0f;
        if (row) coarserow_post<N, C, float>(t, msk, fx, L.cf, P, I, J, K, diag, off, typ);
    } else {
        coarserow_flat(P, diag, off);
    }
    const float ax = diag * xl - off;
    float res;
    if (MODE == 0 || MODE == 3) res = diag > 0.0f ? xl + __fdiv_rn(OMEGA * (bl - ax), diag) : 0.0f;
    else if (MODE == 1) res = bl - ax;
    else res = ax;
    res = row ? res : 0.0f;
    if constexpr (SEARCH) pq = row ? (double)xl * (double)res : 0.0;
    out[l] = sf<T>(res);
    }
    if constexpr (SEARCH) {
        double sum = block_sum_red(pq, lds);
        if (threadIdx.x == 0) S.partials[S.coarse0 + ps] = sum;
    }
}
template <int N, int MODE, class T, class C>
__global__ __launch_bounds__(256) void k_rowcoarse(T* __restrict__ out, const T* __restrict__ x, const T* __restrict__ b,
                                                   const T* __restrict__ xf, CompT<C> L, STopo t, SearchArgs S) {
    // neighbouring rows on one XCD
    rowcoarse_wg<N, MODE, T, C>(xcd_tile(blockIdx.x, gridDim.x), blockIdx.x, out, x, b, xf, L, t, S);
}

// One launch for a whole composite-level pass: the fine part's workgroups (tiles of
// level 0, slots of level 1) and the coarse part's, interleaved in proportion along the
// XCD-ordered workgroup index, so the coarse rows (instruction-bound next to blocks)
// overlap the fine tiles (bandwidth-bound) and the pass drains once. Workgroup g takes
// coarse block floor(g C / G) when that count steps at g, else fine workgroup g minus
// the coarse blocks before it: each index of either part is taken exactly once, in order.
template <int N, int MODE, class T, class C, int NQ>
__global__ __launch_bounds__(256) void k_rowcomp(T* __restrict__ outf, T* __restrict__ outc, const T* __restrict__ xf,
                                                 const T* __restrict__ xc, const T* __restrict__ bf,
                                                 const T* __restrict__ bc, CompT<C> L, STopo t, SearchArgs S, uint32_t ncb) {
    const uint32_t G = gridDim.x, g = xcd_tile(blockIdx.x, G);
    const uint32_t c0 = (uint32_t)(((uint64_t)g * ncb) / G), c1 = (uint32_t)(((uint64_t)(g + 1) * ncb) / G);
    if (c1 > c0) rowcoarse_wg<N, MODE, T, C>(c0, c0, outc, xc, bc, xf, L, t, S);
    else if constexpr (N == 16) rowfine16_wg<MODE, T, C, NQ>(g - c0, outf, xf, bf, xc, L, t, S);
    else rowfine8_wg<MODE, T, C>(g - c0, outf, xf, bf, xc, L, t);
}
// The reduced-precision V-cycle's entry as one launch in the same interleave: level 0's
// fine part through jacobi0w_cvt_wg (r converted and stored, the sweep from zero) and the
// coarse rows of that sweep (their r already converted).
template <class T>
__global__ __launch_bounds__(256) void k_head0(T* __restrict__ outf, T* __restrict__ outc, T* __restrict__ bh,
                                               const float* __restrict__ r, const T* __restrict__ xc, const T* __restrict__ bc,
                                               const T* __restrict__ xf, CompT<T> L, STopo t, int32_t nf,
                                               const double* __restrict__ scal, int slot, uint32_t ncb) {
    const uint32_t G = gridDim.x, g = xcd_tile(blockIdx.x, G);
    const uint32_t c0 = (uint32_t)(((uint64_t)g * ncb) / G), c1 = (uint32_t)(((uint64_t)(g + 1) * ncb) / G);
    if (c1 > c0) rowcoarse_wg<16, 3, T, T>(c0, c0, outc, xc, bc, xf, L, t, SearchArgs{});
    else jacobi0w_cvt_wg<T>(g - c0, outf, bh, r, L.rwf, L.actf, t.where, nf, scal, slot);
}

// ─── dense levels (sflip.jl:1574-1657) ───────────────────────────────────────

template <class C> struct DViewT {
    int3 n;
    const C* c[3];
    const C* dD;
    const uint8_t* act;
};
// A dense level's cell: diag and the neighbours' sum; all loads issued before any use
// (an index off the grid reads the cell itself, and its coupling adds nothing).
template <class T, class C, bool Z = false>
__device__ inline void denserow(const DViewT<C>& L, const T* __restrict__ x, int i, int j, int k, float& diag, float& off) {
    int n0 = L.n.x, n1 = L.n.y, n2 = L.n.z;
    size_t own = cl(i, j, k, n0, n1), sy = (size_t)n0, sz = (size_t)n0 * n1;
    float c0 = lf(L.c[0][cl(i, j, k, n0 + 1, n1)]), c1 = lf(L.c[0][cl(i + 1, j, k, n0 + 1, n1)]);
    float c2 = lf(L.c[1][cl(i, j, k, n0, n1 + 1)]), c3 = lf(L.c[1][cl(i, j + 1, k, n0, n1 + 1)]);
    float c4 = lf(L.c[2][cl(i, j, k, n0, n1)]), c5 = lf(L.c[2][cl(i, j, k + 1, n0, n1)]);
    auto xv = [&](bool ok, size_t idx) -> float {
        if constexpr (Z) return 0.0f;   // x = 0: nothing to read
        return lf(x[ok ? idx : own]);
    };
    bool b0 = i > 0, b1 = i < n0 - 1, b2 = j > 0, b3 = j < n1 - 1, b4 = k > 0, b5 = k < n2 - 1;
    float x0 = xv(b0, own - 1), x1 = xv(b1, own + 1), x2 = xv(b2, own - sy), x3 = xv(b3, own + sy);
    float x4 = xv(b4, own - sz), x5 = xv(b5, own + sz);
    diag = lf(L.dD[own]);
    off = 0.0f;
    if (c0 > 0.0f && b0) { diag += c0; off += c0 * x0; }
    if (c1 > 0.0f && b1) { diag += c1; off += c1 * x1; }
    if (c2 > 0.0f && b2) { diag += c2; off += c2 * x2; }
    if (c3 > 0.0f && b3) { diag += c3; off += c3 * x3; }
    if (c4 > 0.0f && b4) { diag += c4; off += c4 * x4; }
    if (c5 > 0.0f && b5) { diag += c5; off += c5 * x5; }
}
// mode 0 Jacobi, 1 residual, 3 Jacobi from x = 0 (x not read)
template <int MODE, class T, class C>
__global__ __launch_bounds__(256) void k_rowdense(T* __restrict__ out, const T* __restrict__ x, const T* __restrict__ b,
                                                  DViewT<C> L) {
    size_t l = (size_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;   // neighbouring rows on one XCD
    if (l >= (size_t)L.n.x * L.n.y * L.n.z) return;
    int i, j, k;
    unlin(l, L.n.x, L.n.y, i, j, k);
    // the row's loads issued with the activity flag's, the result selected after
    const bool act = L.act[l] != 0;
    float diag, off;
    denserow<T, C, MODE == 3>(L, x, i, j, k, diag, off);
    const float xl = MODE == 3 ? 0.0f : lf(x[l]), bl = lf(b[l]);
    const float ax = diag * xl - off;
    const float res = MODE != 1 ? (diag > 0.0f ? xl + __fdiv_rn(OMEGA * (bl - ax), diag) : 0.0f) : bl - ax;
    out[l] = sf<T>(act ? res : 0.0f);
}
// The coarser level's right-hand side, the residual and its restriction in one pass:
// each coarser cell forms its eight finer cells' residuals (k_rowdense mode 1's
// arithmetic, rounded to T as that kernel would store them), sums them in x-fastest
// order and multiplies by `scale` (1 in the reference's solve; the reduced-precision
// hierarchy carries each dense level's equations divided by 4 per level, so its values
// stay in FP16's range). No residual array is written.
template <class T, class C>
__global__ void k_residrestrict(T* __restrict__ bc, const T* __restrict__ x, const T* __restrict__ b, DViewT<C> L, int3 nc,
                                float scale) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)nc.x * nc.y * nc.z) return;
    int I, J, K;
    unlin(l, nc.x, nc.y, I, J, K);
    float acc = 0.0f;
    for (int c = 0; c < 2; ++c)
        for (int bb = 0; bb < 2; ++bb)
            for (int a = 0; a < 2; ++a) {
                const int i = 2 * I + a, j = 2 * J + bb, k = 2 * K + c;
                const size_t f = cl(i, j, k, L.n.x, L.n.y);
                float res = 0.0f;
                if (L.act[f] != 0) {
                    float diag, off;
                    denserow<T, C, false>(L, x, i, j, k, diag, off);
                    res = lf(b[f]) - (diag * lf(x[f]) - off);
                }
                acc += lf(sf<T>(res));
            }
    bc[l] = sf<T>(acc * scale);
}
template <class T>
__global__ void k_prolongdense(T* __restrict__ xf, const T* __restrict__ xc, const uint8_t* __restrict__ actf, int3 nf, int3 nc,
                               float boost) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)nf.x * nf.y * nf.z) return;
    int i, j, k;
    unlin(l, nf.x, nf.y, i, j, k);
    if (actf[l] != 0) xf[l] = sf<T>(lf(xf[l]) + boost * lf(xc[cl(i >> 1, j >> 1, k >> 1, nc.x, nc.y)]));
}
// sflip.jl:1623-1657 coarsest_kernel!: red-black Gauss-Seidel from zero in one workgroup,
// in arithmetic A (FP64 in the mixed-precision solve).
template <class T, class C, class A>
__global__ __launch_bounds__(1024) void k_scoarsest(T* __restrict__ x, const T* __restrict__ b, DViewT<C> L, int sweeps) {
    __shared__ A xs[1024];
    int t = threadIdx.x;
    int n0 = L.n.x, n1 = L.n.y, n2 = L.n.z, N = n0 * n1 * n2;
    bool inside = t < N;
    int i = t % n0, j = (t / n0) % n1, k = t / (n0 * n1);
    bool act = inside && L.act[t] != 0;
    A bi = inside ? (A)lf(b[t]) : A(0);
    A dg = A(0), cw[6] = {0, 0, 0, 0, 0, 0};
    int nb[6] = {0, 0, 0, 0, 0, 0};
    if (act) {
        dg = (A)lf(L.dD[t]);
        float c;
        c = lf(L.c[0][cl(i, j, k, n0 + 1, n1)]);     if (c > 0.0f && i > 0) { cw[0] = c; nb[0] = t - 1; }
        c = lf(L.c[0][cl(i + 1, j, k, n0 + 1, n1)]); if (c > 0.0f && i < n0 - 1) { cw[1] = c; nb[1] = t + 1; }
        c = lf(L.c[1][cl(i, j, k, n0, n1 + 1)]);     if (c > 0.0f && j > 0) { cw[2] = c; nb[2] = t - n0; }
        c = lf(L.c[1][cl(i, j + 1, k, n0, n1 + 1)]); if (c > 0.0f && j < n1 - 1) { cw[3] = c; nb[3] = t + n0; }
        c = lf(L.c[2][cl(i, j, k, n0, n1)]);         if (c > 0.0f && k > 0) { cw[4] = c; nb[4] = t - n0 * n1; }
        c = lf(L.c[2][cl(i, j, k + 1, n0, n1)]);     if (c > 0.0f && k < n2 - 1) { cw[5] = c; nb[5] = t + n0 * n1; }
    }
    // the diagonal is the same every pass; a pass reads the six neighbours together and
    // adds the coupled ones (an absent neighbour reads cell 0 and adds nothing)
    A diag = dg;
#pragma unroll
    for (int m = 0; m < 6; ++m)
        if (cw[m] > A(0)) diag += cw[m];
    xs[t] = A(0);
    __syncthreads();
    int par = (i + j + k + 1) & 1;
    for (int s = 1; s <= 2 * sweeps; ++s)
        for (int half = 0; half < 2; ++half) {
            int colour = s <= sweeps ? half : 1 - half;
            if (act && par == colour) {
                A v[6];
#pragma unroll
                for (int m = 0; m < 6; ++m) v[m] = xs[nb[m]];
                A acc = bi;
#pragma unroll
                for (int m = 0; m < 6; ++m) acc = cw[m] > A(0) ? acc + cw[m] * v[m] : acc;
                xs[t] = diag > A(0) ? acc / diag : A(0);
            }
            __syncthreads();
        }
    if (inside) x[t] = sf<T>((float)xs[t]);
}

// ─── between the composite levels (sflip.jl:1661-1705) ───────────────────────

template <class T>
__device__ __forceinline__ void restrict01_el(const int32_t n, T* __restrict__ b1, const T* __restrict__ r0, const STopo& t,
                                              int nslots) {
    if (n >= nslots * SBCELLS1) return;
    int s = (n >> 9) + 1, c = n & 511, i = c & 7, j = (c >> 3) & 7, k = c >> 6;
    if (t.where[s - 1].x < 0) return;
    float acc = 0.0f;
    for (int cz = 0; cz < 2; ++cz)
        for (int by = 0; by < 2; ++by)
            for (int ax = 0; ax < 2; ++ax)
                acc += lf(r0[(s - 1) * SBCELLS + (2 * i + ax) + 16 * ((2 * j + by) + 16 * (2 * k + cz))]);
    b1[n] = sf<T>(acc);
}
// Level 1's right-hand side in one launch: the fine part restricted from level 0's
// residual in the first `nfb` workgroups, the coarse part (shared by both levels) copied.
template <class T>
__global__ __launch_bounds__(256) void k_restrictcomp(T* __restrict__ b1f, const T* __restrict__ r0f, T* __restrict__ b1c,
                                                      const T* __restrict__ r0c, size_t nc, STopo t, int nslots, uint32_t nfb) {
    if (blockIdx.x < nfb) {
        restrict01_el<T>((int32_t)(blockIdx.x * 256u + threadIdx.x), b1f, r0f, t, nslots);
    } else {
        const size_t l = (size_t)(blockIdx.x - nfb) * 256u + threadIdx.x;
        if (l < nc) b1c[l] = r0c[l];
    }
}
template <class T>
__global__ void k_restrict12(T* __restrict__ b2, const T* __restrict__ r1f, const T* __restrict__ r1c, STopo t) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    if (covered(t, I, J, K)) {
        float acc = 0.0f;
        for (int c = 0; c < 2; ++c)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) acc += lf(r1f[cellat<8>(t, 2 * I + a, 2 * J + b, 2 * K + c)]);
        b2[l] = sf<T>(acc);
    } else {
        b2[l] = r1c[l];
    }
}
template <class T>
__device__ __forceinline__ void prolong21_el(const int32_t n, T* __restrict__ x1f, const T* __restrict__ x2,
                                             const uint8_t* __restrict__ act1, const STopo& t, float boost, int nslots) {
    if (n >= nslots * SBCELLS1) return;
    int s = (n >> 9) + 1, c = n & 511, i = c & 7, j = (c >> 3) & 7, k = c >> 6;
    int4 w = t.where[s - 1];
    if (w.x < 0 || act1[n] == 0) return;
    x1f[n] = sf<T>(lf(x1f[n]) + boost * lf(x2[cl((w.x * 8 + i) >> 1, (w.y * 8 + j) >> 1, (w.z * 8 + k) >> 1, t.g.nc(0), t.g.nc(1))]));
}
// A composite level's prolongation in one launch: its fine part (level 0 from level 1,
// 1024 cells a workgroup, LVL 0; level 1 from the first dense level, LVL 1) in the first
// `nfb` workgroups, its coarse part (from the coarser grid's coarse part or level) after.
template <class T, int LVL>
__global__ __launch_bounds__(256) void k_prolongcomp(T* __restrict__ xf, const T* __restrict__ fromf,
                                                     const uint8_t* __restrict__ maskf, T* __restrict__ xc,
                                                     const T* __restrict__ fromc, const uint8_t* __restrict__ actc, STopo t,
                                                     float boost, int32_t nfine, uint32_t nfb) {
    if (blockIdx.x < nfb) {
        if constexpr (LVL == 0) prolong10_wg<T>(xcd_tile(blockIdx.x, nfb), xf, fromf, maskf, t, boost, nfine);
        else prolong21_el<T>((int32_t)(blockIdx.x * 256u + threadIdx.x), xf, fromf, maskf, t, boost, nfine);
    } else {
        prolongcoarse_el<T>((size_t)(blockIdx.x - nfb) * 256u + threadIdx.x, xc, fromc, actc, t, boost);
    }
}
template <class T>
__device__ __forceinline__ void prolongcoarse_el(const size_t l, T* __restrict__ xc, const T* __restrict__ from,
                                                 const uint8_t* __restrict__ act, const STopo& t, float boost) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    if (act[l] != 0 && !covered(t, I, J, K)) xc[l] = sf<T>(lf(xc[l]) + boost * lf(from[l]));
}
// Four cells of an x row a thread (two level-1 parents); a workgroup's 1024 cells lie in
// one slot, so the slot's place is uniform.
template <class T>
__device__ __forceinline__ void prolong10_wg(const uint32_t tile, T* __restrict__ x0f, const T* __restrict__ x1f,
                                             const uint8_t* __restrict__ tf, const STopo& t, float boost, int32_t nf) {
    int32_t n = (int32_t)(tile * 1024u) + 4 * (int32_t)threadIdx.x;
    if ((int32_t)(tile * 1024u) >= nf) return;
    int s = (int)((tile * 1024u) >> 12) + 1, c = n & 4095;
    if (t.where[s - 1].x < 0) return;
    int i = c & 15, j = (c >> 4) & 15, k = c >> 8;
    uchar4 ty = *reinterpret_cast<const uchar4*>(tf + n);
    if (ty.x != LIQUID && ty.y != LIQUID && ty.z != LIQUID && ty.w != LIQUID) return;
    const T* par = x1f + (size_t)(s - 1) * SBCELLS1 + (i >> 1) + 8 * ((j >> 1) + 8 * (k >> 1));
    float p0 = boost * lf(par[0]), p1 = boost * lf(par[1]);
    T* xo = x0f + n;
    if (ty.x == LIQUID) xo[0] = sf<T>(lf(xo[0]) + p0);
    if (ty.y == LIQUID) xo[1] = sf<T>(lf(xo[1]) + p0);
    if (ty.z == LIQUID) xo[2] = sf<T>(lf(xo[2]) + p1);
    if (ty.w == LIQUID) xo[3] = sf<T>(lf(xo[3]) + p1);
}

// Conversions between the Krylov vectors (FP32) and the reduced-precision hierarchy:
// y = x * mult * scal[slot] (or * mult when slot < 0), over the fine and the coarse range,
// four values a thread (both parts are multiples of 4 long).
template <class Y, class X>
__global__ void k_convert(Y* __restrict__ yf, const X* __restrict__ xf, size_t nf, Y* __restrict__ yc, const X* __restrict__ xc,
                          size_t nc, const double* __restrict__ scal, int slot, int invert) {
    float c = 1.0f;
    if (slot >= 0) c = invert ? (float)(1.0 / scal[slot]) : (float)scal[slot];
    if (!(c == c) || c == 0.0f || !(fabsf(c) <= 3.0e38f)) c = 1.0f;
    size_t nf4 = nf / 4, ng = nf4 + nc / 4;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (size_t)gridDim.x * blockDim.x) {
        const Four<X>* xs = g < nf4 ? reinterpret_cast<const Four<X>*>(xf) + g : reinterpret_cast<const Four<X>*>(xc) + (g - nf4);
        Four<Y>* ys = g < nf4 ? reinterpret_cast<Four<Y>*>(yf) + g : reinterpret_cast<Four<Y>*>(yc) + (g - nf4);
        Four<X> a = *xs;
        Four<Y> o;
#pragma unroll
        for (int e = 0; e < 4; ++e) o.v[e] = sf<Y>(lf(a.v[e]) * c);
        *ys = o;
    }
}
// A coupling array in reduced precision, times `scale`.
template <class Y>
__global__ void k_castc(Y* __restrict__ y, const float* __restrict__ x, size_t n, float scale) {
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (size_t)gridDim.x * blockDim.x)
        y[l] = sf<Y>(x[l] * scale);
}

// ─── vectors: dot products over the unknowns, axpy and friends (sflip.jl:1766-1822) ──

// error flags (the reference's messages, sflip.jl:1880-1900)
enum { EF_RHO_NONFINITE = 1, EF_RHO_NONPOS = 2, EF_DEN_NONPOS = 4, EF_ALPHA_NONFINITE = 8, EF_RHONEW_NONPOS = 16,
       EF_BETA_NONFINITE = 32 };

constexpr int DOT_BLOCKS = 2048;   // partials: 2 x DOT_BLOCKS doubles at most


// The vector kernels run over the concatenated (fine pool, coarse grid) range in groups of
// four consecutive cells (float4 / uchar4 loads): the fine part is whole slots and the
// coarse grid's sides are multiples of 4, so a group never straddles the two parts or two
// slots. Group g's mask of unknowns (sflip.jl:1772-1805: LIQUID fine cells of live slots,
// LIQUID uncovered coarse cells), one bit per cell.
__device__ inline unsigned unknown4(size_t g, size_t nf4, const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc,
                                    const uint8_t* __restrict__ cm, const STopo& t) {
    if (g < nf4) {
        if (t.where[g >> 10].x < 0) return 0u;
        uchar4 ty = reinterpret_cast<const uchar4*>(tf)[g];
        return (ty.x == LIQUID) | ((ty.y == LIQUID) << 1) | ((ty.z == LIQUID) << 2) | ((ty.w == LIQUID) << 3);
    }
    size_t h = g - nf4;
    uchar4 ty = reinterpret_cast<const uchar4*>(tc)[h], c = reinterpret_cast<const uchar4*>(cm)[h];
    return (ty.x == LIQUID && !(c.x & 1)) | ((ty.y == LIQUID && !(c.y & 1)) << 1) | ((ty.z == LIQUID && !(c.z & 1)) << 2) |
           ((ty.w == LIQUID && !(c.w & 1)) << 3);
}
// Group g of a pressure-sized vector (fine part, coarse part).
__device__ inline float4& grp(float* f, float* c, size_t g, size_t nf4) {
    return g < nf4 ? reinterpret_cast<float4*>(f)[g] : reinterpret_cast<float4*>(c)[g - nf4];
}
__device__ inline float4 grp(const float* f, const float* c, size_t g, size_t nf4) {
    return g < nf4 ? reinterpret_cast<const float4*>(f)[g] : reinterpret_cast<const float4*>(c)[g - nf4];
}
__device__ inline double dot4(unsigned m, float4 a, float4 b) {
    double v = 0;
    if (m & 1) v += (double)a.x * (double)b.x;
    if (m & 2) v += (double)a.y * (double)b.y;
    if (m & 4) v += (double)a.z * (double)b.z;
    if (m & 8) v += (double)a.w * (double)b.w;
    return v;
}

// sflip.jl:1772-1805 inner: the sum over the unknowns, per workgroup.
__global__ __launch_bounds__(256) void k_sdot(double* __restrict__ partials, const float* __restrict__ af,
                                              const float* __restrict__ bf, const float* __restrict__ ac,
                                              const float* __restrict__ bc, const uint8_t* __restrict__ tf,
                                              const uint8_t* __restrict__ tc, const uint8_t* __restrict__ cm, STopo t,
                                              int32_t nf) {
    __shared__ double lds[4];
    size_t nf4 = (size_t)nf / 4, ng = nf4 + (size_t)t.g.nc(0) * t.g.nc(1) * t.g.nc(2) / 4;
    double v = 0;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (size_t)gridDim.x * blockDim.x) {
        unsigned m = unknown4(g, nf4, tf, tc, cm, t);
        if (m) v += dot4(m, grp(af, ac, g, nf4), grp(bf, bc, g, nf4));
    }
    double s = block_sum_red(v, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

enum { SF_SET = 0, SF_NORM, SF_RHO0, SF_ALPHA, SF_BETA };
// The sum of np partials by one workgroup of FINISH_THREADS: each thread adds its strided
// share with four accumulators (loads in flight together), then the fixed tree of
// block_sum_red; the order depends on np only.
constexpr int FINISH_THREADS = 1024;
__device__ inline double finish_sum(const double* __restrict__ p, int np, double* lds) {
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int l = threadIdx.x;
    for (; l + 3 * FINISH_THREADS < np; l += 4 * FINISH_THREADS) {
        a0 += p[l]; a1 += p[l + FINISH_THREADS]; a2 += p[l + 2 * FINISH_THREADS]; a3 += p[l + 3 * FINISH_THREADS];
    }
    for (; l < np; l += FINISH_THREADS) a0 += p[l];
    return block_sum_red((a0 + a1) + (a2 + a3), lds);
}
__global__ __launch_bounds__(FINISH_THREADS) void k_sfinish(const double* __restrict__ partials, int np,
                                                            double* __restrict__ scal, int op, int slot) {
    __shared__ double lds[FINISH_THREADS / 64];
    double s = finish_sum(partials, np, lds);
    if (threadIdx.x != 0) return;
    long long flags = (long long)scal[SS_FLAG];
    switch (op) {
        case SF_SET: scal[slot] = s; break;
        case SF_NORM: scal[slot] = sqrt(s); break;
        case SF_RHO0:   // rho = z.r, with the reference's checks (sflip.jl:1880-1882)
            scal[SS_RHO] = s;
            if (!isfinite(s)) flags |= EF_RHO_NONFINITE;
            else if (!(s > 0)) flags |= EF_RHO_NONPOS;
            break;
        case SF_ALPHA: {   // den = d.q, alpha = rho / den (sflip.jl:1885-1888)
            scal[SS_DEN] = s;
            if (!(s > 0)) flags |= EF_DEN_NONPOS;
            double al = (double)((float)scal[SS_RHO] / (float)s);
            if (!isfinite(al)) flags |= EF_ALPHA_NONFINITE;
            scal[SS_ALPHA] = al;
            break;
        }
        case SF_BETA: {   // z.rold; beta = max(0, (rho_new - z.rold) / rho) (sflip.jl:1895-1900)
            double rn = scal[SS_RHONEW];
            if (!(rn > 0)) flags |= EF_RHONEW_NONPOS;
            double be = (double)fmaxf(0.0f, ((float)rn - (float)s) / (float)scal[SS_RHO]);
            if (!isfinite(be)) flags |= EF_BETA_NONFINITE;
            scal[SS_BETA] = be;
            scal[SS_RHO] = rn;
            break;
        }
    }
    scal[SS_FLAG] = (double)flags;
}

// y = x + coef*y (XPBY) or y += coef*x (AXPY) over n values; coef = mult*scal[slot] (slot >= 0) or mult.
template <int XPBY>
__global__ void k_saxpy(float* __restrict__ y, const float* __restrict__ x, size_t n, const double* __restrict__ scal, int slot,
                        float mult) {
    float c = slot >= 0 ? mult * (float)scal[slot] : mult;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (size_t)gridDim.x * blockDim.x)
        y[l] = XPBY ? x[l] + c * y[l] : y[l] + c * x[l];
}

// The unknowns' mask of the inner product (sflip.jl:1772-1805), for position l of the
// concatenated (fine pool, coarse grid) range.
__device__ inline bool unknown(size_t l, int32_t nf, const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc,
                               const uint8_t* __restrict__ cm, const STopo& t) {
    if (l < (size_t)nf) return tf[l] == LIQUID && t.where[l >> 12].x >= 0;
    size_t m = l - nf;
    return tc[m] == LIQUID && !(cm[m] & 1);
}

// sflip.jl:1889-1892 in one pass: P += alpha d, rold = r, r -= alpha q, and r.r per
// workgroup (alpha from the device scalars).
// P += alpha d and the new residual r - alpha q into the other residual buffer (the old
// one stays where it is, as r_old for the flexible CG's beta), with r.r per workgroup.
__global__ __launch_bounds__(256) void k_fcg_update(double* __restrict__ partials, float* __restrict__ Pf,
                                                    float* __restrict__ Pc, const float* __restrict__ rf,
                                                    const float* __restrict__ rc, float* __restrict__ routf,
                                                    float* __restrict__ routc, const float* __restrict__ df,
                                                    const float* __restrict__ dc,
                                                    const float* __restrict__ qf, const float* __restrict__ qc,
                                                    const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc,
                                                    const uint8_t* __restrict__ cm, STopo t, int32_t nf,
                                                    const double* __restrict__ scal, int aslot) {
    __shared__ double lds[4];
    float al = (float)scal[aslot];
    size_t nf4 = (size_t)nf / 4, ng = nf4 + (size_t)t.g.nc(0) * t.g.nc(1) * t.g.nc(2) / 4;
    double v = 0;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (size_t)gridDim.x * blockDim.x) {
        float4 d = grp(df, dc, g, nf4), q = grp(qf, qc, g, nf4);
        float4& P = grp(Pf, Pc, g, nf4);
        float4 p = P, rv = grp(rf, rc, g, nf4);
        p.x = p.x + al * d.x; p.y = p.y + al * d.y; p.z = p.z + al * d.z; p.w = p.w + al * d.w;
        P = p;
        rv.x = rv.x + -al * q.x; rv.y = rv.y + -al * q.y; rv.z = rv.z + -al * q.z; rv.w = rv.w + -al * q.w;
        grp(routf, routc, g, nf4) = rv;
        unsigned m = unknown4(g, nf4, tf, tc, cm, t);
        if (m) v += dot4(m, rv, rv);
    }
    double s = block_sum_red(v, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
// The reduced-precision V-cycle's correction back to FP32 (z = zh scal[slot], as k_convert)
// fused with k_sdot2 on it: z.r and z.rold summed in k_sdot2's grouping (the same partials
// for the same z).
template <class T>
__global__ __launch_bounds__(256) void k_convert_dot2(double* __restrict__ partials, float* __restrict__ zf,
                                                      float* __restrict__ zc, const T* __restrict__ hf, const T* __restrict__ hc,
                                                      const float* __restrict__ rf, const float* __restrict__ rc,
                                                      const float* __restrict__ of, const float* __restrict__ oc,
                                                      const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc,
                                                      const uint8_t* __restrict__ cm, STopo t, int32_t nf,
                                                      const double* __restrict__ scal, int slot) {
    __shared__ double lds[4];
    float c = (float)scal[slot];
    if (!(c == c) || c == 0.0f || !(fabsf(c) <= 3.0e38f)) c = 1.0f;
    size_t nf4 = (size_t)nf / 4, ng = nf4 + (size_t)t.g.nc(0) * t.g.nc(1) * t.g.nc(2) / 4;
    double v1 = 0, v2 = 0;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (size_t)gridDim.x * blockDim.x) {
        const Four<T> h = g < nf4 ? reinterpret_cast<const Four<T>*>(hf)[g] : reinterpret_cast<const Four<T>*>(hc)[g - nf4];
        const float4 z = make_float4(lf(h.v[0]) * c, lf(h.v[1]) * c, lf(h.v[2]) * c, lf(h.v[3]) * c);
        grp(zf, zc, g, nf4) = z;
        unsigned m = unknown4(g, nf4, tf, tc, cm, t);
        if (!m) continue;
        v1 += dot4(m, z, grp(rf, rc, g, nf4));
        v2 += dot4(m, z, grp(of, oc, g, nf4));
    }
    double s1 = block_sum_red(v1, lds);
    __syncthreads();
    double s2 = block_sum_red(v2, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1;
        partials[gridDim.x + blockIdx.x] = s2;
    }
}
// z.r and z.rold in one pass (sflip.jl:1894-1897): partials[b] and partials[B + b].
__global__ __launch_bounds__(256) void k_sdot2(double* __restrict__ partials, const float* __restrict__ zf,
                                               const float* __restrict__ zc, const float* __restrict__ rf,
                                               const float* __restrict__ rc, const float* __restrict__ of,
                                               const float* __restrict__ oc, const uint8_t* __restrict__ tf,
                                               const uint8_t* __restrict__ tc, const uint8_t* __restrict__ cm, STopo t,
                                               int32_t nf) {
    __shared__ double lds[4];
    size_t nf4 = (size_t)nf / 4, ng = nf4 + (size_t)t.g.nc(0) * t.g.nc(1) * t.g.nc(2) / 4;
    double v1 = 0, v2 = 0;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (size_t)gridDim.x * blockDim.x) {
        unsigned m = unknown4(g, nf4, tf, tc, cm, t);
        if (!m) continue;
        float4 z = grp(zf, zc, g, nf4);
        v1 += dot4(m, z, grp(rf, rc, g, nf4));
        v2 += dot4(m, z, grp(of, oc, g, nf4));
    }
    double s1 = block_sum_red(v1, lds);
    __syncthreads();
    double s2 = block_sum_red(v2, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1;
        partials[gridDim.x + blockIdx.x] = s2;
    }
}
// beta from the two sums (sflip.jl:1895-1900), rho = rho_new.
__global__ __launch_bounds__(FINISH_THREADS) void k_sfinish_beta(const double* __restrict__ partials, int np,
                                                                 double* __restrict__ scal) {
    __shared__ double lds[FINISH_THREADS / 64];
    double rn = finish_sum(partials, np, lds);
    __syncthreads();
    double zo = finish_sum(partials + np, np, lds);
    if (threadIdx.x != 0) return;
    long long flags = (long long)scal[SS_FLAG];
    if (!(rn > 0)) flags |= EF_RHONEW_NONPOS;
    double be = (double)fmaxf(0.0f, ((float)rn - (float)zo) / (float)scal[SS_RHO]);
    if (!isfinite(be)) flags |= EF_BETA_NONFINITE;
    scal[SS_RHONEW] = rn;
    scal[SS_BETA] = be;
    scal[SS_RHO] = rn;
    scal[SS_FLAG] = (double)flags;
}
// y = x + c y (XPBY) or y += c x over the fine and the coarse range in one launch.
template <int XPBY>
__global__ void k_saxpy2(float* __restrict__ yf, const float* __restrict__ xf, size_t nf, float* __restrict__ yc,
                         const float* __restrict__ xc, size_t nc, const double* __restrict__ scal, int slot, float mult) {
    float c = slot >= 0 ? mult * (float)scal[slot] : mult;
    size_t nf4 = nf / 4, ng = nf4 + nc / 4;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (size_t)gridDim.x * blockDim.x) {
        float4 x = grp(xf, xc, g, nf4);
        float4& yr = grp(yf, yc, g, nf4);
        float4 y = yr;
        if (XPBY) { y.x = x.x + c * y.x; y.y = x.y + c * y.y; y.z = x.z + c * y.z; y.w = x.w + c * y.w; }
        else { y.x = y.x + c * x.x; y.y = y.y + c * x.y; y.z = y.z + c * x.z; y.w = y.w + c * x.w; }
        yr = y;
    }
}

// ─── iterative refinement: the solution in FP64, FP32 corrections ────────────

// r = b - A x on the unknowns (0 elsewhere), x in FP64 and each row summed in FP64, then
// rounded to FP32: the residual the FP32 (or FP16) iterations correct.
__global__ __launch_bounds__(256) void k_resid64fine(float* __restrict__ r, const double* __restrict__ x,
                                                     const double* __restrict__ xc, const float* __restrict__ b,
                                                     CompT<float> L, STopo t, int32_t nf) {
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);   // 256 threads a workgroup, as k_rowfine
    int32_t n = tile * 256 + threadIdx.x;
    if ((int32_t)(tile * 256) >= nf) return;
    int s = (int)((tile * 256u) >> 12) + 1, c = n & 4095;
    int4 w = t.where[s - 1];
    if (w.x < 0) return;
    int32_t nb[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) nb[d] = nbslot(t, w, d);
    double res = 0.0;
    if (L.actf[n] == LIQUID) {
        double diag, off;
        finerow<16, double, float, double>(t, nb, x, xc, L.cf, L.dDf, s, c & 15, (c >> 4) & 15, c >> 8, w, diag, off);
        res = (double)b[n] - (diag * x[n] - off);
    }
    r[n] = (float)res;
}
__global__ __launch_bounds__(256) void k_resid64coarse(float* __restrict__ r, const double* __restrict__ x,
                                                       const double* __restrict__ xf, const float* __restrict__ b,
                                                       CompT<float> L, STopo t) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    double res = 0.0;
    unsigned msk = L.cmask[l];
    if (L.actc[l] != 0 && !(msk & 1u)) {
        double diag, off;
        const LevelTypes ty{L.actf, L.tc};
        coarserow<16, double, float, double>(t, msk, xf, L.cf, x, L.cc, L.dDc, I, J, K, diag, off, &ty);
        res = (double)b[l] - (diag * x[l] - off);
    }
    r[l] = (float)res;
}
// The bookkeeping over the fine and the coarse range: mode 0 x64 = x32 and x32 = 0 (the
// start), 1 x64 += x32 and x32 = 0 (a correction folded in), 2 x32 = x64 rounded (the end).
__global__ void k_refine(double* __restrict__ x64f, double* __restrict__ x64c, float* __restrict__ x32f,
                         float* __restrict__ x32c, size_t nf, size_t nc, int mode) {
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < nf + nc; l += (size_t)gridDim.x * blockDim.x) {
        double* d = l < nf ? x64f + l : x64c + (l - nf);
        float* x = l < nf ? x32f + l : x32c + (l - nf);
        if (mode == 0) { *d = (double)*x; *x = 0.0f; }
        else if (mode == 1) { *d += (double)*x; *x = 0.0f; }
        else *x = (float)*d;
    }
}

// ─── projection (sflip.jl:1906-1962) ─────────────────────────────────────────

// The pressure difference over a face, in FP32 from FP32 pressures (the reference's), or
// taken in FP64 from the refined FP64 pressure and then rounded.
__device__ inline float pdiff(float a, float b) { return a - b; }
__device__ inline float pdiff(double a, double b) { return (float)(a - b); }
__device__ inline float pdiv(float a, float b) { return __fdiv_rn(a, b); }
__device__ inline float pdiv(double a, float b) { return (float)(a / (double)b); }

template <class PT>
__global__ __launch_bounds__(256) void k_projectfine(float* __restrict__ u, uint8_t* __restrict__ valid,
                                                     const PT* __restrict__ P, const PT* __restrict__ Pc,
                                                     const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc,
                                                     const float* __restrict__ phi, const float* __restrict__ phic, STopo t, int a,
                                                     const int32_t* __restrict__ order) {
    // a workgroup per live slot (block order), a thread per 17 of its faces along a: a
    // face's two cells lie in the slot's block or in the blocks below and above it along
    // a, looked up once (fineat's index, or -1 where there is no block)
    const int s = order[blockIdx.x];
    const int4 w = t.where[s - 1];
    const SlotsGlobal lk;
    const int32_t slo = lk(t, w.x - (a == 0), w.y - (a == 1), w.z - (a == 2));
    const int32_t shi = lk(t, w.x + (a == 0), w.y + (a == 1), w.z + (a == 2));
    const int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    // fineat's index of the cell at local (li, lj, lk), the coordinate along a in -1..16
    auto idx = [&](int li, int lj, int lk) -> int32_t {
        const int la = a == 0 ? li : (a == 1 ? lj : lk);
        const int32_t sl = la < 0 ? slo : (la > 15 ? shi : s);
        return sl == 0 ? -1 : (sl - 1) * 4096 + (li & 15) + 16 * ((lj & 15) + 16 * (lk & 15));
    };
    auto ctype = [&](int32_t id, int I, int J, int K, bool& isfine) -> uint8_t {   // celltype
        if (id >= 0) { isfine = true; return tf[id]; }
        isfine = false;
        const int ci = I >> 2, cj = J >> 2, ck = K >> 2;
        if (ci < 0 || cj < 0 || ck < 0 || ci >= n0 || cj >= n1 || ck >= n2) return SOLID;
        return tc[cl(ci, cj, ck, n0, n1)];
    };
    auto coarsei = [&](int I, int J, int K) -> size_t {   // fineorcoarse's clamped coarse cell
        return cl(min(max(I >> 2, 0), n0 - 1), min(max(J >> 2, 0), n1 - 1), min(max(K >> 2, 0), n2 - 1), n0, n1);
    };
    for (int r = 0; r < SBFACES / 256; ++r) {
        const int f = r * 256 + (int)threadIdx.x;
        int i, j, k;
        if (a == 0) { i = f % 17; int q = f / 17; j = q & 15; k = q >> 4; }
        else if (a == 1) { i = f & 15; int q = f >> 4; j = q % 17; k = q / 17; }
        else { i = f & 15; j = (f >> 4) & 15; k = f >> 8; }
        const int32_t n = (s - 1) * SBFACES + f;
        const int I = w.x * 16 + i, J = w.y * 16 + j, K = w.z * 16 + k;
        const int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
        const int32_t ip = idx(i, j, k), im = idx(i - (a == 0), j - (a == 1), k - (a == 2));
        bool fp, fm;
        const uint8_t tp = ctype(ip, I, J, K, fp), tm = ctype(im, Im, Jm, Km, fm);
        const bool ok = tp != SOLID && tm != SOLID && (tp == LIQUID || tm == LIQUID);
        if (ok) {
            const float d = (fp && fm ? 1.0f : 2.5f) * t.g.dx;
            const PT Pp = ip >= 0 ? P[ip] : Pc[coarsei(I, J, K)], Pm = im >= 0 ? P[im] : Pc[coarsei(Im, Jm, Km)];
            const float php = ip >= 0 ? phi[ip] : phic[coarsei(I, J, K)], phm = im >= 0 ? phi[im] : phic[coarsei(Im, Jm, Km)];
            float Dl;
            if (tp == LIQUID && tm == LIQUID) Dl = pdiff(Pp, Pm);
            else if (tm == LIQUID) Dl = pdiv(-Pm, theta(phm, php));
            else Dl = pdiv(Pp, theta(php, phm));
            u[n] -= __fdiv_rn(Dl, d);
        }
        valid[n] = ok ? 1 : 0;
    }
}
template <class PT>
__global__ void k_projectcoarse(float* __restrict__ u, uint8_t* __restrict__ valid, const PT* __restrict__ Pc,
                                const uint8_t* __restrict__ tc, const float* __restrict__ phic, STopo t, int a) {
    int nc0 = t.g.nc(0), nc1 = t.g.nc(1), nc2 = t.g.nc(2);
    int n0 = nc0 + (a == 0), n1 = nc1 + (a == 1), n2 = nc2 + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
    bool inside = Im >= 0 && Jm >= 0 && Km >= 0 && I < nc0 && J < nc1 && K < nc2;
    if (inside && !covered(t, I, J, K) && !covered(t, Im, Jm, Km)) {
        size_t cp = cl(I, J, K, nc0, nc1), cm = cl(Im, Jm, Km, nc0, nc1);
        uint8_t tp = tc[cp], tm = tc[cm];
        bool ok = tp != SOLID && tm != SOLID && (tp == LIQUID || tm == LIQUID);
        if (ok) {
            PT Pp = Pc[cp], Pm = Pc[cm];
            float fp = phic[cp], fm = phic[cm];
            float Dl = (tp == LIQUID && tm == LIQUID) ? pdiff(Pp, Pm)
                                                      : (tm == LIQUID ? pdiv(-Pm, theta(fm, fp)) : pdiv(Pp, theta(fp, fm)));
            u[l] -= __fdiv_rn(Dl, t.g.cdx());
        }
        valid[l] = ok ? 1 : 0;
    } else if (!inside) {
        valid[l] = 0;
    }
}

// ─── the projection's adjoint (no counterpart in the reference) ───────────────
// project! maps the faces u to R (u - G A^-1 (D u + gamma terms)): D the scaled
// divergences of k_finerhs / k_coarserhs, G the face differences of k_projectfine /
// k_projectcoarse, R the restriction of the coarse faces under fine blocks. With the
// types and the level set fixed its linear part's transpose is ub -> u'' + D^T A^-1
// (-G^T u''), u'' = R^T ub, A being symmetric: the same solver on the adjoint
// right-hand side.

// The fine cell (I, J, K) as fineorcoarse reads it: its pool index, or -1 and the coarse
// cell's (clamped) index.
__device__ inline int32_t cellref(const STopo& t, int I, int J, int K, size_t& ci) {
    int32_t idx = fineat(t, I, J, K);
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    ci = cl(min(max(I >> 2, 0), n0 - 1), min(max(J >> 2, 0), n1 - 1), min(max(K >> 2, 0), n2 - 1), n0, n1);
    return idx;
}
// The largest |cotangent| over the allocated fine face copies and the coarse faces, as
// float bits (+inf for a nonfinite one).
__global__ __launch_bounds__(256) void k_sadj_absmax(const float* __restrict__ x, size_t n, const int4* __restrict__ where,
                                                     int* __restrict__ bound) {
    __shared__ int red[4];
    size_t l = (size_t)blockIdx.x * 256 + threadIdx.x;
    float m = 0.0f;
    if (l < n && (!where || where[l / SBFACES].x >= 0)) {
        float v = x[l];
        m = isfinite(v) ? fabsf(v) : __int_as_float(0x7f800000);
    }
    int v = __float_as_int(m);
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(bound, max(max(red[0], red[1]), max(red[2], red[3])));
}
// -G^T ub, scaled into fixed-point sums: each valid fine face copy (k_projectfine's test)
// adds -ub dDl/dP / d to the cells whose pressures it differences (a fine cell, or the
// coarse cell under an unallocated block).
__global__ void k_sadj_gradT_fine(unsigned long long* __restrict__ accf, unsigned long long* __restrict__ accc,
                                  const float* __restrict__ ub, const uint8_t* __restrict__ tf, const uint8_t* __restrict__ tc,
                                  const float* __restrict__ phi, const float* __restrict__ phic, STopo t, int a, int32_t nfa,
                                  float scale) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nfa) return;
    SCell s = fineface(t, a, n);
    if (s.w.x < 0) return;
    int I = s.w.x * 16 + s.i, J = s.w.y * 16 + s.j, K = s.w.z * 16 + s.k;
    int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
    bool fp, fm;
    uint8_t tp = celltype(t, tf, tc, I, J, K, fp), tm = celltype(t, tf, tc, Im, Jm, Km, fm);
    if (!(tp != SOLID && tm != SOLID && (tp == LIQUID || tm == LIQUID))) return;
    float d = (fp && fm ? 1.0f : 2.5f) * t.g.dx;
    float x = ub[n] * scale / d;
    size_t cp, cm;
    int32_t ip = cellref(t, I, J, K, cp), im = cellref(t, Im, Jm, Km, cm);
    float gp = 0.0f, gm = 0.0f;   // -ub dDl/dPp / d, -ub dDl/dPm / d
    if (tp == LIQUID && tm == LIQUID) {
        gp = -x;
        gm = x;
    } else if (tm == LIQUID) {
        gm = x / theta(fineorcoarse(t, phi, phic, Im, Jm, Km), fineorcoarse(t, phi, phic, I, J, K));
    } else {
        gp = -x / theta(fineorcoarse(t, phi, phic, I, J, K), fineorcoarse(t, phi, phic, Im, Jm, Km));
    }
    if (gp != 0.0f) atomicAdd(ip >= 0 ? &accf[ip] : &accc[cp], tofix(gp));
    if (gm != 0.0f) atomicAdd(im >= 0 ? &accf[im] : &accc[cm], tofix(gm));
}
// The same for the coarse faces k_projectcoarse moves (both cells uncovered).
__global__ void k_sadj_gradT_coarse(unsigned long long* __restrict__ accc, const float* __restrict__ ub,
                                    const uint8_t* __restrict__ tc, const float* __restrict__ phic, STopo t, int a, float scale) {
    int nc0 = t.g.nc(0), nc1 = t.g.nc(1), nc2 = t.g.nc(2);
    int n0 = nc0 + (a == 0), n1 = nc1 + (a == 1), n2 = nc2 + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
    bool inside = Im >= 0 && Jm >= 0 && Km >= 0 && I < nc0 && J < nc1 && K < nc2;
    if (!inside || covered(t, I, J, K) || covered(t, Im, Jm, Km)) return;
    size_t cp = cl(I, J, K, nc0, nc1), cm = cl(Im, Jm, Km, nc0, nc1);
    uint8_t tp = tc[cp], tm = tc[cm];
    if (!(tp != SOLID && tm != SOLID && (tp == LIQUID || tm == LIQUID))) return;
    float x = ub[l] * scale / t.g.cdx();
    float fp = phic[cp], fm = phic[cm];
    float gp = 0.0f, gm = 0.0f;
    if (tp == LIQUID && tm == LIQUID) {
        gp = -x;
        gm = x;
    } else if (tm == LIQUID) {
        gm = x / theta(fm, fp);
    } else {
        gp = -x / theta(fp, fm);
    }
    if (gp != 0.0f) atomicAdd(&accc[cp], tofix(gp));
    if (gm != 0.0f) atomicAdd(&accc[cm], tofix(gm));
}
// The adjoint right-hand side from the sums on the unknowns (fine LIQUID cells, uncovered
// coarse LIQUID cells), 0 elsewhere, and the solve's start P = 0.
__global__ void k_sadj_rhs_fine(float* __restrict__ b, float* __restrict__ P, const unsigned long long* __restrict__ acc,
                                const uint8_t* __restrict__ tf, const int4* __restrict__ where, int32_t nf, double inv) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    bool u = where[n >> 12].x >= 0 && tf[n] == LIQUID;
    b[n] = u ? (float)((double)(long long)acc[n] * FIX_INV * inv) : 0.0f;
    P[n] = 0.0f;
}
__global__ void k_sadj_rhs_coarse(float* __restrict__ b, float* __restrict__ P, const unsigned long long* __restrict__ acc,
                                  const uint8_t* __restrict__ tc, STopo t, double inv) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    bool u = tc[l] == LIQUID && !covered(t, I, J, K);
    b[l] = u ? (float)((double)(long long)acc[l] * FIX_INV * inv) : 0.0f;
    P[l] = 0.0f;
}
// ub += D^T y per fine face copy: the copy is the low face of its block's cell at its
// index (db/du = +h) and the high face of the cell below it in the same block (-h), for
// LIQUID cells (k_finerhs reads each cell's own block's copies).
template <class PT>
__global__ void k_sadj_divT_fine(float* __restrict__ ub, const PT* __restrict__ y, const uint8_t* __restrict__ tf, STopo t,
                                 int a, int32_t nfa, float h) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nfa) return;
    SCell s = fineface(t, a, n);
    if (s.w.x < 0) return;
    int ia = a == 0 ? s.i : (a == 1 ? s.j : s.k);
    int32_t base = (s.s - 1) * SBCELLS;
    int32_t c = base + (s.i + 16 * (s.j + 16 * s.k)), st = a == 0 ? 1 : (a == 1 ? 16 : 256);
    float add = 0.0f;
    if (ia <= 15 && tf[c] == LIQUID) add += h * (float)y[c];
    if (ia >= 1 && tf[c - st] == LIQUID) add -= h * (float)y[c - st];
    ub[n] += add;
}
// ub += D^T y per coarse face: +16h from the cell above it, -16h from the cell below, for
// uncovered LIQUID cells (k_coarserhs).
template <class PT>
__global__ void k_sadj_divT_coarse(float* __restrict__ ub, const PT* __restrict__ y, const uint8_t* __restrict__ tc, STopo t,
                                   int a, float h16) {
    int nc0 = t.g.nc(0), nc1 = t.g.nc(1), nc2 = t.g.nc(2);
    int n0 = nc0 + (a == 0), n1 = nc1 + (a == 1), n2 = nc2 + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    int Im = I - (a == 0), Jm = J - (a == 1), Km = K - (a == 2);
    float add = 0.0f;
    if (I < nc0 && J < nc1 && K < nc2) {
        size_t cp = cl(I, J, K, nc0, nc1);
        if (tc[cp] == LIQUID && !covered(t, I, J, K)) add += h16 * (float)y[cp];
    }
    if (Im >= 0 && Jm >= 0 && Km >= 0) {
        size_t cm = cl(Im, Jm, Km, nc0, nc1);
        if (tc[cm] == LIQUID && !covered(t, Im, Jm, Km)) add -= h16 * (float)y[cm];
    }
    ub[l] += add;
}

// ─── the solver ──────────────────────────────────────────────────────────────

template <class T> struct SVecT {   // a pressure-sized vector in storage T
    T* f;
    T* c;
};
struct SVec {   // a pressure-sized vector: fine pool part and coarse part
    float* f;
    float* c;
};

struct SparseSolver {
    SparseBasin& s;
    hipStream_t st;
    STopo t;
    int32_t nf;        // fine cells in use (nslots * 4096)
    size_t Cc;

    explicit SparseSolver(SparseBasin& b) : s(b), st(b.st.s), t(b.topo()), nf((int32_t)b.finecells()), Cc(b.ncells()) {}

    SVec vec(SField<float>& f) { return SVec{f.fine.p, f.coarse.p}; }
    static int3 i3(const int n[3]) { return make_int3(n[0], n[1], n[2]); }

    void coefficients() {
        for (int a = 0; a < 3; ++a) {
            if (s.nlive > 0)
                hipLaunchKernelGGL(k_finecoef, dim3((unsigned)s.nlive), dim3(256), 0, st, s.cf[a].p, s.types.fine.p,
                                   s.types.coarse.p, t, a, s.sorder.p);
            hipLaunchKernelGGL(k_coarsecoef, dim3(cdiv(s.nfaces(a), 256)), dim3(256), 0, st, s.cc[a].p, s.types.coarse.p, t, a);
        }
        if (nf)
            hipLaunchKernelGGL(k_finedirichlet, dim3(cdiv(nf, 256)), dim3(256), 0, st, s.dD.fine.p, s.types.fine.p,
                               s.types.coarse.p, s.phi.fine.p, s.phi.coarse.p, t, nf);
        hipLaunchKernelGGL(k_coarsedirichlet, dim3(cdiv(Cc, 256)), dim3(256), 0, st, s.dD.coarse.p, s.types.fine.p,
                           s.types.coarse.p, s.phi.fine.p, s.phi.coarse.p, t);
        hipLaunchKernelGGL(k_activecoarse, dim3(cdiv(Cc, 256)), dim3(256), 0, st, s.D[0].act.p, s.types.coarse.p, t);
        if (!s.cmask.p) s.cmask.alloc(Cc);
        hipLaunchKernelGGL(k_covermask, dim3(cdiv(Cc, 256)), dim3(256), 0, st, s.cmask.p, t);
        if (!s.dgf.p) s.dgf.alloc((size_t)s.blocks * SBCELLS);
        if (s.nslots) {   // level 0's full diagonal (mode 6 of the level-0 rows)
            Hier<float, float> H = hier32();
            hipLaunchKernelGGL((k_rowfine16<6, float, float>), dim3(4 * s.nslots), dim3(256), 0, st, s.dgf.p, (const float*)nullptr,
                               (const float*)nullptr, (const float*)nullptr, H.L0, t, SearchArgs{});
        }
        MAEL_LAUNCH_CHECK();
    }
    void hierarchy() {
        int ns = s.nslots;
        for (int a = 0; a < 3; ++a)
            if (ns)
                hipLaunchKernelGGL(k_coef1, dim3(cdiv(ns * SBFACES1, 256)), dim3(256), 0, st, s.c1[a].p, s.cf[a].p, t, a, ns);
        if (ns)
            hipLaunchKernelGGL(k_cells1, dim3(cdiv(ns * SBCELLS1, 256)), dim3(256), 0, st, s.dD1.p, s.act1.p, s.dD.fine.p,
                               s.types.fine.p, t, ns);
        SLevel& D2 = s.D[0];
        for (int a = 0; a < 3; ++a)
            hipLaunchKernelGGL(k_coef2, dim3(cdiv(D2.faces(a), 256)), dim3(256), 0, st, D2.c[a].p, s.cc[a].p, s.c1[a].p, t, a);
        hipLaunchKernelGGL(k_cells2, dim3(cdiv(Cc, 256)), dim3(256), 0, st, D2.dD.p, D2.act.p, s.dD.coarse.p, s.dD1.p, s.act1.p,
                           t);
        for (size_t l = 1; l < s.D.size(); ++l) {
            SLevel &F = s.D[l - 1], &C = s.D[l];
            for (int a = 0; a < 3; ++a)
                hipLaunchKernelGGL(k_coefdense, dim3(cdiv(C.faces(a), 256)), dim3(256), 0, st, C.c[a].p, F.c[a].p, i3(C.n),
                                   i3(F.n), a);
            hipLaunchKernelGGL(k_cellsdense, dim3(cdiv(C.cells(), 256)), dim3(256), 0, st, C.dD.p, C.act.p, F.dD.p, F.act.p,
                               i3(C.n), i3(F.n));
        }
        MAEL_LAUNCH_CHECK();
    }


    // The hierarchy in one storage precision: its vectors and couplings.
    template <class T, class C> struct Hier {
        SVecT<T> q, tmp, X1x, X1b, X1r, X1t;
        std::vector<T*> Dx, Db, Dt;
        CompT<C> L0, L1;
        std::vector<DViewT<C>> D;
        float rscale = 1.0f;   // dense-level restriction scale
    };
    Hier<float, float> hier32() {
        Hier<float, float> H;
        H.q = {s.q.fine.p, s.q.coarse.p};
        H.tmp = {s.tmp.fine.p, s.tmp.coarse.p};
        H.X1x = {s.X1x.fine.p, s.X1x.coarse.p}; H.X1b = {s.X1b.fine.p, s.X1b.coarse.p};
        H.X1r = {s.X1r.fine.p, s.X1r.coarse.p}; H.X1t = {s.X1t.fine.p, s.X1t.coarse.p};
        for (auto& L : s.D) {
            H.Dx.push_back(L.x.p); H.Db.push_back(L.b.p); H.Dt.push_back(L.t.p);
            H.D.push_back(DViewT<float>{i3(L.n), {L.c[0].p, L.c[1].p, L.c[2].p}, L.dD.p, L.act.p});
        }
        H.L0 = CompT<float>{{{s.cf[0].p, s.cf[1].p, s.cf[2].p}}, s.dD.fine.p, s.types.fine.p,
                            {{s.cc[0].p, s.cc[1].p, s.cc[2].p}}, s.dD.coarse.p, s.D[0].act.p, s.cmask.p, s.types.coarse.p,
                            nullptr, s.dgf.p};
        H.L1 = CompT<float>{{{s.c1[0].p, s.c1[1].p, s.c1[2].p}}, s.dD1.p, s.act1.p,
                            {{s.cc[0].p, s.cc[1].p, s.cc[2].p}}, s.dD.coarse.p, s.D[0].act.p, s.cmask.p, s.types.coarse.p};
        return H;
    }
    // The reduced-precision hierarchy (storage in s.mp): vectors and couplings, the dense
    // levels' equations divided by 4 per level below the coarse grid.
    template <class T> Hier<T, T> hier16() {
        MPStore& m = s.mp;
        auto V = [&](DBuf<uint16_t>& f, DBuf<uint16_t>& c) { return SVecT<T>{(T*)f.p, (T*)c.p}; };
        Hier<T, T> H;
        H.q = V(m.q[0], m.q[1]); H.tmp = V(m.tmp[0], m.tmp[1]);
        H.X1x = V(m.X1[0][0], m.X1[0][1]); H.X1b = V(m.X1[1][0], m.X1[1][1]);
        H.X1r = V(m.X1[2][0], m.X1[2][1]); H.X1t = V(m.X1[3][0], m.X1[3][1]);
        for (size_t l = 0; l < s.D.size(); ++l) {
            H.Dx.push_back((T*)m.Dv[l][0].p); H.Db.push_back((T*)m.Dv[l][1].p);
            H.Dt.push_back((T*)m.Dv[l][2].p);
            H.D.push_back(DViewT<T>{i3(s.D[l].n), {(T*)m.Dc[l][0].p, (T*)m.Dc[l][1].p, (T*)m.Dc[l][2].p}, (T*)m.Dc[l][3].p,
                                    s.D[l].act.p});
        }
        // level 0's rows take their couplings from the cell types: no fine coupling copies
        H.L0 = CompT<T>{{{nullptr, nullptr, nullptr}}, (T*)m.dDf.p, s.types.fine.p,
                        {{(T*)m.cc[0].p, (T*)m.cc[1].p, (T*)m.cc[2].p}}, (T*)m.dDc.p, s.D[0].act.p, s.cmask.p, s.types.coarse.p,
                        (T*)m.rwf.p};
        H.L1 = CompT<T>{{{(T*)m.c1[0].p, (T*)m.c1[1].p, (T*)m.c1[2].p}}, (T*)m.dD1.p, s.act1.p,
                        {{(T*)m.cc[0].p, (T*)m.cc[1].p, (T*)m.cc[2].p}}, (T*)m.dDc.p, s.D[0].act.p, s.cmask.p, s.types.coarse.p};
        H.rscale = 0.25f;
        return H;
    }
    // Round the couplings of the hierarchy into the reduced-precision copies.
    template <class T> void cast_couplings() {
        MPStore& m = s.mp;
        auto cast = [&](DBuf<uint16_t>& y, const float* x, size_t n, float scale) {
            if (n) hipLaunchKernelGGL(k_castc<T>, dim3(std::min<unsigned>(cdiv(n, 256), 8192)), dim3(256), 0, st, (T*)y.p, x, n,
                                      scale);
        };
        for (int a = 0; a < 3; ++a) {
            cast(m.cc[a], s.cc[a].p, s.nfaces(a), 1.0f);
            cast(m.c1[a], s.c1[a].p, (size_t)s.nslots * SBFACES1, 1.0f);
        }
        cast(m.dDf, s.dD.fine.p, s.finecells(), 1.0f);
        cast(m.dDc, s.dD.coarse.p, Cc, 1.0f);
        cast(m.dD1, s.dD1.p, (size_t)s.nslots * SBCELLS1, 1.0f);
        float sc = 1.0f;
        for (size_t l = 0; l < s.D.size(); ++l) {
            for (int a = 0; a < 3; ++a) cast(m.Dc[l][a], s.D[l].c[a].p, s.D[l].faces(a), sc);
            cast(m.Dc[l][3], s.D[l].dD.p, s.D[l].cells(), sc);
            sc *= 0.25f;
        }
        if (s.nslots) {   // level 0's Jacobi weights, omega / diag (mode 5 of the level-0 rows)
            Hier<T, T> H = hier16<T>();
            hipLaunchKernelGGL((k_rowfine16<5, T, T>), dim3(4 * s.nslots), dim3(256), 0, st, (T*)m.rwf.p, (const T*)nullptr,
                               (const T*)nullptr, (const T*)nullptr, H.L0, t, SearchArgs{});
        }
        MAEL_LAUNCH_CHECK();
    }

    template <class T> void zero_t(T* p, size_t n) {
        if (!n) return;
        if (sizeof(T) == 2) fill_async((uint16_t*)p, n, (uint16_t)0, st);
        else fill_async((float*)p, n, 0.0f, st);
    }
    template <class T> void zero_t(SVecT<T> v, size_t nfine) { zero_t(v.f, nfine); zero_t(v.c, Cc); }
    void zero(SVec v, size_t nfine) { zero_t(SVecT<float>{v.f, v.c}, nfine); }

    // One composite-level pass (MODE 0 Jacobi, 1 residual, 2 apply) over the fine part
    // (N^3 a slot) and the coarse part.
    template <int N, int MODE, class T, class C>
    void rows(SVecT<T> out, SVecT<T> x, SVecT<T> b, const CompT<C>& L, SearchArgs S = SearchArgs{}) {
        int32_t nfl = s.nslots * N * N * N;
        // one launch for both parts, except the reduced-precision sweep from zero (pointwise)
        constexpr bool FUSE = !(N == 16 && MODE == 3 && !std::is_same<T, float>::value);
        if constexpr (FUSE) {
            if (nfl) {
                constexpr int NQ = (N == 16 && (std::is_same<T, float>::value || MODE == 0 || MODE == 1)) ? 2 : 1;
                const uint32_t nfw = N == 16 ? (uint32_t)s.nslots * (uint32_t)(4 / NQ) : (uint32_t)s.nslots;
                const uint32_t ncb = (uint32_t)cdiv(Cc, 256);
                if (N == 16 && MODE == 0) s.kt.begin(K_RBGS, st);
                hipLaunchKernelGGL((k_rowcomp<N, MODE, T, C, NQ>), dim3(nfw + ncb), dim3(256), 0, st, out.f, out.c, x.f, x.c,
                                   b.f, b.c, L, t, S, ncb);
                if (N == 16 && MODE == 0) s.kt.end(K_RBGS, st);
                return;
            }
        }
        if (nfl) {
            if (N == 16 && MODE == 0) s.kt.begin(K_RBGS, st);
            if constexpr (N == 16 && MODE == 3 && !std::is_same<T, float>::value)
                hipLaunchKernelGGL((k_jacobi0w<T>), dim3(s.nslots * 4), dim3(256), 0, st, out.f, b.f, L.rwf, L.actf, t.where, nfl);
            else if constexpr (N == 16 && (std::is_same<T, float>::value || MODE == 0 || MODE == 1))   // 8-plane tiles
                hipLaunchKernelGGL((k_rowfine16<MODE, T, C, 2>), dim3(2 * s.nslots), dim3(256), 0, st, out.f, x.f, b.f, x.c, L, t,
                                   S);
            else if constexpr (N == 16)
                hipLaunchKernelGGL((k_rowfine16<MODE, T, C>), dim3(4 * s.nslots), dim3(256), 0, st, out.f, x.f, b.f, x.c, L, t, S);
            else
                hipLaunchKernelGGL((k_rowfine8<MODE, T, C>), dim3(s.nslots), dim3(256), 0, st, out.f, x.f, b.f, x.c, L, t);
            if (N == 16 && MODE == 0) s.kt.end(K_RBGS, st);
        } else if (MODE == 4) {
            fill_async(S.partials, (size_t)1, 0.0, st);   // no fine workgroups: the finish reads one zero
        }
        hipLaunchKernelGGL((k_rowcoarse<N, MODE, T, C>), dim3(cdiv(Cc, 256)), dim3(256), 0, st, out.c, x.c, b.c, x.f, L, t, S);
    }
    // flexible CG's fused step: dn = z + beta d (beta on the device), q = A dn, and d.q
    // (with the reference's checks) into alpha
    void search(SVec q, SVec dn, SVec z, SVec d) {
        Hier<float, float> H = hier32();
        int32_t fine_groups = s.nslots ? 2 * s.nslots : 1;   // the level-0 rows' 8-plane tiles
        SearchArgs S{d.f, d.c, dn.f, dn.c, s.partials.p, fine_groups, s.scal.p};
        s.kt.begin(K_SEARCH, st);
        rows<16, 4>(SVecT<float>{q.f, q.c}, SVecT<float>{z.f, z.c}, SVecT<float>{z.f, z.c}, H.L0, S);
        s.kt.end(K_SEARCH, st);
        hipLaunchKernelGGL(k_sfinish, dim3(1), dim3(FINISH_THREADS), 0, st, s.partials.p, fine_groups + (int)cdiv(Cc, 256), s.scal.p,
                           (int)SF_ALPHA, 0);
        MAEL_LAUNCH_CHECK();
    }
    template <int N, class T, class C> void smooth(SVecT<T> x, SVecT<T> b, SVecT<T> tmp, const CompT<C>& L, int sweeps) {
        for (int k = 0; k < sweeps; ++k) {
            rows<N, 0>(tmp, x, b, L);
            rows<N, 0>(x, tmp, b, L);
        }
    }
    // The same from x = 0: the first sweep reads no x, and x needs no clearing (every cell
    // of the level's slots and coarse part is written).
    template <int N, class T, class C>
    void smooth0(SVecT<T> x, SVecT<T> b, SVecT<T> tmp, const CompT<C>& L, int sweeps, size_t nfine, bool queued = false) {
        if (sweeps < 1) { zero_t(x, nfine); return; }
        if (!queued) rows<N, 3>(tmp, x, b, L);   // else already queued (queue_head)
        rows<N, 0>(x, tmp, b, L);
        smooth<N>(x, b, tmp, L, sweeps - 1);
    }

    // sflip.jl:1722-1762 vcycle!: z = M^-1 r in the hierarchy H (coarsest in arithmetic A).
    template <class T, class C, class A>
    void vcycle_t(Hier<T, C>& H, SVecT<T> z, SVecT<T> r, int sweeps, int coarse, float boost, bool queued = false) {
        int ns = s.nslots;
        smooth0<16>(z, r, H.tmp, H.L0, sweeps, (size_t)nf, queued);
        rows<16, 1>(H.q, z, r, H.L0);
        const uint32_t nfr = (uint32_t)cdiv((size_t)ns * SBCELLS1, 256);
        hipLaunchKernelGGL(k_restrictcomp<T>, dim3(nfr + (uint32_t)cdiv(Cc, 256)), dim3(256), 0, st, H.X1b.f, H.q.f, H.X1b.c, H.q.c,
                           Cc, t, ns, nfr);
        smooth0<8>(H.X1x, H.X1b, H.X1t, H.L1, sweeps, (size_t)ns * SBCELLS1);
        rows<8, 1>(H.X1r, H.X1x, H.X1b, H.L1);
        hipLaunchKernelGGL(k_restrict12<T>, dim3(cdiv(Cc, 256)), dim3(256), 0, st, H.Db[0], H.X1r.f, H.X1r.c, t);
        int nl = (int)s.D.size();
        for (int l = 0; l + 1 < nl; ++l) {
            SLevel& Dl = s.D[l];
            size_t c = Dl.cells();
            if (sweeps < 1) zero_t(H.Dx[l], c);
            for (int k = 0; k < sweeps; ++k) {   // from x = 0: the first sweep reads no x
                if (k == 0)
                    hipLaunchKernelGGL((k_rowdense<3, T, C>), dim3(cdiv(c, 256)), dim3(256), 0, st, H.Dt[l], H.Dx[l], H.Db[l],
                                       H.D[l]);
                else
                    hipLaunchKernelGGL((k_rowdense<0, T, C>), dim3(cdiv(c, 256)), dim3(256), 0, st, H.Dt[l], H.Dx[l], H.Db[l],
                                       H.D[l]);
                hipLaunchKernelGGL((k_rowdense<0, T, C>), dim3(cdiv(c, 256)), dim3(256), 0, st, H.Dx[l], H.Dt[l], H.Db[l], H.D[l]);
            }
            SLevel& Cn = s.D[l + 1];
            hipLaunchKernelGGL((k_residrestrict<T, C>), dim3(cdiv(Cn.cells(), 256)), dim3(256), 0, st, H.Db[l + 1], H.Dx[l],
                               H.Db[l], H.D[l], i3(Cn.n), H.rscale);
        }
        hipLaunchKernelGGL((k_scoarsest<T, C, A>), dim3(1), dim3(1024), 0, st, H.Dx[nl - 1], H.Db[nl - 1], H.D[nl - 1], coarse);
        for (int l = nl - 2; l >= 0; --l) {
            SLevel &Dl = s.D[l], &Cn = s.D[l + 1];
            size_t c = Dl.cells();
            hipLaunchKernelGGL(k_prolongdense<T>, dim3(cdiv(c, 256)), dim3(256), 0, st, H.Dx[l], H.Dx[l + 1], Dl.act.p, i3(Dl.n),
                               i3(Cn.n), boost);
            for (int k = 0; k < sweeps; ++k) {
                hipLaunchKernelGGL((k_rowdense<0, T, C>), dim3(cdiv(c, 256)), dim3(256), 0, st, H.Dt[l], H.Dx[l], H.Db[l], H.D[l]);
                hipLaunchKernelGGL((k_rowdense<0, T, C>), dim3(cdiv(c, 256)), dim3(256), 0, st, H.Dx[l], H.Dt[l], H.Db[l], H.D[l]);
            }
        }
        SLevel& D2 = s.D[0];
        const uint32_t ncb = (uint32_t)cdiv(Cc, 256), nfb1 = (uint32_t)cdiv((size_t)ns * SBCELLS1, 256);
        hipLaunchKernelGGL((k_prolongcomp<T, 1>), dim3(nfb1 + ncb), dim3(256), 0, st, H.X1x.f, H.Dx[0], s.act1.p, H.X1x.c,
                           H.Dx[0], D2.act.p, t, boost, ns, nfb1);
        smooth<8>(H.X1x, H.X1b, H.X1t, H.L1, sweeps);
        const uint32_t nfb0 = (uint32_t)cdiv((size_t)nf, 1024);
        hipLaunchKernelGGL((k_prolongcomp<T, 0>), dim3(nfb0 + ncb), dim3(256), 0, st, z.f, H.X1x.f, s.types.fine.p, z.c, H.X1x.c,
                           D2.act.p, t, boost, nf, nfb0);
        smooth<16>(z, r, H.tmp, H.L0, sweeps);
        MAEL_LAUNCH_CHECK();
    }

    int precision = 0;   // 0 FP32, 1 FP16 storage, 2 BF16 storage
    // z = M^-1 r. In reduced precision r is scaled by 1 / scal[nslot] on the way in (its
    // norm, so its entries are at most 1) and z back on the way out. With `rold`
    // (flexible CG), z.r and z.rold are also summed into the partials as k_sdot2 does
    // (fused with the conversion back in reduced precision).
    // The V-cycle's first launches (the reduced-precision copy of r and the level-0 sweep
    // from zero) queued ahead of a host read that decides whether the V-cycle runs at
    // all, so the device works through the read's round trip; the next vcycle() on the
    // same z and r skips them. Whoever changes r (or leaves the solve) before that clears
    // `head`, and the V-cycle then runs whole (overwriting what the queued launches wrote,
    // all of it V-cycle scratch).
    bool head = false;
    void queue_head(SVec z, SVec r, int sweeps, int nslot) {
        if (sweeps < 1) return;
        if (precision == 0) {
            Hier<float, float> H = hier32();
            rows<16, 3>(H.tmp, SVecT<float>{z.f, z.c}, SVecT<float>{r.f, r.c}, H.L0);
        } else if (precision == 1) {
            queue_head_mp<__half>(r, nslot);
        } else {
            queue_head_mp<bf16>(r, nslot);
        }
        head = true;
    }
    // r into the hierarchy's precision and level 0's sweep from zero on it: the coarse part
    // of r converted, then one launch (k_head0) for the fine part's conversion and sweep
    // and the coarse rows
    template <class T> void queue_head_mp(SVec r, int nslot) {
        Hier<T, T> H = hier16<T>();
        MPStore& m = s.mp;
        SVecT<T> zh{(T*)m.z[0].p, (T*)m.z[1].p}, rh{(T*)m.r[0].p, (T*)m.r[1].p};
        unsigned g = std::min<unsigned>(cdiv(Cc, 256), 8192);
        hipLaunchKernelGGL((k_convert<T, float>), dim3(g), dim3(256), 0, st, rh.f, r.f, (size_t)0, rh.c, r.c, Cc, s.scal.p, nslot,
                           1);
        const uint32_t ncb = (uint32_t)cdiv(Cc, 256), nfw = (uint32_t)s.nslots * 4u;   // 1024 fine cells a workgroup
        hipLaunchKernelGGL((k_head0<T>), dim3(nfw + ncb), dim3(256), 0, st, H.tmp.f, H.tmp.c, rh.f, r.f, zh.c, rh.c, zh.f, H.L0, t,
                           nf, s.scal.p, nslot, ncb);
    }
    void vcycle(SVec z, SVec r, int sweeps, int coarse, float boost, int nslot = -1, const SVec* rold = nullptr) {
        const bool queued = head;
        head = false;
        if (precision == 0) {
            Hier<float, float> H = hier32();
            vcycle_t<float, float, float>(H, SVecT<float>{z.f, z.c}, SVecT<float>{r.f, r.c}, sweeps, coarse, boost, queued);
            if (rold)
                hipLaunchKernelGGL(k_sdot2, dim3(DOT_BLOCKS), dim3(256), 0, st, s.partials.p, z.f, z.c, r.f, r.c, rold->f,
                                   rold->c, s.types.fine.p, s.types.coarse.p, s.cmask.p, t, nf);
            return;
        }
        if (precision == 1) vcycle_mp<__half>(z, r, sweeps, coarse, boost, nslot, rold, queued);
        else vcycle_mp<bf16>(z, r, sweeps, coarse, boost, nslot, rold, queued);
    }
    template <class T>
    void vcycle_mp(SVec z, SVec r, int sweeps, int coarse, float boost, int nslot, const SVec* rold, bool queued) {
        Hier<T, T> H = hier16<T>();
        MPStore& m = s.mp;
        SVecT<T> zh{(T*)m.z[0].p, (T*)m.z[1].p}, rh{(T*)m.r[0].p, (T*)m.r[1].p};
        unsigned g = std::min<unsigned>(cdiv((size_t)nf + Cc, 256), 8192);
        if (!queued && sweeps >= 1) {
            queue_head_mp<T>(r, nslot);
            queued = true;
        } else if (!queued) {
            hipLaunchKernelGGL((k_convert<T, float>), dim3(g), dim3(256), 0, st, rh.f, r.f, (size_t)nf, rh.c, r.c, Cc, s.scal.p,
                               nslot, 1);
        }
        vcycle_t<T, T, double>(H, zh, rh, sweeps, coarse, boost, queued);
        if (rold && nslot >= 0)
            hipLaunchKernelGGL((k_convert_dot2<T>), dim3(DOT_BLOCKS), dim3(256), 0, st, s.partials.p, z.f, z.c, zh.f, zh.c, r.f,
                               r.c, rold->f, rold->c, s.types.fine.p, s.types.coarse.p, s.cmask.p, t, nf, s.scal.p, nslot);
        else {
            hipLaunchKernelGGL((k_convert<float, T>), dim3(g), dim3(256), 0, st, z.f, zh.f, (size_t)nf, z.c, zh.c, Cc, s.scal.p,
                               nslot, 0);
            if (rold)
                hipLaunchKernelGGL(k_sdot2, dim3(DOT_BLOCKS), dim3(256), 0, st, s.partials.p, z.f, z.c, r.f, r.c, rold->f,
                                   rold->c, s.types.fine.p, s.types.coarse.p, s.cmask.p, t, nf);
        }
    }

    void apply(SVec q, SVec x) {
        Hier<float, float> H = hier32();
        rows<16, 2>(SVecT<float>{q.f, q.c}, SVecT<float>{x.f, x.c}, SVecT<float>{x.f, x.c}, H.L0);
    }
    void dot(SVec a, SVec b, int op, int slot) {
        hipLaunchKernelGGL(k_sdot, dim3(DOT_BLOCKS), dim3(256), 0, st, s.partials.p, a.f, b.f, a.c, b.c, s.types.fine.p,
                           s.types.coarse.p, s.cmask.p, t, nf);
        hipLaunchKernelGGL(k_sfinish, dim3(1), dim3(FINISH_THREADS), 0, st, s.partials.p, DOT_BLOCKS, s.scal.p, op, slot);
        MAEL_LAUNCH_CHECK();
    }
    // y += c x (c = mult * scal[slot], or mult when slot < 0), over the used fine cells and the coarse part.
    void axpy(SVec y, SVec x, int slot, float mult) {
        hipLaunchKernelGGL(k_saxpy2<0>, dim3(std::min<unsigned>(cdiv((size_t)nf + Cc, 256), 8192)), dim3(256), 0, st, y.f, x.f,
                           (size_t)nf, y.c, x.c, Cc, s.scal.p, slot, mult);
    }
    // y = x + c y
    void xpby(SVec y, SVec x, int slot, float mult) {
        hipLaunchKernelGGL(k_saxpy2<1>, dim3(std::min<unsigned>(cdiv((size_t)nf + Cc, 256), 8192)), dim3(256), 0, st, y.f, x.f,
                           (size_t)nf, y.c, x.c, Cc, s.scal.p, slot, mult);
    }
    void copy(SVec y, SVec x) {
        if (nf) MAEL_HIP(hipMemcpyAsync(y.f, x.f, (size_t)nf * sizeof(float), hipMemcpyDeviceToDevice, st));
        MAEL_HIP(hipMemcpyAsync(y.c, x.c, Cc * sizeof(float), hipMemcpyDeviceToDevice, st));
    }
    double read(int slot) {
        MAEL_HIP(hipMemcpyAsync(s.st.pinned, s.scal.p + slot, sizeof(double), hipMemcpyDeviceToHost, st));
        s.st.spin();
        return *(double*)s.st.pinned;
    }
    void read2(int slot, double& a, double& b) {   // scal[slot] and the error flags, one copy
        post_read2();
        wait_read2(slot, a, b);
    }
    // the same in two halves, with other work queued between them (other reads too: this
    // copy lands past the words read() uses)
    static constexpr int READ2_AT = 32;
    void post_read2() {
        MAEL_HIP(hipMemcpyAsync((double*)s.st.pinned + READ2_AT, s.scal.p, 16 * sizeof(double), hipMemcpyDeviceToHost, st));
        s.st.mark();
    }
    void wait_read2(int slot, double& a, double& b) {
        s.st.poll();
        a = ((double*)s.st.pinned)[READ2_AT + slot];
        b = ((double*)s.st.pinned)[READ2_AT + SS_FLAG];
    }
    void set(int slot, double v) { fill_async(s.scal.p + slot, (size_t)1, v, st); }

    static void raise_flags(long long f) {
        if (f & EF_RHO_NONFINITE) throw Error(E_NONFINITE, "Nonfinite pressure preconditioner");
        if (f & EF_RHO_NONPOS) throw Error(E_CG_INDEFINITE, "Pressure CG requires a positive preconditioner");
        if (f & EF_DEN_NONPOS) throw Error(E_CG_INDEFINITE, "Pressure CG requires a positive operator");
        if (f & EF_ALPHA_NONFINITE) throw Error(E_NONFINITE, "Nonfinite pressure update");
        if (f & EF_RHONEW_NONPOS) throw Error(E_CG_INDEFINITE, "Pressure CG requires a positive preconditioner");
        if (f & EF_BETA_NONFINITE) throw Error(E_NONFINITE, "Nonfinite pressure direction");
    }

    // Iterative refinement (refine > 0): the solution lives in s.Pd (FP64) and s.P holds
    // the FP32 correction gathered since the last fold.
    int refine = 0;
    void ir(int mode) {
        unsigned g = std::min<unsigned>(cdiv((size_t)nf + Cc, 256), 8192);
        hipLaunchKernelGGL(k_refine, dim3(g), dim3(256), 0, st, s.Pd.fine.p, s.Pd.coarse.p, s.P.fine.p, s.P.coarse.p,
                           (size_t)nf, Cc, mode);
    }
    // Fold the correction in, r = b - A Pd in FP64 rows (stored FP32); returns ||r||.
    double resid64(SVec r) {
        ir(1);
        Hier<float, float> H = hier32();
        if (nf)
            hipLaunchKernelGGL(k_resid64fine, dim3(cdiv(nf, 256)), dim3(256), 0, st, r.f, s.Pd.fine.p, s.Pd.coarse.p,
                               s.b.fine.p, H.L0, t, nf);
        hipLaunchKernelGGL(k_resid64coarse, dim3(cdiv(Cc, 256)), dim3(256), 0, st, r.c, s.Pd.coarse.p, s.Pd.fine.p,
                           s.b.coarse.p, H.L0, t);
        dot(r, r, SF_NORM, SS_RNORM);
        return read(SS_RNORM);
    }
    // ||b - A P|| of the current solution (into r, with q as scratch), in FP64 sums.
    double true_residual(SVec r, SVec q) {
        if (refine > 0) return resid64(r);
        apply(q, vec(s.P));
        copy(r, vec(s.b));
        axpy(r, q, -1, -1.0f);
        dot(r, r, SF_NORM, SS_RNORM);
        return read(SS_RNORM);
    }
    void zero_solution() {
        zero(vec(s.P), (size_t)nf);
        if (refine > 0) {
            fill_async(s.Pd.fine.p, (size_t)nf, 0.0, st);
            fill_async(s.Pd.coarse.p, Cc, 0.0, st);
        }
    }

    // sflip.jl:1867-1904 pressure!: flexible CG from the pressure in s.P. Returns the
    // iterations; throws the reference's errors; `converged` false at maxiter. With
    // refinement the residual is recomputed from the FP64 solution every `refine`
    // iterations and before convergence is accepted.
    int fcg(double tol, int maxiter, int sweeps, int coarse, float boost, bool& converged, double& relres) {
        SVec P = vec(s.P), b = vec(s.b), r = vec(s.r), z = vec(s.z), d = vec(s.d), q = vec(s.q), rold = vec(s.rold);
        head = false;
        converged = false;
        set(SS_FLAG, 0.0);
        dot(b, b, SF_NORM, SS_BB);
        double bnorm = read(SS_BB);
        relres = 0;
        if (!std::isfinite(bnorm)) throw Error(E_NONFINITE, "Nonfinite pressure right-hand side");
        if (bnorm == 0) { converged = true; return 0; }
        double target = (double)((float)tol * (float)bnorm);
        double rnorm;
        if (refine > 0) {
            rnorm = resid64(r);
        } else {
            apply(r, P);
            xpby(r, b, -1, -1.0f);
            dot(r, r, SF_NORM, SS_RNORM);
            rnorm = read(SS_RNORM);
        }
        relres = rnorm / bnorm;
        if (rnorm <= target) { converged = true; return 0; }
        vcycle(z, r, sweeps, coarse, boost, SS_RNORM);
        copy(d, z);
        dot(z, r, SF_RHO0, 0);
        if (!s.d2.fine.p) s.d2.alloc((size_t)s.blocks * SBCELLS, Cc);
        SVec dbuf[2] = {d, vec(s.d2)};
        int cur = 0;
        // r and r_old alternate between two buffers: the update writes the new residual
        // over the older one instead of copying r to r_old
        SVec rbuf[2] = {r, rold};
        int rc = 0;
        for (int it = 1; it <= maxiter; ++it) {
            if (it == 1) {
                apply(q, d);
                dot(d, q, SF_ALPHA, 0);
            } else {   // d = z + beta d fused into q = A d and d.q
                search(q, dbuf[cur ^ 1], z, dbuf[cur]);
                cur ^= 1;
                d = dbuf[cur];
            }
            s.kt.begin(K_ADVANCE, st);
            hipLaunchKernelGGL(k_fcg_update, dim3(DOT_BLOCKS), dim3(256), 0, st, s.partials.p, P.f, P.c, r.f, r.c,
                               rbuf[rc ^ 1].f, rbuf[rc ^ 1].c, d.f, d.c, q.f, q.c, s.types.fine.p, s.types.coarse.p, s.cmask.p,
                               t, nf, s.scal.p, (int)SS_ALPHA);
            s.kt.end(K_ADVANCE, st);
            rc ^= 1;
            r = rbuf[rc];
            rold = rbuf[rc ^ 1];
            hipLaunchKernelGGL(k_sfinish, dim3(1), dim3(FINISH_THREADS), 0, st, s.partials.p, DOT_BLOCKS, s.scal.p, (int)SF_NORM,
                               (int)SS_RNORM);
            double flags;
            post_read2();
            if (refine > 0 && it % refine == 0) {   // r is replaced whatever the read says
                rnorm = resid64(r);
                double recursive;
                wait_read2(SS_RNORM, recursive, flags);
                raise_flags((long long)flags);
            } else {
                // the next V-cycle's head runs during the read
                if (it < maxiter) queue_head(z, r, sweeps, SS_RNORM);
                wait_read2(SS_RNORM, rnorm, flags);
                if (flags != 0.0) head = false;
                raise_flags((long long)flags);
                if (refine > 0 && rnorm <= target) {
                    head = false;   // r is replaced: the V-cycle starts over
                    rnorm = resid64(r);
                }
            }
            relres = rnorm / bnorm;
            if (rnorm <= target) {
                head = false;
                converged = true;
                return it;
            }
            vcycle(z, r, sweeps, coarse, boost, SS_RNORM, &rold);   // and z.r, z.rold
            hipLaunchKernelGGL(k_sfinish_beta, dim3(1), dim3(FINISH_THREADS), 0, st, s.partials.p, DOT_BLOCKS, s.scal.p);
        }
        double flags;
        read2(SS_RNORM, rnorm, flags);
        raise_flags((long long)flags);
        return maxiter;
    }

    SVec vfield(std::vector<SField<float>>& pool, int k) { return SVec{pool[k].fine.p, pool[k].coarse.p}; }
    void workspace(int restart) {
        if ((int)s.Z.size() == restart) return;
        s.V.clear(); s.Z.clear();
        s.V.resize(restart + 1); s.Z.resize(restart);
        for (auto* pool : {&s.V, &s.Z})
            for (auto& f : *pool) {
                f.alloc((size_t)s.blocks * SBCELLS, Cc);
                fill_async(f.fine.p, (size_t)s.blocks * SBCELLS, 0.0f, st);
                fill_async(f.coarse.p, Cc, 0.0f, st);
            }
    }

    // fgmres.jl:26-92 pressure_fgmres!: right-preconditioned restarted FGMRES (Saad 1993),
    // Z_j = M V_j + w D^-1 V_j, two passes of Gram-Schmidt, the small system in FP64.
    int fgmres(double tol, int maxiter, int restart, int sweeps, int coarse, float boost, double& relres) {
        head = false;
        workspace(restart);
        SVec P = vec(s.P), b = vec(s.b), r = vec(s.r), q = vec(s.q), z = vec(s.z), tmp = vec(s.tmp);
        dot(b, b, SF_NORM, SS_BB);
        double bnorm = read(SS_BB);
        if (!std::isfinite(bnorm)) throw Error(E_NONFINITE, "Nonfinite pressure RHS");
        if (bnorm == 0) { zero_solution(); relres = 0; return 0; }
        double target = (double)((float)tol * (float)bnorm);
        auto actual = [&]() { return true_residual(r, q); };
        double residual = actual();
        int iterations = 0;
        while (residual > target && iterations < maxiter) {
            if (!std::isfinite(residual)) throw Error(E_NONFINITE, "Nonfinite pressure residual");
            std::vector<double> H((size_t)(restart + 1) * restart, 0.0), g(restart + 1, 0.0), cs(restart, 0.0), sn(restart, 0.0);
            auto Hs = [&](int i, int j) -> double& { return H[(size_t)i + (size_t)(restart + 1) * j]; };
            g[0] = residual;
            SVec V0 = vfield(s.V, 0);
            zero(V0, (size_t)nf);
            axpy(V0, r, -1, (float)(1.0f / (float)residual));
            int used = 0;
            for (int j = 0; j < std::min(restart, maxiter - iterations); ++j) {
                SVec Vj = vfield(s.V, j), Zj = vfield(s.Z, j);
                vcycle(z, Vj, sweeps, coarse, boost);
                // the diagonal part: one damped Jacobi sweep from zero (fgmres.jl:15-23)
                rows<16, 3>(SVecT<float>{Zj.f, Zj.c}, SVecT<float>{tmp.f, tmp.c}, SVecT<float>{Vj.f, Vj.c}, hier32().L0);
                axpy(Zj, z, -1, 1.0f);
                apply(q, Zj);
                int k = 0;
                for (int pass = 0; pass < 2; ++pass)
                    for (int i = 0; i <= j; ++i, ++k) {
                        dot(q, vfield(s.V, i), SF_SET, SS_H0 + k);
                        axpy(q, vfield(s.V, i), SS_H0 + k, -1.0f);
                    }
                dot(q, q, SF_NORM, SS_TMP);
                std::vector<double> hv(SS_H0 + k);
                MAEL_HIP(hipMemcpyAsync(s.st.pinned, s.scal.p, (SS_H0 + k) * sizeof(double), hipMemcpyDeviceToHost, st));
                MAEL_HIP(hipStreamSynchronize(st));
                std::memcpy(hv.data(), s.st.pinned, (SS_H0 + k) * sizeof(double));
                k = 0;
                for (int pass = 0; pass < 2; ++pass)
                    for (int i = 0; i <= j; ++i, ++k) Hs(i, j) += (double)(float)hv[SS_H0 + k];
                double hnext = (double)(float)hv[SS_TMP];
                if (!std::isfinite(hnext)) throw Error(E_NONFINITE, "Nonfinite pressure Arnoldi vector");
                Hs(j + 1, j) = hnext;
                if (hnext > 0) {
                    SVec Vn = vfield(s.V, j + 1);
                    zero(Vn, (size_t)nf);
                    axpy(Vn, q, -1, (float)(1.0f / (float)hnext));
                }
                for (int i = 0; i < j; ++i) {
                    double a = Hs(i, j), bb = Hs(i + 1, j);
                    Hs(i, j) = cs[i] * a + sn[i] * bb;
                    Hs(i + 1, j) = -sn[i] * a + cs[i] * bb;
                }
                double a = Hs(j, j), bb = Hs(j + 1, j), hyp = std::hypot(a, bb);
                if (!(hyp > 0)) throw Error(E_BREAKDOWN, "Pressure Arnoldi breakdown");
                cs[j] = a / hyp; sn[j] = bb / hyp;
                Hs(j, j) = hyp; Hs(j + 1, j) = 0;
                g[j + 1] = -sn[j] * g[j]; g[j] = cs[j] * g[j];
                used = j + 1;
                ++iterations;
                if (std::fabs(g[j + 1]) <= target || hnext == 0) break;
            }
            std::vector<double> y(used);
            for (int i = used - 1; i >= 0; --i) {
                double acc = g[i];
                for (int kk = i + 1; kk < used; ++kk) acc -= Hs(i, kk) * y[kk];
                y[i] = acc / Hs(i, i);
            }
            for (int j = 0; j < used; ++j) axpy(P, vfield(s.Z, j), -1, (float)y[j]);
            residual = actual();
        }
        relres = residual / bnorm;
        if (!(residual <= target))
            throw Error(E_NOCONVERGE, "Pressure did not converge: " + std::to_string(residual / bnorm) + " after " +
                                          std::to_string(iterations) + " iterations");
        return iterations;
    }
};

// sflip.jl:1883 / 1901 `vcycle!(s, z, r)` alone: the couplings and hierarchy from the
// current types and level set, then z = M b into P. The residual the V-cycle reads is a
// copy of b; in reduced precision it is scaled by its norm as in the solve.
void SparseBasin::precondition(const SparseOpts& o, bool queued) {
    SparseSolver S(*this);
    S.coefficients();
    S.hierarchy();
    S.precision = o.precision;
    if (o.precision != 0) {
        mp_alloc();
        if (o.precision == 1) S.cast_couplings<__half>();
        else S.cast_couplings<bf16>();
    }
    SVec bv = S.vec(b), rv = S.vec(r), z = S.vec(P);
    S.copy(rv, bv);
    S.dot(rv, rv, SF_NORM, SS_RNORM);
    if (queued) S.queue_head(z, rv, o.sweeps, SS_RNORM);
    S.vcycle(z, rv, o.sweeps, o.coarse, o.boost, SS_RNORM);
    MAEL_LAUNCH_CHECK();
    st.sync();
}

// The pressure solve (sflip.jl:1867-1904 with fgmres.jl:26-110) on the right-hand side in
// b, from P (and with refinement the FP64 pressure from P): the projection's and its
// adjoint's. The couplings (SparseSolver::coefficients) must be those of the state.
void SparseBasin::solve_rhs(const SparseOpts& o, SparseStats& stats) {
    SparseSolver S(*this);
    S.hierarchy();
    S.precision = o.precision;
    S.refine = o.refine;
    if (o.refine > 0) {   // the FP64 solution starts from the stored pressure, the correction from 0
        ir_alloc();
        S.ir(0);
    }
    if (o.precision != 0) {
        mp_alloc();
        if (o.precision == 1) S.cast_couplings<__half>();
        else S.cast_couplings<bf16>();
    }
    double rr = 0;
    bool conv = false;
    if (o.policy == 0) {
        stats.fcg_iterations = S.fcg(o.tol, o.maxiter, o.sweeps, o.coarse, o.boost, conv, rr);
    } else if (o.policy == 1) {
        stats.fgmres_iterations = S.fgmres(o.tol, o.maxiter, o.restart, o.sweeps, o.coarse, o.boost, rr);
        stats.used_fgmres = 1;
    } else {
        // fgmres.jl:96-110: CG first; FGMRES when it fails or misses the tolerance
        try {
            stats.fcg_iterations = S.fcg(o.tol, std::min(80, o.maxiter), o.sweeps, o.coarse, o.boost, conv, rr);
        } catch (const Error& e) {
            std::string m = e.what();
            if (m.rfind("Nonfinite pressure", 0) != 0 && m.rfind("Pressure CG ", 0) != 0) throw;
        }
        SVec bv = S.vec(b), rv = S.vec(r), qv = S.vec(q);
        double rn = S.true_residual(rv, qv);
        S.dot(bv, bv, SF_NORM, SS_BB);
        double bn = S.read(SS_BB);
        if (!(rn <= (double)((float)o.tol * (float)bn))) {
            if (!std::isfinite(rn) || rn > 2.0 * bn) S.zero_solution();
            stats.used_fgmres = 1;
            stats.fgmres_iterations = S.fgmres(o.tol, o.maxiter, o.restart, o.sweeps, o.coarse, o.boost, rr);
        }
    }
    stats.iterations = stats.fcg_iterations + stats.fgmres_iterations;
    {   // the true relative residual of the returned pressure, in FP64 sums
        SVec bv = S.vec(b), rv = S.vec(r), qv = S.vec(q);
        double rn = S.true_residual(rv, qv);
        S.dot(bv, bv, SF_NORM, SS_BB);
        double bn = S.read(SS_BB);
        stats.rel_residual = bn > 0 ? rn / bn : 0.0;
    }
    // with refinement the stored pressure is the FP64 one rounded, and the faces take
    // their differences from the FP64 one
    if (S.refine > 0) S.ir(2);
}

void SparseBasin::project(float gamma, const SparseOpts& o, SparseStats& stats) {
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = st.s;
    SparseSolver S(*this);
    STopo t = topo();
    stats = SparseStats{};
    S.coefficients();
    if (S.nf)
        hipLaunchKernelGGL(k_finerhs, dim3(cdiv(S.nf, 256)), dim3(256), 0, s, b.fine.p, P.fine.p, types.fine.p, u[0].fine.p,
                           u[1].fine.p, u[2].fine.p, t, gamma, S.nf);
    hipLaunchKernelGGL(k_coarserhs, dim3(cdiv(S.Cc, 256)), dim3(256), 0, s, b.coarse.p, P.coarse.p, types.coarse.p, u[0].coarse.p,
                       u[1].coarse.p, u[2].coarse.p, t, gamma);
    MAEL_LAUNCH_CHECK();
    solve_rhs(o, stats);
    // sflip.jl:2008-2013 the faces, then the coarse faces under fine blocks from theirs
    for (int a = 0; a < 3; ++a) {
        if (S.nf && nlive > 0) {
            if (o.refine > 0)
                hipLaunchKernelGGL(k_projectfine<double>, dim3((unsigned)nlive), dim3(256), 0, s, u[a].fine.p, fl[a].fine.p,
                                   Pd.fine.p, Pd.coarse.p, types.fine.p, types.coarse.p, phi.fine.p, phi.coarse.p, t, a,
                                   sorder.p);
            else
                hipLaunchKernelGGL(k_projectfine<float>, dim3((unsigned)nlive), dim3(256), 0, s, u[a].fine.p, fl[a].fine.p,
                                   P.fine.p, P.coarse.p, types.fine.p, types.coarse.p, phi.fine.p, phi.coarse.p, t, a,
                                   sorder.p);
        }
        if (o.refine > 0)
            hipLaunchKernelGGL(k_projectcoarse<double>, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, s, u[a].coarse.p,
                               fl[a].coarse.p, Pd.coarse.p, types.coarse.p, phi.coarse.p, t, a);
        else
            hipLaunchKernelGGL(k_projectcoarse<float>, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, s, u[a].coarse.p,
                               fl[a].coarse.p, P.coarse.p, types.coarse.p, phi.coarse.p, t, a);
        launch_restrictfaces_f(*this, u[a].coarse.p, u[a].fine.p, a);
        launch_restrictfaces_u8(*this, fl[a].coarse.p, fl[a].fine.p, a);
    }
    MAEL_LAUNCH_CHECK();
    st.sync();
    stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// The transpose of project!'s linear part, in place on ub (fine pools and coarse grid); P
// is left holding the adjoint pressure y = A^-1 (-G^T R^T ub), solved from zero with the
// same options as the projection.
void SparseBasin::project_adjoint(const SparseOpts& o, SparseStats& stats) {
    auto t0 = std::chrono::steady_clock::now();
    adj_alloc();
    hipStream_t s = st.s;
    SparseSolver S(*this);
    STopo t = topo();
    stats = SparseStats{};
    size_t F = finefaces(), C = finecells();
    for (int a = 0; a < 3; ++a) launch_restrictfaces_T(*this, ub[a].fine.p, ub[a].coarse.p, ub[a].coarse.p, a);
    // the fixed-point scale from the largest cotangent: a cell takes at most 6 coarse and
    // 96 fine faces' terms, each at most |ub| / (0.01 dx) (theta >= 0.01, d >= dx)
    int* bound = (int*)partials.p;
    fill_async(bound, (size_t)1, 0, s);
    for (int a = 0; a < 3; ++a) {
        if (F) hipLaunchKernelGGL(k_sadj_absmax, dim3(cdiv(F, 256)), dim3(256), 0, s, ub[a].fine.p, F, where.p, bound);
        hipLaunchKernelGGL(k_sadj_absmax, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, s, ub[a].coarse.p, nfaces(a),
                           (const int4*)nullptr, bound);
    }
    MAEL_HIP(hipMemcpyAsync(st.pinned, bound, sizeof(int), hipMemcpyDeviceToHost, s));
    st.sync();
    float m;
    std::memcpy(&m, st.pinned, sizeof(float));
    if (!std::isfinite(m)) throw Error(E_NONFINITE, "Nonfinite face cotangent");
    double scale = 1.0;
    if (m > 0.0f) {
        double big = (double)m * 100.0 / (double)g.dx;
        scale = std::ldexp(1.0, std::min(60, std::max(-60, (int)std::floor(std::log2(std::ldexp(1.0, 23) / big)))));
    }
    if (C) fill_async(adjcell[0].p, C, 0ull, s);
    fill_async(adjcell[1].p, ncells(), 0ull, s);
    for (int a = 0; a < 3; ++a) {
        if (F)
            hipLaunchKernelGGL(k_sadj_gradT_fine, dim3(cdiv(F, 256)), dim3(256), 0, s, adjcell[0].p, adjcell[1].p, ub[a].fine.p,
                               types.fine.p, types.coarse.p, phi.fine.p, phi.coarse.p, t, a, (int32_t)F, (float)scale);
        hipLaunchKernelGGL(k_sadj_gradT_coarse, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, s, adjcell[1].p, ub[a].coarse.p,
                           types.coarse.p, phi.coarse.p, t, a, (float)scale);
    }
    if (C)
        hipLaunchKernelGGL(k_sadj_rhs_fine, dim3(cdiv(C, 256)), dim3(256), 0, s, b.fine.p, P.fine.p, adjcell[0].p, types.fine.p,
                           where.p, (int32_t)C, 1.0 / scale);
    hipLaunchKernelGGL(k_sadj_rhs_coarse, dim3(cdiv(ncells(), 256)), dim3(256), 0, s, b.coarse.p, P.coarse.p, adjcell[1].p,
                       types.coarse.p, t, 1.0 / scale);
    MAEL_LAUNCH_CHECK();
    S.coefficients();
    solve_rhs(o, stats);
    float h = g.dx;
    for (int a = 0; a < 3; ++a) {
        if (o.refine > 0) {
            if (F)
                hipLaunchKernelGGL(k_sadj_divT_fine<double>, dim3(cdiv(F, 256)), dim3(256), 0, s, ub[a].fine.p, Pd.fine.p,
                                   types.fine.p, t, a, (int32_t)F, h);
            hipLaunchKernelGGL(k_sadj_divT_coarse<double>, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, s, ub[a].coarse.p,
                               Pd.coarse.p, types.coarse.p, t, a, 16.0f * h);
        } else {
            if (F)
                hipLaunchKernelGGL(k_sadj_divT_fine<float>, dim3(cdiv(F, 256)), dim3(256), 0, s, ub[a].fine.p, P.fine.p,
                                   types.fine.p, t, a, (int32_t)F, h);
            hipLaunchKernelGGL(k_sadj_divT_coarse<float>, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, s, ub[a].coarse.p,
                               P.coarse.p, types.coarse.p, t, a, 16.0f * h);
        }
    }
    MAEL_LAUNCH_CHECK();
    st.sync();
    stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace mael
