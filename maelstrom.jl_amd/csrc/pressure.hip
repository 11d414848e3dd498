// The dense pressure projection (flip.jl:809-1156): ghost-fluid Poisson operator,
// McAdams et al.'s red-black Gauss-Seidel V-cycle as the preconditioner of CG, and
// the velocity update.
//
// MI355X layout: every stencil kernel runs over a 3-D grid of 64x4 workgroups (one
// wavefront per 64 cells of an x row, coalesced), the dot products of CG are summed in
// FP64 per workgroup and then in one fixed-order pass, and the Krylov scalars (alpha,
// beta, rho) stay on the device: the host reads one norm per iteration.
//
// Accuracy: the CG works on a correction in FP32; the pressure is kept in FP64 and the
// true residual b - A P is recomputed in FP64 (with the ghost-fluid diagonal evaluated
// in FP64) whenever the FP32 recursive residual reaches the target, which restarts CG on
// the remaining error (iterative refinement). That reaches relative residuals below the
// FP32 stagnation floor (fgmres.jl:56-58 notes one above 5e-4).
#include "dense.hpp"
#include "scan.hpp"
#include <chrono>
#include <cmath>

namespace mael {

constexpr int BX = 64, BY = 4;

__device__ inline size_t lin(int i, int j, int k, int n1, int n2) { return (size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k); }
__device__ inline uint8_t typeat(const uint8_t* __restrict__ t, int i, int j, int k, int n1, int n2, int n3) {
    return (i < 0 || j < 0 || k < 0 || i >= n1 || j >= n2 || k >= n3) ? SOLID : t[lin(i, j, k, n1, n2)];
}

// The stencil kernels issue all of a cell's neighbour loads before using any: a load
// under a branch that depends on an earlier load (a neighbour's value read only if its
// type is LIQUID) waits for that load, which makes a one-thread-per-cell pass a chain of
// memory round trips per thread (15 in a colour pass, 129 in the restriction). Each
// neighbour is read unconditionally, at the cell's own index where it is off the grid,
// and the sums take it only where the reference's test holds, in the reference's order:
// the same operations on the same values.
// The six neighbours of cell (i, j, k) in the order of NB6 below: their indices and
// whether each is on the grid (typeat's SOLID otherwise).
struct Nb6 {
    size_t e[6];
    bool in[6];
};
__device__ inline Nb6 nb6(int i, int j, int k, int3 n) {
    Nb6 r;
    const size_t c = lin(i, j, k, n.x, n.y), sy = (size_t)n.x, sz = (size_t)n.x * n.y;
    r.in[0] = i + 1 < n.x; r.e[0] = r.in[0] ? c + 1 : c;
    r.in[1] = i > 0;       r.e[1] = r.in[1] ? c - 1 : c;
    r.in[2] = j + 1 < n.y; r.e[2] = r.in[2] ? c + sy : c;
    r.in[3] = j > 0;       r.e[3] = r.in[3] ? c - sy : c;
    r.in[4] = k + 1 < n.z; r.e[4] = r.in[4] ? c + sz : c;
    r.in[5] = k > 0;       r.e[5] = r.in[5] ? c - sz : c;
    return r;
}
// (i, j, k) clamped into the grid, and whether it was inside.
__device__ inline size_t clampidx(int i, int j, int k, int3 n, bool& in) {
    in = i >= 0 && j >= 0 && k >= 0 && i < n.x && j < n.y && k < n.z;
    return lin(min(max(i, 0), n.x - 1), min(max(j, 0), n.y - 1), min(max(k, 0), n.z - 1), n.x, n.y);
}

__device__ inline double block_sum_d(double v, double* lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    int tid = threadIdx.x + blockDim.x * threadIdx.y;
    if ((tid & 63) == 0) lds[tid >> 6] = v;
    __syncthreads();
    double s = 0;
    if (tid == 0) {
        int nw = (blockDim.x * blockDim.y + 63) / 64;
        for (int w = 0; w < nw; ++w) s += lds[w];
    }
    return s;
}
__device__ inline int block_id3() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

// ─── operator pieces ─────────────────────────────────────────────────────────

// flip.jl:879-889 coarsen_kernel!.
__global__ void k_coarsen(uint8_t* __restrict__ tc, const uint8_t* __restrict__ tf, int3 nc, int3 nf) {
    int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y, K = blockIdx.z;
    if (I >= nc.x || J >= nc.y || K >= nc.z) return;
    bool air = false, liq = false;
    uint8_t tv[8];
    bool in[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) tv[m] = tf[clampidx(2 * I + (m & 1), 2 * J + ((m >> 1) & 1), 2 * K + (m >> 2), nf, in[m])];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        uint8_t t = in[m] ? tv[m] : SOLID;
        air |= t == AIR;
        liq |= t == LIQUID;
    }
    tc[lin(I, J, K, nc.x, nc.y)] = air ? AIR : (liq ? LIQUID : SOLID);
}

__constant__ int NB6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

// flip.jl:824-836 diag_kernel!.
__global__ void k_diag(float* __restrict__ diag, const uint8_t* __restrict__ types, const float* __restrict__ phi, int3 n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n.x || j >= n.y || k >= n.z) return;
    size_t c = lin(i, j, k, n.x, n.y);
    float d = 0.0f;
    if (types[c] == LIQUID) {
        const Nb6 nb = nb6(i, j, k, n);
        uint8_t tn[6];
        float pn[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) { tn[m] = types[nb.e[m]]; pn[m] = phi[nb.e[m]]; }
        float pc = phi[c];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            uint8_t t = nb.in[m] ? tn[m] : SOLID;
            if (t == LIQUID) d += 1.0f;
            if (t == AIR) d += __fdiv_rn(1.0f, theta(pc, pn[m]));
        }
    }
    diag[c] = d;
}

// flip.jl:838-857 celldiag / liquidsum; `diag` null on the coarse levels.
struct Row { float d, s; };
template <class X>
__device__ inline Row row(const X* __restrict__ x, const uint8_t* __restrict__ t, const float* __restrict__ diag, int i, int j,
                          int k, int3 n) {
    const Nb6 nb = nb6(i, j, k, n);
    uint8_t tv[6];
    float xv[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) { tv[m] = t[nb.e[m]]; xv[m] = (float)x[nb.e[m]]; }
    const bool fine = diag != nullptr;
    const float dg = fine ? diag[lin(i, j, k, n.x, n.y)] : 0.0f;
    Row r{0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const uint8_t tn = nb.in[m] ? tv[m] : SOLID;
        if (!fine && tn != SOLID) r.d += 1.0f;
        if (tn == LIQUID) r.s += xv[m];
    }
    if (fine) r.d = dg;
    return r;
}

// flip.jl:811-822 rhs_kernel! fused with flip.jl:1132-1135 clearair_kernel!.
__global__ void k_rhs(float* __restrict__ b, float* __restrict__ P, const float* __restrict__ u, const float* __restrict__ v,
                      const float* __restrict__ w, const uint8_t* __restrict__ types, int3 n, float dx, float gamma) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n.x || j >= n.y || k >= n.z) return;
    size_t c = lin(i, j, k, n.x, n.y);
    if (types[c] == LIQUID) {
        size_t iu = lin(i, j, k, n.x + 1, n.y), iv = lin(i, j, k, n.x, n.y + 1);
        float div = u[iu + 1] - u[iu] + v[iv + n.x] - v[iv] + w[c + (size_t)n.x * n.y] - w[c];
        b[c] = -dx * div + dx * dx * gamma;
    } else {
        b[c] = 0.0f;
        P[c] = 0.0f;
    }
}

// flip.jl:859-863 residual_kernel!.
__global__ void k_residual(float* __restrict__ r, const float* __restrict__ x, const float* __restrict__ b,
                           const uint8_t* __restrict__ t, const float* __restrict__ diag, int3 n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n.x || j >= n.y || k >= n.z) return;
    size_t c = lin(i, j, k, n.x, n.y);
    float v = 0.0f;
    if (t[c] == LIQUID) {
        Row rw = row(x, t, diag, i, j, k, n);
        v = b[c] - (rw.d * x[c] - rw.s);
    }
    r[c] = v;
}

// flip.jl:865-877 rbgs_kernel!: one colour, over half the first axis. The colour is
// the reference's 1-based parity: 0-based cells with (i + j + k + 1) % 2 == colour.
__global__ void k_rbgs(float* __restrict__ x, const float* __restrict__ b, const uint8_t* __restrict__ t,
                       const float* __restrict__ diag, int3 n, int colour) {
    int h = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (h >= n.x / 2 || j >= n.y || k >= n.z) return;
    int i = 2 * h + ((colour + j + k + 1) & 1);
    size_t c = lin(i, j, k, n.x, n.y);
    if (t[c] != LIQUID) return;
    Row rw = row(x, t, diag, i, j, k, n);
    x[c] = rw.d > 0.0f ? __fdiv_rn(b[c] + rw.s, rw.d) : 0.0f;
}

// flip.jl:964-973 start_kernel!: the first colour from zero.
__global__ void k_start(float* __restrict__ x, const float* __restrict__ b, const uint8_t* __restrict__ t,
                        const float* __restrict__ diag, int3 n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n.x || j >= n.y || k >= n.z) return;
    size_t c = lin(i, j, k, n.x, n.y);
    float v = 0.0f;
    if (((i + j + k + 1) & 1) == 0 && t[c] == LIQUID) {
        float d;
        if (diag) d = diag[c];
        else {
            d = 0.0f;
            const Nb6 nb = nb6(i, j, k, n);
            uint8_t tv[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) tv[m] = t[nb.e[m]];
#pragma unroll
            for (int m = 0; m < 6; ++m)
                if ((nb.in[m] ? tv[m] : SOLID) != SOLID) d += 1.0f;
        }
        if (d > 0.0f) v = __fdiv_rn(b[c], d);
    }
    x[c] = v;
}

// flip.jl:891-906 restrict_kernel!.
__global__ void k_restrict(float* __restrict__ bc, const float* __restrict__ rf, const uint8_t* __restrict__ tf,
                           const uint8_t* __restrict__ tc, int3 nf, int3 nc) {
    int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y, K = blockIdx.z;
    if (I >= nc.x || J >= nc.y || K >= nc.z) return;
    size_t C = lin(I, J, K, nc.x, nc.y);
    float s = 0.0f;
    if (tc[C] == LIQUID) {
        const float wB[4] = {0.125f, 0.375f, 0.375f, 0.125f};
#pragma unroll
        for (int c = 0; c < 4; ++c) {   // a plane of 16 taps' loads at a time, then their sum
            uint8_t tv[16];
            float rv[16];
            bool in[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const size_t e = clampidx(2 * I - 1 + (m & 3), 2 * J - 1 + (m >> 2), 2 * K - 1 + c, nf, in[m]);
                tv[m] = tf[e];
                rv[m] = rf[e];
            }
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int m = a + 4 * b;
                    if (in[m] && tv[m] == LIQUID) s += wB[a] * wB[b] * wB[c] * rv[m];
                }
        }
    }
    bc[C] = 4.0f * s;
}

// flip.jl:908-929 prolong_kernel!.
__device__ inline void pweights(int i, int& a0, float& w0, int& a1, float& w1) {
    if ((i & 1) == 0) { a0 = i / 2 - 1; w0 = 0.25f; a1 = i / 2; w1 = 0.75f; }
    else { a0 = (i - 1) / 2; w0 = 0.75f; a1 = (i + 1) / 2; w1 = 0.25f; }
}
__global__ void k_prolong(float* __restrict__ xf, const float* __restrict__ xc, const uint8_t* __restrict__ tf,
                          const uint8_t* __restrict__ tc, int3 nf, int3 nc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= nf.x || j >= nf.y || k >= nf.z) return;
    size_t c = lin(i, j, k, nf.x, nf.y);
    if (tf[c] != LIQUID) return;
    int A[2], B[2], Cc[2];
    float WA[2], WB[2], WC[2];
    pweights(i, A[0], WA[0], A[1], WA[1]);
    pweights(j, B[0], WB[0], B[1], WB[1]);
    pweights(k, Cc[0], WC[0], Cc[1], WC[1]);
    uint8_t tv[8];
    float xv[8];
    bool in[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const size_t e = clampidx(A[m & 1], B[(m >> 1) & 1], Cc[(m >> 2) & 1], nc, in[m]);
        tv[m] = tc[e];
        xv[m] = xc[e];
    }
    const float x0 = xf[c];
    float s = 0.0f;
#pragma unroll
    for (int m = 0; m < 8; ++m)
        if (in[m] && tv[m] == LIQUID) s += WA[m & 1] * WB[(m >> 1) & 1] * WC[(m >> 2) & 1] * xv[m];
    xf[c] = x0 + s;
}

// flip.jl:931-962 coarse_kernel!: the coarsest level (<= 1024 cells) in one workgroup,
// `sweeps` RBGS sweeps from zero then as many in reverse colour order. Accumulates in T
// (FP32 as the reference, or FP64).
template <class T>
__global__ __launch_bounds__(1024) void k_coarsest(float* __restrict__ x, const float* __restrict__ b,
                                                   const uint8_t* __restrict__ types, int3 n, int sweeps) {
    __shared__ T xs[1024];
    int t = threadIdx.x;
    int N = n.x * n.y * n.z;
    bool active = t < N;
    int i = t % n.x, j = (t / n.x) % n.y, k = t / (n.x * n.y);
    uint8_t ti = active ? types[t] : SOLID;
    T bi = active ? (T)b[t] : T(0);
    T d = 0;
    int nbr[6];
    bool liq[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        int a = i + NB6[m][0], bb = j + NB6[m][1], c = k + NB6[m][2];
        uint8_t tn = active ? typeat(types, a, bb, c, n.x, n.y, n.z) : SOLID;
        if (tn != SOLID) d += T(1);
        liq[m] = tn == LIQUID;
        nbr[m] = liq[m] ? a + n.x * (bb + n.y * c) : 0;
    }
    xs[t] = T(0);
    __syncthreads();
    int par = (i + j + k + 1) & 1;
    for (int s = 1; s <= 2 * sweeps; ++s)
        for (int half = 0; half < 2; ++half) {
            int colour = s <= sweeps ? half : 1 - half;
            if (active && ti == LIQUID && par == colour) {
                T v[6];   // the six neighbours read together (a non-liquid one reads cell 0)
#pragma unroll
                for (int m = 0; m < 6; ++m) v[m] = xs[nbr[m]];
                T acc = bi;
#pragma unroll
                for (int m = 0; m < 6; ++m) acc = liq[m] ? acc + v[m] : acc;
                xs[t] = d > T(0) ? acc / d : T(0);
            }
            __syncthreads();
        }
    if (active) x[t] = (float)xs[t];
}

// ─── CG passes ───────────────────────────────────────────────────────────────
// Device scalars: [0] rho, [1] p.q, [2] r.r, [3] b.b, [4] alpha, [5] beta, [6] |r|.
enum { S_RHO = 0, S_PQ, S_RR, S_BB, S_ALPHA, S_BETA, S_RNORM, S_N };

// flip.jl:1033-1054 search_kernel!: q = A p for p = z + beta d, and p.q per workgroup.
__global__ __launch_bounds__(256) void k_search(double* __restrict__ partials, float* __restrict__ q,
                                                const float* __restrict__ z, const float* __restrict__ d,
                                                const uint8_t* __restrict__ t, const float* __restrict__ diag, int3 n,
                                                const double* __restrict__ scal) {
    __shared__ double lds[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    float beta = (float)scal[S_BETA];
    double v = 0;
    if (i < n.x && j < n.y && k < n.z) {
        size_t c = lin(i, j, k, n.x, n.y);
        float qv = 0.0f;
        if (t[c] == LIQUID) {
            const Nb6 nb = nb6(i, j, k, n);
            uint8_t tv[6];
            float zv[6], dv[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) { tv[m] = t[nb.e[m]]; zv[m] = z[nb.e[m]]; dv[m] = d[nb.e[m]]; }
            const float dg = diag[c];
            float p = z[c] + beta * d[c];
            float acc = 0.0f;
#pragma unroll
            for (int m = 0; m < 6; ++m)
                if (nb.in[m] && tv[m] == LIQUID) acc += zv[m] + beta * dv[m];
            qv = dg * p - acc;
            v = (double)p * (double)qv;
        }
        q[c] = qv;
    }
    double s = block_sum_d(v, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) partials[block_id3()] = s;
}

// flip.jl:1059-1073 advance_kernel!: d = z + beta d, x += alpha d, r -= alpha q, r.r per workgroup.
__global__ __launch_bounds__(256) void k_advance(double* __restrict__ partials, float* __restrict__ x, float* __restrict__ r,
                                                 float* __restrict__ d, const float* __restrict__ z,
                                                 const float* __restrict__ q, size_t N, const double* __restrict__ scal) {
    __shared__ double lds[4];
    float alpha = (float)scal[S_ALPHA], beta = (float)scal[S_BETA];
    double v = 0;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < N; l += (size_t)gridDim.x * blockDim.x) {
        float p = z[l] + beta * d[l];
        d[l] = p;
        x[l] += alpha * p;
        float rl = r[l] - alpha * q[l];
        r[l] = rl;
        v += (double)rl * (double)rl;
    }
    double s = block_sum_d(v, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_dot(double* __restrict__ partials, const float* __restrict__ a,
                                             const float* __restrict__ b, size_t N) {
    __shared__ double lds[4];
    double v = 0;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < N; l += (size_t)gridDim.x * blockDim.x)
        v += (double)a[l] * (double)b[l];
    double s = block_sum_d(v, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Sum the partials in a fixed order and update the Krylov scalars.
enum { F_SET = 0, F_RHO0, F_ALPHA, F_RR, F_BETA };
// One workgroup of FIN_THREADS, four independent sums a thread (the partials of a fine
// level's search are tens of thousands).
constexpr int FIN_THREADS = 1024;
__global__ __launch_bounds__(FIN_THREADS) void k_finish(const double* __restrict__ partials, int np, double* __restrict__ scal,
                                                        int op, int slot) {
    __shared__ double lds[FIN_THREADS / 64];
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int l = threadIdx.x;
    for (; l + 3 * FIN_THREADS < np; l += 4 * FIN_THREADS) {
        a0 += partials[l]; a1 += partials[l + FIN_THREADS]; a2 += partials[l + 2 * FIN_THREADS]; a3 += partials[l + 3 * FIN_THREADS];
    }
    for (; l < np; l += FIN_THREADS) a0 += partials[l];
    double s = block_sum_d((a0 + a1) + (a2 + a3), lds);
    if (threadIdx.x != 0) return;
    switch (op) {
        case F_SET: scal[slot] = s; break;
        case F_RHO0: scal[S_RHO] = s; scal[S_BETA] = 0.0; break;
        case F_ALPHA: scal[S_PQ] = s; scal[S_ALPHA] = scal[S_RHO] / s; break;
        case F_RR: scal[S_RR] = s; scal[S_RNORM] = sqrt(s); break;
        case F_BETA: scal[S_BETA] = s / scal[S_RHO]; scal[S_RHO] = s; break;
    }
}

// FP64 true residual r = b - A P with the ghost-fluid diagonal in FP64; writes the FP32
// residual for the next CG and the workgroup's share of r.r.
__global__ __launch_bounds__(256) void k_resid64(double* __restrict__ partials, float* __restrict__ r32,
                                                 const double* __restrict__ P, const float* __restrict__ b,
                                                 const uint8_t* __restrict__ t, const float* __restrict__ phi, int3 n) {
    __shared__ double lds[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    double v = 0;
    if (i < n.x && j < n.y && k < n.z) {
        size_t c = lin(i, j, k, n.x, n.y);
        double rr = 0;
        if (t[c] == LIQUID) {
            const Nb6 nb = nb6(i, j, k, n);
            uint8_t tv[6];
            double pv[6];
            float fv[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) { tv[m] = t[nb.e[m]]; pv[m] = P[nb.e[m]]; fv[m] = phi[nb.e[m]]; }
            double d = 0, s = 0, pc = phi[c];
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                const uint8_t tn = nb.in[m] ? tv[m] : SOLID;
                if (tn == LIQUID) { d += 1.0; s += pv[m]; }
                if (tn == AIR) d += 1.0 / theta(pc, (double)fv[m]);
            }
            rr = (double)b[c] - (d * P[c] - s);
        }
        r32[c] = (float)rr;
        v = rr * rr;
    }
    double s = block_sum_d(v, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) partials[block_id3()] = s;
}

__global__ void k_f2d(double* __restrict__ y, const float* __restrict__ x, size_t N) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < N) y[l] = (double)x[l];
}
__global__ void k_addcorr(double* __restrict__ P64, float* __restrict__ P32, const float* __restrict__ e, size_t N) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < N) {
        double p = P64[l] + (double)e[l];
        P64[l] = p;
        P32[l] = (float)p;
    }
}

// flip.jl:1113-1130 project_kernel! for axis A.
template <int A>
__global__ void k_project(float* __restrict__ u, uint8_t* __restrict__ valid, const float* __restrict__ P,
                          const uint8_t* __restrict__ t, const float* __restrict__ phi, int3 n, float dx) {
    int n1 = n.x + (A == 0), n2 = n.y + (A == 1), n3 = n.z + (A == 2);
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n1 || j >= n2 || k >= n3) return;
    int a = i - (A == 0), b = j - (A == 1), c = k - (A == 2);
    uint8_t t0 = typeat(t, a, b, c, n.x, n.y, n.z), t1 = typeat(t, i, j, k, n.x, n.y, n.z);
    bool ok = t0 != SOLID && t1 != SOLID && (t0 == LIQUID || t1 == LIQUID);
    size_t f = lin(i, j, k, n1, n2);
    if (ok) {
        size_t c0 = lin(a, b, c, n.x, n.y), c1 = lin(i, j, k, n.x, n.y);
        float D;
        if (t0 == LIQUID && t1 == LIQUID) D = P[c1] - P[c0];
        else if (t0 == LIQUID) D = __fdiv_rn(-P[c0], theta(phi[c0], phi[c1]));
        else D = __fdiv_rn(P[c1], theta(phi[c1], phi[c0]));
        u[f] -= __fdiv_rn(D, dx);
    }
    valid[f] = ok ? 1 : 0;
}

// ─── host side ───────────────────────────────────────────────────────────────

static inline dim3 grid3(int n1, int n2, int n3) { return dim3(cdiv(n1, BX), cdiv(n2, BY), (unsigned)n3); }
static inline int3 i3(const int n[3]) { return make_int3(n[0], n[1], n[2]); }
constexpr int RED_BLOCKS = 1024;   // workgroups of the 1-D reductions

struct DenseSolver {
    DenseBasin& s;
    hipStream_t st;
    int L;
    DenseSolver(DenseBasin& b) : s(b), st(b.st.s), L(b.levels) {}

    const uint8_t* types(int l) const { return l == 0 ? s.types.p : s.lv[l].types.p; }
    const float* diag(int l) const { return l == 0 ? s.diag.p : nullptr; }
    int3 dims(int l) const { return i3(s.lv[l].n); }

    void residual(int l, float* r, const float* x, const float* b) {
        int3 n = dims(l);
        s.kt.begin(K_RESIDUAL, st);
        hipLaunchKernelGGL(k_residual, grid3(n.x, n.y, n.z), dim3(BX, BY), 0, st, r, x, b, types(l), diag(l), n);
        s.kt.end(K_RESIDUAL, st);
    }
    void smooth(int l, float* x, const float* b, int colour) {
        int3 n = dims(l);
        if (l == 0) s.kt.begin(K_RBGS, st);
        hipLaunchKernelGGL(k_rbgs, grid3(n.x / 2, n.y, n.z), dim3(BX, BY), 0, st, x, b, types(l), diag(l), n, colour);
        if (l == 0) s.kt.end(K_RBGS, st);
    }
    // The V-cycle's first launches (level 0's start and first sweep) queued ahead of a
    // host read that decides whether the V-cycle runs, so the device works through the
    // read's round trip; the next vcycle() skips them. Whoever leaves the solve or changes
    // r first clears `head`.
    bool head = false;
    void queue_head(float* z, const float* r, int sweeps) {
        int3 n = dims(0);
        hipLaunchKernelGGL(k_start, grid3(n.x, n.y, n.z), dim3(BX, BY), 0, st, z, r, types(0), diag(0), n);
        if (2 * sweeps - 1 >= 1) smooth(0, z, r, 1);
        head = true;
    }
    // flip.jl:986-1014 vcycle!.
    void vcycle(float* z, const float* r, int sweeps, int coarse) {
        const bool queued = head;
        head = false;
        std::vector<float*> xs(L), bs(L), rs(L);
        for (int l = 0; l < L; ++l) {
            xs[l] = l == 0 ? z : s.lv[l].x.p;
            bs[l] = l == 0 ? const_cast<float*>(r) : s.lv[l].b.p;
            rs[l] = l == 0 ? s.rv.p : s.lv[l].r.p;
        }
        for (int l = 0; l < L - 1; ++l) {
            int3 n = dims(l), nc = dims(l + 1);
            const bool skip = queued && l == 0;   // start and first sweep already queued
            if (!skip)
                hipLaunchKernelGGL(k_start, grid3(n.x, n.y, n.z), dim3(BX, BY), 0, st, xs[l], bs[l], types(l), diag(l), n);
            for (int c = skip ? 1 : 0; c < 2 * sweeps - 1; ++c) smooth(l, xs[l], bs[l], (c & 1) ? 0 : 1);   // 1,0,1,...
            residual(l, rs[l], xs[l], bs[l]);
            if (l == 0) s.kt.begin(K_RESTRICT, st);
            hipLaunchKernelGGL(k_restrict, grid3(nc.x, nc.y, nc.z), dim3(BX, BY), 0, st, bs[l + 1], rs[l], types(l),
                               types(l + 1), n, nc);
            if (l == 0) s.kt.end(K_RESTRICT, st);
        }
        int3 nL = dims(L - 1);
        s.kt.begin(K_COARSEST, st);
        hipLaunchKernelGGL(k_coarsest<float>, dim3(1), dim3(1024), 0, st, xs[L - 1], bs[L - 1], types(L - 1), nL, coarse);
        s.kt.end(K_COARSEST, st);
        for (int l = L - 2; l >= 0; --l) {
            int3 n = dims(l), nc = dims(l + 1);
            if (l == 0) s.kt.begin(K_PROLONG, st);
            hipLaunchKernelGGL(k_prolong, grid3(n.x, n.y, n.z), dim3(BX, BY), 0, st, xs[l], xs[l + 1], types(l), types(l + 1),
                               n, nc);
            if (l == 0) s.kt.end(K_PROLONG, st);
            for (int c = 0; c < 2 * sweeps; ++c) smooth(l, xs[l], bs[l], (c & 1) ? 0 : 1);        // 1,0,1,0,...
        }
        MAEL_LAUNCH_CHECK();
    }
    double read(int slot) {
        post_read(slot);
        return wait_read();
    }
    // the same in two halves, with other work queued between them
    void post_read(int slot) {
        MAEL_HIP(hipMemcpyAsync(s.st.pinned, s.scal.p + slot, sizeof(double), hipMemcpyDeviceToHost, st));
        s.st.mark();
    }
    double wait_read() {
        s.st.poll();
        return *(double*)s.st.pinned;
    }
    void dot(const float* a, const float* b, int op, int slot = 0) {
        size_t N = s.cells();
        hipLaunchKernelGGL(k_dot, dim3(RED_BLOCKS), dim3(256), 0, st, s.partials.p, a, b, N);
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN_THREADS), 0, st, s.partials.p, RED_BLOCKS, s.scal.p, op, slot);
    }
    // FP64 true residual into r (FP32) and |r|^2 into scal[S_RR].
    double resid64(double* P64, float* r) {
        int3 n = dims(0);
        dim3 gr = grid3(n.x, n.y, n.z);
        s.kt.begin(K_RESID64, st);
        hipLaunchKernelGGL(k_resid64, gr, dim3(BX, BY), 0, st, s.partials.p, r, P64, s.b.p, s.types.p, s.phi.p, n);
        s.kt.end(K_RESID64, st);
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN_THREADS), 0, st, s.partials.p, (int)(gr.x * gr.y * gr.z), s.scal.p, F_RR, 0);
        MAEL_LAUNCH_CHECK();
        return std::sqrt(read(S_RR));
    }
};

// ─── the projection's adjoint (no counterpart in the reference) ───────────────
// project! maps u to u - G A^-1 (D u + dx^2 gamma) (D: the scaled divergence of k_rhs, G:
// the face differences of k_project); with the types and the level set fixed its linear
// part's transpose is ub -> ub + D^T A^-1 (-G^T ub), A being symmetric: the same solver
// on the adjoint right-hand side.

// b = -G^T ub per cell: each valid face of the cell (k_project's test) contributes its
// cotangent times -dD/dP_c / dx (LIQUID both sides: +-1; ghost fluid: 1/theta on the
// LIQUID side).
__global__ void k_adj_gradT(float* __restrict__ b, const float* __restrict__ u, const float* __restrict__ v,
                            const float* __restrict__ w, const uint8_t* __restrict__ t, const float* __restrict__ phi, int3 n,
                            float dx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n.x || j >= n.y || k >= n.z) return;
    size_t c = lin(i, j, k, n.x, n.y);
    float acc = 0.0f;
    if (t[c] == LIQUID) {
        const float* U[3] = {u, v, w};
        for (int A = 0; A < 3; ++A) {
            int n1 = n.x + (A == 0), n2 = n.y + (A == 1);
            int ei = A == 0, ej = A == 1, ek = A == 2;
            // the low face: this cell is its upper cell (c1)
            {
                uint8_t t0 = typeat(t, i - ei, j - ej, k - ek, n.x, n.y, n.z);
                if (t0 != SOLID) {
                    float ubf = U[A][lin(i, j, k, n1, n2)], dD;
                    if (t0 == LIQUID) dD = 1.0f;
                    else dD = 1.0f / theta(phi[c], phi[lin(i - ei, j - ej, k - ek, n.x, n.y)]);
                    acc += -ubf * dD / dx;
                }
            }
            // the high face: this cell is its lower cell (c0)
            {
                uint8_t t1 = typeat(t, i + ei, j + ej, k + ek, n.x, n.y, n.z);
                if (t1 != SOLID) {
                    float ubf = U[A][lin(i + ei, j + ej, k + ek, n1, n2)], dD;
                    if (t1 == LIQUID) dD = -1.0f;
                    else dD = -1.0f / theta(phi[c], phi[lin(i + ei, j + ej, k + ek, n.x, n.y)]);
                    acc += -ubf * dD / dx;
                }
            }
        }
    }
    b[c] = acc;
}
// ub += D^T lambda per face: the face is the high face of its lower cell (db/du = -dx)
// and the low face of its upper cell (db/du = +dx), for LIQUID cells (k_rhs).
template <int A>
__global__ void k_adj_divT(float* __restrict__ ub, const float* __restrict__ lam, const uint8_t* __restrict__ t, int3 n,
                           float dx) {
    int n1 = n.x + (A == 0), n2 = n.y + (A == 1), n3 = n.z + (A == 2);
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n1 || j >= n2 || k >= n3) return;
    int a = i - (A == 0), b = j - (A == 1), c = k - (A == 2);
    uint8_t t0 = typeat(t, a, b, c, n.x, n.y, n.z), t1 = typeat(t, i, j, k, n.x, n.y, n.z);
    float add = 0.0f;
    if (t0 == LIQUID) add += -dx * lam[lin(a, b, c, n.x, n.y)];
    if (t1 == LIQUID) add += dx * lam[lin(i, j, k, n.x, n.y)];
    ub[lin(i, j, k, n1, n2)] += add;
}

// flip.jl:1084-1088: coarse types and the fine diagonal of the operator.
void DenseBasin::operator_setup() {
    hipStream_t s = st.s;
    DenseSolver S(*this);
    int3 n = i3(g.n);
    for (int l = 1; l < levels; ++l) {
        int3 nc = i3(lv[l].n), nf = i3(lv[l - 1].n);
        hipLaunchKernelGGL(k_coarsen, grid3(nc.x, nc.y, nc.z), dim3(BX, BY), 0, s, lv[l].types.p, S.types(l - 1), nc, nf);
    }
    hipLaunchKernelGGL(k_diag, grid3(n.x, n.y, n.z), dim3(BX, BY), 0, s, diag.p, types.p, phi.p, n);
    MAEL_LAUNCH_CHECK();
}

// flip.jl:1075-1111 pressure!: A P = b from the pressure in P (the caller's start), by
// MG-preconditioned CG on the FP32 correction with FP64 restarts (SolveOpts::refine).
void DenseBasin::solve_pressure(const SolveOpts& o, SolveStats& stats) {
    hipStream_t s = st.s;
    DenseSolver S(*this);
    int3 n = i3(g.n);
    dim3 gr = grid3(n.x, n.y, n.z);
    size_t N = cells();
    // FP64 pressure (warm start from P) and its true residual.
    double* P64 = this->P64.p;
    hipLaunchKernelGGL(k_f2d, dim3(cdiv(N, 256)), dim3(256), 0, s, P64, P.p, N);
    S.dot(b.p, b.p, F_SET, S_BB);
    double bnorm = std::sqrt(S.read(S_BB));
    stats = SolveStats{};
    if (!std::isfinite(bnorm)) throw Error(E_NONFINITE, "Nonfinite pressure right-hand side");
    double target = o.tol * bnorm;
    // flip.jl:1091: no solve for a zero right-hand side (the warm start is projected).
    double rnorm = bnorm == 0.0 ? 0.0 : S.resid64(P64, r.p);
    int it = 0;
    float* e = this->e.p;   // the FP32 correction CG solves for
    while (rnorm > target && it < o.maxiter) {
        // CG on A e = r from e = 0 (flip.jl:1095-1108), stopped by the recursive residual.
        fill_async(e, N, 0.0f, s);
        fill_async(d.p, N, 0.0f, s);
        S.vcycle(z.p, r.p, o.sweeps, o.coarse);
        S.dot(z.p, r.p, F_RHO0);
        double rrec = rnorm;
        while (it < o.maxiter) {
            ++it;
            kt.begin(K_SEARCH, s);
            hipLaunchKernelGGL(k_search, gr, dim3(BX, BY), 0, s, partials.p, q.p, z.p, d.p, types.p, diag.p, n, scal.p);
            kt.end(K_SEARCH, s);
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN_THREADS), 0, s, partials.p, (int)(gr.x * gr.y * gr.z), scal.p, F_ALPHA, 0);
            kt.begin(K_ADVANCE, s);
            hipLaunchKernelGGL(k_advance, dim3(RED_BLOCKS), dim3(256), 0, s, partials.p, e, r.p, d.p, z.p, q.p, N, scal.p);
            kt.end(K_ADVANCE, s);
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN_THREADS), 0, s, partials.p, RED_BLOCKS, scal.p, F_RR, 0);
            MAEL_LAUNCH_CHECK();
            S.post_read(S_RNORM);
            if (it < o.maxiter) S.queue_head(z.p, r.p, o.sweeps);   // the next V-cycle's head, during the read
            rrec = S.wait_read();
            if (!std::isfinite(rrec)) throw Error(E_NONFINITE, "Nonfinite pressure residual");
            if (rrec <= target) {
                S.head = false;
                break;
            }
            S.vcycle(z.p, r.p, o.sweeps, o.coarse);
            S.dot(z.p, r.p, F_BETA);
        }
        stats.rel_residual_recursive = rrec / bnorm;
        hipLaunchKernelGGL(k_addcorr, dim3(cdiv(N, 256)), dim3(256), 0, s, P64, P.p, e, N);
        rnorm = S.resid64(P64, r.p);
        if (!o.refine) break;
        if (rnorm > target) ++stats.restarts;
    }
    stats.iterations = it;
    stats.rel_residual = bnorm == 0.0 ? 0.0 : rnorm / bnorm;
}

// flip.jl:1094 / 1106 `vcycle!(s)` alone: the operator from the current types and level
// set, then z = M b into P.
void DenseBasin::precondition(const SolveOpts& o, bool queued) {
    DenseSolver S(*this);
    operator_setup();
    if (queued) S.queue_head(P.p, b.p, o.sweeps);
    S.vcycle(P.p, b.p, o.sweeps, o.coarse);
    st.sync();
}

// The transpose of project!'s linear part, in place on ub; P is left holding the adjoint
// pressure lambda = A^-1 (-G^T ub) (solved from zero).
void DenseBasin::project_adjoint(const SolveOpts& o, SolveStats& stats) {
    auto t0 = std::chrono::steady_clock::now();
    adj_alloc();
    hipStream_t s = st.s;
    int3 n = i3(g.n);
    dim3 gr = grid3(n.x, n.y, n.z);
    hipLaunchKernelGGL(k_adj_gradT, gr, dim3(BX, BY), 0, s, b.p, ub[0].p, ub[1].p, ub[2].p, types.p, phi.p, n, g.dx);
    fill_async(P.p, cells(), 0.0f, s);
    operator_setup();
    solve_pressure(o, stats);
    hipLaunchKernelGGL(k_adj_divT<0>, grid3(n.x + 1, n.y, n.z), dim3(BX, BY), 0, s, ub[0].p, P.p, types.p, n, g.dx);
    hipLaunchKernelGGL(k_adj_divT<1>, grid3(n.x, n.y + 1, n.z), dim3(BX, BY), 0, s, ub[1].p, P.p, types.p, n, g.dx);
    hipLaunchKernelGGL(k_adj_divT<2>, grid3(n.x, n.y, n.z + 1), dim3(BX, BY), 0, s, ub[2].p, P.p, types.p, n, g.dx);
    MAEL_LAUNCH_CHECK();
    st.sync();
    stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void DenseBasin::project(float dt, float gamma, const SolveOpts& o, SolveStats& stats) {
    (void)dt;
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = st.s;
    int3 n = i3(g.n);
    dim3 gr = grid3(n.x, n.y, n.z);
    if (!o.warm) fill_async(P.p, cells(), 0.0f, s);
    hipLaunchKernelGGL(k_rhs, gr, dim3(BX, BY), 0, s, b.p, P.p, u[0].p, u[1].p, u[2].p, types.p, n, g.dx, gamma);
    operator_setup();
    solve_pressure(o, stats);
    // flip.jl:1152-1154 the velocity update and the valid-face flags.
    hipLaunchKernelGGL(k_project<0>, grid3(n.x + 1, n.y, n.z), dim3(BX, BY), 0, s, u[0].p, flag[0].p, P.p, types.p, phi.p, n, g.dx);
    hipLaunchKernelGGL(k_project<1>, grid3(n.x, n.y + 1, n.z), dim3(BX, BY), 0, s, u[1].p, flag[1].p, P.p, types.p, phi.p, n, g.dx);
    hipLaunchKernelGGL(k_project<2>, grid3(n.x, n.y, n.z + 1), dim3(BX, BY), 0, s, u[2].p, flag[2].p, P.p, types.p, phi.p, n, g.dx);
    MAEL_LAUNCH_CHECK();
    st.sync();
    stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace mael
