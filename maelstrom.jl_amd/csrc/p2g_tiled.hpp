// The APIC deposit (flip.jl:513-537) as a workgroup-tiled reduction.
//
// A workgroup takes 256 consecutive particles (resampling leaves them in cell order, so
// they are close in space) and, one velocity component at a time, accumulates their
// 27-node contributions into a box of nodes in LDS with 64-bit integer adds, then adds
// each touched node of the box to HBM once. Contributions are rounded to 2^-32 fixed
// point exactly as the per-particle deposit does and integer addition is associative,
// so the result is bit-identical to depositing every contribution with its own global
// atomic, in any order. A workgroup whose particles spread over more nodes than the box
// holds deposits directly.
#pragma once
#include "common.hpp"

namespace mael {

constexpr int P2G_THREADS = 256;
constexpr int P2G_BOX = 1024;   // nodes per component in LDS (mom + wt: 16 KB)

__device__ inline int wg_min(int v, int* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = red[0];
#pragma unroll
    for (int w = 1; w < P2G_THREADS / 64; ++w) r = min(r, red[w]);
    return r;
}
__device__ inline int wg_max(int v, int* red) { return -wg_min(-v, red); }

struct P2GArgs {
    float lo[3];
    float dx;
    float dt;
    unsigned long long* mom[3];
    unsigned long long* wt[3];
};

// NodeMap: `int32_t/size_t index(int a, int I, int J, int K)` of a face node, < 0 off the
// grid (the dense grid's array bounds, or no fine block holding the face).
template <class NodeMap>
__global__ __launch_bounds__(P2G_THREADS) void k_p2g_tiled(PView ps, int64_t n, NodeMap map, P2GArgs A) {
    __shared__ unsigned long long smom[P2G_BOX], swt[P2G_BOX];
    __shared__ int red[P2G_THREADS / 64];
    int64_t p = (int64_t)blockIdx.x * P2G_THREADS + threadIdx.x;
    bool live = p < n;
    float xp = 0, yp = 0, zp = 0, wT = 0;
    if (live) {
        xp = ps.f[0][p]; yp = ps.f[1][p]; zp = ps.f[2][p];
        wT = temporal(ps.f[15][p], A.dt);
    }
    for (int a = 0; a < 3; ++a) {
        float q0 = ispace(xp, A.lo[0], A.dx, a == 0 ? 0.0f : 0.5f);
        float q1 = ispace(yp, A.lo[1], A.dx, a == 1 ? 0.0f : 0.5f);
        float q2 = ispace(zp, A.lo[2], A.dx, a == 2 ? 0.0f : 0.5f);
        BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
        // finite, in-range bases only shape the box; others take the direct path
        bool ok = live && bx.b > -(1 << 24) && bx.b < (1 << 24) && by.b > -(1 << 24) && by.b < (1 << 24) &&
                  bz.b > -(1 << 24) && bz.b < (1 << 24) && isfinite(q0) && isfinite(q1) && isfinite(q2);
        const int BIG = 1 << 30;
        int x0 = wg_min(ok ? bx.b : BIG, red), y0 = wg_min(ok ? by.b : BIG, red), z0 = wg_min(ok ? bz.b : BIG, red);
        int x1 = wg_max(ok ? bx.b : -BIG, red), y1 = wg_max(ok ? by.b : -BIG, red), z1 = wg_max(ok ? bz.b : -BIG, red);
        int ex = x1 - x0 + 3, ey = y1 - y0 + 3, ez = z1 - z0 + 3;
        bool boxed = x0 != BIG && ex <= P2G_BOX && ey <= P2G_BOX && ez <= P2G_BOX && ex * ey * ez <= P2G_BOX;
        float vp = 0, c0 = 0, c1 = 0, c2 = 0;
        if (live) {
            vp = ps.f[3 + a][p];
            c0 = ps.f[6 + 3 * a][p]; c1 = ps.f[7 + 3 * a][p]; c2 = ps.f[8 + 3 * a][p];
        }
        if (boxed) {
            int V = ex * ey * ez;
            for (int k = threadIdx.x; k < V; k += P2G_THREADS) { smom[k] = 0ull; swt[k] = 0ull; }
            __syncthreads();
        }
        if (live) {
            for (int mz = 0; mz < 3; ++mz) {
                int K = bz.b + mz;
                float wz = pick(bz, mz), d2 = ((float)K - q2) * A.dx;
                for (int my = 0; my < 3; ++my) {
                    int J = by.b + my;
                    float wy = pick(by, my), d1 = ((float)J - q1) * A.dx;
#pragma unroll
                    for (int mx = 0; mx < 3; ++mx) {
                        int I = bx.b + mx;
                        float wgt = wT * pick(bx, mx) * wy * wz;
                        float d0 = ((float)I - q0) * A.dx;
                        unsigned long long m = tofix(wgt * (vp + (c0 * d0 + c1 * d1 + c2 * d2))), w = tofix(wgt);
                        if (boxed && ok) {
                            int l = (I - x0) + ex * ((J - y0) + ey * (K - z0));
                            atomicAdd(&smom[l], m);
                            atomicAdd(&swt[l], w);
                        } else {
                            auto idx = map.index(a, I, J, K);
                            if (idx >= 0) {
                                atomicAdd(&A.mom[a][idx], m);
                                atomicAdd(&A.wt[a][idx], w);
                            }
                        }
                    }
                }
            }
        }
        if (boxed) {
            __syncthreads();
            int V = ex * ey * ez;
            for (int k = threadIdx.x; k < V; k += P2G_THREADS) {
                unsigned long long m = smom[k], w = swt[k];
                if (m == 0ull && w == 0ull) continue;
                int I = x0 + k % ex, J = y0 + (k / ex) % ey, K = z0 + k / (ex * ey);
                auto idx = map.index(a, I, J, K);
                if (idx >= 0) {
                    if (m) atomicAdd(&A.mom[a][idx], m);
                    if (w) atomicAdd(&A.wt[a][idx], w);
                }
            }
            __syncthreads();   // the box is reused by the next component
        }
    }
}

// The particles' level-set splat (flip.jl:539-560) tiled the same way: a workgroup's 256
// particles take the minimum of their order keys over a box of cells in LDS, then each
// touched cell of the box takes one global atomic min. Minimum is order-independent, so
// the result is the direct splat's bit for bit. CellMap: `index(I, J, K)` of a cell (< 0
// off the grid or outside the fine blocks).
constexpr int SPLAT_BOX = 2048;   // cells in LDS (8 KB)

template <class CellMap>
__global__ __launch_bounds__(P2G_THREADS) void k_splat_tiled(PView ps, int64_t n, CellMap map, float lo0, float lo1, float lo2,
                                                            float dx, float dt, int32_t* __restrict__ phip) {
    __shared__ int32_t box[SPLAT_BOX];
    __shared__ int red[P2G_THREADS / 64];
    int64_t p = (int64_t)blockIdx.x * P2G_THREADS + threadIdx.x;
    bool live = p < n;
    float xp = 0, yp = 0, zp = 0;
    int ci = 0, cj = 0, ck = 0;
    bool ok = false;
    if (live) {
        float vx = ps.f[3][p], vy = ps.f[4][p], vz = ps.f[5][p];
        float sp = sqrtf(vx * vx + vy * vy + vz * vz);
        float s = ps.f[15][p] + ahead(strength(sp, dt, dx)) * dt;
        xp = ps.f[0][p] + s * vx; yp = ps.f[1][p] + s * vy; zp = ps.f[2][p] + s * vz;
        float f0 = floorf(__fdiv_rn(xp - lo0, dx)), f1 = floorf(__fdiv_rn(yp - lo1, dx)), f2 = floorf(__fdiv_rn(zp - lo2, dx));
        ci = (int)f0; cj = (int)f1; ck = (int)f2;
        // finite, in-range cells only shape the box; others take the direct path
        ok = fabsf(f0) < 16777216.0f && fabsf(f1) < 16777216.0f && fabsf(f2) < 16777216.0f;
    }
    const int BIG = 1 << 30;
    int x0 = wg_min(ok ? ci : BIG, red), y0 = wg_min(ok ? cj : BIG, red), z0 = wg_min(ok ? ck : BIG, red);
    int x1 = wg_max(ok ? ci : -BIG, red), y1 = wg_max(ok ? cj : -BIG, red), z1 = wg_max(ok ? ck : -BIG, red);
    int ex = x1 - x0 + 3, ey = y1 - y0 + 3, ez = z1 - z0 + 3;
    bool boxed = x0 != BIG && ex <= SPLAT_BOX && ey <= SPLAT_BOX && ez <= SPLAT_BOX && ex * ey * ez <= SPLAT_BOX;
    x0 -= 1; y0 -= 1; z0 -= 1;
    if (boxed) {
        int V = ex * ey * ez;
        for (int k = threadIdx.x; k < V; k += P2G_THREADS) box[k] = 0x7fffffff;
        __syncthreads();
    }
    if (live) {
        float radius = 0.5f * dx;
        for (int dk = -1; dk <= 1; ++dk)
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    int I = ci + di, J = cj + dj, K = ck + dk;
                    float ex_ = lo0 + ((float)I + 0.5f) * dx - xp;
                    float ey_ = lo1 + ((float)J + 0.5f) * dx - yp;
                    float ez_ = lo2 + ((float)K + 0.5f) * dx - zp;
                    float d = sqrtf(ex_ * ex_ + ey_ * ey_ + ez_ * ez_) - radius;
                    if (boxed && ok) {
                        atomicMin(&box[(I - x0) + ex * ((J - y0) + ey * (K - z0))], orderkey(d));
                    } else {
                        auto idx = map.index(I, J, K);
                        if (idx >= 0) atomicMin(&phip[idx], orderkey(d));
                    }
                }
    }
    if (boxed) {
        __syncthreads();
        int V = ex * ey * ez;
        for (int k = threadIdx.x; k < V; k += P2G_THREADS) {
            int32_t v = box[k];
            if (v == 0x7fffffff) continue;
            auto idx = map.index(x0 + k % ex, y0 + (k / ex) % ey, z0 + k / (ex * ey));
            if (idx >= 0) atomicMin(&phip[idx], v);
        }
    }
}

}  // namespace mael
