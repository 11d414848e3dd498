// APIC transfers on the dense grid: particles to grid (flip.jl:501-606) and grid to
// particles (flip.jl:1208-1252).
//
// The deposit is deterministic: each contribution is rounded once to 64-bit fixed point
// (2^-32) and added with an integer atomic. Integer addition is associative, so the
// accumulated momentum and weight are the same whatever order the particles arrive in,
// unlike the reference's float atomics (flip.jl:532-533). The level-set splat keeps the
// reference's integer atomic minimum over order keys, which is order-independent too.
#include "dense.hpp"
#include "scan.hpp"
#include "p2g_tiled.hpp"

namespace mael {

// flip.jl:513-537 p2g_kernel!: the tiled deposit over the dense grid's face arrays.
struct DenseNodes {
    int n[3];
    __device__ int64_t index(int a, int I, int J, int K) const {
        int n1 = n[0] + (a == 0), n2 = n[1] + (a == 1), n3 = n[2] + (a == 2);
        if (I < 0 || I >= n1 || J < 0 || J >= n2 || K < 0 || K >= n3) return -1;
        return (int64_t)I + (int64_t)n1 * ((int64_t)J + (int64_t)n2 * K);
    }
};

struct DenseCells {
    int n[3];
    __device__ int64_t index(int I, int J, int K) const {
        if (I < 0 || I >= n[0] || J < 0 || J >= n[1] || K < 0 || K >= n[2]) return -1;
        return (int64_t)I + (int64_t)n[0] * ((int64_t)J + (int64_t)n[1] * K);
    }
};

// flip.jl:562-565 union_kernel!.
__global__ void k_union(size_t C, float* __restrict__ phi, const float* __restrict__ phi0, const int32_t* __restrict__ phip,
                        float dx) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < C) phi[l] = jmin(phi0[l] + dx, orderfloat(phip[l]));
}

// flip.jl:567-585 combine_kernel! for axis A, from the fixed-point accumulators.
template <int A>
__global__ __launch_bounds__(256) void k_dense_combine(float* __restrict__ u, const float* __restrict__ u0,
                                                       const unsigned long long* __restrict__ mom,
                                                       const unsigned long long* __restrict__ wt,
                                                       const float* __restrict__ phi, int3 n, float r, float full) {
    int n1 = n.x + (A == 0), n2 = n.y + (A == 1), n3 = n.z + (A == 2);
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n1 || j >= n2 || k >= n3) return;
    int bi = i - (A == 0), bj = j - (A == 1), bk = k - (A == 2);
    auto at = [&](int a, int b, int c) { return phi[(size_t)a + (size_t)n.x * ((size_t)b + (size_t)n.y * c)]; };
    bool inb = bi >= 0 && bj >= 0 && bk >= 0;
    bool ina = i < n.x && j < n.y && k < n.z;
    float fa = inb ? at(bi, bj, bk) : at(min(i, n.x - 1), min(j, n.y - 1), min(k, n.z - 1));
    float fb = ina ? at(i, j, k) : fa;
    float phif = 0.5f * (fa + fb);
    size_t idx = (size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k);
    float wgt = fromfix(wt[idx]);
    float m = fromfix(mom[idx]);
    float v0 = u0[idx];
    u[idx] = phif >= -r ? __fdiv_rn(m + jmax(full - wgt, 0.0f) * v0, jmax(full, wgt)) : v0;
}

void DenseBasin::p2g(float dt) {
    hipStream_t s = st.s;
    for (int a = 0; a < 3; ++a) {
        fill_async(mom[a].p, faces(a), 0ull, s);
        fill_async(wt[a].p, faces(a), 0ull, s);
    }
    size_t C = cells();
    int32_t big;
    {
        float f = 1e30f;
        int32_t bits;
        std::memcpy(&bits, &f, 4);
        big = bits;   // orderkey of a positive float is its bit pattern
    }
    fill_async(phip.p, C, big, s);
    if (n > 0) {
        PView ps = view(cur);
        kt.begin(K_P2G, s);
        DenseNodes map{{g.n[0], g.n[1], g.n[2]}};
        P2GArgs A{{g.lo[0], g.lo[1], g.lo[2]}, g.dx, dt, {mom[0].p, mom[1].p, mom[2].p}, {wt[0].p, wt[1].p, wt[2].p}};
        hipLaunchKernelGGL(k_p2g_tiled<DenseNodes>, dim3(cdiv(n, P2G_THREADS)), dim3(P2G_THREADS), 0, s, ps, n, map, A);
        kt.end(K_P2G, s);
        kt.begin(K_SPLAT, s);
        hipLaunchKernelGGL(k_splat_tiled<DenseCells>, dim3(cdiv(n, P2G_THREADS)), dim3(P2G_THREADS), 0, s, ps, n,
                           DenseCells{{g.n[0], g.n[1], g.n[2]}}, g.lo[0], g.lo[1], g.lo[2], g.dx, dt, phip.p);
        kt.end(K_SPLAT, s);
        MAEL_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_union, dim3(cdiv(C, 256)), dim3(256), 0, s, C, phi.p, phi0.p, phip.p, g.dx);
    float r = (float)(band - 1) * g.dx, full = (float)nmin;
    int3 nn = make_int3(g.n[0], g.n[1], g.n[2]);
    dim3 blk(64, 4, 1);
    kt.begin(K_COMBINE, s);
    hipLaunchKernelGGL(k_dense_combine<0>, dim3(cdiv(g.n[0] + 1, 64), cdiv(g.n[1], 4), g.n[2]), blk, 0, s, u[0].p, u0[0].p,
                       mom[0].p, wt[0].p, phi.p, nn, r, full);
    hipLaunchKernelGGL(k_dense_combine<1>, dim3(cdiv(g.n[0], 64), cdiv(g.n[1] + 1, 4), g.n[2]), blk, 0, s, u[1].p, u0[1].p,
                       mom[1].p, wt[1].p, phi.p, nn, r, full);
    hipLaunchKernelGGL(k_dense_combine<2>, dim3(cdiv(g.n[0], 64), cdiv(g.n[1], 4), g.n[2] + 1), blk, 0, s, u[2].p, u0[2].p,
                       mom[2].p, wt[2].p, phi.p, nn, r, full);
    kt.end(K_COMBINE, s);
    MAEL_LAUNCH_CHECK();
}

// The value part of the gather, for the flags (flip.jl:401-405 `gathered`).
template <class V>
__device__ inline float gather_value(const V* __restrict__ u, int n1, int n2, int n3, const DGrid& g, int a, float px,
                                     float py, float pz) {
    float q0 = ispace(px, g.lo[0], g.dx, a == 0 ? 0.0f : 0.5f);
    float q1 = ispace(py, g.lo[1], g.dx, a == 1 ? 0.0f : 0.5f);
    float q2 = ispace(pz, g.lo[2], g.dx, a == 2 ? 0.0f : 0.5f);
    BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
    float vel = 0.0f;
    size_t s1 = (size_t)n1, s2 = (size_t)n1 * n2;
#pragma unroll
    for (int mz = 0; mz < 3; ++mz) {
        int k = min(max(bz.b + mz, 0), n3 - 1);
        float wz = pick(bz, mz);
#pragma unroll
        for (int my = 0; my < 3; ++my) {
            int j = min(max(by.b + my, 0), n2 - 1);
            float wy = pick(by, my);
#pragma unroll
            for (int mx = 0; mx < 3; ++mx) {
                int i = min(max(bx.b + mx, 0), n1 - 1);
                vel += pick(bx, mx) * wy * wz * (float)u[(size_t)i + s1 * j + s2 * k];
            }
        }
    }
    return vel;
}

// flip.jl:1226-1244 g2p_kernel!.
__global__ __launch_bounds__(256) void k_dense_g2p(PView ps, int64_t n, DGrid g, const float* __restrict__ u,
                                                   const float* __restrict__ v, const float* __restrict__ w,
                                                   const uint8_t* __restrict__ fu, const uint8_t* __restrict__ fv,
                                                   const uint8_t* __restrict__ fw, float dt, float gravity) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float xp = ps.f[0][p], yp = ps.f[1][p], zp = ps.f[2][p];
    int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2];
    bool k1 = gather_value(fu, n1 + 1, n2, n3, g, 0, xp, yp, zp) > 0.999f;
    bool k2 = gather_value(fv, n1, n2 + 1, n3, g, 1, xp, yp, zp) > 0.999f;
    bool k3 = gather_value(fw, n1, n2, n3 + 1, g, 2, xp, yp, zp) > 0.999f;
    // v and c (fields 3..14) built in registers and stored unconditionally, so no store's
    // target depends on the divergent branch (see k_dense_spawn).
    float rec[12];
    if (k1 && k2 && k3) {
        gather(u, n1 + 1, n2, n3, g, 0, xp, yp, zp, rec[0], rec[3], rec[4], rec[5]);
        gather(v, n1, n2 + 1, n3, g, 1, xp, yp, zp, rec[1], rec[6], rec[7], rec[8]);
        gather(w, n1, n2, n3 + 1, g, 2, xp, yp, zp, rec[2], rec[9], rec[10], rec[11]);
    } else {
        rec[0] = ps.f[3][p];
        rec[1] = ps.f[4][p];
        rec[2] = ps.f[5][p] - gravity * dt;
#pragma unroll
        for (int f = 3; f < 12; ++f) rec[f] = 0.0f;
    }
#pragma unroll
    for (int f = 0; f < 12; ++f) ps.f[3 + f][p] = rec[f];
}

void DenseBasin::g2p(float dt, float grav) {
    if (n == 0) return;
    kt.begin(K_G2P, st.s);
    hipLaunchKernelGGL(k_dense_g2p, dim3(cdiv(n, 256)), dim3(256), 0, st.s, view(cur), n, g, u[0].p, u[1].p, u[2].p,
                       flag[0].p, flag[1].p, flag[2].p, dt, grav);
    kt.end(K_G2P, st.s);
    MAEL_LAUNCH_CHECK();
}

}  // namespace mael
