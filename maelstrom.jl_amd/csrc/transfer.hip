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
#include <cmath>
#include <cstring>

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
        P2GArgs A{{g.lo[0], g.lo[1], g.lo[2]}, g.dx, {mom[0].p, mom[1].p, mom[2].p}, {wt[0].p, wt[1].p, wt[2].p}, 1.0f};
        hipLaunchKernelGGL((k_p2g_tiled<DenseNodes, P2GSource, P2G_BOX_DENSE>), dim3(cdiv(n, P2G_THREADS)), dim3(P2G_THREADS), 0,
                           s, ps, n,
                           map, P2GSource{ps, dt}, A);
        kt.end(K_P2G, s);
        kt.begin(K_SPLAT, s);
        hipLaunchKernelGGL((k_splat_tiled<DenseCells, SPLAT_BOX_DENSE>), dim3(cdiv(n, P2G_THREADS)), dim3(P2G_THREADS), 0, s, ps, n,
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

// flip.jl:1226-1244 g2p_kernel!, tiled: per component, a workgroup's 256 particles
// (cell-sorted, so close together) gather the flags and the velocity from a box of
// (value, flag) staged in LDS, one read of each face per workgroup instead of two per
// particle and node; a workgroup spread over more nodes than the box holds, or a particle
// outside it, reads the faces directly. Both paths take the flags' sum (gather_value) and
// the APIC sums (gather) in their order of operations, with the indices clamped to the
// array as there, so the result is theirs bit for bit.
constexpr int DG2P_BOX = 4096;   // 32 KB: the cell order runs along whole rows, two runs of a row far apart
__global__ __launch_bounds__(256) void k_dense_g2p(PView ps, int64_t n, DGrid g, const float* __restrict__ u,
                                                   const float* __restrict__ v, const float* __restrict__ w,
                                                   const uint8_t* __restrict__ fu, const uint8_t* __restrict__ fv,
                                                   const uint8_t* __restrict__ fw, float dt, float gravity) {
    __shared__ float2 sbox[DG2P_BOX];   // value, flag
    __shared__ int red[6 * 4];
    const int64_t p = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;   // neighbouring tiles on one XCD
    const bool live = p < n;
    float xp = 0, yp = 0, zp = 0;
    if (live) { xp = ps.f[0][p]; yp = ps.f[1][p]; zp = ps.f[2][p]; }
    const float* U[3] = {u, v, w};
    const uint8_t* F[3] = {fu, fv, fw};
    float fl[3], val[3], c0[3], c1[3], c2[3];
    for (int a = 0; a < 3; ++a) {
        const int n1 = g.n[0] + (a == 0), n2 = g.n[1] + (a == 1), n3 = g.n[2] + (a == 2);
        const float q0 = ispace(xp, g.lo[0], g.dx, a == 0 ? 0.0f : 0.5f);
        const float q1 = ispace(yp, g.lo[1], g.dx, a == 1 ? 0.0f : 0.5f);
        const float q2 = ispace(zp, g.lo[2], g.dx, a == 2 ? 0.0f : 0.5f);
        const BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
        const bool ok = live && bx.b > -(1 << 24) && bx.b < (1 << 24) && by.b > -(1 << 24) && by.b < (1 << 24) &&
                        bz.b > -(1 << 24) && bz.b < (1 << 24) && isfinite(q0) && isfinite(q1) && isfinite(q2);
        const int BIG = 1 << 30;
        int x0, y0, z0, x1, y1, z1;
        wg_bounds(ok, bx.b, by.b, bz.b, red, x0, y0, z0, x1, y1, z1);
        const int ex = x1 - x0 + 3, ey = y1 - y0 + 3, ez = z1 - z0 + 3;
        const bool boxed = x0 != BIG && ex <= DG2P_BOX && ey <= DG2P_BOX && ez <= DG2P_BOX && ex * ey * ez <= DG2P_BOX;
        const float* Ua = U[a];
        const uint8_t* Fa = F[a];
        if (boxed) {
            const int V = ex * ey * ez;
            for (int k = threadIdx.x; k < V; k += 256) {
                const int i = min(max(x0 + k % ex, 0), n1 - 1), j = min(max(y0 + (k / ex) % ey, 0), n2 - 1),
                          kk = min(max(z0 + k / (ex * ey), 0), n3 - 1);
                const size_t idx = (size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * kk);
                sbox[k] = make_float2(Ua[idx], (float)Fa[idx]);
            }
            __syncthreads();
        }
        if (live) {
            if (boxed && ok) {
                const float wx[3] = {bx.w0, bx.w1, bx.w2}, wy[3] = {by.w0, by.w1, by.w2}, wz[3] = {bz.w0, bz.w1, bz.w2};
                float d0[3], d1[3], d2[3];
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    d0[m] = (float)(bx.b + m) - q0;
                    d1[m] = (float)(by.b + m) - q1;
                    d2[m] = (float)(bz.b + m) - q2;
                }
                const float2* bp = sbox + ((bx.b - x0) + ex * ((by.b - y0) + ey * (bz.b - z0)));
                float f_ = 0.0f, v_ = 0.0f, s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
                for (int mz = 0; mz < 3; ++mz)
#pragma unroll
                    for (int my = 0; my < 3; ++my) {
                        const float2* row = bp + ex * (my + ey * mz);
#pragma unroll
                        for (int mx = 0; mx < 3; ++mx) {
                            const float2 e = row[mx];
                            const float wgt = (wx[mx] * wy[my]) * wz[mz];
                            f_ += wgt * e.y;
                            const float t = wgt * e.x;
                            v_ += t;
                            s0 += (t * d0[mx]) * g.dx;
                            s1 += (t * d1[my]) * g.dx;
                            s2 += (t * d2[mz]) * g.dx;
                        }
                    }
                const float c = __fdiv_rn(4.0f, g.dx * g.dx);
                fl[a] = f_; val[a] = v_; c0[a] = s0 * c; c1[a] = s1 * c; c2[a] = s2 * c;
            } else {
                fl[a] = gather_value(Fa, n1, n2, n3, g, a, xp, yp, zp);
                gather(Ua, n1, n2, n3, g, a, xp, yp, zp, val[a], c0[a], c1[a], c2[a]);
            }
        }
        if (boxed) __syncthreads();   // the box is reused by the next component
    }
    if (!live) return;
    // v and c (fields 3..14) built in registers and stored unconditionally, so no store's
    // target depends on the divergent branch (see k_dense_spawn).
    float rec[12];
    if (fl[0] > 0.999f && fl[1] > 0.999f && fl[2] > 0.999f) {
        rec[0] = val[0]; rec[3] = c0[0]; rec[4] = c1[0]; rec[5] = c2[0];
        rec[1] = val[1]; rec[6] = c0[1]; rec[7] = c1[1]; rec[8] = c2[1];
        rec[2] = val[2]; rec[9] = c0[2]; rec[10] = c1[2]; rec[11] = c2[2];
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

namespace mael {

// ─── adjoints of the transfers (no counterpart in the reference) ─────────────
// With the particles' positions and time lags and the level set fixed, P2G is linear in
// the particles' velocities and APIC rows (and in u0), and G2P is linear in the face
// velocities; these are their transposes, for reverse-mode gradients. The G2P adjoint is
// a deposit (the tiled fixed-point scatter, so it is deterministic), the P2G adjoint a
// gather.

// The gather's face nodes: indices clamped to the array (flip.jl:1210-1224).
struct DenseClampNodes {
    int n[3];
    __device__ int64_t index(int a, int I, int J, int K) const {
        int n1 = n[0] + (a == 0), n2 = n[1] + (a == 1), n3 = n[2] + (a == 2);
        I = min(max(I, 0), n1 - 1); J = min(max(J, 0), n2 - 1); K = min(max(K, 0), n3 - 1);
        return (int64_t)I + (int64_t)n1 * ((int64_t)J + (int64_t)n2 * K);
    }
};


// Which particles took the grid's velocity in G2P (the flags' gathers, as k_dense_g2p),
// and the largest contribution one of them makes to a node, as float bits (for the scale).
__global__ __launch_bounds__(256) void k_g2p_adj_reached(PView ps, int64_t n, DGrid g, const uint8_t* __restrict__ fu,
                                                         const uint8_t* __restrict__ fv, const uint8_t* __restrict__ fw,
                                                         uint8_t* __restrict__ reached, PB12 pb, int* __restrict__ bound) {
    __shared__ int red[4];
    int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float m = 0.0f;
    if (p < n) {
        float xp = ps.f[0][p], yp = ps.f[1][p], zp = ps.f[2][p];
        int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2];
        bool k = gather_value(fu, n1 + 1, n2, n3, g, 0, xp, yp, zp) > 0.999f &&
                 gather_value(fv, n1, n2 + 1, n3, g, 1, xp, yp, zp) > 0.999f &&
                 gather_value(fw, n1, n2, n3 + 1, g, 2, xp, yp, zp) > 0.999f;
        reached[p] = k ? 1 : 0;
        if (k) {
            float cd = 6.0f / g.dx;   // |row . d| <= (4 / dx^2)(1.5 dx) sum |row|
            for (int a = 0; a < 3; ++a) {
                float c = fabsf(pb.f[a][p]) + cd * (fabsf(pb.f[3 + 3 * a][p]) + fabsf(pb.f[4 + 3 * a][p]) + fabsf(pb.f[5 + 3 * a][p]));
                m = fmaxf(m, c);
            }
            if (!(m <= 3.0e38f)) m = 3.0e38f;
        }
    }
    int v = __float_as_int(m);   // non-negative: integer order is float order
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(bound, max(max(red[0], red[1]), max(red[2], red[3])));
}
// u-bar = the accumulated deposit / scale.
__global__ void k_adj_fromfix(float* __restrict__ ub, const unsigned long long* __restrict__ acc, size_t n, double inv) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) ub[l] = (float)((double)(long long)acc[l] * FIX_INV * inv);
}
// The particles' input cotangents: v passes through G2P where the grid was not reached
// (flip.jl:1239-1242), the rows never do.
__global__ void k_g2p_adj_post(PB12 pb, const uint8_t* __restrict__ reached, int64_t n) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    bool k = reached[p] != 0;
    for (int f = 0; f < 12; ++f)
        if (k || f >= 3) pb.f[f][p] = 0.0f;
}

void DenseBasin::adj_alloc() {
    if (ub[0].p) return;
    hipStream_t s = st.s;
    for (int a = 0; a < 3; ++a) {
        ub[a].alloc(faces(a)); ub0[a].alloc(faces(a)); adjacc[a].alloc(faces(a));
        fill_async(ub[a].p, faces(a), 0.0f, s);
        fill_async(ub0[a].p, faces(a), 0.0f, s);
    }
    for (int f = 0; f < 12; ++f) {
        pb[f].alloc(capacity);
        fill_async(pb[f].p, (size_t)capacity, 0.0f, s);
    }
    reached.alloc(capacity);
    MAEL_LAUNCH_CHECK();
}

// The transpose of grid2particles! (flip.jl:1226-1252) in u: ub = G2P^T (pb); pb becomes
// the cotangent of the input particle velocities.
void DenseBasin::g2p_adjoint(float dt) {
    (void)dt;
    adj_alloc();
    hipStream_t s = st.s;
    for (int a = 0; a < 3; ++a) fill_async(adjacc[a].p, faces(a), 0ull, s);
    if (n == 0) {
        for (int a = 0; a < 3; ++a) fill_async(ub[a].p, faces(a), 0.0f, s);
        return;
    }
    PB12 P;
    for (int f = 0; f < 12; ++f) P.f[f] = pb[f].p;
    int* bound = (int*)partials.p;
    fill_async(bound, (size_t)1, 0, s);
    hipLaunchKernelGGL(k_g2p_adj_reached, dim3(cdiv(n, 256)), dim3(256), 0, s, view(cur), n, g, flag[0].p, flag[1].p,
                       flag[2].p, reached.p, P, bound);
    int bits = 0;
    MAEL_HIP(hipMemcpyAsync(st.pinned, bound, sizeof(int), hipMemcpyDeviceToHost, s));
    st.sync();
    std::memcpy(&bits, st.pinned, sizeof(int));
    float m;
    std::memcpy(&m, &bits, sizeof(float));
    // a node takes at most 27 cells' particles (2 nmin each): keep the sum within 2^30
    double lim = std::ldexp(1.0, 30) / (27.0 * 2.0 * std::max(nmin, 1)), scale = 1.0;
    if (m > 0.0f) scale = std::ldexp(1.0, std::min(60, std::max(-60, (int)std::floor(std::log2(lim / (double)m)))));
    G2PAdjSource src{{P.f[0], P.f[1], P.f[2], P.f[3], P.f[4], P.f[5], P.f[6], P.f[7], P.f[8], P.f[9], P.f[10], P.f[11]},
                     reached.p, 4.0f / (g.dx * g.dx)};
    P2GArgs A{{g.lo[0], g.lo[1], g.lo[2]}, g.dx, {adjacc[0].p, adjacc[1].p, adjacc[2].p}, {nullptr, nullptr, nullptr},
              (float)scale};
    hipLaunchKernelGGL((k_p2g_tiled<DenseClampNodes, G2PAdjSource, P2G_BOX_DENSE>), dim3(cdiv(n, P2G_THREADS)), dim3(P2G_THREADS),
                       0, s,
                       view(cur), n, DenseClampNodes{{g.n[0], g.n[1], g.n[2]}}, src, A);
    for (int a = 0; a < 3; ++a)
        hipLaunchKernelGGL(k_adj_fromfix, dim3(cdiv(faces(a), 256)), dim3(256), 0, s, ub[a].p, adjacc[a].p, faces(a),
                           1.0 / scale);
    hipLaunchKernelGGL(k_g2p_adj_post, dim3(cdiv(n, 256)), dim3(256), 0, s, P, reached.p, n);
    MAEL_LAUNCH_CHECK();
}

// The P2G adjoint's face part: the cotangent of a face's momentum sum (s = ub / max(nmin,
// W) in the band, else 0) and of u0 (flip.jl:567-585 combine, differentiated).
template <int A>
__global__ __launch_bounds__(256) void k_p2g_adj_faces(float* __restrict__ sface, float* __restrict__ ub0,
                                                       const float* __restrict__ ub, const unsigned long long* __restrict__ wt,
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
    float u = ub[idx];
    if (phif >= -r) {
        float W = fromfix(wt[idx]), den = jmax(full, W);
        sface[idx] = u / den;
        ub0[idx] = u * (jmax(full - W, 0.0f) / den);
    } else {
        sface[idx] = 0.0f;
        ub0[idx] = u;
    }
}
// The transpose of particles2grid! (flip.jl:587-606) in the particles' v and APIC rows and
// in u0, at the positions, lags, level set and weights of the last deposit: pb += P2G^T
// (ub) for the particles (added to what G2P's adjoint passed through), ub0 = its part for
// u0.
void DenseBasin::p2g_adjoint(float dt) {
    adj_alloc();
    hipStream_t s = st.s;
    float r = (float)(band - 1) * g.dx, full = (float)nmin;
    int3 nn = make_int3(g.n[0], g.n[1], g.n[2]);
    dim3 blk(64, 4, 1);
    float* sf[3];
    for (int a = 0; a < 3; ++a) sf[a] = reinterpret_cast<float*>(adjacc[a].p);   // scratch, faces(a) floats
    hipLaunchKernelGGL(k_p2g_adj_faces<0>, dim3(cdiv(g.n[0] + 1, 64), cdiv(g.n[1], 4), g.n[2]), blk, 0, s, sf[0], ub0[0].p,
                       ub[0].p, wt[0].p, phi.p, nn, r, full);
    hipLaunchKernelGGL(k_p2g_adj_faces<1>, dim3(cdiv(g.n[0], 64), cdiv(g.n[1] + 1, 4), g.n[2]), blk, 0, s, sf[1], ub0[1].p,
                       ub[1].p, wt[1].p, phi.p, nn, r, full);
    hipLaunchKernelGGL(k_p2g_adj_faces<2>, dim3(cdiv(g.n[0], 64), cdiv(g.n[1], 4), g.n[2] + 1), blk, 0, s, sf[2], ub0[2].p,
                       ub[2].p, wt[2].p, phi.p, nn, r, full);
    if (n > 0) {
        PB12 P;
        for (int f = 0; f < 12; ++f) P.f[f] = pb[f].p;
        hipLaunchKernelGGL(k_p2g_adj_gather<DenseNodes>, dim3(cdiv(n, 256)), dim3(256), 0, s, view(cur), n,
                           DenseNodes{{g.n[0], g.n[1], g.n[2]}}, g.lo[0], g.lo[1], g.lo[2], g.dx, dt, sf[0], sf[1], sf[2], P);
    }
    MAEL_LAUNCH_CHECK();
}

}  // namespace mael
