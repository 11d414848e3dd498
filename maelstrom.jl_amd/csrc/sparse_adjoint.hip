// Adjoints of the sparse basin's transfers (no counterpart in the reference, which has no
// AD): the transposes of particles2grid! (sflip.jl:808-831) and grid2particles!
// (sflip.jl:2141-2144) at the state the forward phase left. With the particles'
// positions and lags, the level set, the topology, the deposit weights and the valid-face
// flags fixed, P2G is linear in the particles' v and APIC rows and in u0 (fine pools and
// coarse grid), and G2P is linear in the fine face velocities it reads. The G2P adjoint
// is a deposit (the tiled fixed-point scatter, deterministic), the P2G adjoint a gather.
// Cotangents live on every stored copy: a face on the plane between two fine blocks has
// one in each block, as the forward phases read and write each copy.
#include "sparse.hpp"
#include "scan.hpp"
#include "p2g_tiled.hpp"
#include <cmath>
#include <cstring>

namespace mael {

// shared with the dense basin's adjoints (transfer.hip)
__global__ void k_adj_fromfix(float* __restrict__ ub, const unsigned long long* __restrict__ acc, size_t n, double inv);
__global__ void k_g2p_adj_post(PB12 pb, const uint8_t* __restrict__ reached, int64_t n);

// The flag sum of G2P's gather at (px, py, pz) for component a: the terms of
// gather_component in its order, so the same bits as the forward's test.
__device__ inline float flaggather(const STopo& t, const uint8_t* __restrict__ F, int a, float px, float py, float pz) {
    const SGrid& g = t.g;
    float q0 = ispace(px, g.lo[0], g.dx, a == 0 ? 0.0f : 0.5f);
    float q1 = ispace(py, g.lo[1], g.dx, a == 1 ? 0.0f : 0.5f);
    float q2 = ispace(pz, g.lo[2], g.dx, a == 2 ? 0.0f : 0.5f);
    BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
    float flag = 0.0f;
    for (int mz = 0; mz < 3; ++mz) {
        float wz = pick(bz, mz);
        for (int my = 0; my < 3; ++my) {
            float wy = pick(by, my);
            for (int mx = 0; mx < 3; ++mx) {
                int32_t idx = finefaceat(t, a, bx.b + mx, by.b + my, bz.b + mz);
                if (idx >= 0) {
                    float wgt = pick(bx, mx) * wy * wz;
                    flag += wgt * (float)F[idx];
                }
            }
        }
    }
    return flag;
}

// Which particles took the grid's velocity in G2P (all three flag sums > 0.999,
// flip.jl:1239), and the largest contribution one of them makes to a node, as float bits.
__global__ __launch_bounds__(256) void k_sg2p_adj_reached(PView ps, int64_t n, STopo t, const uint8_t* __restrict__ fu,
                                                          const uint8_t* __restrict__ fv, const uint8_t* __restrict__ fw,
                                                          uint8_t* __restrict__ reached, PB12 pb, int* __restrict__ bound) {
    __shared__ int red[4];
    int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float m = 0.0f;
    if (p < n) {
        float xp = ps.f[0][p], yp = ps.f[1][p], zp = ps.f[2][p];
        bool k = flaggather(t, fu, 0, xp, yp, zp) > 0.999f && flaggather(t, fv, 1, xp, yp, zp) > 0.999f &&
                 flaggather(t, fw, 2, xp, yp, zp) > 0.999f;
        reached[p] = k ? 1 : 0;
        if (k) {
            float cd = 6.0f / t.g.dx;   // |row . d| <= (4 / dx^2)(1.5 dx) sum |row|
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

// The transpose of k_restrictfaces (sflip.jl:604-612) on cotangents: a coarse face
// restricted from the 16 fine faces under it (the same faces, found the same way) passes
// a sixteenth of its cotangent to each of them and keeps none (cout = 0); any other coarse
// face keeps its own (cout = cin). cout may be cin.
__global__ void k_sadj_restrictT(float* __restrict__ fine, float* cout, const float* cin, STopo t, int a) {
    int n0 = t.g.nc(0) + (a == 0), n1 = t.g.nc(1) + (a == 1), n2 = t.g.nc(2) + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    const int bi = I >> 2, bj = J >> 2, bk = K >> 2;
    int32_t sb = inblocks(t.g, bi, bj, bk) ? t.slot[blockid(t.g, bi, bj, bk)] : 0;
    bool top = false;
    const float c = cin[l];
    if (sb == 0) {
        const int FA = 4 * (a == 0 ? I : (a == 1 ? J : K));
        const int li = bi - (a == 0), lj = bj - (a == 1), lk = bk - (a == 2);
        sb = (FA & 15) == 0 && inblocks(t.g, li, lj, lk) ? t.slot[blockid(t.g, li, lj, lk)] : 0;
        if (sb == 0) {
            cout[l] = c;
            return;
        }
        top = true;
    }
    cout[l] = 0.0f;
    const float v = c * 0.0625f;
    const int32_t base = (sb - 1) * SBFACES;
    for (int q = 0; q < 4; ++q)
        for (int p = 0; p < 4; ++p) {
            int fi = (4 * I + (a == 0 ? 0 : p)) & 15;
            int fj = (4 * J + (a == 0 ? p : (a == 1 ? 0 : q))) & 15;
            int fk = (4 * K + (a == 2 ? 0 : q)) & 15;
            if (top) { if (a == 0) fi = 16; else if (a == 1) fj = 16; else fk = 16; }
            fine[base + ffn(16, a, fi, fj, fk)] += v;   // distinct coarse faces own distinct fine faces
        }
}
void launch_restrictfaces_T(const SparseBasin& s, float* fine, float* cout, const float* cin, int a) {
    hipLaunchKernelGGL(k_sadj_restrictT, dim3(cdiv(s.nfaces(a), 256)), dim3(256), 0, s.st.s, fine, cout, cin, s.topo(), a);
    MAEL_LAUNCH_CHECK();
}

// The transpose of k_sparse_combine (sflip.jl:796-806) on every fine face copy: in the
// band the momentum sum's cotangent is sm / max(nmin, W) and u0's takes the rest of the
// blend; outside it u0's cotangent is the face's. In place on sm; free slots pass nothing.
__global__ void k_sadj_combineT(float* __restrict__ sm, float* __restrict__ ub0, const unsigned long long* __restrict__ wt,
                                const float* __restrict__ phi, const float* __restrict__ phic, STopo t, float r, float full,
                                int a, int32_t nf) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    SCell c = fineface(t, a, n);
    if (c.w.x < 0) {
        sm[n] = 0.0f;
        ub0[n] = 0.0f;
        return;
    }
    float p[3];
    facepos(t.g, a, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, t.g.dx, p);
    float phif = cellsample(t, phi, phic, p[0], p[1], p[2]);
    float x = sm[n];
    if (phif >= -r) {
        float W = fromfix(wt[n]), den = jmax(full, W);
        sm[n] = x / den;
        ub0[n] = x * (jmax(full - W, 0.0f) / den);
    } else {
        sm[n] = 0.0f;
        ub0[n] = x;
    }
}

// The transpose of k_syncplanes (sflip.jl:620-640): the lower block's copy of a shared
// plane took the upper block's deposit, so its cotangent goes back to the upper copy.
__global__ void k_sadj_syncT(float* __restrict__ sm, STopo t, int a, int nslots) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nslots * 256) return;
    int s = (n >> 8) + 1, f = n & 255, p = f & 15, q = f >> 4;
    int4 w = t.where[s - 1];
    if (w.x < 0) return;
    int b0 = w.x + (a == 0), b1 = w.y + (a == 1), b2 = w.z + (a == 2);
    if (!inblocks(t.g, b0, b1, b2)) return;
    int32_t s2 = t.slot[blockid(t.g, b0, b1, b2)];
    if (s2 == 0) return;
    int hi = a == 0 ? ffn(16, 0, 16, p, q) : (a == 1 ? ffn(16, 1, p, 16, q) : ffn(16, 2, p, q, 16));
    int lo = a == 0 ? ffn(16, 0, 0, p, q) : (a == 1 ? ffn(16, 1, p, 0, q) : ffn(16, 2, p, q, 0));
    size_t h = (size_t)(s - 1) * SBFACES + hi;
    sm[(size_t)(s2 - 1) * SBFACES + lo] += sm[h];
    sm[h] = 0.0f;
}

void SparseBasin::adj_alloc() {
    if (ub[0].fine.p) return;
    hipStream_t s = st.s;
    size_t NF = (size_t)blocks * SBFACES;
    for (int a = 0; a < 3; ++a) {
        ub[a].alloc(NF, nfaces(a));
        ub0[a].alloc(NF, nfaces(a));
        adjacc[a].alloc(NF);
        fill_async(ub[a].fine.p, NF, 0.0f, s);
        fill_async(ub[a].coarse.p, nfaces(a), 0.0f, s);
        fill_async(ub0[a].fine.p, NF, 0.0f, s);
        fill_async(ub0[a].coarse.p, nfaces(a), 0.0f, s);
    }
    adjcell[0].alloc((size_t)blocks * SBCELLS);
    adjcell[1].alloc(ncells());
    for (int f = 0; f < 12; ++f) {
        pb[f].alloc(capacity);
        fill_async(pb[f].p, (size_t)capacity, 0.0f, s);
    }
    reached.alloc(capacity);
    MAEL_LAUNCH_CHECK();
}

// The transpose of grid2particles! in the fine faces: ub = G2P^T (pb) (the coarse part 0:
// G2P reads fine faces only; with the reflected flow selected, the faces are ur's), and pb
// becomes the cotangent of the input particle velocities (v passes through where the grid
// was not reached, the rows never do).
void SparseBasin::g2p_adjoint(float dt) {
    (void)dt;
    adj_alloc();
    hipStream_t s = st.s;
    size_t F = finefaces();
    for (int a = 0; a < 3; ++a) {
        fill_async(adjacc[a].p, F, 0ull, s);
        fill_async(ub[a].coarse.p, nfaces(a), 0.0f, s);
    }
    if (n == 0) {
        for (int a = 0; a < 3; ++a) fill_async(ub[a].fine.p, F, 0.0f, s);
        return;
    }
    STopo t = topo();
    PB12 P;
    for (int f = 0; f < 12; ++f) P.f[f] = pb[f].p;
    int* bound = (int*)partials.p;
    fill_async(bound, (size_t)1, 0, s);
    hipLaunchKernelGGL(k_sg2p_adj_reached, dim3(cdiv(n, 256)), dim3(256), 0, s, view(cur), n, t, fl[0].fine.p, fl[1].fine.p,
                       fl[2].fine.p, reached.p, P, bound);
    MAEL_HIP(hipMemcpyAsync(st.pinned, bound, sizeof(int), hipMemcpyDeviceToHost, s));
    st.sync();
    float m;
    std::memcpy(&m, st.pinned, sizeof(float));
    // a node takes at most 27 cells' particles (2 nmin each): keep the sum within 2^30
    double lim = std::ldexp(1.0, 30) / (27.0 * 2.0 * std::max(nmin, 1)), scale = 1.0;
    if (m > 0.0f) scale = std::ldexp(1.0, std::min(60, std::max(-60, (int)std::floor(std::log2(lim / (double)m)))));
    G2PAdjSource src{{P.f[0], P.f[1], P.f[2], P.f[3], P.f[4], P.f[5], P.f[6], P.f[7], P.f[8], P.f[9], P.f[10], P.f[11]},
                     reached.p, 4.0f / (g.dx * g.dx)};
    P2GArgs A{{g.lo[0], g.lo[1], g.lo[2]}, g.dx, {adjacc[0].p, adjacc[1].p, adjacc[2].p}, {nullptr, nullptr, nullptr},
              (float)scale};
    hipLaunchKernelGGL((k_p2g_tiled<SparseNodes, G2PAdjSource>), dim3(cdiv(n, P2G_THREADS)), dim3(P2G_THREADS), 0, s, view(cur),
                       n, SparseNodes{t}, src, A);
    for (int a = 0; a < 3; ++a)
        if (F) hipLaunchKernelGGL(k_adj_fromfix, dim3(cdiv(F, 256)), dim3(256), 0, s, ub[a].fine.p, adjacc[a].p, F, 1.0 / scale);
    hipLaunchKernelGGL(k_g2p_adj_post, dim3(cdiv(n, 256)), dim3(256), 0, s, P, reached.p, n);
    MAEL_LAUNCH_CHECK();
}

// The transpose of particles2grid! in the particles' v and APIC rows and in u0, at the
// positions, lags, level set and weights of the last deposit: pb += P2G^T (ub) for the
// particles (added to what G2P's adjoint passed through), ub0 = its part for u0. The
// forward's steps are transposed in reverse: the coarse faces' restriction, the blend of
// each fine face copy, the shared planes' copy, the deposit.
void SparseBasin::p2g_adjoint(float dt) {
    adj_alloc();
    hipStream_t s = st.s;
    size_t F = finefaces();
    STopo t = topo();
    float r = (float)(band - 1) * g.dx, full = (float)nmin;
    float* sm[3];
    for (int a = 0; a < 3; ++a) {
        sm[a] = reinterpret_cast<float*>(adjacc[a].p);   // scratch: the momentum sums' cotangents
        if (F) MAEL_HIP(hipMemcpyAsync(sm[a], ub[a].fine.p, F * sizeof(float), hipMemcpyDeviceToDevice, s));
        launch_restrictfaces_T(*this, sm[a], ub0[a].coarse.p, ub[a].coarse.p, a);
        if (F) {
            hipLaunchKernelGGL(k_sadj_combineT, dim3(cdiv(F, 256)), dim3(256), 0, s, sm[a], ub0[a].fine.p, wt[a].p, phi.fine.p,
                               phi.coarse.p, t, r, full, a, (int32_t)F);
            hipLaunchKernelGGL(k_sadj_syncT, dim3(cdiv(nslots * 256, 256)), dim3(256), 0, s, sm[a], t, a, nslots);
        }
    }
    if (n > 0) {
        PB12 P;
        for (int f = 0; f < 12; ++f) P.f[f] = pb[f].p;
        hipLaunchKernelGGL(k_p2g_adj_gather<SparseNodes>, dim3(cdiv(n, 256)), dim3(256), 0, s, view(cur), n, SparseNodes{t},
                           g.lo[0], g.lo[1], g.lo[2], g.dx, dt, sm[0], sm[1], sm[2], P);
    }
    MAEL_LAUNCH_CHECK();
}

}  // namespace mael
