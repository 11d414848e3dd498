// The particle phases of the sparse basin: resampling (sflip.jl:2146-2246), particles
// to grid (sflip.jl:788-831, the transfers of flip.jl:513-560 through the topology) and
// grid to particles (sflip.jl:2141-2144). Particles live in fine blocks only; keys are
// fine pool indices, so a resampled buffer is in the pool's order (slot by slot).
// Ranks are stable and deposits fixed point, as on the dense grid.
#include "sparse.hpp"
#include "scan.hpp"
#include "p2g_tiled.hpp"

namespace mael {

__global__ void k_place(int64_t n, const int32_t* __restrict__ key, const int32_t* __restrict__ rank,
                        const int32_t* __restrict__ cstart, int32_t* __restrict__ seg);
__global__ void k_fixrank(size_t C, const int32_t* __restrict__ counts, const int32_t* __restrict__ cstart,
                          int32_t* __restrict__ seg, int32_t* __restrict__ rank);
__global__ void k_scatter(PView src, PView dst, int64_t n, const int32_t* __restrict__ key, const int32_t* __restrict__ rank,
                          const int32_t* __restrict__ totals, const int32_t* __restrict__ offsets, int32_t nmax);

__device__ inline void scellcounts(int32_t count, float phic, bool solid, int32_t nmin, float deep, float dx, bool seeding,
                                   int32_t& kept, int32_t& add) {   // flip.jl:1289-1297
    kept = min(count, 2 * nmin);
    bool inband = !solid && -deep <= phic && phic <= (seeding ? 0.5f * dx : -dx);
    add = (inband && kept < nmin) ? nmin - kept : 0;
}

// What the scan needs about fine cell n (sflip.jl:2173-2182): free slots count nothing.
struct SparseCells {
    const int32_t* counts;
    const float* phi;
    STopo t;
    int32_t nmin;
    float deep;
    int seeding;
    __device__ void get(size_t n, int32_t& count, int32_t& total) const {
        SCell c = finecell(t, (int32_t)n);
        count = counts[n];
        total = 0;
        if (c.w.x < 0) return;
        float p[3];
        finecentre(t.g, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, p);
        int32_t kept, add;
        scellcounts(count, phi[n], sissolid(t.g, p[0], p[1], p[2]), nmin, deep, t.g.dx, seeding != 0, kept, add);
        total = kept + add;
    }
};

// sflip.jl:2148-2171 sparsebin_kernel!.
__global__ void k_sparse_bin(PView ps, int64_t n, const float* __restrict__ phi, const float* __restrict__ phic, STopo t,
                             float deep, float above, int32_t* __restrict__ key, int32_t* __restrict__ rank,
                             int32_t* __restrict__ counts) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const SGrid& g = t.g;
    float xp = ps.f[0][p], yp = ps.f[1][p], zp = ps.f[2][p];
    float c0 = __fdiv_rn(xp - g.lo[0], g.dx), c1 = __fdiv_rn(yp - g.lo[1], g.dx), c2 = __fdiv_rn(zp - g.lo[2], g.dx);
    int32_t idx = -1;
    if (isfinite(c0) && isfinite(c1) && isfinite(c2) && !sissolid(g, xp, yp, zp)) {
        // outside [0, 16 nb) no block holds the cell
        if (c0 >= 0.0f && c1 >= 0.0f && c2 >= 0.0f && c0 < (float)(16 * g.nb[0]) && c1 < (float)(16 * g.nb[1]) &&
            c2 < (float)(16 * g.nb[2]))
            idx = fineat(t, (int)floorf(c0), (int)floorf(c1), (int)floorf(c2));
        if (idx >= 0) {
            float f = cellsample(t, phi, phic, xp, yp, zp);
            if (!(-deep <= f && f < above)) idx = -1;
        }
    }
    key[p] = idx;
    rank[p] = idx >= 0 ? atomicAdd(&counts[idx], 1) : -1;
}

// sflip.jl:2184-2218 sparsespawn_kernel!.
__global__ void k_sparse_spawn(PView dst, const int32_t* __restrict__ counts, const int32_t* __restrict__ totals,
                               const int32_t* __restrict__ offsets, const float* __restrict__ phi, const float* __restrict__ phic,
                               const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w, STopo t,
                               int32_t nmin, float deep, int seeding, uint32_t seed, int32_t nf) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    SCell c = finecell(t, n);
    if (c.w.x < 0) return;
    const SGrid& g = t.g;
    int I = c.w.x * 16 + c.i, J = c.w.y * 16 + c.j, K = c.w.z * 16 + c.k;
    float pc[3];
    finecentre(g, I, J, K, pc);
    int32_t kept, add;
    scellcounts(counts[n], phi[n], sissolid(g, pc[0], pc[1], pc[2]), nmin, deep, g.dx, seeding != 0, kept, add);
    if (add <= 0) return;
    int64_t base = (int64_t)offsets[n] - totals[n] + kept;
    float cx = g.lo[0] + (float)I * g.dx, cy = g.lo[1] + (float)J * g.dx, cz = g.lo[2] + (float)K * g.dx;
    uint32_t n1 = (uint32_t)n + 1u;   // the reference's 1-based index
    int32_t sd = (int32_t)seed;
    float half = 0.5f * g.dx;
    for (int32_t m = 0; m < add; ++m) {
        int32_t h = (int32_t)(n1 * 31u + (uint32_t)m);
        float jx = hash01(h, sd + 1), jy = hash01(h, sd + 2), jz = hash01(h, sd + 3);
        int32_t o = m % 8;
        float xp = cx + ((float)(o & 1) + jx) * half;
        float yp = cy + ((float)((o >> 1) & 1) + jy) * half;
        float zp = cz + ((float)((o >> 2) & 1) + jz) * half;
        bool air = seeding && cellsample(t, phi, phic, xp, yp, zp) >= 0.0f;
        // built in registers, every field stored unconditionally (see k_dense_spawn)
        float rec[PFIELDS];
        if (air) {
            float nan = __builtin_nanf("");
            rec[0] = nan; rec[1] = nan; rec[2] = nan;
#pragma unroll
            for (int f = 3; f < PFIELDS; ++f) rec[f] = 0.0f;
        } else {
            rec[0] = xp; rec[1] = yp; rec[2] = zp;
            finegather(t, u, 0, xp, yp, zp, rec[3], rec[6], rec[7], rec[8]);
            finegather(t, v, 1, xp, yp, zp, rec[4], rec[9], rec[10], rec[11]);
            finegather(t, w, 2, xp, yp, zp, rec[5], rec[12], rec[13], rec[14]);
            rec[15] = 0.0f;
        }
        int64_t d = base + m;
#pragma unroll
        for (int f = 0; f < PFIELDS; ++f) dst.f[f][d] = rec[f];
    }
}

int64_t SparseBasin::resample(bool seeding, uint32_t seed) {
    hipStream_t s = st.s;
    size_t C = finecells();
    float deep = (float)band * g.dx;
    fill_async(counts.p, C, 0, s);
    PView src = view(cur), dst = view(cur ^ 1);
    STopo t = topo();
    if (n > 0) {
        kt.begin(K_BIN, s);
        hipLaunchKernelGGL(k_sparse_bin, dim3(cdiv(n, 256)), dim3(256), 0, s, src, n, phi.fine.p, phi.coarse.p, t, deep,
                           3.0f * g.dx, key.p, rank.p, counts.p);
        kt.end(K_BIN, s);
        MAEL_LAUNCH_CHECK();
    }
    SparseCells info{counts.p, phi.fine.p, t, nmin, deep, seeding ? 1 : 0};
    int64_t total = C ? scan_cells(info, C, totals.p, offsets.p, cstart.p, bsums.p, (unsigned long long*)st.pinned, s) : 0;
    if (total > capacity)
        throw Error(E_CAPACITY, "the narrow band needs " + std::to_string(total) + " particles, more than the " +
                                    std::to_string(capacity) + " allocated");
    if (n > 0) {
        hipLaunchKernelGGL(k_place, dim3(cdiv(n, 256)), dim3(256), 0, s, n, key.p, rank.p, cstart.p, seg.p);
        kt.begin(K_FIXRANK, s);
        hipLaunchKernelGGL(k_fixrank, dim3(cdiv(C, 256)), dim3(256), 0, s, C, counts.p, cstart.p, seg.p, rank.p);
        kt.end(K_FIXRANK, s);
        kt.begin(K_SCATTER, s);
        hipLaunchKernelGGL(k_scatter, dim3(cdiv(n, 256)), dim3(256), 0, s, src, dst, n, key.p, rank.p, totals.p, offsets.p,
                           2 * nmin);
        kt.end(K_SCATTER, s);
        MAEL_LAUNCH_CHECK();
    }
    if (C) {
        const SField<float>* fv = reflect_flow ? ur : u;   // the flow new particles take (sflip.jl:2312-2316)
        kt.begin(K_SPAWN, s);
        hipLaunchKernelGGL(k_sparse_spawn, dim3(cdiv(C, 256)), dim3(256), 0, s, dst, counts.p, totals.p, offsets.p, phi.fine.p,
                           phi.coarse.p, fv[0].fine.p, fv[1].fine.p, fv[2].fine.p, t, nmin, deep, seeding ? 1 : 0, seed,
                           (int32_t)C);
        kt.end(K_SPAWN, s);
        MAEL_LAUNCH_CHECK();
    }
    cur ^= 1;
    n = total;
    return total;
}

// ─── particles to grid ───────────────────────────────────────────────────────

// flip.jl:513-537 over the topology: nodes are fine faces (the upper block's copy on a
// plane between blocks), deposited by the tiled fixed-point reduction.
struct SparseNodes {
    STopo t;
    __device__ int32_t index(int a, int I, int J, int K) const { return finefaceat(t, a, I, J, K); }
};

// flip.jl:539-560 over the topology.
__global__ __launch_bounds__(256) void k_sparse_splat(PView ps, int64_t n, STopo t, float dt, int32_t* __restrict__ phip) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const SGrid& g = t.g;
    float vx = ps.f[3][p], vy = ps.f[4][p], vz = ps.f[5][p];
    float sp = sqrtf(vx * vx + vy * vy + vz * vz);
    float s = ps.f[15][p] + ahead(strength(sp, dt, g.dx)) * dt;
    float xp = ps.f[0][p] + s * vx, yp = ps.f[1][p] + s * vy, zp = ps.f[2][p] + s * vz;
    int ci = (int)floorf(__fdiv_rn(xp - g.lo[0], g.dx)), cj = (int)floorf(__fdiv_rn(yp - g.lo[1], g.dx)),
        ck = (int)floorf(__fdiv_rn(zp - g.lo[2], g.dx));
    float radius = 0.5f * g.dx;
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                int32_t idx = fineat(t, ci + di, cj + dj, ck + dk);
                if (idx < 0) continue;
                float c[3];
                finecentre(g, ci + di, cj + dj, ck + dk, c);
                float ex = c[0] - xp, ey = c[1] - yp, ez = c[2] - zp;
                float d = sqrtf(ex * ex + ey * ey + ez * ez) - radius;
                atomicMin(&phip[idx], orderkey(d));
            }
}

// sflip.jl:620-640 syncplanes_kernel!, on the fixed-point accumulators.
__global__ void k_syncplanes(unsigned long long* __restrict__ A, STopo t, int a, int nslots) {
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
    A[(size_t)(s - 1) * SBFACES + hi] = A[(size_t)(s2 - 1) * SBFACES + lo];
}

// sflip.jl:790-794 sparseunion_kernel!.
__global__ void k_sparse_union(float* __restrict__ phi, const float* __restrict__ phi0, const int32_t* __restrict__ phip, STopo t,
                               int32_t nf) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    if (t.where[n >> 12].x < 0) return;
    phi[n] = jmin(phi0[n] + t.g.dx, orderfloat(phip[n]));
}

// sflip.jl:796-806 sparsecombine_kernel!.
__global__ void k_sparse_combine(float* __restrict__ u, const float* __restrict__ u0, const unsigned long long* __restrict__ mom,
                                 const unsigned long long* __restrict__ wt, const float* __restrict__ phi,
                                 const float* __restrict__ phic, STopo t, float r, float full, int a, int32_t nf) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    SCell c = fineface(t, a, n);
    if (c.w.x < 0) return;
    float p[3];
    facepos(t.g, a, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, t.g.dx, p);
    float phif = cellsample(t, phi, phic, p[0], p[1], p[2]);
    float wgt = fromfix(wt[n]), m = fromfix(mom[n]), v0 = u0[n];
    u[n] = phif >= -r ? __fdiv_rn(m + jmax(full - wgt, 0.0f) * v0, jmax(full, wgt)) : v0;
}

void SparseBasin::p2g(float dt) {
    hipStream_t s = st.s;
    size_t F = finefaces(), C = finecells();
    STopo t = topo();
    for (int a = 0; a < 3; ++a) {
        fill_async(mom[a].p, F, 0ull, s);
        fill_async(wt[a].p, F, 0ull, s);
    }
    int32_t big;
    {
        float f = 1e30f;
        std::memcpy(&big, &f, 4);   // orderkey of a positive float is its bit pattern
    }
    fill_async(phip.p, C, big, s);
    if (n > 0) {
        PView ps = view(cur);
        kt.begin(K_P2G, s);
        P2GArgs A{{g.lo[0], g.lo[1], g.lo[2]}, g.dx, dt, {mom[0].p, mom[1].p, mom[2].p}, {wt[0].p, wt[1].p, wt[2].p}};
        hipLaunchKernelGGL(k_p2g_tiled<SparseNodes>, dim3(cdiv(n, P2G_THREADS)), dim3(P2G_THREADS), 0, s, ps, n, SparseNodes{t}, A);
        kt.end(K_P2G, s);
        for (int a = 0; a < 3; ++a) {
            hipLaunchKernelGGL(k_syncplanes, dim3(cdiv(nslots * 256, 256)), dim3(256), 0, s, mom[a].p, t, a, nslots);
            hipLaunchKernelGGL(k_syncplanes, dim3(cdiv(nslots * 256, 256)), dim3(256), 0, s, wt[a].p, t, a, nslots);
        }
        kt.begin(K_SPLAT, s);
        hipLaunchKernelGGL(k_sparse_splat, dim3(cdiv(n, 256)), dim3(256), 0, s, ps, n, t, dt, phip.p);
        kt.end(K_SPLAT, s);
        MAEL_LAUNCH_CHECK();
    }
    if (C)
        hipLaunchKernelGGL(k_sparse_union, dim3(cdiv(C, 256)), dim3(256), 0, s, phi.fine.p, phi0.fine.p, phip.p, t, (int32_t)C);
    MAEL_HIP(hipMemcpyAsync(phi.coarse.p, phi0.coarse.p, ncells() * sizeof(float), hipMemcpyDeviceToDevice, s));
    launch_restrictcells(*this, phi.coarse.p, phi.fine.p);
    float r = (float)(band - 1) * g.dx, full = (float)nmin;
    kt.begin(K_COMBINE, s);
    for (int a = 0; a < 3; ++a) {
        if (F)
            hipLaunchKernelGGL(k_sparse_combine, dim3(cdiv(F, 256)), dim3(256), 0, s, u[a].fine.p, u0[a].fine.p, mom[a].p,
                               wt[a].p, phi.fine.p, phi.coarse.p, t, r, full, a, (int32_t)F);
        MAEL_HIP(hipMemcpyAsync(u[a].coarse.p, u0[a].coarse.p, nfaces(a) * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    kt.end(K_COMBINE, s);
    MAEL_LAUNCH_CHECK();
    restrict_velocity();
}

// ─── grid to particles (sflip.jl:2141-2144; flip.jl:1226-1244 over the topology) ──

__global__ __launch_bounds__(256) void k_sparse_g2p(PView ps, int64_t n, STopo t, const float* __restrict__ u,
                                                    const float* __restrict__ v, const float* __restrict__ w,
                                                    const uint8_t* __restrict__ fu, const uint8_t* __restrict__ fv,
                                                    const uint8_t* __restrict__ fw, float dt, float gravity) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float xp = ps.f[0][p], yp = ps.f[1][p], zp = ps.f[2][p];
    bool k1 = finegather_value(t, fu, 0, xp, yp, zp) > 0.999f;
    bool k2 = finegather_value(t, fv, 1, xp, yp, zp) > 0.999f;
    bool k3 = finegather_value(t, fw, 2, xp, yp, zp) > 0.999f;
    float rec[12];   // stored unconditionally (see k_dense_spawn)
    if (k1 && k2 && k3) {
        finegather(t, u, 0, xp, yp, zp, rec[0], rec[3], rec[4], rec[5]);
        finegather(t, v, 1, xp, yp, zp, rec[1], rec[6], rec[7], rec[8]);
        finegather(t, w, 2, xp, yp, zp, rec[2], rec[9], rec[10], rec[11]);
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

void SparseBasin::g2p(float dt, float grav) {
    if (n == 0) return;
    kt.begin(K_G2P, st.s);
    const SField<float>* fv = reflect_flow ? ur : u;   // sflip.jl:2312: the reflected velocity in a first half-step
    hipLaunchKernelGGL(k_sparse_g2p, dim3(cdiv(n, 256)), dim3(256), 0, st.s, view(cur), n, topo(), fv[0].fine.p, fv[1].fine.p,
                       fv[2].fine.p, fl[0].fine.p, fl[1].fine.p, fl[2].fine.p, dt, grav);
    kt.end(K_G2P, st.s);
    MAEL_LAUNCH_CHECK();
}

}  // namespace mael
