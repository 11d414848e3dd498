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
    rank[p] = wave_count_slot(counts, idx);   // arrival: only sizes the segment
}

// sflip.jl:2184-2218 sparsespawn_kernel!, in two passes so that the threads of a launch
// each spawn one particle: the cells that add particles are listed (in any order: where a
// particle goes is fixed by its cell's offsets), then nmin threads a listed cell each take
// one particle m < add. The records are the reference's, bit for bit.
__device__ inline void spawn_counts(const STopo& t, const int32_t* __restrict__ counts, const float* __restrict__ phi, int32_t n,
                                    int32_t nmin, float deep, int seeding, SCell& c, int32_t& kept, int32_t& add) {
    c = finecell(t, n);
    kept = 0; add = 0;
    if (c.w.x < 0) return;
    const SGrid& g = t.g;
    float pc[3];
    finecentre(g, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, pc);
    scellcounts(counts[n], phi[n], sissolid(g, pc[0], pc[1], pc[2]), nmin, deep, g.dx, seeding != 0, kept, add);
}
__global__ void k_spawn_list(const int32_t* __restrict__ counts, const float* __restrict__ phi, STopo t, int32_t nmin, float deep,
                             int seeding, int32_t nf, int32_t* __restrict__ list, int* __restrict__ nlist) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    SCell c;
    int32_t kept, add;
    spawn_counts(t, counts, phi, n, nmin, deep, seeding, c, kept, add);
    if (add > 0) list[atomicAdd(nlist, 1)] = n;
}
__global__ __launch_bounds__(256) void k_sparse_spawn(PView dst, const int32_t* __restrict__ counts,
                                                      const int32_t* __restrict__ totals, const int32_t* __restrict__ offsets,
                                                      const float* __restrict__ phi, const float* __restrict__ phic,
                                                      const float* __restrict__ u, const float* __restrict__ v,
                                                      const float* __restrict__ w, STopo t, int32_t nmin, float deep, int seeding,
                                                      uint32_t seed, const int32_t* __restrict__ list,
                                                      const int* __restrict__ nlist) {
    const SGrid& g = t.g;
    int64_t total = (int64_t)(*nlist) * nmin;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int32_t n = list[e / nmin], m = (int32_t)(e % nmin);
        SCell c;
        int32_t kept, add;
        spawn_counts(t, counts, phi, n, nmin, deep, seeding, c, kept, add);
        if (m >= add) continue;
        int64_t base = (int64_t)offsets[n] - totals[n] + kept;
        int I = c.w.x * 16 + c.i, J = c.w.y * 16 + c.j, K = c.w.z * 16 + c.k;
        float cx = g.lo[0] + (float)I * g.dx, cy = g.lo[1] + (float)J * g.dx, cz = g.lo[2] + (float)K * g.dx;
        uint32_t n1 = (uint32_t)n + 1u;   // the reference's 1-based index
        int32_t sd = (int32_t)seed;
        float half = 0.5f * g.dx;
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
        int* nlist = changed.p;   // a device counter, free during resampling
        fill_async(nlist, (size_t)1, 0, s);
        hipLaunchKernelGGL(k_spawn_list, dim3(cdiv(C, 256)), dim3(256), 0, s, counts.p, phi.fine.p, t, nmin, deep,
                           seeding ? 1 : 0, (int32_t)C, cstart.p, nlist);
        hipLaunchKernelGGL(k_sparse_spawn, dim3(4096), dim3(256), 0, s, dst, counts.p, totals.p, offsets.p, phi.fine.p,
                           phi.coarse.p, fv[0].fine.p, fv[1].fine.p, fv[2].fine.p, t, nmin, deep, seeding ? 1 : 0, seed,
                           cstart.p, nlist);
        kt.end(K_SPAWN, s);
        MAEL_LAUNCH_CHECK();
    }
    cur ^= 1;
    n = total;
    return total;
}

// ─── particles to grid ───────────────────────────────────────────────────────

struct FineCellMap {
    STopo t;
    __device__ int32_t index(int I, int J, int K) const { return fineat(t, I, J, K); }
};

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
// A workgroup's 256 faces lie in one slot (4352 = 17 x 256): the level set's taps find
// their blocks in a table of the 27 blocks around it.
__global__ __launch_bounds__(256) void k_sparse_combine(float* __restrict__ u, const float* __restrict__ u0,
                                                        const unsigned long long* __restrict__ mom,
                                                        const unsigned long long* __restrict__ wt, const float* __restrict__ phi,
                                                        const float* __restrict__ phic, STopo t, float r, float full, int a,
                                                        int32_t nf) {
    __shared__ int32_t s27[27];
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    SCell c = fineface(t, a, n);
    if (c.w.x < 0) return;   // the whole workgroup
    stage27(t, c.w, s27);
    __syncthreads();
    float p[3];
    facepos(t.g, a, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, t.g.dx, p);
    float phif = cellsample(t, phi, phic, p[0], p[1], p[2], Slots27{s27, c.w});
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
        P2GArgs A{{g.lo[0], g.lo[1], g.lo[2]}, g.dx, {mom[0].p, mom[1].p, mom[2].p}, {wt[0].p, wt[1].p, wt[2].p}, 1.0f};
        hipLaunchKernelGGL((k_p2g_tiled<SparseNodes, P2GSource>), dim3(cdiv(n, P2G_THREADS)), dim3(P2G_THREADS), 0, s, ps, n,
                           SparseNodes{t}, P2GSource{ps, dt}, A);
        kt.end(K_P2G, s);
        for (int a = 0; a < 3; ++a) {
            hipLaunchKernelGGL(k_syncplanes, dim3(cdiv(nslots * 256, 256)), dim3(256), 0, s, mom[a].p, t, a, nslots);
            hipLaunchKernelGGL(k_syncplanes, dim3(cdiv(nslots * 256, 256)), dim3(256), 0, s, wt[a].p, t, a, nslots);
        }
        kt.begin(K_SPLAT, s);
        hipLaunchKernelGGL(k_splat_tiled<FineCellMap>, dim3(cdiv(n, P2G_THREADS)), dim3(P2G_THREADS), 0, s, ps, n,
                           FineCellMap{t}, g.lo[0], g.lo[1], g.lo[2], g.dx, dt, phip.p);
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

// One component's gather at a particle: the flag sum (finegather_value) and the APIC
// value and affine row (finegather, sflip.jl:193-215) in one pass over the 27 nodes, in
// the reference's order; `fetch(I, J, K, present, u, flag)` reads a node.
// With ZERO the fetch returns 0 for the value and a negative flag where no face is, and a
// missing node adds zero products instead of being skipped: the sums start at +0 and are
// never -0, so adding a zero leaves them bit for bit as skipping does, without a branch.
template <bool ZERO = false, class Fetch>
__device__ inline void gather_component(const SGrid& g, int a, float px, float py, float pz, Fetch fetch, float& flag,
                                        float& vel, float& a0, float& a1, float& a2) {
    float q0 = ispace(px, g.lo[0], g.dx, a == 0 ? 0.0f : 0.5f);
    float q1 = ispace(py, g.lo[1], g.dx, a == 1 ? 0.0f : 0.5f);
    float q2 = ispace(pz, g.lo[2], g.dx, a == 2 ? 0.0f : 0.5f);
    BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
    flag = 0.0f; vel = 0.0f; a0 = 0.0f; a1 = 0.0f; a2 = 0.0f;
#pragma unroll
    for (int mz = 0; mz < 3; ++mz) {
        float wz = pick(bz, mz), dz = (float)(bz.b + mz) - q2;
#pragma unroll
        for (int my = 0; my < 3; ++my) {
            float wy = pick(by, my), dy = (float)(by.b + my) - q1;
#pragma unroll
            for (int mx = 0; mx < 3; ++mx) {
                bool pres;
                float uv, fv;
                fetch(bx.b + mx, by.b + my, bz.b + mz, pres, uv, fv);
                if (ZERO || pres) {
                    float wgt = pick(bx, mx) * wy * wz;
                    if (ZERO) wgt = pres ? wgt : 0.0f;
                    flag += wgt * fv;
                    float tt = wgt * uv;
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

// flip.jl:1226-1244 g2p_kernel! over the topology, tiled: a workgroup's 256 particles
// (cell-sorted, so close together) gather each component from a box of faces staged in
// LDS (the value, and the flag or -1 where no fine face is there), loaded once per box
// node instead of once per particle and node; a workgroup spread over more nodes than the
// box holds reads the faces directly. Both paths sum the same terms in the same order. The
// box's faces are found through a table of the slots of the blocks under it (and one
// layer below along the component's axis, for a lower block's top faces).
constexpr int G2P_BOX = 2048, G2P_TAB = 64;   // 16 KB: a band sheet through a block spans up to ~2000 nodes
__global__ __launch_bounds__(256) void k_sparse_g2p(PView ps, int64_t n, STopo t, const float* __restrict__ u,
                                                    const float* __restrict__ v, const float* __restrict__ w,
                                                    const uint8_t* __restrict__ fu, const uint8_t* __restrict__ fv,
                                                    const uint8_t* __restrict__ fw, float dt, float gravity) {
    __shared__ float2 sbox[G2P_BOX];      // value, flag (-1: no face)
    __shared__ int32_t stab[G2P_TAB];
    __shared__ int red[6 * 4];
    const SGrid& g = t.g;
    int64_t p = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;   // neighbouring tiles on one XCD
    bool live = p < n;
    float xp = 0, yp = 0, zp = 0;
    if (live) { xp = ps.f[0][p]; yp = ps.f[1][p]; zp = ps.f[2][p]; }
    const float* U[3] = {u, v, w};
    const uint8_t* F[3] = {fu, fv, fw};
    float fl[3], val[3], c0[3], c1[3], c2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float q0 = ispace(xp, g.lo[0], g.dx, a == 0 ? 0.0f : 0.5f);
        float q1 = ispace(yp, g.lo[1], g.dx, a == 1 ? 0.0f : 0.5f);
        float q2 = ispace(zp, g.lo[2], g.dx, a == 2 ? 0.0f : 0.5f);
        BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
        bool ok = live && bx.b > -(1 << 24) && bx.b < (1 << 24) && by.b > -(1 << 24) && by.b < (1 << 24) &&
                  bz.b > -(1 << 24) && bz.b < (1 << 24) && isfinite(q0) && isfinite(q1) && isfinite(q2);
        const int BIG = 1 << 30;
        int x0, y0, z0, x1, y1, z1;
        wg_bounds(ok, bx.b, by.b, bz.b, red, x0, y0, z0, x1, y1, z1);
        int ex = x1 - x0 + 3, ey = y1 - y0 + 3, ez = z1 - z0 + 3;
        bool boxed = x0 != BIG && ex <= G2P_BOX && ey <= G2P_BOX && ez <= G2P_BOX && ex * ey * ez <= G2P_BOX;
        const float* Ua = U[a];
        const uint8_t* Fa = F[a];
        if (boxed) {
            const int tx = (x0 >> 4) - (a == 0), ty = (y0 >> 4) - (a == 1), tz = (z0 >> 4) - (a == 2);
            const int nx = ((x0 + ex - 1) >> 4) - tx + 1, ny = ((y0 + ey - 1) >> 4) - ty + 1, nz = ((z0 + ez - 1) >> 4) - tz + 1;
            const bool tab = nx * ny * nz <= G2P_TAB;
            if (tab && (int)threadIdx.x < nx * ny * nz) {
                int bi = tx + (int)threadIdx.x % nx, bj = ty + ((int)threadIdx.x / nx) % ny, bk = tz + (int)threadIdx.x / (nx * ny);
                stab[threadIdx.x] = inblocks(g, bi, bj, bk) ? t.slot[blockid(g, bi, bj, bk)] : 0;
            }
            __syncthreads();
            int V = ex * ey * ez;
            for (int k = threadIdx.x; k < V; k += 256) {
                const int I = x0 + k % ex, J = y0 + (k / ex) % ey, K = z0 + k / (ex * ey);
                int32_t idx;
                if (tab) {   // finefaceat through the table
                    const int di = (I >> 4) - tx, dj = (J >> 4) - ty, dk = (K >> 4) - tz;
                    const int32_t s1 = stab[di + nx * (dj + ny * dk)];
                    if (s1 != 0) {
                        idx = (s1 - 1) * SBFACES + ffn(16, a, I & 15, J & 15, K & 15);
                    } else {
                        const int GA = a == 0 ? I : (a == 1 ? J : K);
                        const int32_t s2 =
                            (GA & 15) == 0 ? stab[(di - (a == 0)) + nx * ((dj - (a == 1)) + ny * (dk - (a == 2)))] : 0;
                        idx = s2 != 0 ? (s2 - 1) * SBFACES + ffn(16, a, a == 0 ? 16 : (I & 15), a == 1 ? 16 : (J & 15),
                                                                 a == 2 ? 16 : (K & 15))
                                      : -1;
                    }
                } else {
                    idx = finefaceat(t, a, I, J, K);
                }
                sbox[k] = idx >= 0 ? make_float2(Ua[idx], (float)Fa[idx]) : make_float2(0.0f, -1.0f);
            }
            __syncthreads();
        }
        if (live) {
            if (boxed && ok) {
                // gather_component<true> over the box, the factors shared along rows hoisted:
                // the same operations on the same values in the same order, so the same bits
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
                            float wgt = (wx[mx] * wy[my]) * wz[mz];
                            wgt = e.y >= 0.0f ? wgt : 0.0f;
                            f_ += wgt * e.y;
                            const float tt = wgt * e.x;
                            v_ += tt;
                            s0 += (tt * d0[mx]) * g.dx;
                            s1 += (tt * d1[my]) * g.dx;
                            s2 += (tt * d2[mz]) * g.dx;
                        }
                    }
                const float c = __fdiv_rn(4.0f, g.dx * g.dx);
                fl[a] = f_; val[a] = v_; c0[a] = s0 * c; c1[a] = s1 * c; c2[a] = s2 * c;
            } else {
                auto fetch = [&](int I, int J, int K, bool& pres, float& uv, float& fv_) {
                    int32_t idx = finefaceat(t, a, I, J, K);
                    pres = idx >= 0;
                    uv = pres ? Ua[idx] : 0.0f;
                    fv_ = pres ? (float)Fa[idx] : 0.0f;
                };
                gather_component(g, a, xp, yp, zp, fetch, fl[a], val[a], c0[a], c1[a], c2[a]);
            }
        }
        if (boxed) __syncthreads();   // the box is reused by the next component
    }
    if (!live) return;
    bool k1 = fl[0] > 0.999f, k2 = fl[1] > 0.999f, k3 = fl[2] > 0.999f;
    float rec[12];   // stored unconditionally (see k_dense_spawn)
    if (k1 && k2 && k3) {
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
