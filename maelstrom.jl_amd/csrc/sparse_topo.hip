// The sparse basin's allocation, its topology (which blocks are fine: the marks, the
// host block allocator and the device side of `refine!`, sflip.jl:382-576), the
// fine <-> coarse transfers (sflip.jl:578-640) and the support phases that set up and
// close a step: the start (sflip.jl:644-713), redistancing (sflip.jl:833-919),
// classification (sflip.jl:921-948), walls (sflip.jl:1019-1041) and the volume
// (sflip.jl:1975-1996).
#include "sparse.hpp"
#include "scan.hpp"
#include <algorithm>

namespace mael {

__constant__ int NBS[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

SparseBasin::SparseBasin(const SGrid& grid, int nblk, int64_t cap, int32_t bnd, int32_t nm)
    : g(grid), blocks(nblk), band(bnd), nmin(nm), capacity(cap) {
    size_t NB = nblocks(), Cc = ncells();
    size_t nf = (size_t)blocks * SBCELLS, nfa = (size_t)blocks * SBFACES;
    hslot.assign(NB, 0);
    hwhere.assign(blocks, make_int4(-1, -1, -1, 0));
    slot.alloc(NB); where.alloc(blocks); list.alloc(blocks); sorder.alloc(std::max(blocks, 1));
    marks.alloc(NB); marks2.alloc(NB);
    for (auto* f : {&phi, &phi0, &P, &b, &r, &z, &d, &q, &rold, &tmp, &dD}) f->alloc(nf, Cc);
    types.alloc(nf, Cc);
    phip.alloc(nf);
    for (int a = 0; a < 3; ++a) {
        u[a].alloc(nfa, nfaces(a)); u0[a].alloc(nfa, nfaces(a)); fl[a].alloc(nfa, nfaces(a));
        mom[a].alloc(nfa); wt[a].alloc(nfa);
        ur[a].alloc(nfa, nfaces(a)); ur0[a].alloc(nfa, nfaces(a)); fl2[a].alloc(nfa, nfaces(a));
        cf[a].alloc(nfa); cc[a].alloc(nfaces(a));
        c1[a].alloc((size_t)blocks * SBFACES1);
    }
    dD1.alloc((size_t)blocks * SBCELLS1); act1.alloc((size_t)blocks * SBCELLS1);
    for (auto* f : {&X1x, &X1b, &X1r, &X1t}) f->alloc((size_t)blocks * SBCELLS1, Cc);
    // the coarse grid and its halvings until a workgroup solves the coarsest (sflip.jl:319-326)
    int m[3] = {g.nc(0), g.nc(1), g.nc(2)};
    while (true) {
        D.emplace_back();
        SLevel& L = D.back();
        for (int a = 0; a < 3; ++a) L.n[a] = m[a];
        size_t c = L.cells();
        L.x.alloc(c); L.b.alloc(c); L.t.alloc(c); L.dD.alloc(c); L.act.alloc(c);
        for (int a = 0; a < 3; ++a) L.c[a].alloc(L.faces(a));
        if ((long)m[0] * m[1] * m[2] <= 1024 || (m[0] & 1) || (m[1] & 1) || (m[2] & 1)) break;
        for (int a = 0; a < 3; ++a) m[a] /= 2;
    }
    if ((long)m[0] * m[1] * m[2] > 1024)
        throw Error(E_GRID, "the coarse grid (" + std::to_string(g.nc(0)) + ", " + std::to_string(g.nc(1)) + ", " +
                                std::to_string(g.nc(2)) + ") does not halve to at most 1024 cells");
    reach.alloc(Cc); changed.alloc(1);
    stir_waves.alloc(1); stir_heights.alloc(2);
    // per-workgroup sums: the vector kernels' (2 x 2048) or the fused search's (a quarter
    // slot's workgroup each, then the coarse grid's)
    partials.alloc(std::max<size_t>(4096, 4 * (size_t)blocks + (Cc + 255) / 256 + 64));
    scal.alloc(128);
    for (int w = 0; w < 2; ++w)
        for (int f = 0; f < PFIELDS; ++f) pbuf[w][f].alloc(std::max<int64_t>(cap, 1));
    key.alloc(std::max<int64_t>(cap, 1)); rank.alloc(std::max<int64_t>(cap, 1)); seg.alloc(std::max<int64_t>(cap, 1));
    counts.alloc(nf); totals.alloc(nf); offsets.alloc(nf); cstart.alloc(nf);
    bsums.alloc(nf / SCAN_TILE + 2);
    hipStream_t s = st.s;
    fill_async(slot.p, NB, 0, s);
    fill_async(where.p, (size_t)blocks, make_int4(-1, -1, -1, 0), s);
    for (auto* f : {&phi, &phi0, &P, &b, &r, &z, &d, &q, &rold, &tmp, &dD}) {
        fill_async(f->fine.p, nf, 0.0f, s);
        fill_async(f->coarse.p, Cc, 0.0f, s);
    }
    fill_async(types.fine.p, nf, (uint8_t)0, s);
    fill_async(types.coarse.p, Cc, (uint8_t)0, s);
    for (int a = 0; a < 3; ++a) {
        for (auto* f : {&u[a], &u0[a]}) { fill_async(f->fine.p, nfa, 0.0f, s); fill_async(f->coarse.p, nfaces(a), 0.0f, s); }
        fill_async(fl[a].fine.p, nfa, (uint8_t)0, s);
        fill_async(fl[a].coarse.p, nfaces(a), (uint8_t)0, s);
        fill_async(cf[a].p, nfa, 0.0f, s);
        fill_async(cc[a].p, nfaces(a), 0.0f, s);
        for (auto* f : {&ur[a], &ur0[a]}) { fill_async(f->fine.p, nfa, 0.0f, s); fill_async(f->coarse.p, nfaces(a), 0.0f, s); }
        fill_async(fl2[a].fine.p, nfa, (uint8_t)0, s);
        fill_async(fl2[a].coarse.p, nfaces(a), (uint8_t)0, s);
    }
    for (auto* f : {&X1x, &X1b, &X1r, &X1t}) {
        fill_async(f->fine.p, (size_t)blocks * SBCELLS1, 0.0f, s);
        fill_async(f->coarse.p, Cc, 0.0f, s);
    }
    fill_async(counts.p, nf, 0, s); fill_async(totals.p, nf, 0, s); fill_async(offsets.p, nf, 0, s);
    fill_async(scal.p, (size_t)128, 0.0, s);
    st.sync();
}

void SparseBasin::mp_alloc() {
    if (mp.ready) return;
    size_t nf = (size_t)blocks * SBCELLS, nfa = (size_t)blocks * SBFACES, n1 = (size_t)blocks * SBCELLS1, Cc = ncells();
    for (auto* v : {mp.z, mp.r, mp.q, mp.tmp}) { v[0].alloc(nf); v[1].alloc(Cc); }
    for (auto& v : mp.X1) { v[0].alloc(n1); v[1].alloc(Cc); }
    mp.Dv.resize(D.size());
    mp.Dc.resize(D.size());
    for (size_t l = 0; l < D.size(); ++l) {
        for (auto& b : mp.Dv[l]) b.alloc(D[l].cells());
        for (int a = 0; a < 3; ++a) mp.Dc[l][a].alloc(D[l].faces(a));
        mp.Dc[l][3].alloc(D[l].cells());
    }
    for (int a = 0; a < 3; ++a) {
        mp.cc[a].alloc(nfaces(a)); mp.c1[a].alloc((size_t)blocks * SBFACES1);
    }
    mp.dDf.alloc(nf); mp.dDc.alloc(Cc); mp.dD1.alloc(n1); mp.rwf.alloc(nf);
    // every vector starts at zero: free slots are never written and must read as zero
    for (auto* v : {mp.z, mp.r, mp.q, mp.tmp}) { fill_async(v[0].p, nf, (uint16_t)0, st.s); fill_async(v[1].p, Cc, (uint16_t)0, st.s); }
    for (auto& v : mp.X1) { fill_async(v[0].p, n1, (uint16_t)0, st.s); fill_async(v[1].p, Cc, (uint16_t)0, st.s); }
    st.sync();
    mp.ready = true;
}

// ─── marks (sflip.jl:384-455) ────────────────────────────────────────────────

__device__ inline size_t blockof(const SGrid& g, int I, int J, int K) {   // the block of coarse cell (I, J, K)
    return (size_t)(I >> 2) + (size_t)g.nb[0] * ((J >> 2) + (size_t)g.nb[1] * (K >> 2));
}

__global__ void k_marksurface(uint8_t* __restrict__ marks, const float* __restrict__ phic, SGrid g) {
    int n0 = g.nc(0), n1 = g.nc(1), n2 = g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    float p[3];
    coarsecentre(g, I, J, K, p);
    if (sissolid(g, p[0], p[1], p[2])) return;
    float c = phic[l];
    bool surf = fabsf(c) < 1.5f * g.cdx();
    for (int m = 0; m < 6; ++m) {
        int a = I + NBS[m][0], b = J + NBS[m][1], cc = K + NBS[m][2];
        if (a < 0 || b < 0 || cc < 0 || a >= n0 || b >= n1 || cc >= n2) continue;
        coarsecentre(g, a, b, cc, p);
        if (sissolid(g, p[0], p[1], p[2])) continue;
        if ((phic[cl(a, b, cc, n0, n1)] < 0.0f) != (c < 0.0f)) surf = true;
    }
    if (surf) marks[blockof(g, I, J, K)] = 1;
}

__device__ inline void cellvelocity(const float* u, const float* v, const float* w, int i, int j, int k, int n0, int n1,
                                    float c[3]) {
    c[0] = 0.5f * (u[cl(i, j, k, n0 + 1, n1)] + u[cl(i + 1, j, k, n0 + 1, n1)]);
    c[1] = 0.5f * (v[cl(i, j, k, n0, n1 + 1)] + v[cl(i, j + 1, k, n0, n1 + 1)]);
    c[2] = 0.5f * (w[cl(i, j, k, n0, n1)] + w[cl(i, j, k + 1, n0, n1)]);
}

__global__ void k_markflow(uint8_t* __restrict__ marks, const float* __restrict__ u, const float* __restrict__ v,
                           const float* __restrict__ w, const float* __restrict__ phic, SGrid g, float vary) {
    int n0 = g.nc(0), n1 = g.nc(1), n2 = g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    float p[3];
    coarsecentre(g, I, J, K, p);
    if (!(phic[l] < 0.0f) || sissolid(g, p[0], p[1], p[2])) return;
    float c[3], o[3];
    cellvelocity(u, v, w, I, J, K, n0, n1, c);
    float jump = 0.0f;
    for (int m = 0; m < 6; ++m) {
        int a = I + NBS[m][0], b = J + NBS[m][1], cc = K + NBS[m][2];
        if (a < 0 || b < 0 || cc < 0 || a >= n0 || b >= n1 || cc >= n2) continue;
        coarsecentre(g, a, b, cc, p);
        if (!(phic[cl(a, b, cc, n0, n1)] < 0.0f) || sissolid(g, p[0], p[1], p[2])) continue;
        cellvelocity(u, v, w, a, b, cc, n0, n1, o);
        float dx = o[0] - c[0], dy = o[1] - c[1], dz = o[2] - c[2];
        jump = jmax(jump, sqrtf(dx * dx + dy * dy + dz * dz));
    }
    if (jump > vary) marks[blockof(g, I, J, K)] = 1;
}

__global__ void k_markparticles(uint8_t* __restrict__ marks, const float* __restrict__ x, const float* __restrict__ y,
                                const float* __restrict__ z, SGrid g, float reach, int64_t n) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float xp = x[p], yp = y[p], zp = z[p];
    if (!(isfinite(xp) && isfinite(yp) && isfinite(zp))) return;
    float S = 16.0f * g.dx;
    for (int c = 0; c < 2; ++c)
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) {
                float q0 = __fdiv_rn(xp - g.lo[0] + (float)(2 * a - 1) * reach, S);
                float q1 = __fdiv_rn(yp - g.lo[1] + (float)(2 * b - 1) * reach, S);
                float q2 = __fdiv_rn(zp - g.lo[2] + (float)(2 * c - 1) * reach, S);
                int bi = min(max((int)floorf(q0), 0), g.nb[0] - 1);
                int bj = min(max((int)floorf(q1), 0), g.nb[1] - 1);
                int bk = min(max((int)floorf(q2), 0), g.nb[2] - 1);
                marks[(size_t)bi + (size_t)g.nb[0] * (bj + (size_t)g.nb[1] * bk)] = 1;
            }
}

__global__ void k_dilate(uint8_t* __restrict__ out, const uint8_t* __restrict__ marks, SGrid g) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t NB = (size_t)g.nb[0] * g.nb[1] * g.nb[2];
    if (l >= NB) return;
    int i = (int)(l % g.nb[0]), j = (int)((l / g.nb[0]) % g.nb[1]), k = (int)(l / ((size_t)g.nb[0] * g.nb[1]));
    uint8_t m = 0;
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                int a = i + di, b = j + dj, c = k + dk;
                if (a < 0 || b < 0 || c < 0 || a >= g.nb[0] || b >= g.nb[1] || c >= g.nb[2]) continue;
                if (marks[cl(a, b, c, g.nb[0], g.nb[1])]) m = 1;
            }
    out[l] = m;
}

void SparseBasin::marks_compute(float vary, float reach, int rings, bool particles) {
    hipStream_t s = st.s;
    size_t NB = nblocks(), Cc = ncells();
    fill_async(marks.p, NB, (uint8_t)0, s);
    hipLaunchKernelGGL(k_marksurface, dim3(cdiv(Cc, 256)), dim3(256), 0, s, marks.p, phi.coarse.p, g);
    hipLaunchKernelGGL(k_markflow, dim3(cdiv(Cc, 256)), dim3(256), 0, s, marks.p, u[0].coarse.p, u[1].coarse.p, u[2].coarse.p,
                       phi.coarse.p, g, vary);
    if (particles && n > 0)
        hipLaunchKernelGGL(k_markparticles, dim3(cdiv(n, 256)), dim3(256), 0, s, marks.p, pbuf[cur][0].p, pbuf[cur][1].p,
                           pbuf[cur][2].p, g, reach, n);
    for (int r = 0; r < rings; ++r) {
        hipLaunchKernelGGL(k_dilate, dim3(cdiv(NB, 256)), dim3(256), 0, s, marks2.p, marks.p, g);
        std::swap(marks, marks2);
    }
    MAEL_LAUNCH_CHECK();
}

// ─── the block allocator (sflip.jl:528-551) ──────────────────────────────────

void SparseBasin::refine_host(const uint8_t* want, std::vector<int32_t>& added, std::vector<int32_t>& removed) {
    size_t NB = nblocks();
    size_t required = 0;
    for (size_t blk = 0; blk < NB; ++blk) required += want[blk] != 0;
    if (required > (size_t)blocks)
        throw Error(E_BLOCKS, "the sparse grid needs " + std::to_string(required) + " fine blocks, more than its " +
                                  std::to_string(blocks) + " allocated");
    added.clear();
    removed.clear();
    // departures are released before arrivals are placed, both in block order
    for (size_t blk = 0; blk < NB; ++blk)
        if (want[blk] == 0 && hslot[blk] != 0) {
            int32_t sl = hslot[blk];
            hwhere[sl - 1] = make_int4(-1, -1, -1, 0);
            hslot[blk] = 0;
            freel.push_back(sl);
            removed.push_back(sl);
        }
    for (size_t blk = 0; blk < NB; ++blk)
        if (want[blk] != 0 && hslot[blk] == 0) {
            int32_t sl;
            if (freel.empty()) sl = ++nslots;
            else { sl = freel.back(); freel.pop_back(); }   // the most recently freed slot
            hwhere[sl - 1] = make_int4((int)(blk % g.nb[0]), (int)((blk / g.nb[0]) % g.nb[1]),
                                       (int)(blk / ((size_t)g.nb[0] * g.nb[1])), 0);
            hslot[blk] = sl;
            added.push_back(sl);
        }
}

void SparseBasin::upload_topology() {
    // a per-slot launch that walks the live slots in block order runs blocks that are
    // neighbours in x (and, a row later, in y) together, so the layers they read across
    // their sides are still in L2
    std::vector<int32_t> order;
    order.reserve((size_t)blocks);
    for (int32_t sl : hslot)
        if (sl != 0) order.push_back(sl);
    nlive = (int32_t)order.size();
    MAEL_HIP(hipMemcpyAsync(slot.p, hslot.data(), hslot.size() * sizeof(int32_t), hipMemcpyHostToDevice, st.s));
    MAEL_HIP(hipMemcpyAsync(where.p, hwhere.data(), hwhere.size() * sizeof(int4), hipMemcpyHostToDevice, st.s));
    if (nlive > 0)
        MAEL_HIP(hipMemcpyAsync(sorder.p, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice, st.s));
    st.sync();   // the host vectors may change before the copies would otherwise run
}

template <class T> __global__ void k_zeroslots(T* __restrict__ A, const int32_t* __restrict__ lst, int per, int count) {
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)count * per) return;
    int32_t s = lst[n / per];
    A[(int64_t)(s - 1) * per + n % per] = T(0);
}

template <class V> __device__ inline V fromcoarse(float x);
template <> __device__ inline float fromcoarse<float>(float x) { return x; }
template <> __device__ inline uint8_t fromcoarse<uint8_t>(float x) { return x > 0.999f ? 1 : 0; }

// sflip.jl:468-478 prolongcells_kernel!.
template <class V>
__global__ void k_prolongcells(V* __restrict__ fine, const V* __restrict__ coarse, STopo t, const int32_t* __restrict__ lst,
                               int count) {
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)count * SBCELLS) return;
    int32_t s = lst[n >> 12];
    int c = (int)(n & 4095), i = c & 15, j = (c >> 4) & 15, k = c >> 8;
    int4 w = t.where[s - 1];
    float p[3];
    finecentre(t.g, w.x * 16 + i, w.y * 16 + j, w.z * 16 + k, p);
    float cd = t.g.cdx();
    float v = trilinear(coarse, t.g.nc(0), t.g.nc(1), t.g.nc(2), __fdiv_rn(p[0] - t.g.lo[0], cd) - 0.5f,
                        __fdiv_rn(p[1] - t.g.lo[1], cd) - 0.5f, __fdiv_rn(p[2] - t.g.lo[2], cd) - 0.5f);
    fine[(int64_t)(s - 1) * SBCELLS + c] = fromcoarse<V>(v);
}
// sflip.jl:480-496 prolongfaces_kernel!.
template <class V>
__global__ void k_prolongfaces(V* __restrict__ fine, const V* __restrict__ coarse, STopo t, const int32_t* __restrict__ lst,
                               int count, int a) {
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)count * SBFACES) return;
    int32_t s = lst[n / SBFACES];
    int f = (int)(n % SBFACES), i, j, k;
    if (a == 0) { i = f % 17; int r = f / 17; j = r & 15; k = r >> 4; }
    else if (a == 1) { i = f & 15; int r = f >> 4; j = r % 17; k = r / 17; }
    else { i = f & 15; j = (f >> 4) & 15; k = f >> 8; }
    int4 w = t.where[s - 1];
    float p[3];
    facepos(t.g, a, w.x * 16 + i, w.y * 16 + j, w.z * 16 + k, t.g.dx, p);
    float cd = t.g.cdx();
    float o0 = a == 0 ? 0.0f : 0.5f, o1 = a == 1 ? 0.0f : 0.5f, o2 = a == 2 ? 0.0f : 0.5f;
    float v = trilinear(coarse, t.g.nc(0) + (a == 0), t.g.nc(1) + (a == 1), t.g.nc(2) + (a == 2),
                        __fdiv_rn(p[0] - t.g.lo[0], cd) - o0, __fdiv_rn(p[1] - t.g.lo[1], cd) - o1,
                        __fdiv_rn(p[2] - t.g.lo[2], cd) - o2);
    fine[(int64_t)(s - 1) * SBFACES + f] = fromcoarse<V>(v);
}

void SparseBasin::refine(float vary, float reach, int rings, int* nadd, int* nrem) {
    marks_compute(vary, reach, rings, true);
    std::vector<uint8_t> want(nblocks());
    MAEL_HIP(hipMemcpyAsync(want.data(), marks.p, want.size(), hipMemcpyDeviceToHost, st.s));
    st.sync();
    refine_from(want.data(), nadd, nrem);
}

void SparseBasin::refine_from(const uint8_t* want, int* nadd, int* nrem) {
    std::vector<int32_t> added, removed;
    refine_host(want, added, removed);
    *nadd = (int)added.size();
    *nrem = (int)removed.size();
    if (added.empty() && removed.empty()) return;
    upload_topology();
    hipStream_t s = st.s;
    STopo t = topo();
    if (!removed.empty()) {
        MAEL_HIP(hipMemcpyAsync(list.p, removed.data(), removed.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
        int cnt = (int)removed.size();
        unsigned gc = cdiv((int64_t)cnt * SBCELLS, 256), gf = cdiv((int64_t)cnt * SBFACES, 256);
        for (auto* f : {&phi, &phi0, &P, &b, &r, &z, &d, &q, &rold, &tmp, &dD})
            hipLaunchKernelGGL(k_zeroslots<float>, dim3(gc), dim3(256), 0, s, f->fine.p, list.p, SBCELLS, cnt);
        hipLaunchKernelGGL(k_zeroslots<uint8_t>, dim3(gc), dim3(256), 0, s, types.fine.p, list.p, SBCELLS, cnt);
        for (auto* A : {phip.p, counts.p, totals.p})
            hipLaunchKernelGGL(k_zeroslots<int32_t>, dim3(gc), dim3(256), 0, s, A, list.p, SBCELLS, cnt);
        for (int a = 0; a < 3; ++a) {
            for (auto* A : {u[a].fine.p, u0[a].fine.p, cf[a].p})
                hipLaunchKernelGGL(k_zeroslots<float>, dim3(gf), dim3(256), 0, s, A, list.p, SBFACES, cnt);
            for (auto* A : {mom[a].p, wt[a].p})
                hipLaunchKernelGGL(k_zeroslots<unsigned long long>, dim3(gf), dim3(256), 0, s, A, list.p, SBFACES, cnt);
            hipLaunchKernelGGL(k_zeroslots<uint8_t>, dim3(gf), dim3(256), 0, s, fl[a].fine.p, list.p, SBFACES, cnt);
        }
        MAEL_LAUNCH_CHECK();
        st.sync();   // `list` is reused below
    }
    if (!added.empty()) {
        MAEL_HIP(hipMemcpyAsync(list.p, added.data(), added.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
        int cnt = (int)added.size();
        unsigned gc = cdiv((int64_t)cnt * SBCELLS, 256), gf = cdiv((int64_t)cnt * SBFACES, 256);
        for (auto* f : {&phi, &P})
            hipLaunchKernelGGL(k_prolongcells<float>, dim3(gc), dim3(256), 0, s, f->fine.p, f->coarse.p, t, list.p, cnt);
        for (int a = 0; a < 3; ++a) {
            hipLaunchKernelGGL(k_prolongfaces<float>, dim3(gf), dim3(256), 0, s, u[a].fine.p, u[a].coarse.p, t, list.p, cnt, a);
            // the reflected velocity too: a step's second half carries it (sflip.jl:569-572)
            hipLaunchKernelGGL(k_prolongfaces<float>, dim3(gf), dim3(256), 0, s, ur[a].fine.p, ur[a].coarse.p, t, list.p, cnt, a);
            hipLaunchKernelGGL(k_prolongfaces<uint8_t>, dim3(gf), dim3(256), 0, s, fl[a].fine.p, fl[a].coarse.p, t, list.p, cnt,
                               a);
        }
        MAEL_LAUNCH_CHECK();
        st.sync();
    }
}

// ─── fine -> coarse (sflip.jl:583-640) ───────────────────────────────────────

template <class V> __device__ inline V meanof(float acc, float n);
template <> __device__ inline float meanof<float>(float acc, float n) { return acc / n; }
template <> __device__ inline uint8_t meanof<uint8_t>(float acc, float n) { return acc >= n - 0.5f ? 1 : 0; }

template <class V> __global__ void k_restrictcells(V* __restrict__ coarse, const V* __restrict__ fine, STopo t) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    // the 64 fine cells lie in one block, the one covering the coarse cell
    const int bi = I >> 2, bj = J >> 2, bk = K >> 2;
    const int32_t sb = inblocks(t.g, bi, bj, bk) ? t.slot[blockid(t.g, bi, bj, bk)] : 0;
    if (sb == 0) return;
    const int32_t base = (sb - 1) * SBCELLS + ((4 * I) & 15) + 16 * (((4 * J) & 15) + 16 * ((4 * K) & 15));
    float acc = 0.0f;
    for (int c = 0; c < 4; ++c)
        for (int b = 0; b < 4; ++b)
            for (int a = 0; a < 4; ++a) acc += (float)fine[base + a + 16 * b + 256 * c];
    coarse[l] = meanof<V>(acc, 64.0f);
}
template <class V> __global__ void k_restrictfaces(V* __restrict__ coarse, const V* __restrict__ fine, STopo t, int a) {
    int n0 = t.g.nc(0) + (a == 0), n1 = t.g.nc(1) + (a == 1), n2 = t.g.nc(2) + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    // The 16 fine faces lie in one block's plane: finefaceat gives them all from the block
    // holding fine face (4I, 4J, 4K), or all from the top plane of the block below along a,
    // or none.
    const int bi = I >> 2, bj = J >> 2, bk = K >> 2;
    int32_t sb = inblocks(t.g, bi, bj, bk) ? t.slot[blockid(t.g, bi, bj, bk)] : 0;
    bool top = false;
    if (sb == 0) {
        const int FA = 4 * (a == 0 ? I : (a == 1 ? J : K));
        if ((FA & 15) != 0) return;
        const int li = bi - (a == 0), lj = bj - (a == 1), lk = bk - (a == 2);
        sb = inblocks(t.g, li, lj, lk) ? t.slot[blockid(t.g, li, lj, lk)] : 0;
        if (sb == 0) return;
        top = true;
    }
    const int32_t base = (sb - 1) * SBFACES;
    float acc = 0.0f;
    for (int q = 0; q < 4; ++q)
        for (int p = 0; p < 4; ++p) {
            int fi = (4 * I + (a == 0 ? 0 : p)) & 15;
            int fj = (4 * J + (a == 0 ? p : (a == 1 ? 0 : q))) & 15;
            int fk = (4 * K + (a == 2 ? 0 : q)) & 15;
            if (top) { if (a == 0) fi = 16; else if (a == 1) fj = 16; else fk = 16; }
            acc += (float)fine[base + ffn(16, a, fi, fj, fk)];
        }
    coarse[l] = meanof<V>(acc, 16.0f);
}

void launch_restrictcells(const SparseBasin& s, float* coarse, const float* fine) {
    hipLaunchKernelGGL(k_restrictcells<float>, dim3(cdiv(s.ncells(), 256)), dim3(256), 0, s.st.s, coarse, fine, s.topo());
    MAEL_LAUNCH_CHECK();
}
void launch_restrictfaces_f(const SparseBasin& s, float* coarse, const float* fine, int a) {
    hipLaunchKernelGGL(k_restrictfaces<float>, dim3(cdiv(s.nfaces(a), 256)), dim3(256), 0, s.st.s, coarse, fine, s.topo(), a);
    MAEL_LAUNCH_CHECK();
}
void launch_restrictfaces_u8(const SparseBasin& s, uint8_t* coarse, const uint8_t* fine, int a) {
    hipLaunchKernelGGL(k_restrictfaces<uint8_t>, dim3(cdiv(s.nfaces(a), 256)), dim3(256), 0, s.st.s, coarse, fine, s.topo(), a);
    MAEL_LAUNCH_CHECK();
}
void SparseBasin::restrict_velocity() {
    for (int a = 0; a < 3; ++a) launch_restrictfaces_f(*this, u[a].coarse.p, u[a].fine.p, a);
}

// ─── the start (sflip.jl:644-713) ────────────────────────────────────────────

__device__ inline float restphi(const float* __restrict__ eta, int neta, float rmax, float px, float py, float pz) {
    float r = sqrtf(px * px + py * py);
    float tt = jclamp(__fdiv_rn(r, rmax), 0.0f, 1.0f) * (float)(neta - 1);
    int a = min((int)tt, neta - 2);
    float f = tt - (float)a;
    return pz - (eta[a] * (1.0f - f) + eta[a + 1] * f);
}
__device__ inline float swirl(float G, float core, float px, float py, int a) {
    float r = sqrtf(px * px + py * py);
    float turn = r > 1e-4f ? __fdiv_rn(oseen_dev(G, core, r), r) : 0.0f;
    return a == 0 ? -turn * py : (a == 1 ? turn * px : 0.0f);
}

__global__ void k_initcoarse(float* __restrict__ phic, float* __restrict__ uc, float* __restrict__ vc, float* __restrict__ wc,
                             const float* __restrict__ eta, int neta, float rmax, SGrid g, float G, float core) {
    int n0 = g.nc(0), n1 = g.nc(1), n2 = g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t F = (size_t)(n0 + 1) * (n1 + 1) * (n2 + 1);
    if (l >= F) return;
    int I, J, K;
    unlin(l, n0 + 1, n1 + 1, I, J, K);
    float p[3];
    if (I < n0 && J < n1 && K < n2) {
        coarsecentre(g, I, J, K, p);
        phic[cl(I, J, K, n0, n1)] = restphi(eta, neta, rmax, p[0], p[1], p[2]);
    }
    float h = g.cdx();
    if (J < n1 && K < n2) { facepos(g, 0, I, J, K, h, p); uc[cl(I, J, K, n0 + 1, n1)] = swirl(G, core, p[0], p[1], 0); }
    if (I < n0 && K < n2) { facepos(g, 1, I, J, K, h, p); vc[cl(I, J, K, n0, n1 + 1)] = swirl(G, core, p[0], p[1], 1); }
    if (I < n0 && J < n1) { facepos(g, 2, I, J, K, h, p); wc[cl(I, J, K, n0, n1)] = 0.0f; }
}
__global__ void k_initfine(float* __restrict__ phi, float* __restrict__ uf, float* __restrict__ vf, float* __restrict__ wf,
                           const float* __restrict__ eta, int neta, float rmax, STopo t, float G, float core, int nslots) {
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)nslots * SBFACES) return;
    float p[3];
    if (n < (int64_t)nslots * SBCELLS) {
        SCell c = finecell(t, (int32_t)n);
        if (c.w.x >= 0) {
            finecentre(t.g, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, p);
            phi[n] = restphi(eta, neta, rmax, p[0], p[1], p[2]);
        }
    }
    float* U[3] = {uf, vf, wf};
    for (int a = 0; a < 3; ++a) {
        SCell c = fineface(t, a, (int32_t)n);
        if (c.w.x < 0) continue;
        facepos(t.g, a, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, t.g.dx, p);
        U[a][n] = swirl(G, core, p[0], p[1], a);
    }
}

void SparseBasin::init_coarse(const float* eta_h, int neta, float rmax, float G, float core) {
    DBuf<float> eta;
    eta.alloc(neta);
    MAEL_HIP(hipMemcpyAsync(eta.p, eta_h, neta * sizeof(float), hipMemcpyHostToDevice, st.s));
    size_t F = (size_t)(g.nc(0) + 1) * (g.nc(1) + 1) * (g.nc(2) + 1);
    hipLaunchKernelGGL(k_initcoarse, dim3(cdiv(F, 256)), dim3(256), 0, st.s, phi.coarse.p, u[0].coarse.p, u[1].coarse.p,
                       u[2].coarse.p, eta.p, neta, rmax, g, G, core);
    MAEL_LAUNCH_CHECK();
    st.sync();
}
void SparseBasin::init_fine(const float* eta_h, int neta, float rmax, float G, float core) {
    if (nslots == 0) return;
    DBuf<float> eta;
    eta.alloc(neta);
    MAEL_HIP(hipMemcpyAsync(eta.p, eta_h, neta * sizeof(float), hipMemcpyHostToDevice, st.s));
    hipLaunchKernelGGL(k_initfine, dim3(cdiv(finefaces(), 256)), dim3(256), 0, st.s, phi.fine.p, u[0].fine.p, u[1].fine.p,
                       u[2].fine.p, eta.p, neta, rmax, topo(), G, core, nslots);
    MAEL_LAUNCH_CHECK();
    st.sync();
}

// ─── classification (sflip.jl:921-948) ───────────────────────────────────────

__global__ void k_classifyfine(uint8_t* __restrict__ types, const float* __restrict__ phi, STopo t, int32_t nf) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    SCell c = finecell(t, n);
    if (c.w.x < 0) return;
    float p[3];
    finecentre(t.g, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, p);
    types[n] = sissolid(t.g, p[0], p[1], p[2]) ? SOLID : (phi[n] < 0.0f ? LIQUID : AIR);
}
__global__ void k_classifycoarse(uint8_t* __restrict__ types, const float* __restrict__ phic, const uint8_t* __restrict__ tf,
                                 STopo t) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    if (covered(t, I, J, K)) {
        bool air = false, liquid = false;
        for (int c = 0; c < 4; ++c)
            for (int b = 0; b < 4; ++b)
                for (int a = 0; a < 4; ++a) {
                    uint8_t ty = tf[fineat(t, 4 * I + a, 4 * J + b, 4 * K + c)];
                    air |= ty == AIR;
                    liquid |= ty == LIQUID;
                }
        types[l] = air ? AIR : (liquid ? LIQUID : SOLID);
    } else {
        float p[3];
        coarsecentre(t.g, I, J, K, p);
        types[l] = sissolid(t.g, p[0], p[1], p[2]) ? SOLID : (phic[l] < 0.0f ? LIQUID : AIR);
    }
}
void SparseBasin::classify() {
    if (nslots > 0)
        hipLaunchKernelGGL(k_classifyfine, dim3(cdiv(finecells(), 256)), dim3(256), 0, st.s, types.fine.p, phi.fine.p, topo(),
                           (int32_t)finecells());
    hipLaunchKernelGGL(k_classifycoarse, dim3(cdiv(ncells(), 256)), dim3(256), 0, st.s, types.coarse.p, phi.coarse.p,
                       types.fine.p, topo());
    MAEL_LAUNCH_CHECK();
}

// ─── walls (sflip.jl:1019-1041) ──────────────────────────────────────────────

// A workgroup per live slot (block order), a thread per 17 of its faces; a block clear of
// the walls has no walled face and leaves at once.
__global__ __launch_bounds__(256) void k_solidfine(float* __restrict__ u, STopo t, int a, const int32_t* __restrict__ order) {
    const int s = order[blockIdx.x];
    const int4 w = t.where[s - 1];
    if (blockclear(t.g, w)) return;
    for (int r = 0; r < SBFACES / 256; ++r) {
        const int f = r * 256 + (int)threadIdx.x;
        int i, j, k;
        if (a == 0) { i = f % 17; int q = f / 17; j = q & 15; k = q >> 4; }
        else if (a == 1) { i = f & 15; int q = f >> 4; j = q % 17; k = q / 17; }
        else { i = f & 15; j = (f >> 4) & 15; k = f >> 8; }
        float p[3];
        facepos(t.g, a, w.x * 16 + i, w.y * 16 + j, w.z * 16 + k, t.g.dx, p);
        if (walledface_dev(t.g, p, a, t.g.dx)) u[(s - 1) * SBFACES + f] = 0.0f;
    }
}
__global__ void k_solidcoarse(float* __restrict__ u, SGrid g, int a) {
    int nc0 = g.nc(0), nc1 = g.nc(1), nc2 = g.nc(2);
    int n0 = nc0 + (a == 0), n1 = nc1 + (a == 1), n2 = nc2 + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    bool outside = (a == 0 && (I == 0 || I == nc0)) || (a == 1 && (J == 0 || J == nc1)) || (a == 2 && (K == 0 || K == nc2));
    float p[3];
    float h = g.cdx();
    facepos(g, a, I, J, K, h, p);
    if (outside || walledface_dev(g, p, a, h)) u[l] = 0.0f;
}
void SparseBasin::solidify() {
    for (int a = 0; a < 3; ++a) {
        if (nlive > 0)
            hipLaunchKernelGGL(k_solidfine, dim3((unsigned)nlive), dim3(256), 0, st.s, u[a].fine.p, topo(), a, sorder.p);
        hipLaunchKernelGGL(k_solidcoarse, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, st.s, u[a].coarse.p, g, a);
    }
    MAEL_LAUNCH_CHECK();
}

// The gravity part of forces! (sflip.jl:992-1016 with no stirring): w -= g dt where the
// face is out of the walls, then the walls and the coarse faces under fine blocks.
__global__ void k_gravityfine(float* __restrict__ w, STopo t, float dv, int32_t nf) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    SCell c = fineface(t, 2, n);
    if (c.w.x < 0) return;
    float p[3];
    facepos(t.g, 2, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, t.g.dx, p);
    if (!sissolid(t.g, p[0], p[1], p[2])) w[n] += dv;
}
__global__ void k_gravitycoarse(float* __restrict__ w, SGrid g, float dv) {
    int n0 = g.nc(0), n1 = g.nc(1), n2 = g.nc(2) + 1;
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    float p[3];
    facepos(g, 2, I, J, K, g.cdx(), p);
    if (!sissolid(g, p[0], p[1], p[2])) w[l] += dv;
}
void SparseBasin::gravity(float dt, float gacc) {
    float dv = dt * (0.0f - gacc);
    if (nslots > 0)
        hipLaunchKernelGGL(k_gravityfine, dim3(cdiv(finefaces(), 256)), dim3(256), 0, st.s, u[2].fine.p, topo(), dv,
                           (int32_t)finefaces());
    hipLaunchKernelGGL(k_gravitycoarse, dim3(cdiv(nfaces(2), 256)), dim3(256), 0, st.s, u[2].coarse.p, g, dv);
    MAEL_LAUNCH_CHECK();
    solidify();
    restrict_velocity();
}

// ─── redistancing (sflip.jl:833-919) ─────────────────────────────────────────

// A workgroup per slot: the block's 16^3 values and the layers of the six blocks across
// its sides (the coarse cell where there is no block, as in the reference's nphi) are
// staged in LDS once, then each thread takes a 16-cell column along z. A block clear of
// the walls skips the solid tests.
__global__ __launch_bounds__(256) void k_sreinit(float* __restrict__ out, const float* __restrict__ phi,
                                                 const float* __restrict__ phic, STopo t, float phimax,
                                                 const int32_t* __restrict__ order) {
    __shared__ int32_t s27[27];
    __shared__ float tv[4096];       // the block's cells, x fastest
    __shared__ float hv[6][256];     // neighbour m's layer at (p, q), the other two axes in order
    const int32_t slot = order[blockIdx.x] - 1;   // live slots in block order
    const int4 w = t.where[slot];
    stage27(t, w, s27);
    const SGrid& g = t.g;
    const int tid = (int)threadIdx.x, ti = tid & 15, tj = tid >> 4;
    const float* __restrict__ bp = phi + (size_t)slot * 4096;
    float own[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) own[k] = bp[tid + 256 * k];
    __syncthreads();
    const int n0 = g.nc(0), n1 = g.nc(1), n2 = g.nc(2);
    float hl[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {   // entry (p, q) = (ti, tj) of each side's layer
        const int a = m >> 1, e = NBS[m][a] > 0 ? 16 : -1;
        const int li = a == 0 ? e : ti, lj = a == 1 ? e : (a == 0 ? ti : tj), lk = a == 2 ? e : tj;
        const int32_t sn = s27[((li >> 4) + 1) + 3 * ((lj >> 4) + 1) + 9 * ((lk >> 4) + 1)];
        const int I = w.x * 16 + li, J = w.y * 16 + lj, K = w.z * 16 + lk;
        const size_t ci = cl(min(max(I >> 2, 0), n0 - 1), min(max(J >> 2, 0), n1 - 1), min(max(K >> 2, 0), n2 - 1), n0, n1);
        const size_t fi = (size_t)(sn > 0 ? sn - 1 : 0) * 4096 + (li & 15) + 16 * ((lj & 15) + 16 * (lk & 15));
        hl[m] = sn != 0 ? phi[fi] : phic[ci];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) tv[tid + 256 * k] = own[k];
#pragma unroll
    for (int m = 0; m < 6; ++m) hv[m][tid] = hl[m];
    const bool clear = blockclear(g, w);
    // a cell whose three upwind distances are all phimax (deep in the water or the air)
    // takes this one value: the same function of the same arguments
    const float dsat = jmin(eikonal(phimax, phimax, phimax, g.dx), phimax);
    __syncthreads();
    float* __restrict__ op = out + (size_t)slot * 4096;
    for (int k = 0; k < 16; ++k) {
        const int l = tid + 256 * k;
        const int I = w.x * 16 + ti, J = w.y * 16 + tj, K = w.z * 16 + k;
        const float cv = tv[l], h = g.dx;
        float nv[6];   // in the order of NBS: +x, -x, +y, -y, +z, -z
        nv[0] = ti < 15 ? tv[l + 1] : hv[0][tj + 16 * k];
        nv[1] = ti > 0 ? tv[l - 1] : hv[1][tj + 16 * k];
        nv[2] = tj < 15 ? tv[l + 16] : hv[2][ti + 16 * k];
        nv[3] = tj > 0 ? tv[l - 16] : hv[3][ti + 16 * k];
        nv[4] = k < 15 ? tv[l + 256] : hv[4][tid];
        nv[5] = k > 0 ? tv[l - 256] : hv[5][tid];
        bool solid = false, nsolid[6] = {false, false, false, false, false, false};
        if (!clear) {
            float p[3];
            finecentre(g, I, J, K, p);
            solid = sissolid(g, p[0], p[1], p[2]);
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                finecentre(g, I + NBS[m][0], J + NBS[m][1], K + NBS[m][2], p);
                nsolid[m] = sissolid(g, p[0], p[1], p[2]);
            }
        }
        float res;
        if (solid) {
            float lowest = cv;
            bool found = false;
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                if (!nsolid[m]) {
                    float f = nv[m];
                    lowest = found ? jmin(lowest, f) : f;
                    found = true;
                }
            }
            res = found ? lowest : jclamp(cv, -phimax, phimax);
        } else {
            bool inside = cv < 0.0f, frozen = false;
            float ax = phimax, ay = phimax, az = phimax;
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                if (!nsolid[m]) {
                    float f = nv[m];
                    if ((f < 0.0f) != inside) frozen = true;
                    float d = fabsf(f);
                    if (m < 2) ax = jmin(ax, d);
                    else if (m < 4) ay = jmin(ay, d);
                    else az = jmin(az, d);
                }
            }
            if (frozen) res = cv;
            else {
                const bool sat = ax == phimax && ay == phimax && az == phimax;
                float d = dsat;
                if (!sat) d = jmin(eikonal(ax, ay, az, h), phimax);
                res = inside ? -d : d;
            }
        }
        op[l] = res;
    }
}
__global__ void k_coarsewall(float* __restrict__ out, const float* __restrict__ phic, STopo t) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    float c = phic[l], p[3];
    coarsecentre(t.g, I, J, K, p);
    if (!covered(t, I, J, K) && sissolid(t.g, p[0], p[1], p[2])) {
        float lowest = c;
        bool found = false;
        for (int m = 0; m < 6; ++m) {
            int a = I + NBS[m][0], b = J + NBS[m][1], cc = K + NBS[m][2];
            if (a < 0 || b < 0 || cc < 0 || a >= n0 || b >= n1 || cc >= n2) continue;
            coarsecentre(t.g, a, b, cc, p);
            if (sissolid(t.g, p[0], p[1], p[2])) continue;
            float f = phic[cl(a, b, cc, n0, n1)];
            lowest = found ? jmin(lowest, f) : f;
            found = true;
        }
        out[l] = lowest;
    } else {
        out[l] = c;
    }
}
void SparseBasin::redistance(int iterations) {
    float phimax = (float)(iterations + 1) * g.dx;
    hipStream_t s = st.s;
    for (int it = 0; it < iterations && nslots > 0; ++it) {
        if (nlive > 0)   // free slots keep what they hold in both copies: nothing reads them
            hipLaunchKernelGGL(k_sreinit, dim3((unsigned)nlive), dim3(256), 0, s, phi0.fine.p, phi.fine.p, phi.coarse.p,
                               topo(), phimax, sorder.p);
        std::swap(phi.fine, phi0.fine);
    }
    hipLaunchKernelGGL(k_coarsewall, dim3(cdiv(ncells(), 256)), dim3(256), 0, s, phi0.coarse.p, phi.coarse.p, topo());
    std::swap(phi.coarse, phi0.coarse);
    MAEL_LAUNCH_CHECK();
    launch_restrictcells(*this, phi.coarse.p, phi.fine.p);
}

// ─── copies and the volume ───────────────────────────────────────────────────

void SparseBasin::keep() {   // phi0 = phi, u0 = u (sflip.jl:700-703)
    hipStream_t s = st.s;
    MAEL_HIP(hipMemcpyAsync(phi0.fine.p, phi.fine.p, finecells() * sizeof(float), hipMemcpyDeviceToDevice, s));
    MAEL_HIP(hipMemcpyAsync(phi0.coarse.p, phi.coarse.p, ncells() * sizeof(float), hipMemcpyDeviceToDevice, s));
    for (int a = 0; a < 3; ++a) {
        MAEL_HIP(hipMemcpyAsync(u0[a].fine.p, u[a].fine.p, finefaces() * sizeof(float), hipMemcpyDeviceToDevice, s));
        MAEL_HIP(hipMemcpyAsync(u0[a].coarse.p, u[a].coarse.p, nfaces(a) * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
}
void SparseBasin::fill_flags(uint8_t v) {
    for (int a = 0; a < 3; ++a) {
        fill_async(fl[a].fine.p, (size_t)blocks * SBFACES, v, st.s);
        fill_async(fl[a].coarse.p, nfaces(a), v, st.s);
    }
}

__device__ inline double block_sum_dd(double v, double* lds) {
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
// sflip.jl:1975-1996: fine cells and the coarse cells outside fine blocks (x64), in FP64.
__global__ __launch_bounds__(256) void k_svolume(double* __restrict__ partials, const float* __restrict__ phi,
                                                 const float* __restrict__ phic, STopo t, int32_t nf) {
    __shared__ double lds[4];
    size_t Cc = (size_t)t.g.nc(0) * t.g.nc(1) * t.g.nc(2);
    double v = 0;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < (size_t)nf + Cc; l += (size_t)gridDim.x * blockDim.x) {
        float p[3];
        if (l < (size_t)nf) {
            SCell c = finecell(t, (int32_t)l);
            if (c.w.x < 0) continue;
            finecentre(t.g, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, p);
            if (!sissolid(t.g, p[0], p[1], p[2])) v += (double)jclamp(0.5f - __fdiv_rn(phi[l], t.g.dx), 0.0f, 1.0f);
        } else {
            size_t m = l - nf;
            int n0 = t.g.nc(0), n1 = t.g.nc(1);
            int I, J, K;
            unlin(m, n0, n1, I, J, K);
            coarsecentre(t.g, I, J, K, p);
            if (covered(t, I, J, K) || sissolid(t.g, p[0], p[1], p[2])) continue;
            v += 64.0 * (double)jclamp(0.5f - __fdiv_rn(phic[m], t.g.cdx()), 0.0f, 1.0f);
        }
    }
    double s = block_sum_dd(v, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
double SparseBasin::volume() {
    const int NBLK = 1024;
    hipLaunchKernelGGL(k_svolume, dim3(NBLK), dim3(256), 0, st.s, partials.p, phi.fine.p, phi.coarse.p, topo(),
                       (int32_t)finecells());
    MAEL_LAUNCH_CHECK();
    std::vector<double> h(NBLK);
    MAEL_HIP(hipMemcpyAsync(h.data(), partials.p, NBLK * sizeof(double), hipMemcpyDeviceToHost, st.s));
    st.sync();
    double acc = 0;
    for (double x : h) acc += x;
    double dx = g.dx;
    return acc * dx * dx * dx;
}

}  // namespace mael
