// Resampling of the narrow band on the dense grid (flip.jl:1254-1382): bin each
// particle to its cell, count, scan per-cell (count, total) pairs in one pass, give
// every particle a stable rank in its cell, scatter the survivors in cell order and
// spawn new particles where a band cell has too few.
//
// Determinism: the reference ranks particles by atomic arrival (flip.jl:1282), so its
// in-cell order and the survivors of the 2*nmin cap depend on scheduling. Here the
// atomic only sizes the cell; the rank is the particle's position among the cell's
// particles in index order (a counting sort whose segments are then put in index
// order), so keys, ranks, counts, totals, offsets and the resampled arrays are the
// same on every run.
#include "dense.hpp"
#include "scan.hpp"
#include <mutex>

namespace mael {

// ─── per-cell counts (flip.jl:1289-1297) ─────────────────────────────────────

__device__ inline void cellcounts(int32_t count, float phic, bool solid, int32_t nmin, float deep, float dx, bool seeding,
                                  int32_t& kept, int32_t& add) {
    kept = min(count, 2 * nmin);
    bool inband = !solid && -deep <= phic && phic <= (seeding ? 0.5f * dx : -dx);
    add = (inband && kept < nmin) ? nmin - kept : 0;
}

struct DenseCells {   // what the scan needs to know about dense cell l
    const int32_t* counts;
    const float* phi;
    DGrid g;
    int32_t nmin;
    float deep;
    int seeding;
    __device__ void get(size_t l, int32_t& count, int32_t& total) const {
        int n1 = g.n[0], n2 = g.n[1];
        int i, j, k;
        unlin(l, n1, n2, i, j, k);
        count = counts[l];
        int32_t kept, add;
        cellcounts(count, phi[l], issolidcell(g, i, j, k), nmin, deep, g.dx, seeding != 0, kept, add);
        total = kept + add;
    }
};

// ─── binning (flip.jl:1258-1287) ─────────────────────────────────────────────

__global__ void k_dense_bin(PView ps, int64_t n, const float* __restrict__ phi, DGrid g, float deep, float above,
                            int32_t* __restrict__ key, int32_t* __restrict__ rank, int32_t* __restrict__ counts) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float xp = ps.f[0][p], yp = ps.f[1][p], zp = ps.f[2][p];
    float c0 = __fdiv_rn(xp - g.lo[0], g.dx), c1 = __fdiv_rn(yp - g.lo[1], g.dx), c2 = __fdiv_rn(zp - g.lo[2], g.dx);
    bool keep = isfinite(c0) && isfinite(c1) && isfinite(c2);
    // floor(c) in [0, n) is c in [0, n) for finite c
    keep = keep && c0 >= 0.0f && c0 < (float)g.n[0] && c1 >= 0.0f && c1 < (float)g.n[1] && c2 >= 0.0f &&
           c2 < (float)g.n[2] && !issolid(g.radius, g.floor, xp, yp, zp);
    if (keep) {
        float f = trilinear(phi, g.n[0], g.n[1], g.n[2], c0 - 0.5f, c1 - 0.5f, c2 - 0.5f);
        keep = -deep <= f && f < above;
    }
    const int32_t l = keep ? (int32_t)floorf(c0) + g.n[0] * ((int32_t)floorf(c1) + g.n[1] * (int32_t)floorf(c2)) : -1;
    key[p] = l;
    rank[p] = wave_count_slot(counts, l);   // arrival: only sizes the segment
}

// Each kept particle into its cell's segment, at its arrival slot.
__global__ void k_place(int64_t n, const int32_t* __restrict__ key, const int32_t* __restrict__ rank,
                        const int32_t* __restrict__ cstart, int32_t* __restrict__ seg) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int32_t l = key[p];
    if (l >= 0) seg[cstart[l] + rank[p]] = (int32_t)p;
}

// Each cell's particles in index order; the rank is the position there. Up to 16 (nearly
// every cell: nmin is 8, and 2 nmin survive) are sorted in registers by a bitonic network;
// a more crowded cell is sorted in place by insertion (its segment is nearly sorted: the
// previous resample wrote particles cell by cell, and the binning numbers a wavefront's
// particles of one cell in lane order).
__global__ void k_fixrank(size_t C, const int32_t* __restrict__ counts, const int32_t* __restrict__ cstart,
                          int32_t* __restrict__ seg, int32_t* __restrict__ rank) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= C) return;
    const int32_t c = counts[l];
    if (c < 2) return;
    int32_t* s = seg + cstart[l];
    if (c <= 16) {
        int32_t v[16];
#pragma unroll
        for (int a = 0; a < 16; ++a) v[a] = a < c ? s[a] : INT32_MAX;
#pragma unroll
        for (int k = 2; k <= 16; k <<= 1)
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int m = i ^ j;
                    if (m > i) {
                        const int32_t x = v[i], y = v[m];
                        const bool up = (i & k) == 0;
                        v[i] = up ? min(x, y) : max(x, y);
                        v[m] = up ? max(x, y) : min(x, y);
                    }
                }
#pragma unroll
        for (int a = 0; a < 16; ++a)
            if (a < c) rank[v[a]] = a;
        return;
    }
    for (int32_t a = 1; a < c; ++a) {
        int32_t v = s[a];
        int32_t b = a - 1;
        while (b >= 0 && s[b] > v) { s[b + 1] = s[b]; --b; }
        s[b + 1] = v;
    }
    for (int32_t a = 0; a < c; ++a) rank[s[a]] = a;
}

// flip.jl:1307-1319 scatter_kernel!: survivors to their place in cell order.
__global__ void k_scatter(PView src, PView dst, int64_t n, const int32_t* __restrict__ key,
                          const int32_t* __restrict__ rank, const int32_t* __restrict__ totals,
                          const int32_t* __restrict__ offsets, int32_t nmax) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int32_t l = key[p];
    int32_t r = rank[p];
    if (l < 0 || r >= nmax) return;
    int64_t d = (int64_t)offsets[l] - totals[l] + r;
#pragma unroll
    for (int f = 0; f < PFIELDS; ++f) dst.f[f][d] = src.f[f][p];
}

// flip.jl:1321-1354 spawn_kernel!: stratified, jittered new particles carrying the grid's
// APIC velocity; when seeding, those the level set puts in the air get NaN positions.
__global__ void k_dense_spawn(PView dst, const int32_t* __restrict__ counts, const int32_t* __restrict__ totals,
                              const int32_t* __restrict__ offsets, const float* __restrict__ phi,
                              const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                              DGrid g, int32_t nmin, float deep, int seeding, uint32_t seed) {
    size_t C = (size_t)g.n[0] * g.n[1] * g.n[2];
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= C) return;
    int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2];
    int i, j, k;
    unlin(l, n1, n2, i, j, k);
    int32_t kept, add;
    cellcounts(counts[l], phi[l], issolidcell(g, i, j, k), nmin, deep, g.dx, seeding != 0, kept, add);
    if (add <= 0) return;
    int64_t base = (int64_t)offsets[l] - totals[l] + kept;
    float cx = g.lo[0] + (float)i * g.dx, cy = g.lo[1] + (float)j * g.dx, cz = g.lo[2] + (float)k * g.dx;
    uint32_t l1 = (uint32_t)(l + 1);   // the reference's 1-based linear index
    int32_t s = (int32_t)seed;
    float half = 0.5f * g.dx;
    for (int32_t m = 0; m < add; ++m) {
        int32_t h = (int32_t)(l1 * 31u + (uint32_t)m);
        float jx = hash01(h, s + 1), jy = hash01(h, s + 2), jz = hash01(h, s + 3);
        int32_t o = m % 8;
        float xp = cx + ((float)(o & 1) + jx) * half;
        float yp = cy + ((float)((o >> 1) & 1) + jy) * half;
        float zp = cz + ((float)((o >> 2) & 1) + jz) * half;
        int64_t d = base + m;
        bool air = false;
        if (seeding)
            air = trilinear(phi, n1, n2, n3, ispace(xp, g.lo[0], g.dx, 0.5f), ispace(yp, g.lo[1], g.dx, 0.5f),
                            ispace(zp, g.lo[2], g.dx, 0.5f)) >= 0.0f;
        // The record is built in registers and every field stored unconditionally: a
        // branch that picks which field a store goes to lets the compiler merge the two
        // branches' tail stores through one scalar base pointer chosen by a divergent
        // condition, which is wrong for the lanes of the other branch.
        float rec[PFIELDS];
        if (air) {
            float nan = __builtin_nanf("");
            rec[0] = nan; rec[1] = nan; rec[2] = nan;
#pragma unroll
            for (int f = 3; f < PFIELDS; ++f) rec[f] = 0.0f;
        } else {
            rec[0] = xp; rec[1] = yp; rec[2] = zp;
            gather(u, n1 + 1, n2, n3, g, 0, xp, yp, zp, rec[3], rec[6], rec[7], rec[8]);
            gather(v, n1, n2 + 1, n3, g, 1, xp, yp, zp, rec[4], rec[9], rec[10], rec[11]);
            gather(w, n1, n2, n3 + 1, g, 2, xp, yp, zp, rec[5], rec[12], rec[13], rec[14]);
            rec[15] = 0.0f;
        }
#pragma unroll
        for (int f = 0; f < PFIELDS; ++f) dst.f[f][d] = rec[f];
    }
}

// ─── the scan's tile sums without allocation ─────────────────────────────────
// For callers that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops): mael_dense_resample_noalloc keeps the packed scan's tile sums in this array
// instead of the basin's buffer. 16,386 entries (128 KiB) hold the 16,384 tiles of
// 1024 cells of grids up to 16.7 M cells, plus the grand total. In host code `scratch`
// names a host-side stand-in, not a device address: each device holds its own copy, whose
// address hipGetSymbolAddress returns once per device (noalloc_setup, run by mael_create
// with the device current), and only that address is passed to kernels. Every call on a
// device shares its copy: the scan writes every entry it reads before reading it, so
// nothing needs clearing, but two such calls in flight at once on one device race.
static __device__ __align__(256) unsigned long long scratch[NOALLOC_TILE_SUMS];
constexpr int NOALLOC_DEVICES = 64;
static std::mutex noalloc_mu;
static unsigned long long* noalloc_ptr[NOALLOC_DEVICES] = {};

void noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_DEVICES) return;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (noalloc_ptr[device]) return;
    void* p = nullptr;
    MAEL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    noalloc_ptr[device] = (unsigned long long*)p;
}
unsigned long long* noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    return noalloc_ptr[device];
}

int64_t DenseBasin::resample(bool seeding, uint32_t seed, unsigned long long* tilesums) {
    if (!tilesums) tilesums = bsums.p;
    hipStream_t s = st.s;
    size_t C = cells();
    float deep = (float)band * g.dx;
    fill_async(counts.p, C, 0, s);
    PView src = view(cur), dst = view(cur ^ 1);
    if (n > 0) {
        kt.begin(K_BIN, s);
        hipLaunchKernelGGL(k_dense_bin, dim3(cdiv(n, 256)), dim3(256), 0, s, src, n, phi.p, g, deep, 3.0f * g.dx, key.p,
                           rank.p, counts.p);
        kt.end(K_BIN, s);
        MAEL_LAUNCH_CHECK();
    }
    DenseCells cellsinfo{counts.p, phi.p, g, nmin, deep, seeding ? 1 : 0};
    int64_t total = scan_cells(cellsinfo, C, totals.p, offsets.p, cstart.p, tilesums, (unsigned long long*)st.pinned, s);
    if (total > capacity)
        throw Error(E_CAPACITY, "the narrow band needs " + std::to_string(total) + " particles, more than the " +
                                    std::to_string(capacity) + " allocated");
    if (n > 0) {
        hipLaunchKernelGGL(k_place, dim3(cdiv(n, 256)), dim3(256), 0, s, n, key.p, rank.p, cstart.p, seg.p);
        kt.begin(K_FIXRANK, s);
        hipLaunchKernelGGL(k_fixrank, dim3(cdiv(C, 256)), dim3(256), 0, s, C, counts.p, cstart.p, seg.p, rank.p);
        kt.end(K_FIXRANK, s);
        kt.begin(K_SCATTER, s);
        hipLaunchKernelGGL(k_scatter, dim3(cdiv(n, 256)), dim3(256), 0, s, src, dst, n, key.p, rank.p, totals.p,
                           offsets.p, 2 * nmin);
        kt.end(K_SCATTER, s);
        MAEL_LAUNCH_CHECK();
    }
    kt.begin(K_SPAWN, s);
    hipLaunchKernelGGL(k_dense_spawn, dim3(cdiv(C, 256)), dim3(256), 0, s, dst, counts.p, totals.p, offsets.p, phi.p,
                       u[0].p, u[1].p, u[2].p, g, nmin, deep, seeding ? 1 : 0, seed);
    kt.end(K_SPAWN, s);
    MAEL_LAUNCH_CHECK();
    cur ^= 1;
    n = total;
    return total;
}

}  // namespace mael
