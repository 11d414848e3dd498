// The bake writers' device side (flip.jl:1846-1951, sflip.jl:2419-2440, bakefiles.jl):
// the band's particles kept for drawing, packed as Float16 on the device so only what is
// kept crosses to the host (`packband_kernel!`), and each frame's level set averaged and
// quantised to bytes (`quantize_kernel!`, `sparsequantize_kernel!`).
//
// The reference packs kept particles at slots taken by an atomic counter, so their order
// depends on scheduling; here the kept particles are compacted in particle order (a count
// per workgroup, a scan of the counts, then each workgroup writes its kept particles at its
// offset in lane order), which is deterministic and is the reference's order whenever its
// atomics resolve in thread order.
#include "dense.hpp"
#include "scan.hpp"
#include "sparse.hpp"

namespace mael {

namespace {

constexpr int PACK_THREADS = 256, PACK_ITEMS = 8, PACK_TILE = PACK_THREADS * PACK_ITEMS;

// flip.jl:1887-1889: the particle at the frame's time (ST-FLIP's re-synchronisation:
// its lag on) and whether it is kept (the top `keep` metres, or within `throat` of the axis).
struct BandKeep {
    float keep, throat;
    __device__ bool operator()(const PView& ps, int64_t p, float xp[3]) const {
        const float lag = ps.f[15][p];
        xp[0] = ps.f[0][p] + lag * ps.f[3][p];
        xp[1] = ps.f[1][p] + lag * ps.f[4][p];
        xp[2] = ps.f[2][p] + lag * ps.f[5][p];
        return xp[2] > -keep || xp[0] * xp[0] + xp[1] * xp[1] < throat * throat;
    }
};

// Exclusive prefix of `flag` over the workgroup, in thread order; `total` the count.
__device__ inline int wg_prefix(int flag, int* lds, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long ball = __ballot(flag);
    const int below = __popcll(ball & ((1ull << lane) - 1ull));
    if (lane == 0) lds[wave] = __popcll(ball);
    __syncthreads();
    int base = 0;
    total = 0;
    for (int w = 0; w < PACK_THREADS / 64; ++w) {
        if (w < wave) base += lds[w];
        total += lds[w];
    }
    __syncthreads();
    return base + below;
}

}  // namespace

// Per workgroup tile of PACK_TILE particles: how many are kept.
__global__ __launch_bounds__(PACK_THREADS) void k_band_count(PView ps, int64_t n, BandKeep K,
                                                             unsigned long long* __restrict__ counts) {
    __shared__ int lds[PACK_THREADS / 64];
    int mine = 0;
    const int64_t base = (int64_t)blockIdx.x * PACK_TILE;
    for (int e = 0; e < PACK_ITEMS; ++e) {
        const int64_t p = base + (int64_t)e * PACK_THREADS + threadIdx.x;
        float xp[3];
        if (p < n && K(ps, p, xp)) ++mine;
    }
    int total = 0;
    for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < PACK_THREADS / 64; ++w) total += lds[w];
        counts[blockIdx.x] = (unsigned long long)total;
    }
}

// Each workgroup's kept particles, in particle order, from its offset: positions at the
// frame's time and velocities as Float16 (round to nearest even, as Julia's Float16()).
__global__ __launch_bounds__(PACK_THREADS) void k_band_pack(PView ps, int64_t n, BandKeep K,
                                                            const unsigned long long* __restrict__ offsets,
                                                            __half* __restrict__ pos, __half* __restrict__ vel) {
    __shared__ int lds[PACK_THREADS / 64];
    const int64_t base = (int64_t)blockIdx.x * PACK_TILE;
    int64_t at = (int64_t)offsets[blockIdx.x];
    for (int e = 0; e < PACK_ITEMS; ++e) {
        const int64_t p = base + (int64_t)e * PACK_THREADS + threadIdx.x;
        float xp[3] = {0.0f, 0.0f, 0.0f};
        const int kept = p < n && K(ps, p, xp);
        int total;
        const int m = wg_prefix(kept, lds, total);
        if (kept) {
            const int64_t d = at + m;
            pos[3 * d] = __float2half_rn(xp[0]);
            pos[3 * d + 1] = __float2half_rn(xp[1]);
            pos[3 * d + 2] = __float2half_rn(xp[2]);
            vel[3 * d] = __float2half_rn(ps.f[3][p]);
            vel[3 * d + 1] = __float2half_rn(ps.f[4][p]);
            vel[3 * d + 2] = __float2half_rn(ps.f[5][p]);
        }
        at += total;
    }
}

// The packing on the basin's stream into device buffers (grown to the particle count on
// first use), then the kept columns to the host. Returns how many were kept.
int64_t pack_band(const PView& ps, int64_t n, float keep, float throat, DBuf<__half>& dpos, DBuf<__half>& dvel,
                  DBuf<unsigned long long>& counts, uint16_t* hpos, uint16_t* hvel, Stream& st) {
    hipStream_t s = st.s;
    const int nb = (int)cdiv(n, PACK_TILE);
    if (counts.n < (size_t)nb + 2) counts.alloc((size_t)nb + 2);
    if (dpos.n < (size_t)3 * n) {
        dpos.alloc((size_t)3 * std::max<int64_t>(n, 1));
        dvel.alloc((size_t)3 * std::max<int64_t>(n, 1));
    }
    BandKeep K{keep, throat};
    if (nb > 0) hipLaunchKernelGGL(k_band_count, dim3(nb), dim3(PACK_THREADS), 0, s, ps, n, K, counts.p);
    launch_scan_bsums(counts.p, nb, s);   // exclusive offsets; counts[nb] the total
    MAEL_HIP(hipMemcpyAsync(st.pinned, counts.p + nb, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    if (nb > 0) hipLaunchKernelGGL(k_band_pack, dim3(nb), dim3(PACK_THREADS), 0, s, ps, n, K, counts.p, dpos.p, dvel.p);
    MAEL_LAUNCH_CHECK();
    st.sync();
    const int64_t m = (int64_t)*(unsigned long long*)st.pinned;
    if (m > 0) {
        MAEL_HIP(hipMemcpyAsync(hpos, dpos.p, (size_t)m * 3 * sizeof(__half), hipMemcpyDeviceToHost, s));
        MAEL_HIP(hipMemcpyAsync(hvel, dvel.p, (size_t)m * 3 * sizeof(__half), hipMemcpyDeviceToHost, s));
        st.sync();
    }
    return m;
}

// flip.jl:1859-1867 quantize_kernel!: the level set averaged over cells twice as wide,
// clamped to +-range, as a byte (0 the deepest, 255 the highest; round to nearest even).
__global__ void k_quantize(uint8_t* __restrict__ q, const float* __restrict__ phi, int3 nf, int3 nq, float range) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y, K = blockIdx.z;
    if (I >= nq.x || J >= nq.y || K >= nq.z) return;
    float s = 0.0f;
    for (int c = 0; c < 2; ++c)
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a)
                s += phi[(size_t)(2 * I + a) + (size_t)nf.x * ((size_t)(2 * J + b) + (size_t)nf.y * (2 * K + c))];
    const float f = jclamp(__fdiv_rn(0.125f * s, range), -1.0f, 1.0f);
    q[(size_t)I + (size_t)nq.x * ((size_t)J + (size_t)nq.y * K)] = (uint8_t)(int)rintf(127.5f * (f + 1.0f));
}

// sflip.jl:2427-2432 sparsequantize_kernel!: the level set sampled at the centres of cells
// `cell` wide (fine where the taps are fine), clamped to +-range, as a byte.
__global__ void k_squantize(uint8_t* __restrict__ q, const float* __restrict__ phif, const float* __restrict__ phic, STopo t,
                            int3 nq, float cell, float range) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= nq.x || j >= nq.y || k >= nq.z) return;
    // p = lo + (Vec3f(i, j, k) - 0.5) * cell with the reference's 1-based (i, j, k)
    const float px = t.g.lo[0] + ((float)(i + 1) - 0.5f) * cell, py = t.g.lo[1] + ((float)(j + 1) - 0.5f) * cell,
                pz = t.g.lo[2] + ((float)(k + 1) - 0.5f) * cell;
    const float f = jclamp(__fdiv_rn(cellsample(t, phif, phic, px, py, pz), range), -1.0f, 1.0f);
    q[(size_t)i + (size_t)nq.x * ((size_t)j + (size_t)nq.y * k)] = (uint8_t)(int)rintf(127.5f * (f + 1.0f));
}

int64_t DenseBasin::pack_band(float keep, float throat, uint16_t* pos, uint16_t* vel) {
    return mael::pack_band(view(cur), n, keep, throat, bake_pos, bake_vel, bake_counts, pos, vel, st);
}
void DenseBasin::quantize_levelset(float range, uint8_t* out) {
    int3 nf = make_int3(g.n[0], g.n[1], g.n[2]), nq = make_int3(g.n[0] / 2, g.n[1] / 2, g.n[2] / 2);
    const size_t N = (size_t)nq.x * nq.y * nq.z;
    if (bake_q.n < N) bake_q.alloc(N);
    hipLaunchKernelGGL(k_quantize, dim3(cdiv(nq.x, 64), cdiv(nq.y, 4), nq.z), dim3(64, 4), 0, st.s, bake_q.p, phi.p, nf, nq, range);
    MAEL_LAUNCH_CHECK();
    MAEL_HIP(hipMemcpyAsync(out, bake_q.p, N, hipMemcpyDeviceToHost, st.s));
    st.sync();
}

int64_t SparseBasin::pack_band(float keep, float throat, uint16_t* pos, uint16_t* vel) {
    return mael::pack_band(view(cur), n, keep, throat, bake_pos, bake_vel, bake_counts, pos, vel, st);
}
void SparseBasin::quantize_levelset(float range, uint8_t* out) {
    int3 nq = make_int3(g.nb[0] * 8, g.nb[1] * 8, g.nb[2] * 8);
    const size_t N = (size_t)nq.x * nq.y * nq.z;
    if (bake_q.n < N) bake_q.alloc(N);
    hipLaunchKernelGGL(k_squantize, dim3(cdiv(nq.x, 64), cdiv(nq.y, 4), nq.z), dim3(64, 4), 0, st.s, bake_q.p, phi.fine.p,
                       phi.coarse.p, topo(), nq, 2.0f * g.dx, range);
    MAEL_LAUNCH_CHECK();
    MAEL_HIP(hipMemcpyAsync(out, bake_q.p, N, hipMemcpyDeviceToHost, st.s));
    st.sync();
}

}  // namespace mael
