// The per-cell scan of resampling, for the dense and the sparse grid: each cell's
// (count, total) packed in one 64-bit word (count in the high half, total in the low)
// and scanned in one pass, so the counting sort's segment starts and the inclusive
// offsets the reference keeps (`AK.accumulate!`, flip.jl:1371) come out together.
// Reduce-then-scan over tiles of 1024 cells: per-tile sums, one workgroup scans them,
// each tile then scans itself from its prefix.
#pragma once
#include "common.hpp"

namespace mael {

constexpr int SCAN_THREADS = 256, SCAN_ITEMS = 4, SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ inline unsigned long long pack_ct(int32_t count, int32_t total) {
    return ((unsigned long long)(uint32_t)count << 32) | (unsigned long long)(uint32_t)total;
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Exclusive scan of one value per thread over a 256-thread block; `tot` gets the block sum.
__device__ inline unsigned long long block_excl_scan(unsigned long long v, unsigned long long* lds,
                                                     unsigned long long& tot) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    unsigned long long inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        unsigned long long t = __shfl_up(inc, (unsigned)o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    unsigned long long pre = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        unsigned long long x = lds[w];
        if (w < wid) pre += x;
        all += x;
    }
    __syncthreads();
    tot = all;
    return pre + inc - v;
}

template <class Info>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(Info info, size_t C, unsigned long long* __restrict__ bsums) {
    __shared__ unsigned long long lds[SCAN_THREADS / 64];
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    unsigned long long s = 0;
#pragma unroll
    for (int e = 0; e < SCAN_ITEMS; ++e) {
        size_t l = base + e;
        if (l < C) {
            int32_t c, t;
            info.get(l, c, t);
            s += pack_ct(c, t);
        }
    }
    s = wave_sum_u64(s);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0;
        for (int w = 0; w < SCAN_THREADS / 64; ++w) a += lds[w];
        bsums[blockIdx.x] = a;
    }
}

// One workgroup: exclusive scan of the tile sums in place; bsums[nb] = the grand total.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(unsigned long long* __restrict__ bsums, int nb);

template <class Info>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(Info info, size_t C, const unsigned long long* __restrict__ bsums,
                                                            int32_t* __restrict__ totals, int32_t* __restrict__ offsets,
                                                            int32_t* __restrict__ cstart) {
    __shared__ unsigned long long lds[SCAN_THREADS / 64];
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    int32_t cs[SCAN_ITEMS], ts[SCAN_ITEMS];
    unsigned long long mine = 0;
#pragma unroll
    for (int e = 0; e < SCAN_ITEMS; ++e) {
        size_t l = base + e;
        cs[e] = 0; ts[e] = 0;
        if (l < C) info.get(l, cs[e], ts[e]);
        mine += pack_ct(cs[e], ts[e]);
    }
    unsigned long long tot;
    unsigned long long run = bsums[blockIdx.x] + block_excl_scan(mine, lds, tot);
#pragma unroll
    for (int e = 0; e < SCAN_ITEMS; ++e) {
        size_t l = base + e;
        if (l < C) {
            int32_t lo = (int32_t)(uint32_t)(run & 0xffffffffull);
            cstart[l] = (int32_t)(uint32_t)(run >> 32);
            totals[l] = ts[e];
            offsets[l] = lo + ts[e];   // inclusive, as the reference's offsets
        }
        run += pack_ct(cs[e], ts[e]);
    }
}

// Fill `n` elements with `v` (a kernel, not a memset, so it can sit in a captured graph).
template <class T> __global__ void k_fill(T* __restrict__ a, size_t n, T v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) a[i] = v;
}
template <class T> inline void fill_async(T* a, size_t n, T v, hipStream_t s) {
    if (!n) return;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, s, a, n, v);
    MAEL_LAUNCH_CHECK();
}

// The whole scan; returns the grand total (one host sync, as the reference's
// `Array(offsets[end])`, flip.jl:1372).
template <class Info>
inline int64_t scan_cells(const Info& info, size_t C, int32_t* totals, int32_t* offsets, int32_t* cstart,
                          unsigned long long* bsums, unsigned long long* pinned, hipStream_t s) {
    int nb = (int)((C + SCAN_TILE - 1) / SCAN_TILE);
    if (nb > 0) {
        hipLaunchKernelGGL(k_scan_reduce<Info>, dim3(nb), dim3(SCAN_THREADS), 0, s, info, C, bsums);
        MAEL_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, s, bsums, nb);
    MAEL_LAUNCH_CHECK();
    if (nb > 0) {
        hipLaunchKernelGGL(k_scan_down<Info>, dim3(nb), dim3(SCAN_THREADS), 0, s, info, C, bsums, totals, offsets, cstart);
        MAEL_LAUNCH_CHECK();
    }
    MAEL_HIP(hipMemcpyAsync(pinned, bsums + nb, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    MAEL_HIP(hipStreamSynchronize(s));
    return (int64_t)(uint32_t)(*pinned & 0xffffffffull);
}

}  // namespace mael
