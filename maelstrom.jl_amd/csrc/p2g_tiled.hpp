// The APIC deposit (flip.jl:513-537) as a tiled reduction.
//
// A workgroup takes 256 consecutive particles (resampling leaves them in cell order, so
// they are close in space) and, one velocity component at a time, accumulates their
// 27-node contributions into a box of nodes in LDS with 64-bit integer adds, then adds
// each touched node of the box to HBM once. When the workgroup's particles spread over
// more nodes than its box holds (typically when its 256 particles straddle two blocks
// that are far apart in the pool), each wavefront takes a quarter of the LDS for a box of
// its own 64 particles, and only a wavefront that does not fit either deposits directly.
// Contributions are rounded to 2^-32 fixed point exactly as the per-particle deposit does
// and integer addition is associative, so the result is bit-identical to depositing every
// contribution with its own global atomic, in any order.
#pragma once
#include "common.hpp"

namespace mael {

constexpr int P2G_THREADS = 256;
// Nodes per component in LDS (mom + wt: 16 B a node). 2048 (32 KB, four workgroups a CU)
// suits the sparse pools, whose workgroups stay inside a block or two; the dense grid
// takes 4096 (64 KB, two a CU): its cell order runs along whole rows, where a row crossing
// the funnel's ring of band cells splits a workgroup's particles into two runs far apart.
constexpr int P2G_BOX = 2048, P2G_BOX_DENSE = 4096;

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
// Six workgroup minima at once (one pair of barriers); red holds 6 x (P2G_THREADS / 64).
__device__ inline void wg_min6(int v[6], int* red) {
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[k] = min(v[k], __shfl_xor(v[k], o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) red[k * (P2G_THREADS / 64) + (threadIdx.x >> 6)] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int r = red[k * (P2G_THREADS / 64)];
#pragma unroll
        for (int w = 1; w < P2G_THREADS / 64; ++w) r = min(r, red[k * (P2G_THREADS / 64) + w]);
        v[k] = r;
    }
}
// The workgroup's bounds of (b0, b1, b2) over the `ok` lanes: x0 y0 z0 (BIG if none), x1 y1 z1.
__device__ inline void wg_bounds(bool ok, int b0, int b1, int b2, int* red, int& x0, int& y0, int& z0, int& x1, int& y1,
                                 int& z1) {
    const int BIG = 1 << 30;
    int v[6] = {ok ? b0 : BIG, ok ? b1 : BIG, ok ? b2 : BIG, ok ? -b0 : BIG, ok ? -b1 : BIG, ok ? -b2 : BIG};
    wg_min6(v, red);
    x0 = v[0]; y0 = v[1]; z0 = v[2];
    x1 = v[3] == BIG ? -BIG : -v[3]; y1 = v[4] == BIG ? -BIG : -v[4]; z1 = v[5] == BIG ? -BIG : -v[5];
}
__device__ inline int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_max(int v) { return -wave_min(-v); }

// A box of nodes: origin, extents, and where it starts in LDS.
struct TileBox {
    int x0, y0, z0, ex, ey, ez, base;
    bool use;
    __device__ int at(int I, int J, int K) const { return base + (I - x0) + ex * ((J - y0) + ey * (K - z0)); }
    __device__ int size() const { return ex * ey * ez; }
};

// The box of the workgroup (if its particles' nodes fit in `cap`), else of the thread's
// wavefront (if they fit in `cap` / 4, in the wavefront's quarter). `ok` lanes are the
// ones whose base nodes (b0, b1, b2) shape it; nodes span b .. b + `span` - 1.
__device__ inline TileBox tile_box(bool ok, int b0, int b1, int b2, int span, int cap, int* red, bool& wgbox) {
    const int BIG = 1 << 30;
    TileBox t;
    int x0, y0, z0, x1, y1, z1;
    wg_bounds(ok, b0, b1, b2, red, x0, y0, z0, x1, y1, z1);
    int ex = x1 - x0 + span, ey = y1 - y0 + span, ez = z1 - z0 + span;
    wgbox = x0 != BIG && ex <= cap && ey <= cap && ez <= cap && ex * ey * ez <= cap;
    if (wgbox) {
        t.x0 = x0; t.y0 = y0; t.z0 = z0; t.ex = ex; t.ey = ey; t.ez = ez; t.base = 0; t.use = true;
        return t;
    }
    int wcap = cap / (P2G_THREADS / 64);
    x0 = wave_min(ok ? b0 : BIG); y0 = wave_min(ok ? b1 : BIG); z0 = wave_min(ok ? b2 : BIG);
    x1 = wave_max(ok ? b0 : -BIG); y1 = wave_max(ok ? b1 : -BIG); z1 = wave_max(ok ? b2 : -BIG);
    ex = x1 - x0 + span; ey = y1 - y0 + span; ez = z1 - z0 + span;
    t.use = x0 != BIG && ex <= wcap && ey <= wcap && ez <= wcap && ex * ey * ez <= wcap;
    t.x0 = x0; t.y0 = y0; t.z0 = z0; t.ex = ex; t.ey = ey; t.ez = ez;
    t.base = (threadIdx.x >> 6) * wcap;
    return t;
}
// The lanes that clear and flush box t: the workgroup's threads, or the wavefront's lanes.
__device__ inline int tile_first(bool wgbox) { return wgbox ? threadIdx.x : (threadIdx.x & 63); }
__device__ inline int tile_stride(bool wgbox) { return wgbox ? P2G_THREADS : 64; }

struct P2GArgs {
    float lo[3];
    float dx;
    unsigned long long* mom[3];
    unsigned long long* wt[3];   // null: no weights accumulated
    float scale;                 // every momentum contribution is multiplied by it before rounding
};

// The particle side of the deposit: the temporal weight of particle p, and the value and
// affine row it deposits for component a.
struct P2GSource {   // flip.jl:513-537: v_a and the APIC row c_a, weighted by temporal(lag, dt)
    PView ps;
    float dt;
    __device__ float weight(int64_t p) const { return temporal(ps.f[15][p], dt); }
    __device__ void comp(int64_t p, int a, float& v, float& c0, float& c1, float& c2) const {
        v = ps.f[3 + a][p];
        c0 = ps.f[6 + 3 * a][p]; c1 = ps.f[7 + 3 * a][p]; c2 = ps.f[8 + 3 * a][p];
    }
};

// NodeMap: `int32_t/size_t index(int a, int I, int J, int K)` of a face node, < 0 off the
// grid (the dense grid's array bounds, or no fine block holding the face). `ps` gives the
// positions (fields 0-2).
template <class NodeMap, class Source, int BOX = P2G_BOX>
__global__ __launch_bounds__(P2G_THREADS, BOX > 2048 ? 2 : 4) void k_p2g_tiled(PView ps, int64_t n, NodeMap map, Source src,
                                                                              P2GArgs A) {
    __shared__ unsigned long long smom[BOX], swt[BOX];
    __shared__ int red[6 * (P2G_THREADS / 64)];
    const int64_t p0 = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * P2G_THREADS;   // neighbouring tiles on one XCD
    for (int a = 0; a < 3; ++a) {
        // the box from the particles in order (a wavefront's own box needs its 64 in order)
        int64_t p = p0 + threadIdx.x;
        bool live = p < n;
        float xp = 0, yp = 0, zp = 0;
        if (live) { xp = ps.f[0][p]; yp = ps.f[1][p]; zp = ps.f[2][p]; }
        float q0 = ispace(xp, A.lo[0], A.dx, a == 0 ? 0.0f : 0.5f);
        float q1 = ispace(yp, A.lo[1], A.dx, a == 1 ? 0.0f : 0.5f);
        float q2 = ispace(zp, A.lo[2], A.dx, a == 2 ? 0.0f : 0.5f);
        BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
        // finite, in-range bases only shape the box; others take the direct path
        bool ok = live && bx.b > -(1 << 24) && bx.b < (1 << 24) && by.b > -(1 << 24) && by.b < (1 << 24) &&
                  bz.b > -(1 << 24) && bz.b < (1 << 24) && isfinite(q0) && isfinite(q1) && isfinite(q2);
        bool wgbox;
        TileBox T = tile_box(ok, bx.b, by.b, bz.b, 3, BOX, red, wgbox);
        int first = tile_first(wgbox), stride = tile_stride(wgbox);
        if (wgbox) {
            // Depositing into the workgroup's box, a wavefront takes every eighth particle:
            // particles of one cell (consecutive in cell order) would otherwise add to the
            // same LDS nodes in the same instruction, which serialises the atomics. The box
            // holds every particle's nodes, and integer sums do not depend on the order.
            p = p0 + (int64_t)((threadIdx.x & 31) * 8 + (threadIdx.x >> 5));
            live = p < n;
            xp = yp = zp = 0;
            if (live) { xp = ps.f[0][p]; yp = ps.f[1][p]; zp = ps.f[2][p]; }
            q0 = ispace(xp, A.lo[0], A.dx, a == 0 ? 0.0f : 0.5f);
            q1 = ispace(yp, A.lo[1], A.dx, a == 1 ? 0.0f : 0.5f);
            q2 = ispace(zp, A.lo[2], A.dx, a == 2 ? 0.0f : 0.5f);
            bx = bspline(q0); by = bspline(q1); bz = bspline(q2);
            ok = live && bx.b > -(1 << 24) && bx.b < (1 << 24) && by.b > -(1 << 24) && by.b < (1 << 24) &&
                 bz.b > -(1 << 24) && bz.b < (1 << 24) && isfinite(q0) && isfinite(q1) && isfinite(q2);
        }
        float vp = 0, c0 = 0, c1 = 0, c2 = 0, wT = 0;
        if (live) {
            src.comp(p, a, vp, c0, c1, c2);
            wT = src.weight(p);
        }
        if (T.use)
            for (int k = first; k < T.size(); k += stride) { smom[T.base + k] = 0ull; swt[T.base + k] = 0ull; }
        __syncthreads();
        // The 27 contributions in the order of operations of flip.jl:524-533, with the
        // factors shared along rows hoisted: wgt = ((wT w_x) w_y) w_z and
        // v = vp + ((c0 d0 + c1 d1) + c2 d2) are the same operations on the same values, so
        // the same bits. A particle whose contributions all lie within 2^18 in magnitude
        // takes the short fixed-point rounding (tofix_near); others take tofix.
        const float bound = wT * (fabsf(vp) + (fabsf(c0) + fabsf(c1) + fabsf(c2)) * (2.0f * A.dx)) * fabsf(A.scale);
        if (live && T.use && ok && bound < 262144.0f) {
            const float wx[3] = {bx.w0, bx.w1, bx.w2}, wy[3] = {by.w0, by.w1, by.w2}, wz[3] = {bz.w0, bz.w1, bz.w2};
            float cd0[3], cd1[3], cd2[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                cd0[m] = c0 * (((float)(bx.b + m) - q0) * A.dx);
                cd1[m] = c1 * (((float)(by.b + m) - q1) * A.dx);
                cd2[m] = c2 * (((float)(bz.b + m) - q2) * A.dx);
            }
            const int l0 = T.at(bx.b, by.b, bz.b);
#pragma unroll
            for (int my = 0; my < 3; ++my)
#pragma unroll
                for (int mx = 0; mx < 3; ++mx) {
                    const float wxy = (wT * wx[mx]) * wy[my], cxy = cd0[mx] + cd1[my];
#pragma unroll
                    for (int mz = 0; mz < 3; ++mz) {
                        const float wgt = wxy * wz[mz];
                        const int l = l0 + mx + T.ex * (my + T.ey * mz);
                        atomicAdd(&smom[l], tofix_near((wgt * (vp + (cxy + cd2[mz]))) * A.scale));
                        atomicAdd(&swt[l], tofix_near(wgt));
                    }
                }
        } else if (live) {   // the rare path: large values, or a particle outside the box
            for (int mz = 0; mz < 3; ++mz) {
                int K = bz.b + mz;
                float wz = pick(bz, mz), d2 = ((float)K - q2) * A.dx;
                for (int my = 0; my < 3; ++my) {
                    int J = by.b + my;
                    float wy = pick(by, my), d1 = ((float)J - q1) * A.dx;
                    for (int mx = 0; mx < 3; ++mx) {
                        int I = bx.b + mx;
                        float wgt = wT * pick(bx, mx) * wy * wz;
                        float d0 = ((float)I - q0) * A.dx;
                        unsigned long long m = tofix((wgt * (vp + (c0 * d0 + c1 * d1 + c2 * d2))) * A.scale),
                                           w = tofix(wgt);
                        if (T.use && ok) {
                            int l = T.at(I, J, K);
                            atomicAdd(&smom[l], m);
                            atomicAdd(&swt[l], w);
                        } else {
                            auto idx = map.index(a, I, J, K);
                            if (idx >= 0) {
                                atomicAdd(&A.mom[a][idx], m);
                                if (A.wt[a]) atomicAdd(&A.wt[a][idx], w);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (T.use)
            for (int k = first; k < T.size(); k += stride) {
                unsigned long long m = smom[T.base + k], w = swt[T.base + k];
                if (m == 0ull && w == 0ull) continue;
                int I = T.x0 + k % T.ex, J = T.y0 + (k / T.ex) % T.ey, K = T.z0 + k / (T.ex * T.ey);
                auto idx = map.index(a, I, J, K);
                if (idx >= 0) {
                    if (m) atomicAdd(&A.mom[a][idx], m);
                    if (w && A.wt[a]) atomicAdd(&A.wt[a][idx], w);
                }
            }
        __syncthreads();   // the boxes are reused by the next component
    }
}

// The particles' level-set splat (flip.jl:539-560) tiled the same way: a workgroup's 256
// particles take the minimum of their order keys over a box of cells in LDS, then each
// touched cell of the box takes one global atomic min. Minimum is order-independent, so
// the result is the direct splat's bit for bit. CellMap: `index(I, J, K)` of a cell (< 0
// off the grid or outside the fine blocks).
// Cells in LDS (4 B a cell): 4096 for the sparse pools, 8192 for the dense grid's rows (as
// the deposit's boxes); a workgroup whose cells do not fit splats directly.
constexpr int SPLAT_BOX = 4096, SPLAT_BOX_DENSE = 8192;

template <class CellMap, int BOX = SPLAT_BOX>
__global__ __launch_bounds__(P2G_THREADS) void k_splat_tiled(PView ps, int64_t n, CellMap map, float lo0, float lo1, float lo2,
                                                            float dx, float dt, int32_t* __restrict__ phip) {
    __shared__ int32_t box[BOX];
    __shared__ int red[6 * (P2G_THREADS / 64)];
    const int64_t p0 = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * P2G_THREADS;   // neighbouring tiles on one XCD
    int64_t p = p0 + threadIdx.x;
    bool live = false, ok = false;
    float xp = 0, yp = 0, zp = 0;
    int ci = 0, cj = 0, ck = 0;
    auto prep = [&]() {   // particle p's advanced position and cell
        live = p < n;
        xp = yp = zp = 0.0f; ci = cj = ck = 0; ok = false;
        if (!live) return;
        float vx = ps.f[3][p], vy = ps.f[4][p], vz = ps.f[5][p];
        float sp = sqrtf(vx * vx + vy * vy + vz * vz);
        float s = ps.f[15][p] + ahead(strength(sp, dt, dx)) * dt;
        xp = ps.f[0][p] + s * vx; yp = ps.f[1][p] + s * vy; zp = ps.f[2][p] + s * vz;
        float f0 = floorf(__fdiv_rn(xp - lo0, dx)), f1 = floorf(__fdiv_rn(yp - lo1, dx)), f2 = floorf(__fdiv_rn(zp - lo2, dx));
        ci = (int)f0; cj = (int)f1; ck = (int)f2;
        // finite, in-range cells only shape the box; others take the direct path
        ok = fabsf(f0) < 16777216.0f && fabsf(f1) < 16777216.0f && fabsf(f2) < 16777216.0f;
    };
    prep();
    const int BIG = 1 << 30;
    int x0, y0, z0, x1, y1, z1;
    wg_bounds(ok, ci, cj, ck, red, x0, y0, z0, x1, y1, z1);
    int ex = x1 - x0 + 3, ey = y1 - y0 + 3, ez = z1 - z0 + 3;
    bool boxed = x0 != BIG && ex <= BOX && ey <= BOX && ez <= BOX && ex * ey * ez <= BOX;
    x0 -= 1; y0 -= 1; z0 -= 1;
    if (boxed) {
        int V = ex * ey * ez;
        for (int k = threadIdx.x; k < V; k += P2G_THREADS) box[k] = 0x7fffffff;
        __syncthreads();
        // every eighth particle per lane group, as in k_p2g_tiled (the minimum does not
        // depend on the order)
        p = p0 + (int64_t)((threadIdx.x & 31) * 8 + (threadIdx.x >> 5));
        prep();
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

// ─── adjoints of the transfers (no counterpart in the reference) ─────────────
// With positions, lags, level set and weights fixed, P2G is linear in the particles' v
// and APIC rows and G2P in the face velocities. The G2P adjoint deposits through
// k_p2g_tiled with this source; the P2G adjoint gathers with k_p2g_adj_gather.

// The G2P adjoint's particle side: the output cotangents, the rows times the gather's
// 4 / dx^2, for the particles that took the grid's velocity (the others deposit 0).
struct G2PAdjSource {
    const float* pb[12];
    const uint8_t* reached;
    float cs;
    __device__ float weight(int64_t p) const { return reached[p] ? 1.0f : 0.0f; }
    __device__ void comp(int64_t p, int a, float& v, float& c0, float& c1, float& c2) const {
        if (!reached[p]) { v = c0 = c1 = c2 = 0.0f; return; }
        v = pb[a][p];
        c0 = pb[3 + 3 * a][p] * cs; c1 = pb[4 + 3 * a][p] * cs; c2 = pb[5 + 3 * a][p] * cs;
    }
};
struct PB12 { float* f[12]; };

// The P2G adjoint's particle part: each particle gathers the momentum cotangents s of the
// nodes it deposited to (NodeMap as k_p2g_tiled's), with the deposit's weights and
// offsets, and adds them to its v and row cotangents.
template <class NodeMap>
__global__ __launch_bounds__(256) void k_p2g_adj_gather(PView ps, int64_t n, NodeMap map, float lo0, float lo1, float lo2,
                                                        float dx, float dt, const float* __restrict__ s0,
                                                        const float* __restrict__ s1, const float* __restrict__ s2, PB12 pb) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float xp = ps.f[0][p], yp = ps.f[1][p], zp = ps.f[2][p];
    float wT = temporal(ps.f[15][p], dt);
    const float* S[3] = {s0, s1, s2};
    for (int a = 0; a < 3; ++a) {
        float q0 = ispace(xp, lo0, dx, a == 0 ? 0.0f : 0.5f);
        float q1 = ispace(yp, lo1, dx, a == 1 ? 0.0f : 0.5f);
        float q2 = ispace(zp, lo2, dx, a == 2 ? 0.0f : 0.5f);
        BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
        float v = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
        for (int mz = 0; mz < 3; ++mz) {
            int K = bz.b + mz;
            float wz = pick(bz, mz), d2 = ((float)K - q2) * dx;
            for (int my = 0; my < 3; ++my) {
                int J = by.b + my;
                float wy = pick(by, my), d1 = ((float)J - q1) * dx;
                for (int mx = 0; mx < 3; ++mx) {
                    int I = bx.b + mx;
                    int64_t idx = (int64_t)map.index(a, I, J, K);
                    if (idx < 0) continue;
                    float t = wT * pick(bx, mx) * wy * wz * S[a][idx];
                    v += t;
                    c0 += t * (((float)I - q0) * dx);
                    c1 += t * d1;
                    c2 += t * d2;
                }
            }
        }
        pb.f[a][p] += v;
        pb.f[3 + 3 * a][p] += c0; pb.f[4 + 3 * a][p] += c1; pb.f[5 + 3 * a][p] += c2;
    }
}

}  // namespace mael
