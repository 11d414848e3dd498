// Whitewater from the narrow band (flip.jl:1411-1511, sflip.jl:2320-2417, whitewater.jl:
// 11-352; Ihmsen, Akinci, Akinci & Teschner 2012): the first consumer of the resampled,
// cell-sorted particle order. A fluid particle's neighbours are the particles of the 27
// cells round it, found through each cell's run in that order (its start is
// offsets - totals), so no sort or cell list is built; the normals are the level set's
// gradient; diffuse particles are spray, foam or bubbles by where they are in the level
// set and ride the grid's velocity.
//
// The reference appends new diffuse particles at ring-buffer slots taken by an atomic
// counter, so which slot a particle gets depends on scheduling. Here a first pass counts
// each fluid particle's new ones (the neighbour sums done once, their results kept), a
// scan turns the counts into offsets, and a second pass writes each fluid particle's new
// ones at head + offset in particle order: deterministic, and the reference's result
// whenever its atomics resolve in thread order.
#include "dense.hpp"
#include "scan.hpp"
#include "sparse.hpp"
#include "whitewater.hpp"

namespace mael {

namespace {

// whitewater.jl:45 clampfn: I scaled to [0, 1] over (lo, hi).
__device__ inline float clampfn(float I, float lo, float hi) { return __fdiv_rn(jmin(I, hi) - jmin(I, lo), hi - lo); }
__device__ inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline float sumsq3(const float a[3]) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

// whitewater.jl:306-310 carried: the steady flow the waves ride on (the basin's swirl, or none).
__device__ inline void carried(const WhiteParams& P, const float x[3], float c[3]) {
    c[0] = c[1] = c[2] = 0.0f;
    if (!P.stream) return;
    float r = sqrtf(x[0] * x[0] + x[1] * x[1]);
    float turn = r > 1e-4f ? __fdiv_rn(oseen_dev(P.gamma, P.core, r), r) : 0.0f;
    c[0] = -turn * x[1];
    c[1] = turn * x[0];
}

// The fluid particles' cells, in the resampled order: the dense grid (cells dx wide,
// flip.jl:1497) or the sparse grid's fine cells (sflip.jl:2330-2344).
struct DenseCells {
    DGrid g;
    const int32_t* offsets;   // inclusive, per cell
    const int32_t* totals;
    // whitewater.jl:84 fluidcell with interpolation.jl:20-25 cellcoords: clamped to the grid
    __device__ void cell(const float x[3], int c[3]) const {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            int v = (int)floorf(__fdiv_rn(x[a] - g.lo[a], g.dx));
            c[a] = v < 0 ? 0 : (v > g.n[a] - 1 ? g.n[a] - 1 : v);
        }
    }
    // whitewater.jl:87-92 cellrange: [start, end) of cell c's particles, empty off the grid
    __device__ bool range(int i, int j, int k, int64_t& s, int64_t& e) const {
        if (i < 0 || j < 0 || k < 0 || i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) return false;
        size_t l = (size_t)i + (size_t)g.n[0] * ((size_t)j + (size_t)g.n[1] * k);
        e = offsets[l];
        s = e - totals[l];
        return true;
    }
    __host__ __device__ float h() const { return g.dx; }
};
struct SparseCells {
    STopo t;
    const int32_t* offsets;   // inclusive, per fine cell of the slots in use
    const int32_t* totals;
    // sflip.jl:2334-2337 fluidcell: floor((x - lo) / h), not clamped
    __device__ void cell(const float x[3], int c[3]) const {
#pragma unroll
        for (int a = 0; a < 3; ++a) c[a] = (int)floorf(__fdiv_rn(x[a] - t.g.lo[a], t.g.dx));
    }
    // sflip.jl:2338-2342 cellrange: the fine cell's run, empty where the cell is not fine
    __device__ bool range(int i, int j, int k, int64_t& s, int64_t& e) const {
        int32_t idx = fineat(t, i, j, k);
        if (idx < 0) return false;
        e = offsets[idx];
        s = e - totals[idx];
        return true;
    }
    __host__ __device__ float h() const { return t.g.dx; }
};

// flip.jl:1433-1447 levelnormals_kernel! / sflip.jl:2347-2362 sparsenormals_kernel!: the
// level set's normalised gradient within `near` of the surface, zero deeper.
struct DensePhi {
    DGrid g;
    const float* phi;
    __device__ float at(float x, float y, float z) const {   // trilinear at index space q
        return trilinear(phi, g.n[0], g.n[1], g.n[2], x, y, z);
    }
    __device__ void normal(const float xp[3], float near, float nrm[3]) const {
        const float q0 = ispace(xp[0], g.lo[0], g.dx, 0.5f), q1 = ispace(xp[1], g.lo[1], g.dx, 0.5f),
                    q2 = ispace(xp[2], g.lo[2], g.dx, 0.5f);
        nrm[0] = nrm[1] = nrm[2] = 0.0f;
        if (at(q0, q1, q2) > -near) {
            float gx = at(q0 + 0.5f, q1, q2) - at(q0 - 0.5f, q1, q2);
            float gy = at(q0, q1 + 0.5f, q2) - at(q0, q1 - 0.5f, q2);
            float gz = at(q0, q1, q2 + 0.5f) - at(q0, q1, q2 - 0.5f);
            float l = sqrtf(gx * gx + gy * gy + gz * gz);
            if (l > 0.0f) { nrm[0] = __fdiv_rn(gx, l); nrm[1] = __fdiv_rn(gy, l); nrm[2] = __fdiv_rn(gz, l); }
        }
    }
    __device__ float sample(const float x[3]) const {
        return at(ispace(x[0], g.lo[0], g.dx, 0.5f), ispace(x[1], g.lo[1], g.dx, 0.5f), ispace(x[2], g.lo[2], g.dx, 0.5f));
    }
};
struct SparsePhi {
    STopo t;
    const float* fine;
    const float* coarse;
    __device__ float sample(const float x[3]) const { return cellsample(t, fine, coarse, x[0], x[1], x[2]); }
    __device__ void normal(const float xp[3], float near, float nrm[3]) const {
        nrm[0] = nrm[1] = nrm[2] = 0.0f;
        if (sample(xp) > -near) {
            const float h = 0.5f * t.g.dx;
            float gx = cellsample(t, fine, coarse, xp[0] + h, xp[1], xp[2]) - cellsample(t, fine, coarse, xp[0] - h, xp[1], xp[2]);
            float gy = cellsample(t, fine, coarse, xp[0], xp[1] + h, xp[2]) - cellsample(t, fine, coarse, xp[0], xp[1] - h, xp[2]);
            float gz = cellsample(t, fine, coarse, xp[0], xp[1], xp[2] + h) - cellsample(t, fine, coarse, xp[0], xp[1], xp[2] - h);
            float l = sqrtf(gx * gx + gy * gy + gz * gz);
            if (l > 0.0f) { nrm[0] = __fdiv_rn(gx, l); nrm[1] = __fdiv_rn(gy, l); nrm[2] = __fdiv_rn(gz, l); }
        }
    }
};

// The grid's velocity at a point: flip.jl:80-81 velocity / sflip.jl:177-183.
struct DenseVel {
    DGrid g;
    const float* u[3];
    __device__ void at(const float x[3], float v[3]) const {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            v[a] = trilinear(u[a], g.n[0] + (a == 0), g.n[1] + (a == 1), g.n[2] + (a == 2),
                             ispace(x[0], g.lo[0], g.dx, a == 0 ? 0.0f : 0.5f), ispace(x[1], g.lo[1], g.dx, a == 1 ? 0.0f : 0.5f),
                             ispace(x[2], g.lo[2], g.dx, a == 2 ? 0.0f : 0.5f));
    }
    __host__ __device__ float radius() const { return g.radius; }
    __host__ __device__ float floorz() const { return g.floor; }
};
struct SparseVel {
    STopo t;
    const float* uf[3];
    const float* uc[3];
    __device__ void at(const float x[3], float v[3]) const {
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = facesample(t, uf[a], uc[a], a, x[0], x[1], x[2]);
    }
    __host__ __device__ float radius() const { return t.g.radius; }
    __host__ __device__ float floorz() const { return t.g.floor; }
};

}  // namespace

// flip.jl:1433-1447 / sflip.jl:2347-2362: the normals of the fluid particles.
template <class PHI>
__global__ __launch_bounds__(256) void k_wnormals(float* __restrict__ nrm, PView ps, int64_t n, PHI phi, float near) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float xp[3] = {ps.f[0][p], ps.f[1][p], ps.f[2][p]}, nv[3];
    phi.normal(xp, near, nv);
    nrm[3 * p] = nv[0]; nrm[3 * p + 1] = nv[1]; nrm[3 * p + 2] = nv[2];
}

// whitewater.jl:131-175, the first half of emit_kernel!: how many diffuse particles fluid
// particle p makes this frame, and the numbers its second half needs (aired + crested,
// crested, energy).
template <class CELLS>
__global__ __launch_bounds__(256) void k_wcount(unsigned long long* __restrict__ counts, float4* __restrict__ keep, PView ps,
                                                const float* __restrict__ nrm, int64_t n, CELLS cells, WhiteParams P, float dt,
                                                uint32_t seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    counts[i] = 0;
    const float h = cells.h();
    const float x[3] = {ps.f[0][i], ps.f[1][i], ps.f[2][i]}, v[3] = {ps.f[3][i], ps.f[4][i], ps.f[5][i]};
    const float ni[3] = {nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]};
    if (x[2] < P.floor + 2.0f * h || sumsq3(ni) == 0.0f || x[0] * x[0] + x[1] * x[1] > P.within * P.within) return;
    float c[3];
    carried(P, x, c);
    const float vr[3] = {v[0] - c[0], v[1] - c[1], v[2] - c[2]};
    const float rs = sqrtf(sumsq3(vr));
    const float energy = clampfn(0.5f * rs * rs, P.energy[0], P.energy[1]);
    const int32_t i1 = (int32_t)(i + 1);   // the reference's 1-based index
    const float draw = hash01(i1, (int32_t)seed);
    float upper = energy * P.kair * dt + energy * P.kcrest * dt;
    if (0.0f <= upper && upper < 1.0f) {   // exact rejection before the neighbour search
        upper = __uint_as_float(__float_as_uint(upper) + 8u);
        if (draw >= upper) return;
    }
    int cc[3];
    cells.cell(x, cc);
    float air = 0.0f, crest = 0.0f;
    const float ni2 = sumsq3(ni);
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                int64_t s, e;
                if (!cells.range(cc[0] + di, cc[1] + dj, cc[2] + dk, s, e)) continue;
                for (int64_t q = s; q < e; ++q) {
                    const float dx[3] = {x[0] - ps.f[0][q], x[1] - ps.f[1][q], x[2] - ps.f[2][q]};
                    const float d = sqrtf(sumsq3(dx));
                    if (0.0f < d && d < h) {
                        const float w = 1.0f - __fdiv_rn(d, h);
                        const float vij[3] = {v[0] - ps.f[3][q], v[1] - ps.f[4][q], v[2] - ps.f[5][q]};
                        const float xij[3] = {__fdiv_rn(dx[0], d), __fdiv_rn(dx[1], d), __fdiv_rn(dx[2], d)};
                        const float lv = sqrtf(sumsq3(vij));
                        if (lv > 0.0f) {
                            const float vh[3] = {__fdiv_rn(vij[0], lv), __fdiv_rn(vij[1], lv), __fdiv_rn(vij[2], lv)};
                            air += lv * (1.0f - dot3(vh, xij)) * w;
                        }
                        const float nq[3] = {nrm[3 * q], nrm[3 * q + 1], nrm[3 * q + 2]};
                        const float mx[3] = {-xij[0], -xij[1], -xij[2]};
                        if (ni2 > 0.0f && sumsq3(nq) > 0.0f && dot3(mx, ni) < 0.0f) crest += (1.0f - dot3(ni, nq)) * w;
                    }
                }
            }
    bool facing = false;
    if (rs > 0.0f && ni2 > 0.0f) {
        const float vh[3] = {__fdiv_rn(vr[0], rs), __fdiv_rn(vr[1], rs), __fdiv_rn(vr[2], rs)};
        facing = dot3(vh, ni) >= 0.6f;
    }
    const float aired = energy * P.kair * clampfn(air, P.air[0], P.air[1]) * dt;
    const float crested = energy * P.kcrest * (facing ? clampfn(crest, P.crest[0], P.crest[1]) : 0.0f) * dt;
    const float tot = aired + crested;
    int32_t whole = (int32_t)floorf(tot);
    whole += draw < tot - (float)whole ? 1 : 0;
    counts[i] = (unsigned long long)(whole > 0 ? whole : 0);
    keep[i] = make_float4(tot, crested, energy, 0.0f);
}

// whitewater.jl:176-198, the second half: fluid particle p's new diffuse particles at the
// ring buffer's slots head + offset + m (mod its capacity).
__global__ __launch_bounds__(256) void k_wemit(WhiteView W, const unsigned long long* __restrict__ offsets,
                                               const float4* __restrict__ keep, PView ps, const float* __restrict__ nrm,
                                               int64_t n, float h, WhiteParams P, uint32_t seed, int64_t head) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o0 = (int64_t)offsets[i], whole = (int64_t)offsets[i + 1] - o0;
    if (whole <= 0) return;
    // A frame that makes more than the buffer holds wraps round it: as in the sequential
    // order, only its last `capacity` particles stay, so the earlier ones are not written
    // (two writes to one slot in one launch would race).
    const int64_t total = (int64_t)offsets[n], first = total > W.capacity ? total - W.capacity : 0;
    const float4 k = keep[i];
    const float tot = k.x, crested = k.y, energy = k.z;
    const float x[3] = {ps.f[0][i], ps.f[1][i], ps.f[2][i]}, v[3] = {ps.f[3][i], ps.f[4][i], ps.f[5][i]};
    const float ni[3] = {nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]};
    const float ln = sqrtf(sumsq3(ni));
    const float nh[3] = {__fdiv_rn(ni[0], ln), __fdiv_rn(ni[1], ln), __fdiv_rn(ni[2], ln)};
    const int32_t i1 = (int32_t)(i + 1), sd = (int32_t)seed;
    for (int32_t m = 1; m <= (int32_t)whole; ++m) {
        if (o0 + m - 1 < first) continue;
        // whitewater.jl:178: slot mod(head += 1, cap) + 1 in the reference's 1-based terms
        const int64_t slot = (head + o0 + m) % W.capacity;
        const float a = hash01((int32_t)((uint32_t)i1 + (uint32_t)m * 7919u), sd + 31);
        const float bb = hash01((int32_t)((uint32_t)i1 - (uint32_t)m * 104729u), sd + 77);
        const float cc = hash01(m, (int32_t)((uint32_t)i1 + (uint32_t)sd));
        const float jh = 0.5f * h;
        const float jit[3] = {(a - 0.5f) * jh, (bb - 0.5f) * jh, (cc - 0.5f) * jh};
        const float u = hash01((int32_t)((uint32_t)i1 + (uint32_t)m * 613u), sd + 5);
        const bool thrown =
            P.kick > 0.0f && hash01((int32_t)((uint32_t)i1 - (uint32_t)m * 331u), sd + 9) * tot < P.tossed * crested;
        float wx[3], wv[3];
        if (thrown) {
            const float sp = P.kick * energy * (0.5f + u);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                wx[c] = x[c] + h * nh[c] + 0.5f * jit[c];
                wv[c] = v[c] + sp * nh[c];
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                wx[c] = x[c] + jit[c];
                wv[c] = v[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            W.x[c][slot] = wx[c];
            W.v[c][slot] = wv[c];
        }
        W.life[slot] = P.life[0] + (P.life[1] - P.life[0]) * hash01(sd - m, i1);
    }
}

// flip.jl:1456-1483 leveldiffuse_kernel! / sflip.jl:2365-2391 sparsediffuse_kernel!: every
// live diffuse particle a step on, by where it is in the level set.
template <class PHI, class VEL>
__global__ __launch_bounds__(256) void k_wdiffuse(WhiteView W, PHI phi, VEL vel, float dt, WhiteParams P, float above,
                                                  float below) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W.capacity) return;
    float life = W.life[i];
    if (!(life > 0.0f)) return;
    float x[3] = {W.x[0][i], W.x[1][i], W.x[2][i]}, v[3] = {W.v[0][i], W.v[1][i], W.v[2][i]};
    const float f = phi.sample(x);
    float vf[3];
    vel.at(x, vf);
    uint8_t kind;
    if (f > above) {
        v[2] = v[2] + (-P.gravity * dt);
        kind = WW_SPRAY;
    } else if (f < -below) {
        const float up = dt * (P.buoyancy * P.gravity);
        v[0] = (v[0] + dt * 0.0f) + P.drag * (vf[0] - v[0]);
        v[1] = (v[1] + dt * 0.0f) + P.drag * (vf[1] - v[1]);
        v[2] = (v[2] + up) + P.drag * (vf[2] - v[2]);
        life -= dt;
        kind = WW_BUBBLE;
    } else {
        v[0] = vf[0]; v[1] = vf[1]; v[2] = vf[2];
        life -= dt;
        kind = WW_FOAM;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c] = x[c] + dt * v[c];
    const float reach = jmin(P.within, vel.radius());
    if (x[2] < vel.floorz() || x[0] * x[0] + x[1] * x[1] > reach * reach) life = 0.0f;
    W.kind[i] = kind;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        W.x[c][i] = x[c];
        W.v[c][i] = v[c];
    }
    W.life[i] = life;
}

// whitewater.jl:366-379 packwhite_kernel!: the live diffuse particles (x, y, z, kind) and
// velocities as Float16, in slot order.
__global__ __launch_bounds__(256) void k_wlive_count(unsigned long long* __restrict__ counts, WhiteView W) {
    __shared__ int lds[4];
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int mine = i < W.capacity && W.life[i] > 0.0f ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mine;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = (unsigned long long)(lds[0] + lds[1] + lds[2] + lds[3]);
}
__global__ __launch_bounds__(256) void k_wlive_pack(WhiteView W, const unsigned long long* __restrict__ offsets,
                                                    __half* __restrict__ out, __half* __restrict__ vel) {
    __shared__ int lds[4];
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int live = i < W.capacity && W.life[i] > 0.0f ? 1 : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long ball = __ballot(live);
    if (lane == 0) lds[wave] = __popcll(ball);
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += lds[w];
    if (!live) return;
    const int64_t d = (int64_t)offsets[blockIdx.x] + base + __popcll(ball & ((1ull << lane) - 1ull));
    out[4 * d] = __float2half_rn(W.x[0][i]);
    out[4 * d + 1] = __float2half_rn(W.x[1][i]);
    out[4 * d + 2] = __float2half_rn(W.x[2][i]);
    out[4 * d + 3] = __float2half_rn((float)W.kind[i]);
    vel[3 * d] = __float2half_rn(W.v[0][i]);
    vel[3 * d + 1] = __float2half_rn(W.v[1][i]);
    vel[3 * d + 2] = __float2half_rn(W.v[2][i]);
}

// ─── host ────────────────────────────────────────────────────────────────────

Whitewater::Whitewater(int64_t cap, const WhiteParams& p) : P(p), capacity(cap) {
    for (int c = 0; c < 3; ++c) {
        x[c].alloc(std::max<int64_t>(cap, 1));
        v[c].alloc(std::max<int64_t>(cap, 1));
        fill_async(x[c].p, (size_t)cap, 0.0f, st.s);
        fill_async(v[c].p, (size_t)cap, 0.0f, st.s);
    }
    life.alloc(std::max<int64_t>(cap, 1));
    kind.alloc(std::max<int64_t>(cap, 1));
    fill_async(life.p, (size_t)cap, 0.0f, st.s);
    fill_async(kind.p, (size_t)cap, (uint8_t)0, st.s);
    st.sync();
}

WhiteView Whitewater::view() {
    WhiteView W;
    for (int c = 0; c < 3; ++c) { W.x[c] = x[c].p; W.v[c] = v[c].p; }
    W.life = life.p;
    W.kind = kind.p;
    W.capacity = capacity;
    return W;
}

// The emission of one frame on `s` (the basin's stream): normals, counts, their scan, the
// new particles. The fluid particles are the basin's live ones, in the resampled order.
template <class CELLS, class PHI>
static void emit(Whitewater& w, const PView& ps, int64_t n, CELLS cells, PHI phi, float h, float dt, float near, Stream& st) {
    hipStream_t s = st.s;
    w.frames += 1;
    if (n <= 0) return;
    if (w.nrm.n < (size_t)(3 * n)) w.nrm.alloc((size_t)(3 * n));
    if (w.counts.n < (size_t)(n + 2)) {
        w.counts.alloc((size_t)(n + 2));
        w.keep.alloc((size_t)n);
    }
    const unsigned g = (unsigned)cdiv(n, 256);
    hipLaunchKernelGGL((k_wnormals<PHI>), dim3(g), dim3(256), 0, s, w.nrm.p, ps, n, phi, near);
    hipLaunchKernelGGL((k_wcount<CELLS>), dim3(g), dim3(256), 0, s, w.counts.p, w.keep.p, ps, w.nrm.p, n, cells, w.P, dt,
                       (uint32_t)w.frames);
    launch_scan_bsums(w.counts.p, (int)n, s);   // exclusive offsets, counts[n] the total
    hipLaunchKernelGGL(k_wemit, dim3(g), dim3(256), 0, s, w.view(), w.counts.p, w.keep.p, ps, w.nrm.p, n, h, w.P,
                       (uint32_t)w.frames, w.head);
    MAEL_LAUNCH_CHECK();
    MAEL_HIP(hipMemcpyAsync(st.pinned, w.counts.p + n, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    st.sync();
    w.head += (int64_t)*(unsigned long long*)st.pinned;
}

// flip.jl:1496-1511 whitewater!(w, s::FlipBasin, u, v, dt; substeps, near, above, below):
// emitted with Float32(dt), moved in sub-steps of Float32(dt / substeps).
void dense_whitewater(DenseBasin& b, Whitewater& w, double dt, int substeps, float near, float above, float below) {
    DenseCells cells{b.g, b.offsets.p, b.totals.p};
    DensePhi phi{b.g, b.phi.p};
    emit(w, b.view(b.cur), b.n, cells, phi, b.g.dx, (float)dt, near * b.g.dx, b.st);
    DenseVel vel{b.g, {b.u[0].p, b.u[1].p, b.u[2].p}};
    const float hs = (float)(dt / substeps);
    for (int k = 0; k < substeps; ++k)
        hipLaunchKernelGGL((k_wdiffuse<DensePhi, DenseVel>), dim3(cdiv(w.capacity, 256)), dim3(256), 0, b.st.s, w.view(), phi, vel,
                           hs, w.P, above * b.g.dx, below * b.g.dx);
    MAEL_LAUNCH_CHECK();
    b.st.sync();
}

// sflip.jl:2401-2417 whitewater!(w, s::SparseBasin, u, v, dt; ...).
void sparse_whitewater(SparseBasin& b, Whitewater& w, double dt, int substeps, float near, float above, float below) {
    STopo t = b.topo();
    SparseCells cells{t, b.offsets.p, b.totals.p};
    SparsePhi phi{t, b.phi.fine.p, b.phi.coarse.p};
    emit(w, b.view(b.cur), b.n, cells, phi, b.g.dx, (float)dt, near * b.g.dx, b.st);
    SparseVel vel{t, {b.u[0].fine.p, b.u[1].fine.p, b.u[2].fine.p}, {b.u[0].coarse.p, b.u[1].coarse.p, b.u[2].coarse.p}};
    const float hs = (float)(dt / substeps);
    for (int k = 0; k < substeps; ++k)
        hipLaunchKernelGGL((k_wdiffuse<SparsePhi, SparseVel>), dim3(cdiv(w.capacity, 256)), dim3(256), 0, b.st.s, w.view(), phi,
                           vel, hs, w.P, above * b.g.dx, below * b.g.dx);
    MAEL_LAUNCH_CHECK();
    b.st.sync();
}

int64_t Whitewater::pack(uint16_t* out, uint16_t* vel) {
    hipStream_t s = st.s;
    const int nb = (int)cdiv(capacity, 256);
    if (pcounts.n < (size_t)nb + 2) pcounts.alloc((size_t)nb + 2);
    if (pout.n < (size_t)(4 * capacity)) {
        pout.alloc((size_t)(4 * std::max<int64_t>(capacity, 1)));
        pvel.alloc((size_t)(3 * std::max<int64_t>(capacity, 1)));
    }
    if (nb) hipLaunchKernelGGL(k_wlive_count, dim3(nb), dim3(256), 0, s, pcounts.p, view());
    launch_scan_bsums(pcounts.p, nb, s);
    if (nb) hipLaunchKernelGGL(k_wlive_pack, dim3(nb), dim3(256), 0, s, view(), pcounts.p, pout.p, pvel.p);
    MAEL_LAUNCH_CHECK();
    MAEL_HIP(hipMemcpyAsync(st.pinned, pcounts.p + nb, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    st.sync();
    const int64_t m = (int64_t)*(unsigned long long*)st.pinned;
    if (m > 0) {
        MAEL_HIP(hipMemcpyAsync(out, pout.p, (size_t)m * 4 * sizeof(__half), hipMemcpyDeviceToHost, s));
        MAEL_HIP(hipMemcpyAsync(vel, pvel.p, (size_t)m * 3 * sizeof(__half), hipMemcpyDeviceToHost, s));
        st.sync();
    }
    return m;
}

}  // namespace mael
