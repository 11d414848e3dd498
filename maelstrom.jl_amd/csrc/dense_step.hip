// The dense basin's support phases around the hot path, so that a step runs the
// reference's `step!(::FlipBasin)` (flip.jl:1396-1410): ST-FLIP advection of the particles
// and semi-Lagrangian advection of the grid (flip.jl:350-499), trapped air
// (flip.jl:688-760), the stirring forces (flip.jl:762-798, basin.jl:105-181),
// extrapolation (flip.jl:1158-1206) and the fastest speed (flip.jl:1387).
//
// Every kernel takes the reference's unfused FP32 operation order (the library is built
// with -ffp-contract=off, divisions are IEEE), so the advected positions, fields and
// extrapolated velocities are the reference's bit for bit where the inputs are.
#include "dense.hpp"
#include "scan.hpp"
#include "stir.hpp"

namespace mael {

namespace {

constexpr int BX = 64, BY = 4;
inline dim3 grid3(int n1, int n2, int n3) { return dim3(cdiv(n1, BX), cdiv(n2, BY), (unsigned)n3); }

// The grid's velocity and the flags of the faces it was carried out to (flip.jl:401-405 `Flow`).
struct DFlow {
    DGrid g;
    const float* u[3];
    const uint8_t* f[3];
};

__host__ __device__ inline float faceoffset(int a, int axis) { return a == axis ? 0.0f : 0.5f; }

// flip.jl:80-81 velocity: the three components sampled trilinearly at their faces.
__device__ inline void dvelocity(const DFlow& f, float px, float py, float pz, float v[3]) {
    const DGrid& g = f.g;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        v[a] = trilinear(f.u[a], g.n[0] + (a == 0), g.n[1] + (a == 1), g.n[2] + (a == 2), ispace(px, g.lo[0], g.dx, faceoffset(0, a)),
                         ispace(py, g.lo[1], g.dx, faceoffset(1, a)), ispace(pz, g.lo[2], g.dx, faceoffset(2, a)));
}
// flip.jl:370-372 reached: every face the samples at p read was projected or extrapolated.
__device__ inline bool dreached(const DFlow& f, float px, float py, float pz) {
    const DGrid& g = f.g;
    float m[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        m[a] = trilinear(f.f[a], g.n[0] + (a == 0), g.n[1] + (a == 1), g.n[2] + (a == 2), ispace(px, g.lo[0], g.dx, faceoffset(0, a)),
                         ispace(py, g.lo[1], g.dx, faceoffset(1, a)), ispace(pz, g.lo[2], g.dx, faceoffset(2, a)));
    return jmin(jmin(m[0], m[1]), m[2]) > 0.999f;
}
// flip.jl:367 substeps: how many sub-steps of at most a cell a move of `cells` takes (1..32).
__device__ inline int dsubsteps(float cells) {
    float c = ceilf(cells);
    return c != c ? 1 : (c >= 32.0f ? 32 : (c <= 1.0f ? 1 : (int)c));
}
// flip.jl:374-380 inwalls: back inside the wall and above the floor, a quarter cell in.
__device__ inline void dinwalls(const DGrid& g, float& x, float& y, float& z) {
    float m = 0.25f * g.dx;
    float r = sqrtf(x * x + y * y);
    if (r > g.radius - m) {
        x = __fdiv_rn(x * (g.radius - m), r);
        y = __fdiv_rn(y * (g.radius - m), r);
    }
    if (z < g.floor + m) z = g.floor + m;
}
__device__ inline float norm3(const float v[3]) { return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

}  // namespace

// flip.jl:438-464 advectparticles_kernel!: each particle by its own time (dt, its lag and
// a jitter, at most 2 dt) along the flow in RK2 sub-steps of at most a cell; where the flow
// was not carried out to, ballistic under gravity.
__global__ __launch_bounds__(256) void k_dadvect_particles(PView ps, int64_t n, DFlow flow, float dt, float gravity,
                                                           int32_t seed) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float y0 = ps.f[0][p], y1 = ps.f[1][p], y2 = ps.f[2][p];
    float b0 = ps.f[3][p], b1 = ps.f[4][p], b2 = ps.f[5][p];
    const float h0 = flow.g.dx;
    const float speed = sqrtf(b0 * b0 + b1 * b1 + b2 * b2);
    const float gm = strength(speed, dt, h0);
    const float xi = hash01((int32_t)(p + 1), seed) - 0.5f;   // the reference's 1-based index
    const float tau = ps.f[15][p];
    const float act = jclamp(dt + tau + gm * xi * dt, 0.0f, 2.0f * dt);
    ps.f[15][p] = dt + tau - act;
    float vf[3];
    dvelocity(flow, y0, y1, y2, vf);
    const int m = dsubsteps(__fdiv_rn(jmax(speed, norm3(vf)) * act, h0));
    const float h = __fdiv_rn(act, (float)m);
    for (int s = 0; s < m; ++s) {
        if (dreached(flow, y0, y1, y2)) {
            float v1[3];
            dvelocity(flow, y0, y1, y2, v1);
            const float hh = 0.5f * h;
            dvelocity(flow, y0 + hh * v1[0], y1 + hh * v1[1], y2 + hh * v1[2], vf);
            b0 = vf[0]; b1 = vf[1]; b2 = vf[2];
            y0 = y0 + h * b0; y1 = y1 + h * b1; y2 = y2 + h * b2;
        } else {
            const float dz = 0.5f * gravity * h;
            y0 = y0 + h * b0; y1 = y1 + h * b1; y2 = y2 + h * (b2 - dz);
            b2 = b2 - gravity * h;
        }
        dinwalls(flow.g, y0, y1, y2);
    }
    ps.f[0][p] = y0; ps.f[1][p] = y1; ps.f[2][p] = y2;
}

// flip.jl:356-369 advect_kernel!: the sample of `src` at node (i, j, k) (offset `off` in
// cells) traced back along the flow in RK2 sub-steps of at most a cell. `n` is src's shape.
__global__ void k_dadvect_grid(float* __restrict__ out, const float* __restrict__ src, DFlow flow, int3 n, float o0, float o1,
                               float o2, float dt) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= n.x || j >= n.y || k >= n.z) return;
    const DGrid& g = flow.g;
    float px = g.lo[0] + ((float)i + o0) * g.dx, py = g.lo[1] + ((float)j + o1) * g.dx, pz = g.lo[2] + ((float)k + o2) * g.dx;
    float v[3];
    dvelocity(flow, px, py, pz, v);
    const int m = dsubsteps(__fdiv_rn(norm3(v) * dt, g.dx));
    const float h = __fdiv_rn(dt, (float)m);
    for (int s = 0; s < m; ++s) {
        float v1[3], v2[3];
        dvelocity(flow, px, py, pz, v1);
        const float hh = 0.5f * h;
        dvelocity(flow, px - hh * v1[0], py - hh * v1[1], pz - hh * v1[2], v2);
        px = px - h * v2[0]; py = py - h * v2[1]; pz = pz - h * v2[2];
    }
    out[(size_t)i + (size_t)n.x * ((size_t)j + (size_t)n.y * k)] =
        trilinear(src, n.x, n.y, n.z, ispace(px, g.lo[0], g.dx, o0), ispace(py, g.lo[1], g.dx, o1), ispace(pz, g.lo[2], g.dx, o2));
}

void DenseBasin::advect(float dt, float gacc, int32_t seed) {
    hipStream_t s = st.s;
    DFlow f;
    f.g = g;
    for (int a = 0; a < 3; ++a) {
        f.u[a] = u[a].p;
        f.f[a] = flag[a].p;
    }
    if (n > 0)
        hipLaunchKernelGGL(k_dadvect_particles, dim3(cdiv(n, 256)), dim3(256), 0, s, view(cur), n, f, dt, gacc, seed);
    for (int a = 0; a < 3; ++a) {   // u -> u0, v -> v0, w -> w0 at their faces
        int3 nn = make_int3(g.n[0] + (a == 0), g.n[1] + (a == 1), g.n[2] + (a == 2));
        hipLaunchKernelGGL(k_dadvect_grid, grid3(nn.x, nn.y, nn.z), dim3(BX, BY), 0, s, u0[a].p, u[a].p, f, nn,
                           faceoffset(0, a), faceoffset(1, a), faceoffset(2, a), dt);
    }
    hipLaunchKernelGGL(k_dadvect_grid, grid3(g.n[0], g.n[1], g.n[2]), dim3(BX, BY), 0, s, phi0.p, phi.p, f,
                       make_int3(g.n[0], g.n[1], g.n[2]), 0.5f, 0.5f, 0.5f, dt);
    MAEL_LAUNCH_CHECK();
}

// flip.jl:738-745 fillair_kernel!: AIR the atmosphere does not reach becomes water, at
// least `depth` deep in the level set.
__global__ void k_dfillair(uint8_t* __restrict__ types, float* __restrict__ phi, const uint8_t* __restrict__ reach, size_t C,
                           float depth) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= C) return;
    if (types[l] == AIR && reach[l] == 0) {
        types[l] = LIQUID;
        phi[l] = jmin(phi[l], -depth);
    }
}

void DenseBasin::filltrapped(int maxrounds) {
    if (!reach.p) {
        reach.alloc(cells());
        changed.alloc(1);
    }
    sweep_trapped_air(reach.p, changed.p, types.p, make_int3(g.n[0], g.n[1], g.n[2]), maxrounds, st);
    hipLaunchKernelGGL(k_dfillair, dim3(cdiv(cells(), 256)), dim3(256), 0, st.s, types.p, phi.p, reach.p, cells(), 2.0f * g.dx);
    MAEL_LAUNCH_CHECK();
}

// flip.jl:764-775 forces_kernel!: u += dt (stir - g z) on the faces of axis A out of the
// walls, the water's velocity read from the copy `old` (the stir relaxes what it reads).
template <int A>
__global__ void k_dforces(float* __restrict__ ua, const float* __restrict__ old, DGrid g, StirParams P, float dt, float time,
                          float gravity) {
    const int n1 = g.n[0] + (A == 0), n2 = g.n[1] + (A == 1), n3 = g.n[2] + (A == 2);
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= n1 || j >= n2 || k >= n3) return;
    const float o0 = faceoffset(0, A), o1 = faceoffset(1, A), o2 = faceoffset(2, A);
    const float px = g.lo[0] + ((float)i + o0) * g.dx, py = g.lo[1] + ((float)j + o1) * g.dx, pz = g.lo[2] + ((float)k + o2) * g.dx;
    if (issolid(g.radius, g.floor, px, py, pz)) return;
    // The stir's component A reads the water's velocity along A alone.
    const float vel = trilinear(old, n1, n2, n3, ispace(px, g.lo[0], g.dx, o0), ispace(py, g.lo[1], g.dx, o1),
                                ispace(pz, g.lo[2], g.dx, o2));
    const float acc = stir<A>(P, px, py, pz, vel, time);
    ua[(size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k)] += dt * (A == 2 ? acc - gravity : acc);
}

void DenseBasin::set_stir(const StirParams& p, const float* waves, int nwaves) {
    stirp = p;
    stirp.nwaves = nwaves;
    stirp.surface = 0;
    stir_waves.alloc(std::max(7 * nwaves, 1));
    if (nwaves) MAEL_HIP(hipMemcpyAsync(stir_waves.p, waves, 7 * nwaves * sizeof(float), hipMemcpyHostToDevice, st.s));
    st.sync();
    has_stir = true;
}

void DenseBasin::forces(float dt, float time, float gacc) {
    if (!has_stir) {   // no stirring model: the gravity term alone
        gravity(dt, gacc);
        return;
    }
    hipStream_t s = st.s;
    for (int a = 0; a < 3; ++a)   // flip.jl:784: the forces read a copy of the velocity they change
        MAEL_HIP(hipMemcpyAsync(u0[a].p, u[a].p, faces(a) * sizeof(float), hipMemcpyDeviceToDevice, s));
    StirParams P = stirp;
    P.waves = stir_waves.p;
    hipLaunchKernelGGL(k_dforces<0>, grid3(g.n[0] + 1, g.n[1], g.n[2]), dim3(BX, BY), 0, s, u[0].p, u0[0].p, g, P, dt, time, gacc);
    hipLaunchKernelGGL(k_dforces<1>, grid3(g.n[0], g.n[1] + 1, g.n[2]), dim3(BX, BY), 0, s, u[1].p, u0[1].p, g, P, dt, time, gacc);
    hipLaunchKernelGGL(k_dforces<2>, grid3(g.n[0], g.n[1], g.n[2] + 1), dim3(BX, BY), 0, s, u[2].p, u0[2].p, g, P, dt, time, gacc);
    MAEL_LAUNCH_CHECK();
    solidify();
}

// flip.jl:1160-1179 extrapolate_kernel!: one layer; a face not yet flagged takes the mean
// of its flagged neighbours of the same component (none across a wall).
template <int A>
__global__ void k_dextrapolate(float* __restrict__ out, uint8_t* __restrict__ fout, const float* __restrict__ src,
                               const uint8_t* __restrict__ fin, DGrid g) {
    const int n1 = g.n[0] + (A == 0), n2 = g.n[1] + (A == 1), n3 = g.n[2] + (A == 2);
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= n1 || j >= n2 || k >= n3) return;
    const size_t s1 = (size_t)n1, s2 = (size_t)n1 * n2;
    const size_t l = (size_t)i + s1 * j + s2 * k;
    if (fin[l] != 0) {
        out[l] = src[l];
        fout[l] = 1;
        return;
    }
    float s = 0.0f;
    int m = 0;
    // flip.jl:779-784 walled: a wall cell on either side, or the grid's own edge
    const int a = i - (A == 0), b = j - (A == 1), c = k - (A == 2);
    const bool walled = !(a >= 0 && b >= 0 && c >= 0 && i < g.n[0] && j < g.n[1] && k < g.n[2]) || issolidcell(g, a, b, c) ||
                        issolidcell(g, i, j, k);
    if (!walled) {
        // the six neighbours in the reference's order: +x, -x, +y, -y, +z, -z, their flags
        // and values read together (one off the grid reads this face's own and is skipped)
        const bool in[6] = {i + 1 < n1, i > 0, j + 1 < n2, j > 0, k + 1 < n3, k > 0};
        const size_t e[6] = {in[0] ? l + 1 : l, in[1] ? l - 1 : l, in[2] ? l + s1 : l,
                             in[3] ? l - s1 : l, in[4] ? l + s2 : l, in[5] ? l - s2 : l};
        uint8_t fv[6];
        float sv[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) { fv[q] = fin[e[q]]; sv[q] = src[e[q]]; }
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (in[q] && fv[q]) { s += sv[q]; ++m; }
    }
    out[l] = m > 0 ? __fdiv_rn(s, (float)m) : src[l];
    fout[l] = m > 0 ? 1 : 0;
}
// flip.jl:1182-1186 unreached_kernel!: zero where the extrapolation did not reach.
__global__ void k_dunreached(float* __restrict__ u, const uint8_t* __restrict__ reached, size_t N) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < N && reached[l] == 0) u[l] = 0.0f;
}

void DenseBasin::extrapolate(int layers) {
    hipStream_t s = st.s;
    if (!flag2[0].p)
        for (int a = 0; a < 3; ++a) flag2[a].alloc(faces(a));
    for (int a = 0; a < 3; ++a) {
        float *src = u[a].p, *dst = u0[a].p;
        uint8_t *fin = flag[a].p, *fout = flag2[a].p;
        auto kern = a == 0 ? k_dextrapolate<0> : (a == 1 ? k_dextrapolate<1> : k_dextrapolate<2>);
        dim3 gr = grid3(g.n[0] + (a == 0), g.n[1] + (a == 1), g.n[2] + (a == 2));
        for (int layer = 0; layer < layers; ++layer) {
            hipLaunchKernelGGL(kern, gr, dim3(BX, BY), 0, s, dst, fout, src, fin, g);
            std::swap(src, dst);
            std::swap(fin, fout);
        }
        if (src != u[a].p) MAEL_HIP(hipMemcpyAsync(u[a].p, src, faces(a) * sizeof(float), hipMemcpyDeviceToDevice, s));
        // the flags of what was reached stay in `flag`: the particles' advection reads them
        if (fin != flag[a].p) MAEL_HIP(hipMemcpyAsync(flag[a].p, fin, faces(a), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_dunreached, dim3(cdiv(faces(a), 256)), dim3(256), 0, s, u[a].p, flag[a].p, faces(a));
    }
    MAEL_LAUNCH_CHECK();
    solidify();
}

// flip.jl:230-294 maxabs over the three face arrays: |u| maxima as order-preserving bits
// (non-negative floats order as their bit patterns; a NaN wins, as Julia's max lets it).
__global__ __launch_bounds__(256) void k_dmaxabs(unsigned int* __restrict__ out, const float* __restrict__ u, size_t N) {
    unsigned int m = 0;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < N; l += (size_t)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(fabsf(u[l])));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

float DenseBasin::maxspeed() {
    if (!changed.p) {
        reach.alloc(cells());
        changed.alloc(1);
    }
    unsigned int* out = (unsigned int*)changed.p;
    fill_async(out, (size_t)1, 0u, st.s);
    for (int a = 0; a < 3; ++a)
        hipLaunchKernelGGL(k_dmaxabs, dim3(std::min<size_t>(cdiv(faces(a), 256), 2048)), dim3(256), 0, st.s, out, u[a].p,
                           faces(a));
    MAEL_LAUNCH_CHECK();
    MAEL_HIP(hipMemcpyAsync(st.pinned, out, sizeof(unsigned int), hipMemcpyDeviceToHost, st.s));
    st.sync();
    return __builtin_bit_cast(float, *(unsigned int*)st.pinned);
}

}  // namespace mael
