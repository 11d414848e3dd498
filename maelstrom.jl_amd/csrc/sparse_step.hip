// The sparse basin's support phases around the hot path, so a step runs the reference's
// `substep!` (sflip.jl:2295-2318): advection of the particles and the grid
// (sflip.jl:715-788, flip.jl:426-499), trapped air (sflip.jl:950-990, flip.jl:690-760),
// the stirring forces (sflip.jl:992-1056, basin.jl:105-235), and the projection's kick,
// extrapolation and reflection (sflip.jl:2017-2139).
#include "sparse.hpp"
#include "scan.hpp"

namespace mael {

__constant__ int NBX[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

// A velocity field of the topology: face arrays (fine pool, coarse) and their flags.
struct SFlow {
    STopo t;
    const float* uf[3];
    const float* uc[3];
    const uint8_t* ff[3];
    const uint8_t* fc[3];
};

// sflip.jl:177-183 velocity and reached through the topology (block slots through `lk`).
template <class L = SlotsGlobal>
__device__ inline void svelocity(const SFlow& f, float px, float py, float pz, float v[3], const L& lk = L()) {
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = facesample(f.t, f.uf[a], f.uc[a], a, px, py, pz, lk);
}
template <class L = SlotsGlobal>
__device__ inline bool sreached(const SFlow& f, float px, float py, float pz, const L& lk = L()) {
    float m = facesample(f.t, f.ff[0], f.fc[0], 0, px, py, pz, lk);
    m = jmin(m, facesample(f.t, f.ff[1], f.fc[1], 1, px, py, pz, lk));
    m = jmin(m, facesample(f.t, f.ff[2], f.fc[2], 2, px, py, pz, lk));
    return m > 0.999f;
}
// flip.jl:367 substeps.
__device__ inline int substeps(float cells) {
    float c = ceilf(cells);
    int m = c != c ? 1 : (c >= 32.0f ? 32 : (c <= 1.0f ? 1 : (int)c));
    return m;
}
// sflip.jl:185-191 inwalls.
__device__ inline void sinwalls(const SGrid& g, float& x, float& y, float& z) {
    float m = 0.25f * g.dx;
    float r = sqrtf(x * x + y * y);
    if (r > g.radius - m) {
        x = __fdiv_rn(x * (g.radius - m), r);
        y = __fdiv_rn(y * (g.radius - m), r);
    }
    if (z < g.floor + m) z = g.floor + m;
}

// flip.jl:438-464 advectparticles_kernel! over the topology.
// The workgroup's particles are cell-sorted, so their samples' blocks are looked up in a
// table of the 27 blocks around the first particle's block (others in the global table).
__global__ __launch_bounds__(256) void k_sadvect_particles(PView ps, int64_t n, SFlow flow, float dt, float gravity,
                                                           int32_t seed) {
    __shared__ int32_t s27[27];
    const int64_t p0 = (int64_t)blockIdx.x * blockDim.x;
    int4 home;
    {
        const SGrid& g = flow.t.g;
        const float f0 = floorf(__fdiv_rn(ps.f[0][p0] - g.lo[0], g.dx)), f1 = floorf(__fdiv_rn(ps.f[1][p0] - g.lo[1], g.dx)),
                    f2 = floorf(__fdiv_rn(ps.f[2][p0] - g.lo[2], g.dx));
        const bool ok = fabsf(f0) < 1e8f && fabsf(f1) < 1e8f && fabsf(f2) < 1e8f;   // NaN fails too
        home = ok ? make_int4((int)f0 >> 4, (int)f1 >> 4, (int)f2 >> 4, 0) : make_int4(-4, -4, -4, 0);
    }
    stage27(flow.t, home, s27);
    __syncthreads();
    const Slots27 lk{s27, home};
    int64_t p = p0 + threadIdx.x;
    if (p >= n) return;
    float y0 = ps.f[0][p], y1 = ps.f[1][p], y2 = ps.f[2][p];
    float b0 = ps.f[3][p], b1 = ps.f[4][p], b2 = ps.f[5][p];
    float h0 = flow.t.g.dx;
    float speed = sqrtf(b0 * b0 + b1 * b1 + b2 * b2);
    float gm = strength(speed, dt, h0);
    float xi = hash01((int32_t)(p + 1), seed) - 0.5f;
    float tau = ps.f[15][p];
    float act = jclamp(dt + tau + gm * xi * dt, 0.0f, 2.0f * dt);
    ps.f[15][p] = dt + tau - act;
    float vf[3];
    svelocity(flow, y0, y1, y2, vf, lk);
    int m = substeps(__fdiv_rn(jmax(speed, sqrtf(vf[0] * vf[0] + vf[1] * vf[1] + vf[2] * vf[2])) * act, h0));
    float h = __fdiv_rn(act, (float)m);
    const float vs0 = vf[0], vs1 = vf[1], vs2 = vf[2];   // the flow at the start point
    for (int s = 0; s < m; ++s) {
        if (sreached(flow, y0, y1, y2, lk)) {
            float v1[3] = {vs0, vs1, vs2};
            if (s > 0) svelocity(flow, y0, y1, y2, v1, lk);
            float hh = 0.5f * h;
            svelocity(flow, y0 + hh * v1[0], y1 + hh * v1[1], y2 + hh * v1[2], vf, lk);
            b0 = vf[0]; b1 = vf[1]; b2 = vf[2];
            y0 = y0 + h * b0; y1 = y1 + h * b1; y2 = y2 + h * b2;
        } else {
            float dz = 0.5f * gravity * h;
            y0 = y0 + h * b0; y1 = y1 + h * b1; y2 = y2 + h * (b2 - dz);
            b2 = b2 - gravity * h;
        }
        sinwalls(flow.t.g, y0, y1, y2);
    }
    ps.f[0][p] = y0; ps.f[1][p] = y1; ps.f[2][p] = y2;
}

// sflip.jl:715-731 backtrace: RK2 sub-steps of at most `cells` cells `h` wide.
template <class L = SlotsGlobal>
__device__ inline void backtrace(const SFlow& f, float& px, float& py, float& pz, float dt, float h, const L& lk = L()) {
    float v[3];
    svelocity(f, px, py, pz, v, lk);
    int m = substeps(__fdiv_rn(sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * dt, 4.0f * h));
    float k = __fdiv_rn(dt, (float)m);
    for (int s = 0; s < m; ++s) {
        float v1[3] = {v[0], v[1], v[2]}, v2[3];   // the first sub-step starts where v was taken
        if (s > 0) svelocity(f, px, py, pz, v1, lk);
        float hk = 0.5f * k;
        svelocity(f, px - hk * v1[0], py - hk * v1[1], pz - hk * v1[2], v2, lk);
        px = px - k * v2[0]; py = py - k * v2[1]; pz = pz - k * v2[2];
    }
}

// sflip.jl:733-763: cells (phi) and faces (the carried velocity) along the flow, into
// phi0 and u0; fine parts over the pool, coarse cells outside blocks, all coarse faces.
// A workgroup's 256 cells (or faces) lie in one slot: the blocks of its samples are looked
// up in a table of the 27 blocks around it.
__global__ __launch_bounds__(256) void k_sadvect_finecells(float* __restrict__ out, const float* __restrict__ src,
                                                           const float* __restrict__ srcc, SFlow f, float dt, int32_t nf) {
    __shared__ int32_t s27[27];
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    SCell c = finecell(f.t, n);
    if (c.w.x < 0) return;   // the whole workgroup
    stage27(f.t, c.w, s27);
    __syncthreads();
    const Slots27 lk{s27, c.w};
    float p[3];
    finecentre(f.t.g, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, p);
    backtrace(f, p[0], p[1], p[2], dt, f.t.g.dx, lk);
    out[n] = cellsample(f.t, src, srcc, p[0], p[1], p[2], lk);
}
template <int A>
__global__ __launch_bounds__(256) void k_sadvect_finefaces(float* __restrict__ out, const float* __restrict__ src,
                                                           const float* __restrict__ srcc, SFlow f, float dt, int32_t nfa) {
    __shared__ int32_t s27[27];
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    SCell c = fineface(f.t, A, n);
    if (c.w.x < 0) return;   // the whole workgroup (4352 = 17 x 256 faces a slot)
    stage27(f.t, c.w, s27);
    __syncthreads();
    const Slots27 lk{s27, c.w};
    float p[3];
    facepos(f.t.g, A, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, f.t.g.dx, p);
    backtrace(f, p[0], p[1], p[2], dt, f.t.g.dx, lk);
    out[n] = facesample_a<A>(f.t, src, srcc, p[0], p[1], p[2], lk);
}
__global__ void k_sadvect_coarsecells(float* __restrict__ out, const float* __restrict__ src, const float* __restrict__ srcc,
                                      SFlow f, float dt) {
    int n0 = f.t.g.nc(0), n1 = f.t.g.nc(1), n2 = f.t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    if (covered(f.t, I, J, K)) return;
    float p[3];
    coarsecentre(f.t.g, I, J, K, p);
    backtrace(f, p[0], p[1], p[2], dt, f.t.g.cdx());
    out[l] = cellsample(f.t, src, srcc, p[0], p[1], p[2]);
}
__global__ void k_sadvect_coarsefaces(float* __restrict__ out, const float* __restrict__ src, const float* __restrict__ srcc,
                                      SFlow f, int a, float dt) {
    int n0 = f.t.g.nc(0) + (a == 0), n1 = f.t.g.nc(1) + (a == 1), n2 = f.t.g.nc(2) + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    float p[3];
    facepos(f.t.g, a, I, J, K, f.t.g.cdx(), p);
    backtrace(f, p[0], p[1], p[2], dt, f.t.g.cdx());
    out[l] = facesample(f.t, src, srcc, a, p[0], p[1], p[2]);
}

SFlow SparseBasin::flow(bool reflected) const {
    SFlow f;
    f.t = topo();
    for (int a = 0; a < 3; ++a) {
        f.uf[a] = reflected ? ur[a].fine.p : u[a].fine.p;
        f.uc[a] = reflected ? ur[a].coarse.p : u[a].coarse.p;
        f.ff[a] = fl[a].fine.p;
        f.fc[a] = fl[a].coarse.p;
    }
    return f;
}

void SparseBasin::advect(float dt, bool carry_reflected, float gacc, int32_t seed) {
    hipStream_t s = st.s;
    SFlow f = flow(false);
    if (n > 0) {
        hipLaunchKernelGGL(k_sadvect_particles, dim3(cdiv(n, 256)), dim3(256), 0, s, view(cur), n, f, dt, gacc, seed);
    }
    if (nslots)
        hipLaunchKernelGGL(k_sadvect_finecells, dim3(cdiv(finecells(), 256)), dim3(256), 0, s, phi0.fine.p, phi.fine.p,
                           phi.coarse.p, f, dt, (int32_t)finecells());
    hipLaunchKernelGGL(k_sadvect_coarsecells, dim3(cdiv(ncells(), 256)), dim3(256), 0, s, phi0.coarse.p, phi.fine.p,
                       phi.coarse.p, f, dt);
    for (int a = 0; a < 3; ++a) {
        const SField<float>& c = carry_reflected ? ur[a] : u[a];
        if (nslots) {
            auto kern = a == 0 ? k_sadvect_finefaces<0> : (a == 1 ? k_sadvect_finefaces<1> : k_sadvect_finefaces<2>);
            hipLaunchKernelGGL(kern, dim3(cdiv(finefaces(), 256)), dim3(256), 0, s, u0[a].fine.p, c.fine.p, c.coarse.p, f, dt,
                               (int32_t)finefaces());
        }
        hipLaunchKernelGGL(k_sadvect_coarsefaces, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, s, u0[a].coarse.p, c.fine.p,
                           c.coarse.p, f, a, dt);
    }
    MAEL_LAUNCH_CHECK();
}

// ─── trapped air (flip.jl:698-745, sflip.jl:950-990) ─────────────────────────

__global__ void k_seedair(uint8_t* __restrict__ reach, const uint8_t* __restrict__ types, int3 n) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n.x * n.y * n.z) return;
    int k = (int)(l / ((size_t)n.x * n.y));
    reach[l] = (k == n.z - 1 && types[l] == AIR) ? 1 : 0;
}
__global__ void k_sweepair(uint8_t* __restrict__ reach, int* __restrict__ changed, const uint8_t* __restrict__ types, int3 n,
                           int axis) {
    int na = axis == 0 ? n.y : n.x, nb = axis == 2 ? n.y : n.z;
    int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y * blockDim.y + threadIdx.y;
    if (a >= na || b >= nb) return;
    int len = axis == 0 ? n.x : (axis == 1 ? n.y : n.z);
    int count = 0;
    for (int pass = 0; pass < 2; ++pass) {
        bool carry = false;
        for (int m = 0; m < len; ++m) {
            int l = pass == 0 ? m : len - 1 - m;
            int i, j, k;
            if (axis == 0) { i = l; j = a; k = b; }
            else if (axis == 1) { i = a; j = l; k = b; }
            else { i = a; j = b; k = l; }
            size_t c = cl(i, j, k, n.x, n.y);
            if (types[c] == AIR) {
                if (reach[c]) carry = true;
                else if (carry) { reach[c] = 1; ++count; }
            } else {
                carry = false;
            }
        }
    }
    if (count > 0) atomicAdd(changed, count);
}
// The same sweep with a wavefront per line: in the serial passes, a reached AIR cell
// reaches every AIR cell after it in its run (forward) and then before it (backward). Each
// lane takes up to SWEEP_CHUNK consecutive cells; a lane's effect on the carry is "pass it
// on" (all its cells AIR) or a fixed value, and a scan over the wavefront gives every lane
// the carry the serial pass brings into its cells, so the reached cells and the count of
// newly reached ones are the serial sweep's.
constexpr int SWEEP_CHUNK = 8;   // lines of at most 64 x 8 cells
__device__ inline void carry_scan(int& pass, int& val, bool reverse) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int p2 = reverse ? __shfl_down(pass, o, 64) : __shfl_up(pass, o, 64);
        int v2 = reverse ? __shfl_down(val, o, 64) : __shfl_up(val, o, 64);
        bool has = reverse ? lane + o < 64 : lane >= o;
        if (has) {   // the earlier lanes' effect (p2, v2), then this lane's (pass, val)
            val = pass ? (v2 | val) : val;
            pass = pass & p2;
        }
    }
}
__global__ __launch_bounds__(256) void k_sweepairw(uint8_t* __restrict__ reach, int* __restrict__ changed,
                                                   const uint8_t* __restrict__ types, int3 n, int axis) {
    const int lane = threadIdx.x & 63;
    const int na = axis == 0 ? n.y : n.x, nb = axis == 2 ? n.y : n.z;
    const int line = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (line >= na * nb) return;   // the whole wavefront
    const int a = line % na, b = line / na;
    const int len = axis == 0 ? n.x : (axis == 1 ? n.y : n.z), ch = (len + 63) / 64;
    const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)n.x : (size_t)n.x * n.y);
    const size_t c0 = axis == 0 ? cl(0, a, b, n.x, n.y) : (axis == 1 ? cl(a, 0, b, n.x, n.y) : cl(a, b, 0, n.x, n.y));
    bool air[SWEEP_CHUNK], r0[SWEEP_CHUNK], r[SWEEP_CHUNK];
    int allair = 1;
#pragma unroll
    for (int e = 0; e < SWEEP_CHUNK; ++e) {
        const int m = lane * ch + e;
        const bool in = e < ch && m < len;
        air[e] = in && types[c0 + (size_t)m * stride] == AIR;
        r0[e] = in && reach[c0 + (size_t)m * stride] != 0;
        r[e] = r0[e];
        if (e < ch && !air[e]) allair = 0;
    }
    for (int pass = 0; pass < 2; ++pass) {
        const bool back = pass == 1;
        // this lane's effect from no carry
        int val = 0;
#pragma unroll
        for (int e0 = 0; e0 < SWEEP_CHUNK; ++e0) {
            const int e = back ? SWEEP_CHUNK - 1 - e0 : e0;
            if (e >= ch) continue;
            val = air[e] ? (val | (int)r[e]) : 0;
        }
        int pz = allair, vz = val;
        carry_scan(pz, vz, back);   // inclusive over this lane and the earlier ones
        int cin = back ? __shfl_down(vz, 1, 64) : __shfl_up(vz, 1, 64);
        if (back ? lane == 63 : lane == 0) cin = 0;
        int c = cin;
#pragma unroll
        for (int e0 = 0; e0 < SWEEP_CHUNK; ++e0) {
            const int e = back ? SWEEP_CHUNK - 1 - e0 : e0;
            if (e >= ch) continue;
            if (air[e]) {
                if (r[e]) c = 1;
                else if (c) r[e] = true;
            } else {
                c = 0;
            }
        }
    }
    int count = 0;
#pragma unroll
    for (int e = 0; e < SWEEP_CHUNK; ++e)
        if (r[e] && !r0[e]) {
            reach[c0 + (size_t)(lane * ch + e) * stride] = 1;
            ++count;
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) count += __shfl_xor(count, o, 64);
    if (lane == 0 && count > 0) atomicAdd(changed, count);
}
__global__ void k_fillfine(uint8_t* __restrict__ tf, float* __restrict__ phi, const uint8_t* __restrict__ reach,
                           const uint8_t* __restrict__ tc, STopo t, float depth, int32_t nf) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    SCell c = finecell(t, n);
    if (c.w.x < 0 || tf[n] != AIR) return;
    int I = c.w.x * 16 + c.i, J = c.w.y * 16 + c.j, K = c.w.z * 16 + c.k;
    size_t cc = cl(I >> 2, J >> 2, K >> 2, t.g.nc(0), t.g.nc(1));
    if (reach[cc] == 0 && tc[cc] == AIR) {
        tf[n] = LIQUID;
        phi[n] = jmin(phi[n], -depth);
    }
}
__global__ void k_fillcoarse(uint8_t* __restrict__ tc, float* __restrict__ phic, const uint8_t* __restrict__ reach, STopo t,
                             float depth) {
    int n0 = t.g.nc(0), n1 = t.g.nc(1), n2 = t.g.nc(2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    if (!covered(t, I, J, K) && tc[l] == AIR && reach[l] == 0) {
        tc[l] = LIQUID;
        phic[l] = jmin(phic[l], -depth);
    }
}

void sweep_trapped_air(uint8_t* reach, int* changed, const uint8_t* types, int3 nn, int maxrounds, Stream& st) {
    hipStream_t s = st.s;
    const size_t C = (size_t)nn.x * nn.y * nn.z;
    hipLaunchKernelGGL(k_seedair, dim3(cdiv(C, 256)), dim3(256), 0, s, reach, types, nn);
    for (int r = 0; r < maxrounds; ++r) {
        fill_async(changed, (size_t)1, 0, s);
        if (std::max(nn.x, std::max(nn.y, nn.z)) <= 64 * SWEEP_CHUNK) {   // a wavefront per line
            hipLaunchKernelGGL(k_sweepairw, dim3(cdiv((size_t)nn.x * nn.y, 4)), dim3(256), 0, s, reach, changed, types, nn, 2);
            hipLaunchKernelGGL(k_sweepairw, dim3(cdiv((size_t)nn.y * nn.z, 4)), dim3(256), 0, s, reach, changed, types, nn, 0);
            hipLaunchKernelGGL(k_sweepairw, dim3(cdiv((size_t)nn.x * nn.z, 4)), dim3(256), 0, s, reach, changed, types, nn, 1);
        } else {
            hipLaunchKernelGGL(k_sweepair, dim3(cdiv(nn.x, 16), cdiv(nn.y, 16)), dim3(16, 16), 0, s, reach, changed, types, nn, 2);
            hipLaunchKernelGGL(k_sweepair, dim3(cdiv(nn.y, 16), cdiv(nn.z, 16)), dim3(16, 16), 0, s, reach, changed, types, nn, 0);
            hipLaunchKernelGGL(k_sweepair, dim3(cdiv(nn.x, 16), cdiv(nn.z, 16)), dim3(16, 16), 0, s, reach, changed, types, nn, 1);
        }
        MAEL_LAUNCH_CHECK();
        MAEL_HIP(hipMemcpyAsync(st.pinned, changed, sizeof(int), hipMemcpyDeviceToHost, s));
        st.sync();
        if (*(int*)st.pinned == 0) break;
    }
}

void SparseBasin::filltrapped(int maxrounds) {
    hipStream_t s = st.s;
    sweep_trapped_air(reach.p, changed.p, types.coarse.p, make_int3(g.nc(0), g.nc(1), g.nc(2)), maxrounds, st);
    if (nslots)
        hipLaunchKernelGGL(k_fillfine, dim3(cdiv(finecells(), 256)), dim3(256), 0, s, types.fine.p, phi.fine.p, reach.p,
                           types.coarse.p, topo(), 2.0f * g.dx, (int32_t)finecells());
    hipLaunchKernelGGL(k_fillcoarse, dim3(cdiv(ncells(), 256)), dim3(256), 0, s, types.coarse.p, phi.coarse.p, reach.p, topo(),
                       2.0f * g.cdx());
    MAEL_LAUNCH_CHECK();
}

// ─── the stirring forces (basin.jl:105-235) ──────────────────────────────────

// sflip.jl:992-1016: u += dt (stir - g z) on the faces out of the walls, the flow read from `old`.
template <int A>
__global__ void k_sforcesfine(float* __restrict__ u, SFlow old, StirParams P, float dt, float time, float gravity,
                              int32_t nfa) {
    int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nfa) return;
    SCell c = fineface(old.t, A, n);
    if (c.w.x < 0) return;
    float p[3];
    facepos(old.t.g, A, c.w.x * 16 + c.i, c.w.y * 16 + c.j, c.w.z * 16 + c.k, old.t.g.dx, p);
    if (sissolid(old.t.g, p[0], p[1], p[2])) return;
    float vel = facesample(old.t, old.uf[A], old.uc[A], A, p[0], p[1], p[2]);
    float acc = stir<A>(P, p[0], p[1], p[2], vel, time);
    u[n] += dt * (A == 2 ? acc - gravity : acc);
}
template <int A>
__global__ void k_sforcescoarse(float* __restrict__ u, SFlow old, StirParams P, float dt, float time, float gravity) {
    const SGrid& g = old.t.g;
    int n0 = g.nc(0) + (A == 0), n1 = g.nc(1) + (A == 1), n2 = g.nc(2) + (A == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    float p[3];
    facepos(g, A, I, J, K, g.cdx(), p);
    if (sissolid(g, p[0], p[1], p[2])) return;
    float vel = facesample(old.t, old.uf[A], old.uc[A], A, p[0], p[1], p[2]);
    float acc = stir<A>(P, p[0], p[1], p[2], vel, time);
    u[l] += dt * (A == 2 ? acc - gravity : acc);
}
template <int A>
static void launch_forces(SparseBasin& b, const SFlow& old, const StirParams& P, float dt, float time, float gacc) {
    if (b.nslots)
        hipLaunchKernelGGL(k_sforcesfine<A>, dim3(cdiv(b.finefaces(), 256)), dim3(256), 0, b.st.s, b.u[A].fine.p, old, P, dt,
                           time, gacc, (int32_t)b.finefaces());
    hipLaunchKernelGGL(k_sforcescoarse<A>, dim3(cdiv(b.nfaces(A), 256)), dim3(256), 0, b.st.s, b.u[A].coarse.p, old, P, dt,
                       time, gacc);
}

void SparseBasin::forces(float dt, float time, float gacc) {
    hipStream_t s = st.s;
    for (int a = 0; a < 3; ++a) {   // u0 = u: the old flow the stir reads
        MAEL_HIP(hipMemcpyAsync(u0[a].fine.p, u[a].fine.p, finefaces() * sizeof(float), hipMemcpyDeviceToDevice, s));
        MAEL_HIP(hipMemcpyAsync(u0[a].coarse.p, u[a].coarse.p, nfaces(a) * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    SFlow old;
    old.t = topo();
    for (int a = 0; a < 3; ++a) {
        old.uf[a] = u0[a].fine.p; old.uc[a] = u0[a].coarse.p;
        old.ff[a] = fl[a].fine.p; old.fc[a] = fl[a].coarse.p;
    }
    StirParams P = stirp;
    P.waves = stir_waves.p;
    P.heights = stir_heights.p;
    launch_forces<0>(*this, old, P, dt, time, gacc);
    launch_forces<1>(*this, old, P, dt, time, gacc);
    launch_forces<2>(*this, old, P, dt, time, gacc);
    MAEL_LAUNCH_CHECK();
    solidify();
    restrict_velocity();
}

void SparseBasin::set_stir(const StirParams& p, const float* waves, int nwaves, const float* heights, int nheights) {
    stirp = p;
    stirp.nwaves = nwaves;
    stirp.nheights = nheights;
    stir_waves.alloc(std::max(7 * nwaves, 1));
    stir_heights.alloc(std::max(nheights, 2));
    if (nwaves)
        MAEL_HIP(hipMemcpyAsync(stir_waves.p, waves, 7 * nwaves * sizeof(float), hipMemcpyHostToDevice, st.s));
    if (nheights)
        MAEL_HIP(hipMemcpyAsync(stir_heights.p, heights, nheights * sizeof(float), hipMemcpyHostToDevice, st.s));
    st.sync();
}

// ─── kick, extrapolation, reflection (sflip.jl:2017-2139) ────────────────────

// One layer: an unflagged face takes the mean of its flagged neighbours (fine faces;
// neighbours through finefaceat's rule, the blocks read from a table of the 27 around
// the slot). A workgroup per live slot (block order), a thread per 17 of its faces. The
// output arrays hold the layer before the input's (or, with `fresh`, nothing yet). A
// flagged face keeps its value from layer to layer, and so does an unflagged one until it
// is flagged: a face whose flag is the same in both and that takes no neighbour's mean is
// already this layer's there and is skipped. Neighbours' values are read only where
// flagged.
template <int A>
__global__ __launch_bounds__(256) void k_sextrapfine(float* __restrict__ out, uint8_t* __restrict__ fout,
                                                     const float* __restrict__ src, const uint8_t* __restrict__ fin,
                                                     float* __restrict__ out2, const float* __restrict__ src2, STopo t,
                                                     int fresh, const int32_t* __restrict__ order) {
    __shared__ int32_t s27[27];
    const int s = order[blockIdx.x];   // 1-based, live
    const int4 w = t.where[s - 1];
    stage27(t, w, s27);
    const bool clear = blockclear(t.g, w);
    __syncthreads();
    // finefaceat for a neighbour at global (GI, GJ, GK): its own block's copy, else the
    // top face of the block below along A when it lies on that plane
    auto face = [&](int GI, int GJ, int GK) -> int32_t {
        const int di = (GI >> 4) - w.x + 1, dj = (GJ >> 4) - w.y + 1, dk = (GK >> 4) - w.z + 1;
        const int32_t s1 = s27[di + 3 * dj + 9 * dk];
        if (s1 != 0) return (s1 - 1) * SBFACES + ffn(16, A, GI & 15, GJ & 15, GK & 15);
        const int GA = A == 0 ? GI : (A == 1 ? GJ : GK);
        if ((GA & 15) != 0) return -1;
        const int32_t s2 = s27[(di - (A == 0)) + 3 * (dj - (A == 1)) + 9 * (dk - (A == 2))];
        if (s2 == 0) return -1;
        return (s2 - 1) * SBFACES + ffn(16, A, A == 0 ? 16 : (GI & 15), A == 1 ? 16 : (GJ & 15), A == 2 ? 16 : (GK & 15));
    };
    for (int r = 0; r < SBFACES / 256; ++r) {
        const int f = r * 256 + (int)threadIdx.x;
        const int32_t n = (s - 1) * SBFACES + f;
        const uint8_t f0 = fin[n];
        const uint8_t fo = fresh ? 2 : fout[n];   // 2: not known
        if (f0) {
            if (fo == 1) continue;
            out[n] = src[n];
            if (out2) out2[n] = src2[n];
            fout[n] = 1;
            continue;
        }
        int i, j, k;
        if (A == 0) { i = f % 17; int q = f / 17; j = q & 15; k = q >> 4; }
        else if (A == 1) { i = f & 15; int q = f >> 4; j = q % 17; k = q / 17; }
        else { i = f & 15; j = (f >> 4) & 15; k = f >> 8; }
        const int I = w.x * 16 + i, J = w.y * 16 + j, K = w.z * 16 + k;
        bool walled = false;
        if (!clear) {
            float p[3];
            facepos(t.g, A, I, J, K, t.g.dx, p);
            walled = walledface_dev(t.g, p, A, t.g.dx);
        }
        int32_t nidx[6];
        uint8_t nf[6];
        bool any = false;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            nidx[q] = walled ? -1 : face(I + NBX[q][0], J + NBX[q][1], K + NBX[q][2]);
            nf[q] = nidx[q] >= 0 ? fin[nidx[q]] : 0;
            any = any || nf[q] != 0;
        }
        if (!any) {   // stays as it is
            if (fo == 0) continue;
            out[n] = src[n];
            if (out2) out2[n] = src2[n];
            fout[n] = 0;
            continue;
        }
        float acc = 0.0f, acc2 = 0.0f;
        int cnt = 0;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (nf[q]) {
                acc += src[nidx[q]];
                if (out2) acc2 += src2[nidx[q]];
                ++cnt;
            }
        out[n] = __fdiv_rn(acc, (float)cnt);
        if (out2) out2[n] = __fdiv_rn(acc2, (float)cnt);
        fout[n] = 1;
    }
}
__global__ void k_sextrapcoarse(float* __restrict__ out, uint8_t* __restrict__ fout, const float* __restrict__ src,
                                const uint8_t* __restrict__ fin, float* __restrict__ out2, const float* __restrict__ src2,
                                SGrid g, int a) {
    int n0 = g.nc(0) + (a == 0), n1 = g.nc(1) + (a == 1), n2 = g.nc(2) + (a == 2);
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (size_t)n0 * n1 * n2) return;
    if (fin[l]) {
        out[l] = src[l];
        if (out2) out2[l] = src2[l];
        fout[l] = 1;
        return;
    }
    int I, J, K;
    unlin(l, n0, n1, I, J, K);
    float h = g.cdx(), p[3];
    facepos(g, a, I, J, K, h, p);
    float acc = 0.0f, acc2 = 0.0f;
    int m = 0;
    if (!walledface_dev(g, p, a, h)) {
        for (int q = 0; q < 6; ++q) {
            int x = I + NBX[q][0], y = J + NBX[q][1], z = K + NBX[q][2];
            if (x < 0 || y < 0 || z < 0 || x >= n0 || y >= n1 || z >= n2) continue;
            size_t e = cl(x, y, z, n0, n1);
            if (fin[e]) {
                acc += src[e];
                if (out2) acc2 += src2[e];
                ++m;
            }
        }
    }
    out[l] = m > 0 ? __fdiv_rn(acc, (float)m) : src[l];
    if (out2) out2[l] = m > 0 ? __fdiv_rn(acc2, (float)m) : src2[l];
    fout[l] = m > 0 ? 1 : 0;
}
__global__ void k_unreached(float* __restrict__ u, const uint8_t* __restrict__ reached, size_t n) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n && reached[l] == 0) u[l] = 0.0f;
}
// sflip.jl:2118-2122: k = valid ? u - k : 0.
__global__ void k_kick(float* __restrict__ k, const float* __restrict__ u, const uint8_t* __restrict__ valid, size_t n) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) k[l] = valid[l] ? u[l] - k[l] : 0.0f;
}
__global__ void k_addto(float* __restrict__ k, const float* __restrict__ u, size_t n) {
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) k[l] += u[l];
}

void SparseBasin::save_reflect() {   // ur = u before the projection (sflip.jl:2306-2309)
    for (int a = 0; a < 3; ++a) {
        MAEL_HIP(hipMemcpyAsync(ur[a].fine.p, u[a].fine.p, finefaces() * sizeof(float), hipMemcpyDeviceToDevice, st.s));
        MAEL_HIP(hipMemcpyAsync(ur[a].coarse.p, u[a].coarse.p, nfaces(a) * sizeof(float), hipMemcpyDeviceToDevice, st.s));
    }
}
void SparseBasin::kick() {
    for (int a = 0; a < 3; ++a) {
        hipLaunchKernelGGL(k_kick, dim3(cdiv(finefaces(), 256)), dim3(256), 0, st.s, ur[a].fine.p, u[a].fine.p, fl[a].fine.p,
                           finefaces());
        hipLaunchKernelGGL(k_kick, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, st.s, ur[a].coarse.p, u[a].coarse.p,
                           fl[a].coarse.p, nfaces(a));
    }
    MAEL_LAUNCH_CHECK();
}
void SparseBasin::reflected() {
    for (int a = 0; a < 3; ++a) {
        hipLaunchKernelGGL(k_addto, dim3(cdiv(finefaces(), 256)), dim3(256), 0, st.s, ur[a].fine.p, u[a].fine.p, finefaces());
        hipLaunchKernelGGL(k_addto, dim3(cdiv(nfaces(a), 256)), dim3(256), 0, st.s, ur[a].coarse.p, u[a].coarse.p, nfaces(a));
    }
    MAEL_LAUNCH_CHECK();
}

// sflip.jl:2097-2116 extrapolate!: `layers` fine layers and `coarselayers` coarse ones
// (even counts leave the results in place), the kick along with the velocity.
void SparseBasin::extrapolate(int layers, int coarselayers, bool withkick) {
    hipStream_t s = st.s;
    STopo t = topo();
    size_t F = finefaces();
    for (int a = 0; a < 3; ++a) {
        float *f = u[a].fine.p, *tmp = u0[a].fine.p, *k = withkick ? ur[a].fine.p : nullptr, *ktmp = withkick ? ur0[a].fine.p : nullptr;
        uint8_t *fin = fl[a].fine.p, *fout = fl2[a].fine.p;
        auto kern = a == 0 ? k_sextrapfine<0> : (a == 1 ? k_sextrapfine<1> : k_sextrapfine<2>);
        for (int l = 0; l < layers; ++l) {   // layer 1 writes over the input of layer 0
            if (nlive > 0)   // free slots: nothing reads their faces
                hipLaunchKernelGGL(kern, dim3((unsigned)nlive), dim3(256), 0, s, tmp, fout, f, fin, ktmp, k, t, l == 0 ? 1 : 0,
                                   sorder.p);
            std::swap(f, tmp); std::swap(k, ktmp); std::swap(fin, fout);
        }
        if (f != u[a].fine.p) {   // an odd count: copy back
            MAEL_HIP(hipMemcpyAsync(u[a].fine.p, f, F * sizeof(float), hipMemcpyDeviceToDevice, s));
            if (withkick) MAEL_HIP(hipMemcpyAsync(ur[a].fine.p, k, F * sizeof(float), hipMemcpyDeviceToDevice, s));
            MAEL_HIP(hipMemcpyAsync(fl[a].fine.p, fin, F, hipMemcpyDeviceToDevice, s));
        }
        launch_restrictfaces_f(*this, u[a].coarse.p, u[a].fine.p, a);
        if (withkick) launch_restrictfaces_f(*this, ur[a].coarse.p, ur[a].fine.p, a);
        launch_restrictfaces_u8(*this, fl[a].coarse.p, fl[a].fine.p, a);
        size_t Fc = nfaces(a);
        float *fc = u[a].coarse.p, *tc = u0[a].coarse.p, *kc = withkick ? ur[a].coarse.p : nullptr,
              *ktc = withkick ? ur0[a].coarse.p : nullptr;
        uint8_t *finc = fl[a].coarse.p, *foutc = fl2[a].coarse.p;
        for (int l = 0; l < coarselayers; ++l) {
            hipLaunchKernelGGL(k_sextrapcoarse, dim3(cdiv(Fc, 256)), dim3(256), 0, s, tc, foutc, fc, finc, ktc, kc, g, a);
            std::swap(fc, tc); std::swap(kc, ktc); std::swap(finc, foutc);
        }
        if (fc != u[a].coarse.p) {
            MAEL_HIP(hipMemcpyAsync(u[a].coarse.p, fc, Fc * sizeof(float), hipMemcpyDeviceToDevice, s));
            if (withkick) MAEL_HIP(hipMemcpyAsync(ur[a].coarse.p, kc, Fc * sizeof(float), hipMemcpyDeviceToDevice, s));
            MAEL_HIP(hipMemcpyAsync(fl[a].coarse.p, finc, Fc, hipMemcpyDeviceToDevice, s));
        }
        for (int w = 0; w < (withkick ? 2 : 1); ++w) {
            SField<float>& x = w == 0 ? u[a] : ur[a];
            if (F) hipLaunchKernelGGL(k_unreached, dim3(cdiv(F, 256)), dim3(256), 0, s, x.fine.p, fl[a].fine.p, F);
            hipLaunchKernelGGL(k_unreached, dim3(cdiv(Fc, 256)), dim3(256), 0, s, x.coarse.p, fl[a].coarse.p, Fc);
        }
        MAEL_LAUNCH_CHECK();
    }
    solidify();
    restrict_velocity();
}

// ─── the fastest water (sflip.jl:2250-2272) ──────────────────────────────────

__global__ __launch_bounds__(256) void k_maxabs(float* __restrict__ partials, const float* __restrict__ a, size_t n) {
    __shared__ float lds[4];
    float v = 0.0f;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (size_t)gridDim.x * blockDim.x)
        v = jmax(v, fabsf(a[l]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = jmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = jmax(jmax(lds[0], lds[1]), jmax(lds[2], lds[3]));
}
float SparseBasin::maxspeed() {
    const int NB = 512;
    std::vector<float> h(NB);
    float m = 0.0f;
    for (int a = 0; a < 3; ++a)
        for (int part = 0; part < 2; ++part) {
            size_t cnt = part == 0 ? finefaces() : nfaces(a);
            const float* A = part == 0 ? u[a].fine.p : u[a].coarse.p;
            if (!cnt) continue;
            hipLaunchKernelGGL(k_maxabs, dim3(NB), dim3(256), 0, st.s, (float*)partials.p, A, cnt);
            MAEL_LAUNCH_CHECK();
            MAEL_HIP(hipMemcpyAsync(h.data(), partials.p, NB * sizeof(float), hipMemcpyDeviceToHost, st.s));
            st.sync();
            for (float x : h) m = std::max(m, x);
        }
    return m;
}

}  // namespace mael
