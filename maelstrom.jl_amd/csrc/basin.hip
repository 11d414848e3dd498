// The dense basin's allocation and its support phases: the initial balanced vortex
// (flip.jl:298-314), classification (flip.jl:674-677), walls (flip.jl:793-807), the
// water's volume (flip.jl:262-293) and gravity (the gravity term of flip.jl:765-775).
// These set up and close a step around the hot path; they are not part of it.
#include "dense.hpp"
#include "scan.hpp"

namespace mael {

// One workgroup: exclusive scan of the tile sums in place; bsums[nb] = the grand total.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(unsigned long long* __restrict__ bsums, int nb) {
    __shared__ unsigned long long lds[SCAN_THREADS / 64];
    unsigned long long carry = 0;
    for (int base = 0; base < nb; base += SCAN_TILE) {
        unsigned long long v[SCAN_ITEMS], mine = 0;
#pragma unroll
        for (int e = 0; e < SCAN_ITEMS; ++e) {
            int idx = base + threadIdx.x * SCAN_ITEMS + e;
            v[e] = idx < nb ? bsums[idx] : 0ull;
            mine += v[e];
        }
        unsigned long long tot;
        unsigned long long run = carry + block_excl_scan(mine, lds, tot);
#pragma unroll
        for (int e = 0; e < SCAN_ITEMS; ++e) {
            int idx = base + threadIdx.x * SCAN_ITEMS + e;
            if (idx < nb) bsums[idx] = run;
            run += v[e];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) bsums[nb] = carry;
}

void launch_scan_bsums(unsigned long long* bsums, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, s, bsums, nb);
    MAEL_LAUNCH_CHECK();
}

DenseBasin::DenseBasin(const DGrid& grid, int lev, int64_t cap, int32_t bnd, int32_t nm)
    : g(grid), levels(lev), band(bnd), nmin(nm), capacity(cap) {
    size_t C = cells();
    for (int a = 0; a < 3; ++a) {
        u[a].alloc(faces(a)); u0[a].alloc(faces(a)); flag[a].alloc(faces(a));
        mom[a].alloc(faces(a)); wt[a].alloc(faces(a));
    }
    phi.alloc(C); phi0.alloc(C); phip.alloc(C); types.alloc(C);
    P.alloc(C); b.alloc(C); r.alloc(C); z.alloc(C); d.alloc(C); q.alloc(C); diag.alloc(C);
    P64.alloc(C); e.alloc(C); rv.alloc(C);
    lv.resize(levels);
    for (int l = 0; l < levels; ++l) {
        for (int a = 0; a < 3; ++a) lv[l].n[a] = g.n[a] >> l;
        if (l > 0) {
            size_t c = lv[l].cells();
            lv[l].types.alloc(c); lv[l].x.alloc(c); lv[l].b.alloc(c); lv[l].r.alloc(c);
        }
    }
    size_t nparts = std::max<size_t>(1024, (size_t)cdiv(g.n[0] + 1, 64) * cdiv(g.n[1] + 1, 4) * (g.n[2] + 1));
    partials.alloc(nparts);
    scal.alloc(16);
    for (int w = 0; w < 2; ++w)
        for (int f = 0; f < PFIELDS; ++f) pbuf[w][f].alloc(std::max<int64_t>(cap, 1));
    key.alloc(std::max<int64_t>(cap, 1)); rank.alloc(std::max<int64_t>(cap, 1)); seg.alloc(std::max<int64_t>(cap, 1));
    counts.alloc(C); totals.alloc(C); offsets.alloc(C); cstart.alloc(C);
    bsums.alloc(C / SCAN_TILE + 2);
    hipStream_t s = st.s;
    for (int a = 0; a < 3; ++a) {
        fill_async(u[a].p, faces(a), 0.0f, s); fill_async(u0[a].p, faces(a), 0.0f, s);
        fill_async(flag[a].p, faces(a), (uint8_t)0, s);
    }
    for (auto* f : {&phi, &phi0, &P, &b, &r, &z, &d, &q, &diag, &e, &rv}) fill_async(f->p, C, 0.0f, s);
    fill_async(P64.p, C, 0.0, s);
    fill_async(types.p, C, (uint8_t)0, s);
    fill_async(counts.p, C, 0, s); fill_async(totals.p, C, 0, s); fill_async(offsets.p, C, 0, s);
    fill_async(scal.p, (size_t)16, 0.0, s);
    st.sync();
}

// flip.jl:298-308 initphi_kernel!: the balanced surface as the level set.
__global__ void k_initphi(float* __restrict__ phi, const float* __restrict__ eta, int neta, DGrid g, float rmax) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) return;
    float px = g.lo[0] + ((float)i + 0.5f) * g.dx, py = g.lo[1] + ((float)j + 0.5f) * g.dx,
          pz = g.lo[2] + ((float)k + 0.5f) * g.dx;
    float r = sqrtf(px * px + py * py);
    float t = jclamp(__fdiv_rn(r, rmax), 0.0f, 1.0f) * (float)(neta - 1);
    int a = min((int)t, neta - 2);
    float f = t - (float)a;
    float h = eta[a] * (1.0f - f) + eta[a + 1] * f;
    phi[(size_t)i + (size_t)g.n[0] * ((size_t)j + (size_t)g.n[1] * k)] = pz - h;
}

// flip.jl:310-316 initvelocity_kernel!: the Lamb-Oseen swirl on the faces of axis A.
template <int A>
__global__ void k_initvel(float* __restrict__ u, DGrid g, float Gamma, float core) {
    int n1 = g.n[0] + (A == 0), n2 = g.n[1] + (A == 1), n3 = g.n[2] + (A == 2);
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n1 || j >= n2 || k >= n3) return;
    float px = g.lo[0] + ((float)i + (A == 0 ? 0.0f : 0.5f)) * g.dx;
    float py = g.lo[1] + ((float)j + (A == 1 ? 0.0f : 0.5f)) * g.dx;
    float r = sqrtf(px * px + py * py);
    float turn = 0.0f;
    if (r > 1e-4f) {
        float q = __fdiv_rn(r, core);
        float sw = __fdiv_rn(Gamma, 2.0f * 3.14159265f * fmaxf(r, 1e-4f)) * (1.0f - expf(-(q * q)));
        turn = __fdiv_rn(sw, r);
    }
    u[(size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k)] = A == 0 ? -turn * py : (A == 1 ? turn * px : 0.0f);
}

void DenseBasin::init_vortex(const float* eta_host, int neta, float rmax, float Gamma, float core) {
    hipStream_t s = st.s;
    DBuf<float> eta;
    eta.alloc(neta);
    MAEL_HIP(hipMemcpyAsync(eta.p, eta_host, sizeof(float) * neta, hipMemcpyHostToDevice, s));
    dim3 blk(64, 4);
    hipLaunchKernelGGL(k_initphi, dim3(cdiv(g.n[0], 64), cdiv(g.n[1], 4), g.n[2]), blk, 0, s, phi.p, eta.p, neta, g, rmax);
    hipLaunchKernelGGL(k_initvel<0>, dim3(cdiv(g.n[0] + 1, 64), cdiv(g.n[1], 4), g.n[2]), blk, 0, s, u[0].p, g, Gamma, core);
    hipLaunchKernelGGL(k_initvel<1>, dim3(cdiv(g.n[0], 64), cdiv(g.n[1] + 1, 4), g.n[2]), blk, 0, s, u[1].p, g, Gamma, core);
    hipLaunchKernelGGL(k_initvel<2>, dim3(cdiv(g.n[0], 64), cdiv(g.n[1], 4), g.n[2] + 1), blk, 0, s, u[2].p, g, Gamma, core);
    MAEL_LAUNCH_CHECK();
    st.sync();   // `eta` is freed on return
}

// flip.jl:674-677 classify_kernel!.
__global__ void k_classify(uint8_t* __restrict__ types, const float* __restrict__ phi, DGrid g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) return;
    size_t l = (size_t)i + (size_t)g.n[0] * ((size_t)j + (size_t)g.n[1] * k);
    types[l] = issolidcell(g, i, j, k) ? SOLID : (phi[l] < 0.0f ? LIQUID : AIR);
}
void DenseBasin::classify() {
    hipLaunchKernelGGL(k_classify, dim3(cdiv(g.n[0], 64), cdiv(g.n[1], 4), g.n[2]), dim3(64, 4), 0, st.s, types.p, phi.p, g);
    MAEL_LAUNCH_CHECK();
}

// flip.jl:779-797 walled / solid_kernel!: zero on every face with a wall on a side.
template <int A> __global__ void k_solid(float* __restrict__ u, DGrid g) {
    int n1 = g.n[0] + (A == 0), n2 = g.n[1] + (A == 1), n3 = g.n[2] + (A == 2);
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= n1 || j >= n2 || k >= n3) return;
    int a = i - (A == 0), b = j - (A == 1), c = k - (A == 2);
    bool walled = !(a >= 0 && b >= 0 && c >= 0 && i < g.n[0] && j < g.n[1] && k < g.n[2]) || issolidcell(g, a, b, c) ||
                  issolidcell(g, i, j, k);
    if (walled) u[(size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k)] = 0.0f;
}
void DenseBasin::solidify() {
    dim3 blk(64, 4);
    hipLaunchKernelGGL(k_solid<0>, dim3(cdiv(g.n[0] + 1, 64), cdiv(g.n[1], 4), g.n[2]), blk, 0, st.s, u[0].p, g);
    hipLaunchKernelGGL(k_solid<1>, dim3(cdiv(g.n[0], 64), cdiv(g.n[1] + 1, 4), g.n[2]), blk, 0, st.s, u[1].p, g);
    hipLaunchKernelGGL(k_solid<2>, dim3(cdiv(g.n[0], 64), cdiv(g.n[1], 4), g.n[2] + 1), blk, 0, st.s, u[2].p, g);
    MAEL_LAUNCH_CHECK();
}

// The gravity term of flip.jl:765-775 forces_kernel! (the stir is not on this path):
// w -= dt g on every non-solid z face.
__global__ void k_gravity(float* __restrict__ w, DGrid g, float dt, float grav) {
    int n3 = g.n[2] + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= g.n[0] || j >= g.n[1] || k >= n3) return;
    float px = g.lo[0] + ((float)i + 0.5f) * g.dx, py = g.lo[1] + ((float)j + 0.5f) * g.dx, pz = g.lo[2] + (float)k * g.dx;
    if (!issolid(g.radius, g.floor, px, py, pz)) w[(size_t)i + (size_t)g.n[0] * ((size_t)j + (size_t)g.n[1] * k)] += dt * -grav;
}
void DenseBasin::gravity(float dt, float grav) {
    hipLaunchKernelGGL(k_gravity, dim3(cdiv(g.n[0], 64), cdiv(g.n[1], 4), g.n[2] + 1), dim3(64, 4), 0, st.s, u[2].p, g, dt, grav);
    MAEL_LAUNCH_CHECK();
    solidify();
}

// flip.jl:624-664 reinit_kernel!: one Jacobi sweep of redistancing.
__global__ void k_reinit(float* __restrict__ out, const float* __restrict__ phi, DGrid g, float phimax) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    int n1 = g.n[0], n2 = g.n[1], n3 = g.n[2];
    if (i >= n1 || j >= n2 || k >= n3) return;
    const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    size_t c0 = (size_t)i + (size_t)n1 * ((size_t)j + (size_t)n2 * k);
    float c = phi[c0];
    float res;
    if (issolidcell(g, i, j, k)) {
        float lowest = c;
        bool found = false;
        for (int m = 0; m < 6; ++m) {
            int a = i + nb[m][0], b = j + nb[m][1], cc = k + nb[m][2];
            if (a < 0 || b < 0 || cc < 0 || a >= n1 || b >= n2 || cc >= n3 || issolidcell(g, a, b, cc)) continue;
            float f = phi[(size_t)a + (size_t)n1 * ((size_t)b + (size_t)n2 * cc)];
            lowest = found ? fminf(lowest, f) : f;
            found = true;
        }
        res = found ? lowest : jclamp(c, -phimax, phimax);
    } else {
        bool inside = c < 0.0f, frozen = false;
        float ax = phimax, ay = phimax, az = phimax;
        for (int m = 0; m < 6; ++m) {
            int a = i + nb[m][0], b = j + nb[m][1], cc = k + nb[m][2];
            if (a < 0 || b < 0 || cc < 0 || a >= n1 || b >= n2 || cc >= n3 || issolidcell(g, a, b, cc)) continue;
            float f = phi[(size_t)a + (size_t)n1 * ((size_t)b + (size_t)n2 * cc)];
            if ((f < 0.0f) != inside) frozen = true;
            float d = fabsf(f);
            if (m < 2) ax = fminf(ax, d);
            else if (m < 4) ay = fminf(ay, d);
            else az = fminf(az, d);
        }
        if (frozen) res = c;
        else {
            float d = fminf(eikonal(ax, ay, az, g.dx), phimax);
            res = inside ? -d : d;
        }
    }
    out[c0] = res;
}

// flip.jl:672-679 redistance!: `iterations` Jacobi sweeps, phi and phi0 ping-ponged (phi0
// is the step's scratch here, as in the reference).
void DenseBasin::redistance(int iterations) {
    float phimax = (float)(iterations + 1) * g.dx;
    for (int it = 0; it < iterations; ++it) {
        hipLaunchKernelGGL(k_reinit, dim3(cdiv(g.n[0], 64), cdiv(g.n[1], 4), g.n[2]), dim3(64, 4), 0, st.s, phi0.p, phi.p,
                           g, phimax);
        std::swap(phi.p, phi0.p);
    }
    MAEL_LAUNCH_CHECK();
}

// flip.jl:262-279 volume_kernel!, summed in FP64 in a fixed order.
__global__ __launch_bounds__(256) void k_volume(double* __restrict__ partials, const float* __restrict__ phi, DGrid g) {
    __shared__ double lds[4];
    size_t C = (size_t)g.n[0] * g.n[1] * g.n[2];
    double v = 0;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < C; l += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(l % (size_t)g.n[0]), j = (int)((l / (size_t)g.n[0]) % (size_t)g.n[1]), k = (int)(l / ((size_t)g.n[0] * g.n[1]));
        if (!issolidcell(g, i, j, k)) v += (double)jclamp(0.5f - __fdiv_rn(phi[l], g.dx), 0.0f, 1.0f);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}
double DenseBasin::volume() {
    hipLaunchKernelGGL(k_volume, dim3(1024), dim3(256), 0, st.s, partials.p, phi.p, g);
    MAEL_LAUNCH_CHECK();
    std::vector<double> h(1024);
    MAEL_HIP(hipMemcpyAsync(h.data(), partials.p, sizeof(double) * 1024, hipMemcpyDeviceToHost, st.s));
    st.sync();
    double s = 0;
    for (double x : h) s += x;
    return s * (double)g.dx * g.dx * g.dx;
}

}  // namespace mael
