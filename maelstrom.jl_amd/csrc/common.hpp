// Shared device-side helpers and host plumbing of the MI355X (gfx950) hot-path library.
//
// Layout conventions (all 0-based):
//  * dense cells: l = i + n1 (j + n2 k), x fastest (the reference's column-major order);
//  * face arrays along axis a have one more sample along a: u (n1+1, n2, n3) etc.;
//  * particles are SoA: x, y, z; vx, vy, vz; the APIC rows c[3a+b]; lag.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <string>
#include <stdexcept>

namespace mael {

constexpr uint8_t SOLID = 0, LIQUID = 1, AIR = 2;

// Status codes of the C ABI (mirrored in include/maelstrom_hip.h).
enum Status : int {
    OK = 0, E_ARG = -1, E_CAPACITY = -2, E_BLOCKS = -3, E_NONFINITE = -4, E_CG_INDEFINITE = -5,
    E_NOCONVERGE = -6, E_BREAKDOWN = -7, E_HIP = -8, E_GRID = -9, E_TIMEOUT = -10,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MAEL_HIP(call)                                                                                 \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::mael::Error(::mael::E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

#define MAEL_LAUNCH_CHECK() MAEL_HIP(hipGetLastError())

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// ─── grid descriptors passed by value to kernels ─────────────────────────────

struct DGrid {            // a dense MAC grid (the reference's FlipGrid)
    float lo[3];
    float dx;
    int n[3];
    float radius, floor;
};

// ─── device math (reference semantics, unfused; the library is built with -ffp-contract=off) ──

__host__ __device__ inline float jclamp(float x, float lo, float hi) { return x > hi ? hi : (x < lo ? lo : x); }

// Julia's min/max on Float32: NaN-propagating, -0 below +0.
__device__ inline float jmin(float x, float y) {
    if (x != x || y != y) return __builtin_nanf("");
    if (y < x) return y;
    if (x < y) return x;
    return (signbit(y) && !signbit(x)) ? y : x;
}
__device__ inline float jmax(float x, float y) {
    if (x != x || y != y) return __builtin_nanf("");
    if (y > x) return y;
    if (x > y) return x;
    return (signbit(x) && !signbit(y)) ? y : x;
}

// The reference's hash01 (sea.jl:61-66): a 32-bit mix scaled by 2^-32 with one
// correctly rounded conversion and one IEEE division.
__device__ inline float hash01(int32_t i, int32_t j) {
    uint32_t k = ((uint32_t)i * 0x27d4eb2du) ^ ((uint32_t)j * 0x165667b1u);
    k = (k ^ (k >> 15)) * 0x85ebca6bu;
    k = (k ^ (k >> 13)) * 0xc2b2ae35u;
    return __fdiv_rn(__uint2float_rn(k ^ (k >> 16)), 4294967296.0f);
}

struct BSpline { int b; float w0, w1, w2; };
// Quadratic B-spline weights (flip.jl:99-105).
__device__ inline BSpline bspline(float q) {
    float b = floorf(q - 0.5f);
    float f = q - b;
    float a0 = 1.5f - f, a1 = f - 1.0f, a2 = f - 0.5f;
    BSpline r;
    r.b = (int)b;
    r.w0 = 0.5f * (a0 * a0);
    r.w1 = 0.75f - a1 * a1;
    r.w2 = 0.5f * (a2 * a2);
    return r;
}
__device__ inline float pick(const BSpline& s, int m) { return m == 0 ? s.w0 : (m == 1 ? s.w1 : s.w2); }

__device__ inline int32_t orderkey(float f) {
    int32_t b = __float_as_int(f);
    return b >= 0 ? b : (b ^ 0x7fffffff);
}
__device__ inline float orderfloat(int32_t k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff)); }

template <class T> __device__ inline T theta(T pi, T pj) {
    T t = pi / (pi - pj);
    return t > T(1) ? T(1) : (t < T(0.01f) ? T(0.01f) : t);
}

__device__ inline float temporal(float lag, float dt) {
    float r = jclamp(__fdiv_rn(-lag, dt), -0.5f, 0.5f) - 0.5f;
    float a = 1.0f - r * r;
    return 2.1875f * (a * a * a);
}
__device__ inline float strength(float speed, float dt, float h) {
    float c = jclamp(__fdiv_rn(speed * dt, h), 0.0f, 1.0f);
    return c * c * (3.0f - 2.0f * c);
}
__device__ inline float ahead(float g) {
    float h2 = 0.25f * g * g;
    float num = h2 * (0.5625f - h2 * (0.7f - __fdiv_rn(h2 * 3.0f, 7.0f)));
    float den = 0.421875f + h2 * (0.1875f - h2 * (0.15f + __fdiv_rn(h2, 7.0f)));
    return __fdiv_rn(num, den);
}

__device__ inline bool issolid(float radius, float fl, float px, float py, float pz) {
    return px * px + py * py > radius * radius || pz < fl;
}
__device__ inline bool issolidcell(const DGrid& g, int i, int j, int k) {
    return issolid(g.radius, g.floor, g.lo[0] + ((float)i + 0.5f) * g.dx, g.lo[1] + ((float)j + 0.5f) * g.dx,
                   g.lo[2] + ((float)k + 0.5f) * g.dx);
}

// Index space of a sample point: (p - lo)/dx - offset, IEEE division.
__device__ inline float ispace(float p, float lo, float dx, float off) { return __fdiv_rn(p - lo, dx) - off; }

// Trilinear sample of a 3-D array at index-space q, clamped (interpolation.jl:4-17).
template <class V>
__device__ inline float trilinear(const V* __restrict__ A, int n1, int n2, int n3, float q0, float q1, float q2) {
    float x = jclamp(q0, 0.0f, (float)(n1 - 1)), y = jclamp(q1, 0.0f, (float)(n2 - 1)), z = jclamp(q2, 0.0f, (float)(n3 - 1));
    int i = min((int)x, n1 - 2), j = min((int)y, n2 - 2), k = min((int)z, n3 - 2);
    float fx = x - (float)i, fy = y - (float)j, fz = z - (float)k;
    float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    size_t s1 = (size_t)n1, s2 = (size_t)n1 * n2;
    size_t b = (size_t)i + s1 * j + s2 * k;
    float c00 = (float)A[b] * gx + (float)A[b + 1] * fx;
    float c10 = (float)A[b + s1] * gx + (float)A[b + s1 + 1] * fx;
    float c01 = (float)A[b + s2] * gx + (float)A[b + s2 + 1] * fx;
    float c11 = (float)A[b + s2 + s1] * gx + (float)A[b + s2 + s1 + 1] * fx;
    return (c00 * gy + c10 * fy) * gz + (c01 * gy + c11 * fy) * fz;
}

// APIC gather of one face component at xp over clamped nodes (flip.jl:1210-1224):
// velocity and affine part (4/dx^2 sum w u (x_i - x_p)). Same summation order as the
// reference, so results are reproducible bit for bit.
template <class V>
__device__ inline void gather(const V* __restrict__ u, int n1, int n2, int n3, const DGrid& g, int a, float px, float py,
                              float pz, float& vel, float& a0, float& a1, float& a2) {
    float q0 = ispace(px, g.lo[0], g.dx, a == 0 ? 0.0f : 0.5f);
    float q1 = ispace(py, g.lo[1], g.dx, a == 1 ? 0.0f : 0.5f);
    float q2 = ispace(pz, g.lo[2], g.dx, a == 2 ? 0.0f : 0.5f);
    BSpline bx = bspline(q0), by = bspline(q1), bz = bspline(q2);
    vel = 0.0f; a0 = 0.0f; a1 = 0.0f; a2 = 0.0f;
    size_t s1 = (size_t)n1, s2 = (size_t)n1 * n2;
#pragma unroll
    for (int mz = 0; mz < 3; ++mz) {
        int k = min(max(bz.b + mz, 0), n3 - 1);
        float wz = pick(bz, mz), dz = (float)(bz.b + mz) - q2;
#pragma unroll
        for (int my = 0; my < 3; ++my) {
            int j = min(max(by.b + my, 0), n2 - 1);
            float wy = pick(by, my), dy = (float)(by.b + my) - q1;
#pragma unroll
            for (int mx = 0; mx < 3; ++mx) {
                int i = min(max(bx.b + mx, 0), n1 - 1);
                float wgt = pick(bx, mx) * wy * wz;
                float val = (float)u[(size_t)i + s1 * j + s2 * k];
                float t = wgt * val;
                vel += t;
                a0 += (t * ((float)(bx.b + mx) - q0)) * g.dx;
                a1 += (t * dy) * g.dx;
                a2 += (t * dz) * g.dx;
            }
        }
    }
    float c = __fdiv_rn(4.0f, g.dx * g.dx);
    a0 *= c; a1 *= c; a2 *= c;
}

// flip.jl:611-622 eikonal: the distance at which |grad phi| = 1 from the smallest
// neighbour distances along each axis (Godunov upwind).
__device__ inline float eikonal(float a, float b, float c, float h) {
    float t;
    if (b < a) { t = a; a = b; b = t; }
    if (c < b) { t = b; b = c; c = t; }
    if (b < a) { t = a; a = b; b = t; }
    float d = a + h;
    if (d > b) {
        float ab = a - b;
        d = 0.5f * (a + b + sqrtf(fmaxf(2.0f * h * h - ab * ab, 0.0f)));
        if (d > c) {
            float s = a + b + c;
            d = (s + sqrtf(fmaxf(s * s - 3.0f * (a * a + b * b + c * c - h * h), 0.0f))) / 3.0f;
        }
    }
    return d;
}

// sea.jl:39 oseen: the Lamb-Oseen vortex's azimuthal speed at radius r.
__device__ inline float oseen_dev(float G, float core, float r) {
    float q = __fdiv_rn(r, core);
    return __fdiv_rn(G, 2.0f * 3.14159265f * fmaxf(r, 1e-4f)) * (1.0f - expf(-(q * q)));
}

// ─── particles ───────────────────────────────────────────────────────────────

constexpr int PFIELDS = 16;   // x y z vx vy vz c0..c8 lag
struct PView {                 // device pointers of one particle buffer (SoA)
    float* f[PFIELDS];
};

// MI355X deals a launch's workgroups round-robin over its 8 XCDs, each with its own L2.
// Stencil kernels whose neighbouring workgroups share data take their tile from this
// bijection of the workgroup id instead, so that consecutive tiles run on one XCD:
// XCD x (= b mod 8) gets the x-th contiguous run of tile ids.
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t G) {
    uint32_t q = G / 8, r = G % 8, x = b % 8, i = b / 8;
    return x * q + min(x, r) + i;
}

// (I, J, K) of linear index l of an n0 x n1 x n2 array (x fastest), in 32-bit arithmetic
// (every grid here has fewer than 2^31 entries).
__device__ inline void unlin(size_t l, int n0, int n1, int& I, int& J, int& K) {
    const uint32_t l32 = (uint32_t)l, q = l32 / (uint32_t)n0;
    I = (int)(l32 - q * (uint32_t)n0);
    K = (int)(q / (uint32_t)n1);
    J = (int)(q - (uint32_t)K * (uint32_t)n1);
}

// atomicAdd(&counts[idx], 1) for every lane with idx >= 0, with one atomic per distinct
// index among the wavefront's lanes (cell-sorted particles share a few cells a wavefront,
// and same-address atomics serialise): returns the lane's slot, the old count plus its
// place among the lanes of that index in lane order (-1 for idx < 0). Every active lane of
// the wavefront must call it (the loop runs on wavefront-uniform values).
__device__ inline int32_t wave_count_slot(int32_t* __restrict__ counts, int32_t idx) {
    const int lane = (int)(threadIdx.x & 63);
    unsigned long long todo = __ballot(idx >= 0);
    int32_t slot = -1;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int32_t k = __shfl(idx, leader, 64);
        const unsigned long long grp = __ballot(idx == k) & todo;
        int32_t base = 0;
        if (lane == leader) base = atomicAdd(&counts[k], (int32_t)__popcll(grp));
        base = __shfl(base, leader, 64);
        if ((grp >> lane) & 1ull) slot = base + (int32_t)__popcll(grp & ((1ull << lane) - 1ull));
        todo &= ~grp;
    }
    return slot;
}

// ─── fixed-point accumulation (deterministic P2G) ────────────────────────────
// A contribution c is accumulated as round(c * 2^32) in a 64-bit integer: integer
// addition is associative, so the sum does not depend on the order the particles
// arrive in. Range +-2^31, resolution 2^-32.
constexpr double FIX_SCALE = 4294967296.0;
constexpr double FIX_INV = 1.0 / 4294967296.0;
__device__ inline unsigned long long tofix(float c) {
    return (unsigned long long)(long long)__double2ll_rn((double)c * FIX_SCALE);
}
// tofix for |c| < 2^19, in fewer instructions: c 2^32 is exact in FP64, and adding
// 1.5 * 2^52 rounds it to an integer (to nearest even, as tofix does) whose two's
// complement is the sum's low bits.
__device__ inline unsigned long long tofix_near(float c) {
    constexpr double MAGIC = 6755399441055744.0;   // 1.5 * 2^52
    const double x = (double)c * FIX_SCALE + MAGIC;
    return (unsigned long long)(__double_as_longlong(x) - __double_as_longlong(MAGIC));
}
__device__ inline float fromfix(unsigned long long v) { return (float)((double)(long long)v * FIX_INV); }

}  // namespace mael
