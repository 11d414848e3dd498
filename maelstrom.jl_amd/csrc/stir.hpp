// The stirring model of `forces!` (basin.jl:13-30 Basin, :69-80 Race, :92-181 Stir,
// :191-235 SurfaceStir; sea.jl:60-76 hash and value noise): its parameters and the
// acceleration it gives a face of one axis, shared by the dense and sparse basins.
#pragma once
#include "common.hpp"

namespace mael {

// The stirring model's parameters (basin.jl:13-30 Basin, :69-80 Race, :92-103 Stir,
// :191-205 SurfaceStir); the waves (7 floats each: direction x, y, wave number, angular
// frequency, amplitude already scaled by the race, steepness, phase) and the surface
// table live in device arrays.
struct StirParams {
    float radius = 18, band = 3, relax = 0.3f, core = 2.5f, Gamma = 225;
    float current[2] = {0, 0};
    float eddy = 0.6f, eddyscale = 3.5f, eddytime = 4, guide = 4, free = 0.6f, turbulence = 0, boil = 0.5f, reach = 2.5f;
    float origin[2] = {0, 0};
    float offset = 0;
    int surface = 0;                 // 0 Stir, 1 SurfaceStir
    float hradius = 0, ramp = 2;     // SurfaceStir's table radius and onset ramp
    const float* waves = nullptr;
    int nwaves = 0;
    const float* heights = nullptr;
    int nheights = 0;
};

// sea.jl:68-76 value noise.
__device__ inline float vnoise(float x, float y) {
    float xi = floorf(x), yi = floorf(y);
    float fx = x - xi, fy = y - yi;
    float sx = fx * fx * (3.0f - 2.0f * fx), sy = fy * fy * (3.0f - 2.0f * fy);
    int32_t i = (int32_t)xi, j = (int32_t)yi;
    float a = hash01(i, j), b = hash01(i + 1, j), c = hash01(i, j + 1), d = hash01(i + 1, j + 1);
    return (a + (b - a) * sx) + ((c + (d - c) * sx) - (a + (b - a) * sx)) * sy;
}
// basin.jl:110-121 eddies.
__device__ inline void eddies(const StirParams& P, float x, float y, float t, float e[2]) {
    float s = P.eddyscale;
    float a = __fdiv_rn(x - P.current[0] * t, s), b = __fdiv_rn(y - P.current[1] * t, s);
    float tau = __fdiv_rn(t, P.eddytime);
    auto psi = [&](float p, float q) {
        return vnoise(p + 17.0f, q + tau) + 0.5f * vnoise(2.1f * p - 5.0f, 2.1f * q - 1.7f * tau);
    };
    float dl = 0.25f;
    float dpdy = __fdiv_rn(psi(a, b + dl) - psi(a, b - dl), 2.0f * dl);
    float dpdx = __fdiv_rn(psi(a + dl, b) - psi(a - dl, b), 2.0f * dl);
    e[0] = P.eddy * dpdy;
    e[1] = -P.eddy * dpdx;
}
// basin.jl:124-127 boils.
__device__ inline float boils(const StirParams& P, float x, float y, float t) {
    float nn = vnoise(__fdiv_rn(0.6f * (x - P.current[0] * t), P.eddyscale) + 41.0f,
                      __fdiv_rn(0.6f * (y - P.current[1] * t), P.eddyscale) + __fdiv_rn(0.2f * t, P.eddytime));
    return P.boil * (__fdiv_rn(jmax(nn - 0.6f, 0.0f), 0.4f) - 0.15f);
}
// basin.jl:139-151 orbital: component A of the waves' velocity (A = 0, 1 horizontal, 2
// vertical) at sea coordinates (x, y), depth z, time t.
template <int A>
__device__ inline float orbital(const StirParams& P, float x, float y, float z, float t) {
    float o = 0.0f;
    for (int k = 0; k < P.nwaves; ++k) {
        const float* W = P.waves + 7 * k;   // d0 d1 k omega a q phase
        float th = W[2] * (W[0] * x + W[1] * y) - W[3] * t + W[6];
        float e = W[4] * W[3] * expf(W[2] * jmin(z, 0.0f));
        if (A < 2) o += (e * sinf(th)) * W[A];
        else o -= e * cosf(th);
    }
    return o;
}
__device__ inline float swirl_dev(const StirParams& P, float r) { return oseen_dev(P.Gamma, P.core, r); }

// Component A of the stirring acceleration at (x, y, z) for the water's velocity vel
// along A (basin.jl:153-181 `Stir`, or basin.jl:208-235 `SurfaceStir` when P.surface).
// Each component reads only the water's velocity along its own axis, so a face of axis
// A needs that one sample; the horizontal parts need the eddies, the vertical the boils.
template <int A>
__device__ inline float stir(const StirParams& P, float x, float y, float z, float vel, float t) {
    float r = sqrtf(x * x + y * y);
    float k = jclamp(__fdiv_rn(r - (P.radius - P.band), P.band), 0.0f, 1.0f);
    k = k * k * (3.0f - 2.0f * k);
    float v = swirl_dev(P, r);
    float turn = r > 1e-4f ? __fdiv_rn(v, r) : 0.0f;
    float eddy[2] = {0.0f, 0.0f};
    if (A < 2) eddies(P, x, y, t, eddy);
    float acc;
    if (!P.surface) {
        float churn = jclamp(__fdiv_rn(r - 2.0f, 2.0f), 0.0f, 1.0f);
        if (A < 2) {
            float core = jclamp(2.0f - __fdiv_rn(r, P.core), 0.0f, 1.0f);
            float hold = (1.0f - k) * jmax(jclamp(__fdiv_rn(-z - P.free, P.free), 0.0f, 1.0f), core);
            float aim = A == 0 ? -turn * y + P.current[0] + churn * eddy[0] : turn * x + P.current[1] + churn * eddy[1];
            float gh = __fdiv_rn(hold, P.guide);
            float guide = gh * (aim - vel);
            float top = (1.0f - k) * churn * jclamp(1.0f + __fdiv_rn(z, P.free), 0.0f, 1.0f);
            float pf = P.eddy > 0.0f ? __fdiv_rn(top * P.turbulence, P.eddy) : 0.0f;
            float push = P.eddy > 0.0f ? pf * eddy[A] : 0.0f;
            acc = guide + push;
        } else {
            float rising = (1.0f - k) * expf(__fdiv_rn(jmin(z, 0.0f), P.reach));
            acc = rising * churn * boils(P, x, y, t);
        }
        if (k == 0.0f) return acc;
        float sea = orbital<A>(P, P.origin[0] + x, P.origin[1] + y, z, t - P.offset);
        float target = A == 0 ? __fdiv_rn(-v * y, r) + P.current[0] + sea
                              : (A == 1 ? __fdiv_rn(v * x, r) + P.current[1] + sea : sea);
        float kr = __fdiv_rn(k, P.relax);
        return acc + kr * (target - vel);
    }
    int nh = P.nheights;
    float hh = jclamp(__fdiv_rn(r, P.hradius) * (float)(nh - 1), 0.0f, (float)(nh - 1));
    int i = min((int)floorf(hh), nh - 2);
    float eta = P.heights[i] + (hh - (float)i) * (P.heights[i + 1] - P.heights[i]);
    float relative = z - eta;
    float onset = jclamp(__fdiv_rn(t, P.ramp), 0.0f, 1.0f);
    onset = onset * onset * (3.0f - 2.0f * onset);
    float churn = jclamp(__fdiv_rn(r, 1.3f * P.core), 0.0f, 1.0f);
    churn = churn * churn * (3.0f - 2.0f * churn);
    if (A < 2) {
        float aim = A == 0 ? -turn * y + onset * churn * eddy[0] : turn * x + onset * churn * eddy[1];
        float hold = (1.0f - k) * jclamp(__fdiv_rn(-relative - P.free, P.free), 0.0f, 1.0f);
        float gh = __fdiv_rn(hold, P.guide);
        float guide = gh * (aim - vel);
        float top = (1.0f - k) * churn * expf(__fdiv_rn(jmin(relative, 0.0f), P.free)) * onset;
        float pf = P.eddy > 0.0f ? __fdiv_rn(top * P.turbulence, P.eddy) : 0.0f;
        float push = P.eddy > 0.0f ? pf * eddy[A] : 0.0f;
        acc = guide + push;
    } else {
        float rising = (1.0f - k) * churn * expf(__fdiv_rn(jmin(relative, 0.0f), P.reach)) * onset;
        acc = rising * boils(P, x, y, t);
    }
    if (k == 0.0f) return acc;
    float sea = orbital<A>(P, P.origin[0] + x, P.origin[1] + y, relative, t - P.offset);
    float target = A == 0 ? -turn * y + sea : (A == 1 ? turn * x + sea : sea);
    float kr = __fdiv_rn(k, P.relax);
    return acc + kr * (target - vel);
}

}  // namespace mael
