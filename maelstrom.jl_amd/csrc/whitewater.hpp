// Whitewater's device state (whitewater.jl:210-300 `Whitewater`, `WhitewaterRules`): a
// ring buffer of diffuse particles and the rules they are made and moved by.
#pragma once
#include "dense.hpp"

namespace mael {

constexpr uint8_t WW_SPRAY = 1, WW_FOAM = 2, WW_BUBBLE = 3;   // bakefiles.jl:71-73

// whitewater.jl:30-44 WhitewaterRules, with the whitewater's floor, gravity and the stream
// it rides (whitewater.jl:306-310 `carried`: the basin's swirl, or none).
struct WhiteParams {
    float air[2] = {5, 20}, crest[2] = {2, 8}, energy[2] = {2, 20};
    float kair = 40, kcrest = 80;
    int32_t spray = 6, bubble = 20;
    float life[2] = {2, 8};
    float buoyancy = 2, drag = 0.5f, kick = 0, tossed = 1;
    float within = __builtin_inff();
    float floor = 0, gravity = 9.81f;
    int32_t stream = 0;          // 1: the waves ride the basin's swirl
    float gamma = 0, core = 1;   // its circulation and core
};

struct WhiteView {
    float* x[3];
    float* v[3];
    float* life;
    uint8_t* kind;
    int64_t capacity;
};

struct Whitewater {
    WhiteParams P;
    int64_t capacity;
    int64_t head = 0;     // particles made so far (the next one goes to slot (head + 1) mod capacity)
    int64_t frames = 0;   // frames made (each frame's emission seed)
    Stream st;
    DBuf<float> x[3], v[3], life;
    DBuf<uint8_t> kind;
    // scratch of an emission: the fluid particles' normals, new-particle counts (then their
    // offsets) and the numbers the writing pass needs; of a packing: counts and Float16 output
    DBuf<float> nrm;
    DBuf<unsigned long long> counts, pcounts;
    DBuf<float4> keep;
    DBuf<__half> pout, pvel;
    Whitewater(int64_t capacity, const WhiteParams& p);
    WhiteView view();
    int64_t pack(uint16_t* out, uint16_t* vel);   // whitewater.jl:366-385: the live ones as Float16
};

struct SparseBasin;
// flip.jl:1496-1511 / sflip.jl:2401-2417 whitewater!(w, s, u, v, dt; substeps, near, above,
// below): the frame's emission (seed: the frame count) and `substeps` diffuse steps.
void dense_whitewater(DenseBasin& b, Whitewater& w, double dt, int substeps, float near, float above, float below);
void sparse_whitewater(SparseBasin& b, Whitewater& w, double dt, int substeps, float near, float above, float below);

}  // namespace mael
