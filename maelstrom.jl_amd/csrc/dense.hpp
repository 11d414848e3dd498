// Device state of one dense basin (the reference's `FlipBasin`, flip.jl:156-226) and
// its phases. One HIP stream per basin; every phase is asynchronous on it except
// where a host value is returned (the resample total, solver statistics).
#pragma once
#include "common.hpp"
#include "stir.hpp"
#include <hip/hip_fp16.h>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

namespace mael {

// A device allocation that frees itself.
template <class T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p; n = o.n; o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t count) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        if (count) MAEL_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    }
};

struct SolveOpts {
    double tol = 1e-4;
    int maxiter = 60, sweeps = 2, coarse = 40;
    int precision = 0;        // 0 FP32 V-cycle, 1 FP16 storage, 2 BF16 storage
    int refine = 0;           // 1: FP64 true-residual recomputation with restarts (0: the reference's FP32 PCG)
    int warm = 1;             // start from the last pressure (else from zero)
};
struct SolveStats {
    int iterations = 0, restarts = 0;
    double rel_residual = 0, rel_residual_recursive = 0;
    double ms_total = 0;
};

struct Stream {
    hipStream_t s = nullptr;
    hipEvent_t rd = nullptr;
    void* pinned = nullptr;   // small pinned host scratch for scalars read back
    Stream() {
        MAEL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        MAEL_HIP(hipEventCreateWithFlags(&rd, hipEventDisableTiming));
        MAEL_HIP(hipHostMalloc(&pinned, 4096, hipHostMallocDefault));
    }
    ~Stream() {
        if (rd) (void)hipEventDestroy(rd);
        if (s) (void)hipStreamDestroy(s);
        if (pinned) (void)hipHostFree(pinned);
    }
    void sync() { MAEL_HIP(hipStreamSynchronize(s)); }
    // The same wait by polling an event: the Krylov loops read a scalar back once an
    // iteration, on the critical path, and a blocking wait wakes the host later.
    void spin() {
        mark();
        poll();
    }
    void mark() { MAEL_HIP(hipEventRecord(rd, s)); }   // what poll() waits for
    // Spins for the first POLL_SPIN queries (a Krylov read-back lands in tens of
    // microseconds), then yields the core between queries; a wait longer than
    // POLL_LIMIT_S seconds is reported as E_TIMEOUT instead of spinning on forever.
    static constexpr int POLL_SPIN = 20000;
    static constexpr double POLL_LIMIT_S = 300.0;
    void poll() {
        hipError_t e;
        long long polls = 0;
        std::chrono::steady_clock::time_point t0;
        while ((e = hipEventQuery(rd)) == hipErrorNotReady) {
            if (++polls < POLL_SPIN) continue;
            if (polls == POLL_SPIN) t0 = std::chrono::steady_clock::now();
            std::this_thread::yield();
            if ((polls & 1023) == 0 &&
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > POLL_LIMIT_S)
                throw Error(E_TIMEOUT, "a device wait exceeded " + std::to_string((int)POLL_LIMIT_S) + " s");
        }
        MAEL_HIP(e);
    }
};

// Kernel ids for the launch timer (mirrored as MAEL_K_* in include/maelstrom_hip.h).
enum KernelId {
    K_P2G = 0, K_SPLAT, K_COMBINE, K_G2P, K_BIN, K_FIXRANK, K_SCATTER, K_SPAWN, K_SEARCH, K_ADVANCE, K_RBGS,
    K_RESIDUAL, K_RESTRICT, K_PROLONG, K_COARSEST, K_RESID64, K_NKERNELS
};

// HIP events around every launch of one chosen kernel, on the stream it runs on, so a
// benchmark can read that kernel's launch durations over a timed region.
// Times every STRIDE-th launch of the target kernel: each recorded event leaves the
// stream idle for a few microseconds, so an event pair around every launch would slow
// the run it measures.
struct KTimer {
    static constexpr size_t STRIDE = 8;
    int target = -1;
    std::vector<hipEvent_t> ev;
    size_t used = 0, seen = 0;
    bool open = false;
    ~KTimer() { for (auto e : ev) (void)hipEventDestroy(e); }
    void enable(int id, size_t launches) {
        while (ev.size() < 2 * launches) {
            hipEvent_t e;
            MAEL_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
        target = id;
        used = 0;
        seen = 0;
        open = false;
    }
    bool on(int id) const { return id == target && used + 2 <= ev.size(); }
    void begin(int id, hipStream_t s) {
        if (!on(id)) return;
        open = seen++ % STRIDE == 0;
        if (open) MAEL_HIP(hipEventRecord(ev[used], s));
    }
    void end(int id, hipStream_t s) {
        if (on(id) && open) {
            MAEL_HIP(hipEventRecord(ev[used + 1], s));
            used += 2;
            open = false;
        }
    }
    double read(int* launches) {   // total ms of the timed launches; disables the timer
        double t = 0;
        for (size_t i = 0; i < used; i += 2) {
            MAEL_HIP(hipEventSynchronize(ev[i + 1]));
            float ms = 0;
            MAEL_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            t += ms;
        }
        if (launches) *launches = (int)(used / 2);
        target = -1;
        used = 0;
        return t;
    }
};

// Levels of the dense multigrid (flip.jl:117-123, 195-202): types and x, b, r per level.
struct DLevel {
    int n[3];
    DBuf<uint8_t> types;
    DBuf<float> x, b, r;
    size_t cells() const { return (size_t)n[0] * n[1] * n[2]; }
};

struct DenseBasin {
    DGrid g;
    int levels;
    int32_t band, nmin;
    Stream st;
    KTimer kt;

    // grid state (flip.jl:160-172)
    DBuf<float> u[3], u0[3];
    DBuf<float> phi, phi0;
    DBuf<int32_t> phip;
    DBuf<uint8_t> types, flag[3];
    DBuf<float> P, b, r, z, d, q, diag;
    DBuf<double> P64;                         // the pressure, accumulated in FP64
    DBuf<float> e, rv;                        // CG's FP32 correction; level 0's V-cycle residual
    DBuf<unsigned long long> mom[3], wt[3];   // fixed-point P2G accumulators
    std::vector<DLevel> lv;                   // lv[0] mirrors the fine grid (only n is used)
    DBuf<double> partials;                    // per-block partial dot products
    DBuf<double> scal;                        // device scalars of the Krylov loop

    // particles: two SoA buffers (flip.jl:131-151), `cur` is the live one
    DBuf<float> pbuf[2][PFIELDS];
    int cur = 0;
    int64_t n = 0, capacity = 0;

    // resampling (flip.jl:216-219)
    DBuf<int32_t> key, rank, seg;
    DBuf<int32_t> counts, totals, offsets, cstart;
    DBuf<unsigned long long> bsums;

    size_t cells() const { return (size_t)g.n[0] * g.n[1] * g.n[2]; }
    size_t faces(int a) const { return (size_t)(g.n[0] + (a == 0)) * (g.n[1] + (a == 1)) * (g.n[2] + (a == 2)); }
    PView view(int which) const {
        PView v;
        for (int f = 0; f < PFIELDS; ++f) v.f[f] = pbuf[which][f].p;
        return v;
    }

    DenseBasin(const DGrid& grid, int levels, int64_t capacity, int32_t band, int32_t nmin);

    // phases
    // flip.jl:1364-1382; `tilesums` (nullptr: the basin's bsums) receives the packed
    // scan's tile sums, cells() / SCAN_TILE + 1 of them
    int64_t resample(bool seeding, uint32_t seed, unsigned long long* tilesums = nullptr);
    void p2g(float dt);                                      // flip.jl:587-606
    void g2p(float dt, float gravity);                       // flip.jl:1246-1252
    void project(float dt, float gamma, const SolveOpts& o, SolveStats& s);   // flip.jl:1144-1156
    void operator_setup();                                   // coarse types and the fine diagonal
    void solve_pressure(const SolveOpts& o, SolveStats& s);  // A P = b from P
    // z = M b: PCG's V-cycle preconditioner applied once to the right-hand side, into P
    // (with `queued`, through the path the CG loop takes when the V-cycle's first launches
    // were queued ahead of a host read)
    void precondition(const SolveOpts& o, bool queued);
    // adjoints of the hot path's linear maps (no counterpart in the reference): cotangents
    // of the face velocities in ub (u0's in ub0) and of the particles' v, c in pb
    DBuf<float> ub[3], ub0[3], pb[12];
    DBuf<unsigned long long> adjacc[3];
    DBuf<uint8_t> reached;
    void adj_alloc();
    void g2p_adjoint(float dt);
    void project_adjoint(const SolveOpts& o, SolveStats& s);
    void p2g_adjoint(float dt);
    // support phases for initialisation and a whole step
    void classify();                                         // flip.jl:674-677
    void redistance(int iterations);                         // flip.jl:672-679
    void solidify();                                         // flip.jl:800-807
    double volume();                                         // flip.jl:293
    void init_vortex(const float* eta, int neta, float rmax, float Gamma, float core);   // flip.jl:298-314
    void gravity(float dt, float g);                         // the gravity part of forces!, flip.jl:765-775
    // the rest of a step (dense_step.hip)
    DBuf<uint8_t> reach, flag2[3];                           // trapped air's reach; extrapolation's flags
    DBuf<int32_t> changed;                                   // a device counter (trapped air, maxspeed)
    StirParams stirp;                                        // the stirring model (basin.jl:92-181)
    DBuf<float> stir_waves;
    bool has_stir = false;
    void advect(float dt, float gravity, int32_t seed);      // flip.jl:490-499
    void filltrapped(int maxrounds);                         // flip.jl:754-760
    void set_stir(const StirParams& p, const float* waves, int nwaves);
    void forces(float dt, float time, float gravity);        // flip.jl:786-798
    void extrapolate(int layers);                            // flip.jl:1194-1206
    float maxspeed();                                        // flip.jl:1387
    // the bake writers' device side (bake.hip): the kept band particles as Float16 into
    // host `pos`, `vel` (3 a particle), returning how many; the level set as bytes
    DBuf<__half> bake_pos, bake_vel;
    DBuf<unsigned long long> bake_counts;
    DBuf<uint8_t> bake_q;
    int64_t pack_band(float keep, float throat, uint16_t* pos, uint16_t* vel);   // flip.jl:1880-1904
    void quantize_levelset(float range, uint8_t* out);                          // flip.jl:1859-1874
};

// The resampling scan's tile sums for callers that cannot allocate: a file-scope device
// array (resample.hip) of NOALLOC_TILE_SUMS entries, one copy per device.
constexpr size_t NOALLOC_TILE_SUMS = 16386;
// Resolve the array's address on `device` (current); once per device, at context set-up.
void noalloc_setup(int device);
// Its address on `device`, or nullptr if noalloc_setup has not run there.
unsigned long long* noalloc_scratch(int device);

// filltrapped!'s reach (flip.jl:690-745) on a dense grid of types: seed the top layer's
// AIR, then sweep the lines along z, x and y until a round reaches nothing new (at most
// `maxrounds`); `reach` then marks the AIR joined to the atmosphere.
void sweep_trapped_air(uint8_t* reach, int* changed, const uint8_t* types, int3 n, int maxrounds, Stream& st);

// scan helper shared by the dense and sparse resampling
void launch_scan_bsums(unsigned long long* bsums, int nblocks, hipStream_t s);

}  // namespace mael
