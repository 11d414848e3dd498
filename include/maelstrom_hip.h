/*
 * maelstrom_hip.h — the C ABI of the MI355X (gfx950) hot path of Maelstrom.jl.
 *
 * The reference runs its hot path as KernelAbstractions kernels launched from Julia
 * (`k!(backend, wg)(args...; ndrange)`, e.g. flip.jl:596) with device arrays from
 * `KA.zeros`/`Mantle.devicearray`. This library replaces that layer at phase
 * granularity: one handle per basin owns the device state, and each entry point runs
 * one phase of a step (resample, P2G, projection, G2P) on the basin's HIP stream, so
 * that the phase's kernels can be fused and the Krylov loop can keep its scalars on the
 * device. The Julia package calls these with `ccall`; see INTEGRATION.md.
 *
 * Conventions
 *  - Every function returns a status (MAEL_OK = 0, negative on error) unless stated.
 *    `mael_last_error(ctx)` then holds the message; where the reference raises an
 *    error the text is the reference's (e.g. "the narrow band needs N particles, more
 *    than the C allocated", flip.jl:1373).
 *  - Indices are 0-based. Grid arrays are x-fastest (the reference's column-major
 *    order). Particles are structure-of-arrays (x, y, z, vx, vy, vz, the nine entries
 *    of the APIC rows cx, cy, cz, lag), unlike the reference's interleaved Point3f.
 *  - Calls on one basin are issued in order on its stream; a call that returns a host
 *    value (a count, statistics) synchronises. A basin is not re-entrant.
 */
#ifndef MAELSTROM_HIP_H
#define MAELSTROM_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define MAEL_OK               0
#define MAEL_E_ARG           -1  /* bad argument, shape or handle */
#define MAEL_E_CAPACITY      -2  /* particle capacity exceeded, flip.jl:1373, sflip.jl:2237 */
#define MAEL_E_BLOCKS        -3  /* fine-block pool exceeded, sflip.jl:530 */
#define MAEL_E_NONFINITE     -4  /* "Nonfinite pressure ...", sflip.jl:1873-1898, fgmres.jl:31-65 */
#define MAEL_E_CG_INDEFINITE -5  /* "Pressure CG requires a positive ...", sflip.jl:1882-1896 */
#define MAEL_E_NOCONVERGE    -6  /* "Pressure did not converge ...", fgmres.jl:90 */
#define MAEL_E_BREAKDOWN     -7  /* "Pressure Arnoldi breakdown", fgmres.jl:77 */
#define MAEL_E_HIP           -8  /* a HIP runtime error */
#define MAEL_E_GRID          -9  /* "the coarse grid ... does not halve ...", sflip.jl:326 */
#define MAEL_E_TIMEOUT      -10  /* a device wait ran past the library's limit (no counterpart in the reference) */

/* ── context ──────────────────────────────────────────────────────────────── */

typedef struct mael_ctx mael_ctx;

/* Open `device` (a HIP device ordinal). Replaces `Mantle.defaultbackend()`
 * (test/runtests.jl:20): the backend every basin is built on. */
int mael_create(int device, mael_ctx **out);
void mael_destroy(mael_ctx *ctx);
/* The message of the last failed call on `ctx` (or on a basin built on it). */
const char *mael_last_error(const mael_ctx *ctx);
/* The library's version string. */
const char *mael_version(void);

/* ── the dense grid (`FlipGrid`, flip.jl:30-62) ───────────────────────────── */

typedef struct {
    float lo[3];      /* the grid's low corner, m */
    float dx;         /* cell width, m */
    int32_t n[3];     /* cells per axis */
    float radius;     /* the basin's wall radius, m */
    float floor;      /* the basin's floor height, m (= -depth) */
    int32_t levels;   /* multigrid levels: the coarsest has at most 1024 cells */
} mael_dense_grid;

/* `FlipGrid(b; dx, top = b.rimwall, pad)` (flip.jl:48-62): the sizing rule. */
int mael_flipgrid(float radius, float depth, float top, float dx, int32_t pad, mael_dense_grid *out);

/* ── a dense basin (`FlipBasin`, flip.jl:156-226) ─────────────────────────── */

typedef struct mael_dense mael_dense;

/* `FlipBasin(b, backend; dx, band, nmin, capacity)` (flip.jl:189-223): allocates the
 * grid fields, the multigrid levels and two particle buffers of `capacity`. */
int mael_dense_create(mael_ctx *ctx, const mael_dense_grid *grid, int64_t capacity, int32_t band, int32_t nmin,
                      mael_dense **out);
void mael_dense_destroy(mael_dense *d);

/* Arrays of a basin, for `Array(s.u)`, `copyto!(s.φ, …)` and the like (the reference reads
 * solver arrays this way in checkpoint.jl:10-13 and live.jl:36-37). Counts are elements. */
enum {
    MAEL_U = 0, MAEL_V, MAEL_W,           /* face velocities (float) */
    MAEL_U0, MAEL_V0, MAEL_W0,            /* the grid's own velocity combine falls back to (float) */
    MAEL_PHI, MAEL_PHI0,                  /* level set (float) */
    MAEL_TYPES,                           /* SOLID 0 / LIQUID 1 / AIR 2 (uint8) */
    MAEL_P,                               /* pressure P = p dt / rho (float) */
    MAEL_FU, MAEL_FV, MAEL_FW,            /* valid-face flags (uint8) */
    MAEL_RHS,                             /* the projection's right-hand side (float) */
    MAEL_PX, MAEL_PY, MAEL_PZ,            /* particle positions (float, live buffer) */
    MAEL_VX, MAEL_VY, MAEL_VZ,            /* particle velocities */
    MAEL_C0, MAEL_C1, MAEL_C2, MAEL_C3, MAEL_C4, MAEL_C5, MAEL_C6, MAEL_C7, MAEL_C8,  /* APIC rows cx, cy, cz */
    MAEL_LAG,                             /* particle time lag */
    MAEL_KEY, MAEL_RANK,                  /* last resample: each particle's cell (-1 dropped) and rank */
    MAEL_COUNTS, MAEL_TOTALS, MAEL_OFFSETS, /* last resample: per cell (offsets inclusive) */
    MAEL_MOM_U, MAEL_MOM_V, MAEL_MOM_W,   /* last P2G: momentum accumulators (int64 fixed point, 2^-32) */
    MAEL_WT_U, MAEL_WT_V, MAEL_WT_W,      /* last P2G: weight accumulators (int64 fixed point, 2^-32) */
    MAEL_UB_U, MAEL_UB_V, MAEL_UB_W,      /* adjoints: cotangents of the face velocities (float) */
    MAEL_UB0_U, MAEL_UB0_V, MAEL_UB0_W,   /* adjoints: cotangents of u0 (float) */
    MAEL_PB_VX, MAEL_PB_VY, MAEL_PB_VZ,   /* adjoints: cotangents of the particles' v (float, capacity) */
    MAEL_PB_C0, MAEL_PB_C1, MAEL_PB_C2, MAEL_PB_C3, MAEL_PB_C4, MAEL_PB_C5, MAEL_PB_C6, MAEL_PB_C7,
    MAEL_PB_C8,                           /* ... and of their APIC rows */
    MAEL_NARRAYS
};
/* Device pointer, element count and element size of array `which`. */
int mael_dense_array(mael_dense *d, int32_t which, void **dptr, int64_t *count, int32_t *elsize);
/* Copy `count` elements between host memory and array `which` (synchronous). */
int mael_dense_upload(mael_dense *d, int32_t which, const void *host, int64_t count);
int mael_dense_download(mael_dense *d, int32_t which, void *host, int64_t count);
/* `s.particles.n`. */
int mael_dense_count(mael_dense *d, int64_t *n);
int mael_dense_set_count(mael_dense *d, int64_t n);

/* `resample!(s; seeding)` (flip.jl:1364-1382). `seed` is `s.steps + 1`. On success
 * `*total` is the new particle count. Ranks are stable (particle-index order within a
 * cell), so the result is deterministic. MAEL_E_CAPACITY if the band needs more
 * particles than the capacity (then the particles are unchanged). */
int mael_dense_resample(mael_dense *d, int32_t seeding, uint32_t seed, int64_t *total);

/* mael_dense_resample for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). It has no
 * counterpart in the reference. The same result as mael_dense_resample, bit for bit; the
 * packed scan's tile sums live in a static device array of the library (128 KiB per device,
 * held until the process ends) instead of the basin's buffer, so the call allocates
 * nothing. Limit: grids of at most 16,384 x 1024 cells (16.7 M); larger ones get
 * MAEL_E_ARG. All such calls on one device share that array: calls in flight at once on
 * one device (basins on different streams) race. */
int mael_dense_resample_noalloc(mael_dense *d, int32_t seeding, uint32_t seed, int64_t *total);

/* `particles2grid!(s, dt)` (flip.jl:587-606): APIC deposit (deterministic), splat,
 * union with φ0 + dx, combine with u0 within (band - 1) dx of the surface. */
int mael_dense_p2g(mael_dense *d, float dt);

/* `grid2particles!(s, dt)` (flip.jl:1246-1252). */
int mael_dense_g2p(mael_dense *d, float dt, float gravity);

/* Adjoints of the hot path's linear maps, for reverse-mode gradients (a ChainRulesCore
 * `rrule` on the Julia side). They have no counterpart in the reference, which has no AD.
 * Each is the transpose of its phase's map at the state that phase left (positions, lags,
 * level set, types, flags and deposit weights held fixed); cotangents live in the
 * MAEL_UB_*, MAEL_UB0_* and MAEL_PB_* arrays.
 *   g2p_adjoint:     UB = G2P^T (PB); PB becomes the cotangent of G2P's input particles
 *                    (v passes through where the grid was not reached, rows never do).
 *   project_adjoint: UB <- UB + D^T A^-1 (-G^T UB) in place (the transposed MG-PCG solve;
 *                    A is symmetric), P is left holding the adjoint pressure.
 *   p2g_adjoint:     PB += P2G^T (UB) for the particles' v and rows (added to what
 *                    g2p_adjoint passed through), UB0 = the part for u0. */
int mael_dense_g2p_adjoint(mael_dense *d, float dt);
int mael_dense_p2g_adjoint(mael_dense *d, float dt);

typedef struct {
    double tol;          /* relative residual target (the reference's default 1e-4) */
    int32_t maxiter;     /* CG iterations in all (60) */
    int32_t sweeps;      /* RBGS sweeps per level (2) */
    int32_t coarse;      /* coarsest-level sweeps (40) */
    int32_t precision;   /* 0: FP32 V-cycle (the reference's) */
    int32_t refine;      /* 1: FP64 true-residual restarts until tol is met */
    int32_t warm;        /* 1: start from the last pressure (the reference's warm start); 0: from zero */
} mael_solve_opts;

typedef struct {
    int32_t iterations;        /* CG iterations */
    int32_t restarts;          /* FP64 residual restarts */
    double rel_residual;       /* ||b - A P|| / ||b||, FP64 */
    double rel_residual_rec;   /* the FP32 recursive residual at the last CG stop */
    double ms;                 /* wall time of the solve */
} mael_solve_stats;

/* Defaults of `pressure!(s::FlipBasin)` (flip.jl:1082): tol 1e-4, maxiter 60, sweeps 2, coarse 40,
 * warm start, no refinement (the reference's FP32 PCG). */
void mael_solve_defaults(mael_solve_opts *o);

/* `project!(s, dt)` (flip.jl:1144-1156) after the caller's γ (the volume control,
 * flip.jl:1146-1147): rhs, clear air, `pressure!`, project the faces and flag them. */
int mael_dense_project(mael_dense *d, float dt, float gamma, const mael_solve_opts *opts, mael_solve_stats *stats);
/* `vcycle!(s)` (flip.jl:986-1014) alone, as PCG applies it (flip.jl:1094, 1106): the operator
 * from the basin's types and level set, then P = M RHS (the preconditioner applied once to
 * the RHS array). `queued` nonzero runs it through the path the CG loop takes when the
 * V-cycle's first launches were queued ahead of a host read; the result is the same. For
 * tests of the preconditioner; `opts` NULL takes the defaults. */
int mael_dense_vcycle(mael_dense *d, const mael_solve_opts *opts, int32_t queued);
/* The transpose of `project!`'s linear part (see mael_dense_g2p_adjoint). */
int mael_dense_project_adjoint(mael_dense *d, const mael_solve_opts *opts, mael_solve_stats *stats);

/* Support phases around the hot path (not part of it): */
/* `classify!(s)` (flip.jl:674-677). */
int mael_dense_classify(mael_dense *d);
/* `redistance!(s; iterations)` (flip.jl:672-679): Jacobi redistancing sweeps of φ (φ0 is scratch). */
int mael_dense_redistance(mael_dense *d, int32_t iterations);
/* `volume(s)` (flip.jl:293), m^3. */
int mael_dense_volume(mael_dense *d, double *volume);
/* The balanced vortex of `initialize!` (flip.jl:325-330): φ from the surface table
 * `eta` (neta samples of `surface(b, r)` over r in [0, rmax]) and the Lamb-Oseen swirl. */
int mael_dense_init_vortex(mael_dense *d, const float *eta, int32_t neta, float rmax, float gamma_circ, float core);
/* `solidify!(s)` (flip.jl:800-807). */
int mael_dense_solidify(mael_dense *d);
/* The gravity term of `forces!` (flip.jl:765-775) followed by `solidify!`. */
int mael_dense_gravity(mael_dense *d, float dt, float gravity);
/* Copy array `src` into array `dst` of the same size and type on the device (e.g. the
 * `copyto!(s.φ0, s.φ)` of flip.jl:342). */
int mael_dense_copy(mael_dense *d, int32_t dst, int32_t src);
/* Wait for the basin's stream. */
int mael_dense_sync(mael_dense *d);

/* ── the sparse grid (`SparseGrid`, sflip.jl:31-53) ───────────────────────── */

typedef struct {
    float lo[3];      /* the grid's low corner, m */
    float dx;         /* fine cell width, m (coarse cells are 4 dx, blocks 16 dx) */
    int32_t nb[3];    /* blocks per axis (multiples of 4) */
    float radius;     /* the basin's wall radius, m */
    float floor;      /* the basin's floor height, m (= -depth) */
} mael_sparse_grid;

/* `SparseGrid(b; dx, top = b.rimwall, pad)` (sflip.jl:46-53): the sizing rule. */
int mael_sparsegrid(float radius, float depth, float top, float dx, int32_t pad, mael_sparse_grid *out);

/* ── a sparse basin (`SparseBasin`, sflip.jl:242-347) ─────────────────────── */

typedef struct mael_sparse mael_sparse;

/* `SparseBasin(b, backend; dx, band, nmin, capacity, blocks)` (sflip.jl:301-347): the fine
 * pools for `blocks` 16^3 blocks, the coarse grid and its multigrid halvings, two
 * particle buffers of `capacity`. MAEL_E_GRID if the coarse grid does not halve to at
 * most 1024 cells (sflip.jl:326). */
int mael_sparse_create(mael_ctx *ctx, const mael_sparse_grid *grid, int32_t blocks, int64_t capacity, int32_t band,
                       int32_t nmin, mael_sparse **out);
void mael_sparse_destroy(mael_sparse *s);

/* Arrays of a sparse basin. A field has a fine part (the pool: 4096 cells or 4352 faces
 * per slot, slot-major; see the layout notes in the library) and a coarse part (dense,
 * x fastest). Counts are elements; fine parts count the whole pool (`blocks` slots). */
enum {
    MAEL_S_PHI_F = 0, MAEL_S_PHI_C, MAEL_S_PHI0_F, MAEL_S_PHI0_C, MAEL_S_P_F, MAEL_S_P_C,
    MAEL_S_TYPES_F, MAEL_S_TYPES_C,                                   /* uint8 */
    MAEL_S_U_F, MAEL_S_U_C, MAEL_S_V_F, MAEL_S_V_C, MAEL_S_W_F, MAEL_S_W_C,
    MAEL_S_U0_F, MAEL_S_U0_C, MAEL_S_V0_F, MAEL_S_V0_C, MAEL_S_W0_F, MAEL_S_W0_C,
    MAEL_S_FU_F, MAEL_S_FU_C, MAEL_S_FV_F, MAEL_S_FV_C, MAEL_S_FW_F, MAEL_S_FW_C,   /* uint8 */
    MAEL_S_RHS_F, MAEL_S_RHS_C,
    MAEL_S_PX, MAEL_S_PY, MAEL_S_PZ, MAEL_S_VX, MAEL_S_VY, MAEL_S_VZ,
    MAEL_S_C0, MAEL_S_C1, MAEL_S_C2, MAEL_S_C3, MAEL_S_C4, MAEL_S_C5, MAEL_S_C6, MAEL_S_C7, MAEL_S_C8, MAEL_S_LAG,
    MAEL_S_KEY, MAEL_S_RANK,                                          /* last resample, per particle (int32) */
    MAEL_S_COUNTS, MAEL_S_TOTALS, MAEL_S_OFFSETS,                     /* last resample, per fine cell (int32) */
    MAEL_S_MOM_U, MAEL_S_MOM_V, MAEL_S_MOM_W, MAEL_S_WT_U, MAEL_S_WT_V, MAEL_S_WT_W,   /* int64 fixed point 2^-32 */
    MAEL_S_SLOT,                                                      /* per block: 0 or the 1-based slot (int32) */
    MAEL_S_WHERE,                                                     /* per slot: block x, y, z, 0 (int32 x4; -1 free) */
    MAEL_S_UB_U_F, MAEL_S_UB_U_C, MAEL_S_UB_V_F, MAEL_S_UB_V_C, MAEL_S_UB_W_F, MAEL_S_UB_W_C,   /* adjoints: face cotangents */
    MAEL_S_UB0_U_F, MAEL_S_UB0_U_C, MAEL_S_UB0_V_F, MAEL_S_UB0_V_C, MAEL_S_UB0_W_F, MAEL_S_UB0_W_C, /* ... of u0 */
    MAEL_S_PB_VX, MAEL_S_PB_VY, MAEL_S_PB_VZ,                         /* ... of the particles' v (capacity) */
    MAEL_S_PB_C0, MAEL_S_PB_C1, MAEL_S_PB_C2, MAEL_S_PB_C3, MAEL_S_PB_C4, MAEL_S_PB_C5, MAEL_S_PB_C6, MAEL_S_PB_C7,
    MAEL_S_PB_C8,                                                     /* ... and of their APIC rows */
    MAEL_S_NARRAYS
};
int mael_sparse_array(mael_sparse *s, int32_t which, void **dptr, int64_t *count, int32_t *elsize);
int mael_sparse_upload(mael_sparse *s, int32_t which, const void *host, int64_t count);
int mael_sparse_download(mael_sparse *s, int32_t which, void *host, int64_t count);
int mael_sparse_count(mael_sparse *s, int64_t *n);
int mael_sparse_set_count(mael_sparse *s, int64_t n);

/* The host side of the topology (sflip.jl:252-290 `slots`, `wheres`, `free`, `nslots`):
 * `slot` has nb0*nb1*nb2 entries, `where` 3 per pool slot, `freelist` up to `blocks`. */
int mael_sparse_get_topology(mael_sparse *s, int32_t *slot, int32_t *where, int32_t *nslots, int32_t *freelist,
                             int32_t *nfree);
int mael_sparse_set_topology(mael_sparse *s, const int32_t *slot, const int32_t *where, int32_t nslots,
                             const int32_t *freelist, int32_t nfree);

/* The marks of `refine!` (sflip.jl:518-526): the surface's blocks, those where the coarse
 * velocity varies by more than `vary`, those within `reach` of a particle (if
 * `particles`), grown by `rings`; written to `marks` (one byte a block). */
int mael_sparse_marks(mael_sparse *s, float vary, float reach, int32_t rings, int32_t particles, uint8_t *marks);
/* `refine!`'s allocation from given marks (sflip.jl:527-575): release unmarked blocks, then
 * give marked ones a slot (last freed first), zero the released slots and start new
 * blocks from the coarse grid. Bit-exact: the same marks give the same slots.
 * MAEL_E_BLOCKS with the reference's text if more blocks are wanted than the pool holds. */
int mael_sparse_refine_marks(mael_sparse *s, const uint8_t *marks, int32_t *added, int32_t *removed);
/* `refine!(s)` (sflip.jl:516-576): the marks, then the allocation. */
int mael_sparse_refine(mael_sparse *s, float vary, float reach, int32_t rings, int32_t *added, int32_t *removed);

/* `resample!(s; seeding)` (sflip.jl:2227-2246), new particles from (u, v, w). */
int mael_sparse_resample(mael_sparse *s, int32_t seeding, uint32_t seed, int64_t *total);
/* `particles2grid!(s, dt)` (sflip.jl:808-831). */
int mael_sparse_p2g(mael_sparse *s, float dt);
/* `grid2particles!(s, dt)` (sflip.jl:2141-2144). */
int mael_sparse_g2p(mael_sparse *s, float dt, float gravity);

typedef struct {
    double tol;          /* relative residual target */
    int32_t maxiter;     /* iterations in all */
    int32_t sweeps;      /* damped Jacobi sweep pairs per level */
    int32_t coarse;      /* coarsest-level RBGS sweeps */
    int32_t restart;     /* FGMRES restart length (at most 40) */
    int32_t policy;      /* 0 flexible CG (sflip.jl:1867), 1 FGMRES (fgmres.jl:26), 2 CG then FGMRES (fgmres.jl:96) */
    int32_t precision;   /* V-cycle storage: 0 FP32 (the reference's), 1 FP16, 2 BF16 (FP32 arithmetic, FP64 coarsest) */
    float boost;         /* coarse-correction scale */
    int32_t refine;      /* 0: the reference's FP32 solve; k > 0: iterative refinement, the solution
                          * kept in FP64 and its residual recomputed in FP64 every k iterations and
                          * before convergence is accepted (no counterpart in the reference) */
} mael_sparse_opts;
typedef struct {
    int32_t iterations, fcg_iterations, fgmres_iterations, used_fgmres;
    double rel_residual;   /* ||b - A P|| / ||b|| of the returned pressure (FP64 sums) */
    double ms;
} mael_sparse_stats;
/* The defaults of `pressure!(::SparseBasin{B,<:SurfaceStir})` (fgmres.jl:96): tol 5e-4,
 * maxiter 240, sweeps 1, coarse 40, restart 8, policy 2, boost 1.5; precision 0, refine 0. */
void mael_sparse_defaults(mael_sparse_opts *o);
/* `project!(s, dt)` (sflip.jl:1998-2015) after the caller's γ: couplings, rhs, clear,
 * the solve by `opts->policy`, the faces and flags. Errors carry the reference's texts. */
int mael_sparse_project(mael_sparse *s, float gamma, const mael_sparse_opts *opts, mael_sparse_stats *stats);

/* Adjoints of the sparse hot path's linear maps, as mael_dense_g2p_adjoint's (no
 * counterpart in the reference, which has no AD), at the state the forward phase left
 * (positions, lags, level set, topology, types, flags and deposit weights held fixed).
 * Cotangents live in MAEL_S_UB_*, MAEL_S_UB0_* (every stored face copy: a face on the
 * plane between two fine blocks has one in each, as the forward phases read and write
 * each) and MAEL_S_PB_*.
 *   g2p_adjoint:     UB = G2P^T (PB) (coarse parts 0: G2P reads fine faces; with
 *                    mael_sparse_set_flow(1) the faces are the reflected flow's); PB
 *                    becomes the cotangent of G2P's input particles.
 *   project_adjoint: UB <- R^T UB, then UB += D^T A^-1 (-G^T UB) (the transposed solve with
 *                    `opts`; A is symmetric); P is left holding the adjoint pressure.
 *   p2g_adjoint:     PB += P2G^T (UB) for the particles' v and rows, UB0 = the part for u0. */
int mael_sparse_g2p_adjoint(mael_sparse *s, float dt);
int mael_sparse_p2g_adjoint(mael_sparse *s, float dt);
int mael_sparse_project_adjoint(mael_sparse *s, const mael_sparse_opts *opts, mael_sparse_stats *stats);

/* The flexible CG's preconditioner alone (sflip.jl:1883, 1901 `vcycle!`, the V-cycle of
 * sflip.jl:1218-1311 with opts->sweeps, coarse, boost and precision): the couplings and
 * hierarchy from the basin's types and level set, then P = M RHS over the fine pool and the
 * coarse grid. `queued` as in mael_dense_vcycle. For tests of the preconditioner. */
int mael_sparse_vcycle(mael_sparse *s, const mael_sparse_opts *opts, int32_t queued);

/* Support phases around the hot path (not part of it): */
/* The start of `initialize!` (sflip.jl:681-694): φ from the surface table and the swirl,
 * on the coarse grid, then on the fine blocks. */
int mael_sparse_init_coarse(mael_sparse *s, const float *eta, int32_t neta, float rmax, float gamma_circ, float core);
int mael_sparse_init_fine(mael_sparse *s, const float *eta, int32_t neta, float rmax, float gamma_circ, float core);
int mael_sparse_classify(mael_sparse *s);                        /* sflip.jl:942-946 */
int mael_sparse_solidify(mael_sparse *s);                        /* sflip.jl:1034-1041 */
int mael_sparse_redistance(mael_sparse *s, int32_t iterations);  /* sflip.jl:910-919 */
int mael_sparse_restrict_velocity(mael_sparse *s);               /* sflip.jl:618 */
int mael_sparse_keep(mael_sparse *s);                            /* φ0 = φ, u0 = u (sflip.jl:700-703) */
int mael_sparse_fill_flags(mael_sparse *s, uint8_t value);       /* sflip.jl:696-698 */
int mael_sparse_volume(mael_sparse *s, double *volume);          /* sflip.jl:1990-1996 */
int mael_sparse_gravity(mael_sparse *s, float dt, float gravity); /* forces! without stirring, sflip.jl:1043-1056 */
int mael_sparse_sync(mael_sparse *s);
int mael_sparse_ktimer(mael_sparse *s, int32_t kernel, int32_t launches);
int mael_sparse_ktime(mael_sparse *s, double *ms, int32_t *launches);

/* The rest of a sparse step (sflip.jl:2295-2318), support phases around the hot path: */
/* `advect!(s, dt; carried)` (sflip.jl:773-788): the particles along (u, v, w) with ST-FLIP's
 * lag (seed `s.steps`), φ into φ0 and the carried velocity (u, or the reflected ur when
 * `carry_reflected`) into u0. */
int mael_sparse_advect(mael_sparse *s, float dt, int32_t carry_reflected, float gravity, int32_t seed);
/* `filltrapped!(s; maxrounds)` (sflip.jl:972-990). */
int mael_sparse_filltrapped(mael_sparse *s, int32_t maxrounds);
/* The stirring model of `forces!` (basin.jl: Basin, Race, Stir / SurfaceStir). */
typedef struct {
    float radius, band, relax, core, gamma_circ;   /* the basin's */
    float current[2];                              /* the race's */
    float eddy, eddyscale, eddytime, guide, free, turbulence, boil, reach;
    float origin[2], offset;                       /* the sea's coordinates and film time offset */
    int32_t surface;                               /* 0 Stir, 1 SurfaceStir */
    float hradius, ramp;                           /* SurfaceStir: the heights table's radius, the onset ramp */
} mael_stir;
/* `waves`: 7 floats a wave (direction x, y, wave number, angular frequency, amplitude
 * scaled by the race, steepness, phase), those the reference keeps (basin.jl:101);
 * `heights`: SurfaceStir's surface table (basin.jl:198). */
int mael_sparse_set_stir(mael_sparse *s, const mael_stir *stir, const float *waves, int32_t nwaves, const float *heights,
                         int32_t nheights);
/* `forces!(s, dt)` (sflip.jl:1043-1056) at time `time`. */
int mael_sparse_forces(mael_sparse *s, float dt, float time, float gravity);
/* Advection-reflection (sflip.jl:2118-2139): ur = u before the projection; the kick
 * ur = valid ? u - ur : 0 after it; ur += u after the extrapolation. */
int mael_sparse_save_reflect(mael_sparse *s);
int mael_sparse_kick(mael_sparse *s);
int mael_sparse_reflected(mael_sparse *s);
/* `extrapolate!(s; layers, coarselayers, kick)` (sflip.jl:2097-2116). */
int mael_sparse_extrapolate(mael_sparse *s, int32_t layers, int32_t coarselayers, int32_t kick);
/* Which flow G2P and resampling read: 0 (u, v, w), 1 the reflected (ur, vr, wr). */
int mael_sparse_set_flow(mael_sparse *s, int32_t reflected);
/* `maxspeed(s)` (sflip.jl:2264-2272). */
int mael_sparse_maxspeed(mael_sparse *s, float *speed);

/* The rest of a dense step (flip.jl:1396-1410), support phases around the hot path: */
/* `advect!(s, dt)` (flip.jl:490-499): the particles along (u, v, w) by ST-FLIP's own times
 * (lag, jitter from hash01 with seed `s.steps`; ballistic where the flow was not carried
 * out to), then u, v, w and φ semi-Lagrangian into u0, v0, w0 and φ0. */
int mael_dense_advect(mael_dense *d, float dt, float gravity, int32_t seed);
/* `filltrapped!(s; maxrounds)` (flip.jl:754-760): AIR not joined to the top layer's AIR
 * becomes LIQUID, at least 2 dx deep in φ. */
int mael_dense_filltrapped(mael_dense *d, int32_t maxrounds);
/* The stirring model `Stir` (basin.jl:92-181; `surface` must be 0: the dense basin takes
 * `Stir`) and its waves, as mael_sparse_set_stir. Without one, forces is gravity alone. */
int mael_dense_set_stir(mael_dense *d, const mael_stir *stir, const float *waves, int32_t nwaves);
/* `forces!(s, dt)` (flip.jl:786-798) at time `time`: gravity and the stir, then the walls. */
int mael_dense_forces(mael_dense *d, float dt, float time, float gravity);
/* `extrapolate!(s; layers)` (flip.jl:1194-1206): the valid faces' velocity `layers` faces
 * out, the flags of the faces reached, zero beyond, then the walls. */
int mael_dense_extrapolate(mael_dense *d, int32_t layers);
/* `maxspeed(s)` (flip.jl:1387): the largest |u|, |v|, |w| on the faces. */
int mael_dense_maxspeed(mael_dense *d, float *speed);

/* ── the bake writers' device side (flip.jl:1846-1951, sflip.jl:2419-2440) ─── */
/* `packband_kernel!` (flip.jl:1880-1893) for `record!(::BandParticles)`: each particle at
 * the frame's time (x + lag v) kept if above -keep or within `throat` of the axis; the
 * kept ones' positions and velocities as Float16 (IEEE binary16 bits, round to nearest
 * even), 3 a particle, into host `pos` and `vel` (room for `cap` >= n particles), in
 * particle order; `*count` how many. */
int mael_dense_pack_band(mael_dense *d, float keep, float throat, uint16_t *pos, uint16_t *vel, int64_t cap, int64_t *count);
int mael_sparse_pack_band(mael_sparse *s, float keep, float throat, uint16_t *pos, uint16_t *vel, int64_t cap,
                          int64_t *count);
/* `record!(::LevelSets, s)`: the dense level set averaged over cells twice as wide
 * (flip.jl:1859-1867, (n1/2) x (n2/2) x (n3/2) bytes), or the sparse one sampled at the
 * centres of cells 2 dx wide (sflip.jl:2427-2432, (8 nb0) x (8 nb1) x (8 nb2) bytes),
 * clamped to +-range and quantised to 0..255, into host `out` (x fastest). */
int mael_dense_quantize_levelset(mael_dense *d, float range, uint8_t *out);
int mael_sparse_quantize_levelset(mael_sparse *s, float range, uint8_t *out);

/* ── whitewater (flip.jl:1411-1511, sflip.jl:2320-2417, whitewater.jl) ───── */
/* `WhitewaterRules` (whitewater.jl:30-44) and the whitewater's floor, gravity and the
 * stream its waves ride (`carried`, whitewater.jl:306-310: stream 1 = the basin's swirl of
 * circulation gamma_circ and core `core`, 0 = none). */
typedef struct {
    float air[2], crest[2], energy[2];   /* the potentials' clamp ranges */
    float kair, kcrest;                  /* diffuse particles per second at full potential */
    int32_t spray, bubble;               /* neighbour counts (the SPH path's classification) */
    float life[2];                       /* lifetime range, s */
    float buoyancy, drag, kick, tossed, within;
    float floor, gravity;
    int32_t stream;
    float gamma_circ, core;
} mael_white_rules;
/* `Whitewater(b, backend; capacity, rules)` (whitewater.jl:255-300): a ring buffer of
 * `capacity` diffuse particles (x, y, z, vx, vy, vz, life, kind), all dead. */
typedef struct mael_white mael_white;
int mael_white_create(mael_ctx *ctx, int64_t capacity, const mael_white_rules *rules, mael_white **out);
void mael_white_destroy(mael_white *w);
enum { MAEL_W_X = 0, MAEL_W_Y, MAEL_W_Z, MAEL_W_VX, MAEL_W_VY, MAEL_W_VZ, MAEL_W_LIFE /* float */,
       MAEL_W_KIND /* uint8: 1 spray, 2 foam, 3 bubble */ };
int mael_white_array(mael_white *w, int32_t which, void **dptr, int64_t *count, int32_t *elsize);
int mael_white_upload(mael_white *w, int32_t which, const void *host, int64_t count);
int mael_white_download(mael_white *w, int32_t which, void *host, int64_t count);
/* The ring buffer's head (diffuse particles made so far) and the frames made. */
int mael_white_state(mael_white *w, int64_t *head, int64_t *frames);
int mael_white_set_state(mael_white *w, int64_t head, int64_t frames);
/* `whitewater!(w, s, u, v, dt; substeps, near, above, below)` (flip.jl:1496-1511,
 * sflip.jl:2401-2417): the frame's emission from the basin's particles through the cell
 * runs of its last resample (offsets - totals), normals from the level set within
 * `near` cells of the surface, the new particles in particle order from the head (the
 * reference takes slots by an atomic counter), then `substeps` steps of the diffuse
 * particles, spray above `above` cells, bubbles below `below`. */
int mael_dense_whitewater(mael_dense *d, mael_white *w, double dt, int32_t substeps, float near, float above, float below);
int mael_sparse_whitewater(mael_sparse *s, mael_white *w, double dt, int32_t substeps, float near, float above, float below);
/* `packwhite_kernel!` (whitewater.jl:366-379) for `record!(::WhitewaterFrames)`: the live
 * diffuse particles in slot order, (x, y, z, kind) and (vx, vy, vz) as Float16, into host
 * `out` (4 a particle) and `vel` (3), room for `cap` >= capacity; `*count` how many. */
int mael_white_pack(mael_white *w, uint16_t *out, uint16_t *vel, int64_t cap, int64_t *count);

/* ── checkpoint and restore (checkpoint.jl:7-66) ─────────────────────────── */
/* A versioned raw file of a basin's physical state (the reference writes Julia's
 * `Serialization`, which only the same Julia version reads): a header, the host's
 * metadata (JSON text, `meta`: the basin, grid, stirring model and clocks the reference
 * keeps on its host), and the arrays a step reads (φ, types, faces and flags, P, the live
 * particles; for the sparse basin the slot table, the live slots' blocks and the free
 * list, the fine pools cut to the live slots). Written to `path`.tmp and renamed over
 * `path`. `save_sparse_checkpoint(s, w, path)` (checkpoint.jl:7-35) is the sparse one. */
int mael_dense_checkpoint(mael_dense *d, const char *path, const char *meta);
int mael_sparse_checkpoint(mael_sparse *s, const char *path, const char *meta);
/* The kind (0 dense, 1 sparse) and metadata of a checkpoint, to build the basin it is
 * restored into; `buf` receives at most cap - 1 bytes and a NUL, `*len` the full length.
 * Needs no device. */
int mael_checkpoint_meta(const char *path, int32_t *kind, char *buf, int64_t cap, int64_t *len);
/* Restore into a basin of the checkpoint's grid (a sparse one may have a larger pool and
 * a dense or sparse one a larger particle capacity): the arrays, then φ0 = φ and u0 = u as
 * `load_sparse_checkpoint` (checkpoint.jl:37-66) does. */
int mael_dense_restore(mael_dense *d, const char *path);
int mael_sparse_restore(mael_sparse *s, const char *path);

/* ── launch timing (no counterpart in the reference, which has no profiler hooks) ── */
/* Kernel ids for the launch timer. */
enum {
    MAEL_K_P2G = 0, MAEL_K_SPLAT, MAEL_K_COMBINE, MAEL_K_G2P, MAEL_K_BIN, MAEL_K_FIXRANK, MAEL_K_SCATTER,
    MAEL_K_SPAWN, MAEL_K_SEARCH, MAEL_K_ADVANCE, MAEL_K_RBGS, MAEL_K_RESIDUAL, MAEL_K_RESTRICT, MAEL_K_PROLONG,
    MAEL_K_COARSEST, MAEL_K_RESID64, MAEL_K_COUNT
};
/* Record HIP events around every 8th of the next launches of `kernel` (fine-level
 * launches only for the multigrid kernels; up to `launches` timed ones), on the basin's
 * stream: an event pair leaves the stream idle for microseconds, so timing every launch
 * would slow the run it measures. `*_ktime` returns their total ms and count. */
int mael_dense_ktimer(mael_dense *d, int32_t kernel, int32_t launches);
/* Total milliseconds and number of the recorded launches; stops recording. */
int mael_dense_ktime(mael_dense *d, double *ms, int32_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* MAELSTROM_HIP_H */
