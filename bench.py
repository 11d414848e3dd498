#!/usr/bin/env python3
"""Benchmark of the hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sparse5cm]

A step is one pass of the hot path in the reference's order over the basin's live
particles. On the sparse basin (the default, 5 cm fine cells): refine the blocks,
particles to grid, redistance, classify, gravity, pressure projection (the composite
multigrid-preconditioned flexible CG, FGMRES where CG fails, to a 1e-6 relative
residual), grid to particles, resample. On the dense basin the same without
refinement (MG-PCG). Inputs are synthetic (the balanced Lamb-Oseen vortex of
`initialize!`) and resident on the GPU before the timed region. Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time


def log(msg):
    """A progress line on stderr (stdout carries only the result line)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "maelstrom.jl_amd"))

METRIC = "ms/step at 5cm sparse, MP-MGPCG vs FP32 @1e-6 rel. residual; % HBM peak"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (basin kwargs, dx, description)
    "sparse5cm": (dict(radius=18.0, depth=12.0, Γ=225.0, core=2.5), 0.05,
                  "sparse SparseBasin at 5 cm: 48x48x20 blocks of 16^3 fine cells over a 192x192x80 coarse grid "
                  "(20 cm), Lamb-Oseen vortex Γ 225 m^2/s core 2.5 m, 8 particles/cell in the 4-cell band, "
                  "composite MG flexible CG then FGMRES(20) to 1e-6, CFL 4"),
    "sparse60cm": (dict(radius=18.0, depth=12.0, Γ=225.0, core=2.5), 0.6,
                   "sparse SparseBasin at 60 cm: 4x4x4 blocks over a 16x16x16 coarse grid, the same vortex and solve"),
    "dense256": (dict(radius=18.0, depth=12.0, Γ=225.0, core=2.5), 0.15,
                 "dense FlipBasin 256x256x128, dx 0.15 m, Lamb-Oseen vortex Γ 225 m^2/s core 2.5 m, 8 particles/cell "
                 "in the 3-cell band, MG-PCG (6 levels, 2 RBGS sweeps) to 1e-6, CFL 4"),
    "dense64": (dict(radius=18.0, depth=12.0, Γ=225.0, core=2.5), 0.6,
                "dense FlipBasin 64x64x32, dx 0.6 m, the same vortex, MG-PCG (4 levels) to 1e-6"),
}


# the library's kernel timer brackets every 8th launch of the timed kernel with HIP events
KTIMER_STRIDE = 8


def sparse_bytes(kernel, live_cells, n_particles, precision=0, coarse_cells=0, covered=0):
    """Compulsory bytes of one launch of `kernel` on the sparse basin, at the element sizes
    the launch reads: FP32 (the reference's), or 2 B for the reduced-precision V-cycle's
    vectors and couplings. The level-0 sweep is one launch over the live fine cells and
    the coarse grid (`covered` of its cells lie under live blocks)."""
    e = 4 if precision == 0 else 2
    return {
        # per fine cell: x, b, the Dirichlet diagonal (FP32) or the Jacobi weight
        # omega/diag (FP16/BF16) and the result (e B each), the type (1 B); level 0's
        # couplings follow from the types (the row is matrix-free). Per coarse cell outside
        # the blocks: x, b, the Dirichlet diagonal and the result (e B each), the type, the
        # activity and the cover mask (1 B each); under a block: activity, mask, result.
        "jacobi": (4 * e + 1) * live_cells + (4 * e + 3) * (coarse_cells - covered) + (e + 2) * covered,
        # the flexible CG's search (FP32): z, d read, d, q written, the full diagonal read
        # (4 B each) and the type (1 B) per fine cell; per coarse cell outside the blocks the
        # same and the activity and cover mask; under a block activity, mask and q
        "search": 21 * live_cells + 23 * (coarse_cells - covered) + 6 * covered,
        # the flexible CG's update: P read and written, r read, the new r written (over
        # the older residual buffer), d, q read (4 B each) and the type (1 B) per unknown;
        # the cover mask too on the coarse grid
        "advance": 25 * live_cells + 26 * coarse_cells,
        # x (12 B) read, v and the APIC rows (48 B) written per particle; u, v, w and their
        # valid flags (15 B) read per fine cell's face triple
        "g2p": 60 * n_particles + 15 * live_cells,
        # the deposit: x, v, the rows and the lag (64 B) read per particle; the momentum
        # and weight sums (8 B each) added to each of a fine cell's three faces
        "p2g": 64 * n_particles + 48 * live_cells,
        # x (12 B) read, key and rank (8 B) written per particle
        "bin": 20 * n_particles,
        "scatter": 136 * n_particles,
    }[kernel]


def measured_traffic(key, units):
    """HBM bytes per launch from a counter measurement committed under profiles/: the line
    `<key>: <bytes per unit>` (FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for
    kernels of known traffic), times this launch's units; None when no such line exists."""
    for name in sorted(os.listdir(os.path.join(ROOT, "profiles"))):
        with open(os.path.join(ROOT, "profiles", name), encoding="utf-8") as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(round(float(line.split(":", 1)[1]) * units))
    return None


def algorithmic_bytes(kernel, s, n_particles):
    """Compulsory bytes of one launch of `kernel` at the reference's element sizes."""
    n1, n2, n3 = s.n_cells
    C = n1 * n2 * n3
    return {
        # one colour: x of every cell (neighbours), types of every cell, b/diag/x-write of half
        "rbgs": 11 * C,
        # z, d, diag, q (4 B each) and types (1 B) per cell
        "search": 17 * C,
        # z, q read; d, x, r read and written (4 B each) per cell
        "advance": 32 * C,
        # x (12 B) read, v (12 B) and c (36 B) written per particle
        "g2p": 60 * n_particles,
        # x, v, c, lag (64 B) read per particle
        "p2g": 64 * n_particles,
        # 64 B read and 64 B written per particle, key and rank (8 B) read
        "scatter": 136 * n_particles,
        # r, x (4 B) read, types (1 B), r (4 B) written per cell
        "residual": 13 * C,
    }[kernel]


def cpu_baseline(s, iters_per_step, dt):
    """The oracle's FP32 CPU port on the same state, on the host's cores: resample, P2G and
    G2P over all particles, and a few MG-PCG iterations, scaled to the GPU's iterations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    g = orc.Grid()
    for a in range(3):
        g.lo[a] = s.grid.lo[a]
        g.n[a] = s.grid.n[a]
    g.dx, g.radius, g.floor, g.levels = s.grid.dx, s.grid.radius, s.grid.floor, s.grid.levels
    n = s.n
    ps = {f: s.get(f) for f in orc.PFIELDS}
    phi, u, v, w = (s.get(k) for k in ("PHI", "U", "V", "W"))
    types = s.get("TYPES")
    flags = [s.get(k) for k in ("FU", "FV", "FW")]
    t = {}
    t0 = time.perf_counter()
    orc.dense_p2g(g, ps, n, dt, phi, u, v, w, fp32=True)
    t["p2g"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.dense_g2p(g, ps, n, u, v, w, *flags, dt)
    t["g2p"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.dense_resample(g, ps, n, s.capacity, phi, u, v, w, seed=s.steps + 1)
    t["resample"] = time.perf_counter() - t0
    b = orc.dense_rhs(g, u, v, w - np.float32(9.81 * dt), types, 0.0).astype(np.float32)
    t0 = time.perf_counter()
    orc.dense_pressure(g, types, phi, b, tol=1e-30, maxiter=1, fp32=True)
    t1 = time.perf_counter()
    orc.dense_pressure(g, types, phi, b, tol=1e-30, maxiter=4, fp32=True)
    t2 = time.perf_counter()
    per_iter = (t2 - t1 - (t1 - t0)) / 3
    setup = (t1 - t0) - per_iter
    ms = 1e3 * (t["p2g"] + t["g2p"] + t["resample"] + setup + per_iter * iters_per_step)
    cores = host_cores()
    sample = ("oracle FP32 port on %d host threads (OpenMP stencils and particle loops, sequential resampling): "
              "P2G, G2P and resample over all %d particles timed once; MG-PCG set-up plus 1 and 4 iterations timed, "
              "per-iteration cost scaled to the GPU's %.1f iterations/step (%.1f s of CPU work)"
              % (cores, n, iters_per_step, (t2 - t0) + t["p2g"] + t["g2p"] + t["resample"]))
    return {"value": round(ms, 1), "unit": "ms/step", "cores": cores, "kind": "port", "sample": sample}


def host_cores():
    """The threads the oracle's OpenMP loops run on: OMP_NUM_THREADS if set, else every core."""
    try:
        return int(os.environ.get("OMP_NUM_THREADS", "")) or os.cpu_count()
    except ValueError:
        return os.cpu_count()


def sparse_cpu_baseline(s, iters_per_step, dt, sample_particles=2_000_000):
    """The oracle's FP32 CPU port on the same sparse state, on the host's cores (its
    stencil, V-cycle and particle loops are OpenMP-parallel; the resampling is sequential):
    P2G, G2P and resample on a sample of the particles (scaled to all), the pressure set-up
    and flexible CG iterations on the whole grid (scaled to the GPU's iterations per step)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    from maelstrom_amd.sparse import BCELLS, BFACES
    slot, where, ns, free = s.topology()
    n = s.n
    m = min(n, sample_particles)
    o = orc.SparseOracle(s.basin.radius, s.basin.depth, s.basin.rimwall, s.grid.dx, 2, max(ns, 1), int(1.5 * n) + 4096,
                         s.band, s.nmin)
    # the topology and the fields from the GPU
    o.array("slot")[:] = slot
    o.array("where")[:3 * ns] = where[:ns].ravel()
    _set_nslots(o, ns, free)
    for gname, oname in [("PHI", "phi"), ("PHI0", "phi0"), ("TYPES", "types"), ("U", "u"), ("V", "v"), ("W", "w"),
                         ("U0", "u0"), ("V0", "v0"), ("W0", "w0"), ("FU", "fu"), ("FV", "fv"), ("FW", "fw")]:
        o.array(oname + "_f")[:] = s.get(gname + "_F")
        o.array(oname + "_c")[:] = s.get(gname + "_C").ravel()
    names = ["px", "py", "pz", "vx", "vy", "vz"] + ["c%d" % k for k in range(9)] + ["lag"]
    for nm in names:
        o.array(nm)[:m] = s.get(nm.upper(), count=m)
    o.n = m
    t = {}
    t0 = time.perf_counter()
    o.p2g(dt, fp32=True)
    t["p2g"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.g2p(dt)
    t["g2p"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.resample(seed=s.steps + 1)
    t["resample"] = time.perf_counter() - t0
    scale = n / max(m, 1)
    t0 = time.perf_counter()
    o.project(gamma=0.0, tol=1e-30, maxiter=1, policy=0, fp32=True)
    t1 = time.perf_counter()
    o.project(gamma=0.0, tol=1e-30, maxiter=5, policy=0, fp32=True)
    t2 = time.perf_counter()
    per_iter = (t2 - t1 - (t1 - t0)) / 4
    setup = max((t1 - t0) - per_iter, 0.0)
    halves = 2 if s.reflect else 1
    ms = 1e3 * (halves * ((t["p2g"] + t["g2p"] + t["resample"]) * scale + setup) + per_iter * iters_per_step)
    cores = host_cores()
    sample = ("oracle FP32 port on %d host threads (OpenMP stencils, V-cycle and particle loops; sequential "
              "resampling and dot products), the hot path only: P2G, G2P and resample on %d of the %d particles "
              "(scaled by %.1f), the pressure set-up and 1 and 5 flexible-CG iterations on the whole grid, the "
              "per-iteration cost scaled to the GPU's %.1f iterations a step, %d half-steps a step (%.1f s of CPU "
              "work)" % (cores, m, n, scale, iters_per_step, halves, (t2 - t0) + t["p2g"] + t["g2p"] + t["resample"]))
    return {"value": round(ms, 1), "unit": "ms/step", "cores": cores, "kind": "port", "sample": sample}


def sparse_mp_error(s, dt, tol, refine, chunk=8):
    """The error of the GPU solves (FP32, FP16 and BF16 V-cycles, all with the FP64
    refinement) against the FP64 oracle's solve of the same system to 1e-10: the state of
    a first half-step's projection, the same pressure to start from, gamma = 0. The oracle
    (OpenMP on the host's cores) solves in chunks of `chunk` flexible-CG iterations, each
    warm-started from the last chunk's pressure and the velocities before the projection,
    with a progress line after each, until 1e-10 or its floor (three chunks without a 1 %
    gain). Reports the relative L2 error of P over the unknowns
    and the relative divergence left after each projection (||b(u)|| / ||b(u_before)||
    over the unknowns, b the oracle's FP64 right-hand side, evaluated on the host)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    from maelstrom_amd import sparse as sp
    sp.before_projection(s, dt / 2, reflect=True)
    keep = {f: s.get(f) for f in ("U_F", "V_F", "W_F", "U_C", "V_C", "W_C", "P_F", "P_C")}
    slot, where, ns, free = s.topology()
    o = orc.SparseOracle(s.basin.radius, s.basin.depth, s.basin.rimwall, s.grid.dx, 2, max(ns, 1), 1, s.band, s.nmin)
    o.array("slot")[:] = slot
    o.array("where")[:3 * ns] = where[:ns].ravel()
    _set_nslots(o, ns, free)
    for gname, oname in [("PHI", "phi"), ("TYPES", "types"), ("U", "u"), ("V", "v"), ("W", "w"), ("P", "P")]:
        o.array(oname + "_f")[:] = s.get(gname + "_F")
        o.array(oname + "_c")[:] = s.get(gname + "_C").ravel()
    # the unknowns: LIQUID fine cells of live slots, LIQUID coarse cells outside the blocks
    nb0, nb1, nb2 = (int(x) for x in s.grid.nb)
    covered = (slot.reshape(nb2, nb1, nb0) != 0).repeat(4, 0).repeat(4, 1).repeat(4, 2).ravel()
    uf = (o.array("types_f")[:ns * 4096] == 1) & np.repeat(where[:ns, 0] >= 0, 4096)
    uc = (o.array("types_c") == 1) & ~covered

    def bnorm():
        bf, bc = o.rhs(0.0)
        return float(np.sqrt(np.sum(bf[uf] ** 2) + np.sum(bc[uc] ** 2)))

    def gpu_solve(p, tol_, refine_, maxiter):
        for f, v in keep.items():
            s.set(f, v)
        st = sp.project(s, dt / 2, gamma=0.0, tol=tol_, maxiter=maxiter, restart=20, policy=2, boost=1.5, precision=p,
                        refine=refine_)
        Pg = np.concatenate([s.get("P_F")[uf].astype(np.float64), s.get("P_C").ravel()[uc].astype(np.float64)])
        for gname, oname in [("U", "u"), ("V", "v"), ("W", "w")]:
            o.array(oname + "_f")[:] = s.get(gname + "_F")
            o.array(oname + "_c")[:] = s.get(gname + "_C").ravel()
        return st, Pg

    b0 = bnorm()
    t0 = time.perf_counter()
    before = {nm: (o.array(nm + "_f").copy(), o.array(nm + "_c").copy()) for nm in ("u", "v", "w")}
    iters, ref, best, flat = 0, None, float("inf"), 0
    while iters < 600:
        for nm, (fv, cv) in before.items():
            o.array(nm + "_f")[:] = fv
            o.array(nm + "_c")[:] = cv
        ref = o.project(gamma=0.0, tol=1e-10, maxiter=chunk, policy=0)
        iters += ref["iterations"]
        log("mp_error: oracle FP64 solve, %d iterations, relative residual %.3e (%.0f s)"
            % (iters, ref["rel_residual"], time.perf_counter() - t0))
        if ref["rel_residual"] <= 1e-10:
            break
        # stop at the FP64 solve's floor: three chunks without a 1 % gain
        flat = flat + 1 if ref["rel_residual"] > 0.99 * best else 0
        best = min(best, ref["rel_residual"])
        if flat >= 3:
            break
    P64 = np.concatenate([o.array("P_f")[:ns * 4096][uf].astype(np.float64), o.array("P_c")[uc].astype(np.float64)])
    rep = {"reference": "oracle FP64 composite MG flexible CG to 1e-10 (chunks of %d iterations, warm-started)" % chunk,
           "tight": {"iterations": iters, "rel_residual": ref["rel_residual"]}}
    rep["tight"].update(seconds=round(time.perf_counter() - t0, 1), div_rel=bnorm() / b0)
    for p, name in ((0, "fp32"), (1, "fp16"), (2, "bf16")):
        st, Pg = gpu_solve(p, tol, refine, 240)
        rep[name] = {"iterations": st.iterations, "rel_residual": st.rel_residual,
                     "P_rel_l2": float(np.linalg.norm(Pg - P64) / np.linalg.norm(P64)), "div_rel": bnorm() / b0}
        log("mp_error: %s done" % name)
    for f, v in keep.items():
        s.set(f, v)
    return rep


def _set_nslots(o, ns, free):
    """Give the oracle handle the GPU's slot count and free list (its where/slot arrays are
    already filled)."""
    import oracle as orc
    import numpy as np
    free = np.ascontiguousarray(free, np.int32)
    orc._slib().orc_sparse_set_topology_counts(o.h, int(ns), free.ctypes.data if free.size else None, free.size)


def sparse_main(args, desc, kw, dx):
    import ctypes as C
    import numpy as np
    import maelstrom_amd as m
    from maelstrom_amd import sparse as sp
    from maelstrom_amd._lib import KERNEL_ID
    b = m.Basin(**kw)
    # README.md:86-90: SurfaceStir(b, narutowaves(); race = Race(; eddy = 3)), cfl 4
    stir = m.Stir(b, waves=m.narutowaves(), race=m.Race(eddy=3.0), surface=True)
    # a 12000-block pool: the stirred surface's band outgrows the default 8000 within ~20 steps
    s = m.SparseBasin(b, dx=dx, stir=stir, band=4, nmin=8, cfl=4.0, capacity=30_000_000, blocks=12000)
    sp.initialize(s)
    log("initialized: %d particles" % s.n)
    umax = max(float(np.abs(s.get(f)).max()) for f in ("U_F", "V_F", "W_F", "U_C", "V_C", "W_C"))
    left = 1.0 / 24
    dt = left / math.ceil(left / (s.cfl * dx / max(umax, 1e-3)))
    def solve_for(precision):
        return dict(tol=args.tol, maxiter=240, restart=20, policy=2, boost=1.5, precision=precision, refine=args.refine)

    def timed(s, precision, steps, kernel_id):
        """`steps` whole steps between two synchronisations; the kernel timer on `kernel_id`.
        Each step ends in the resample's host read of its total, so the per-step wall times
        (for the median) need no extra synchronisation."""
        solve = solve_for(precision)
        s.sync()
        s.ctx.check(L.mael_sparse_ktimer(s.h, kernel_id, 64 * 1024))
        iters, fg, pms, nparts, each = [], [], [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            k0, j0 = len(s.iterations), len(s.solve_ms)
            ts = time.perf_counter()
            st = sp.step(s, dt, **solve)
            each.append(time.perf_counter() - ts)
            iters.append(sum(s.iterations[k0:]))   # both half-steps' solves
            fg.append(st.used_fgmres)
            pms.append(sum(s.solve_ms[j0:]))
            nparts.append(s.n)
        s.sync()
        t1 = time.perf_counter()
        kms, kl = C.c_double(), C.c_int32()
        s.ctx.check(L.mael_sparse_ktime(s.h, C.byref(kms), C.byref(kl)))
        return dict(ms=1e3 * (t1 - t0) / steps, median_ms=1e3 * float(np.median(each)), pressure_ms=sum(pms) / steps,
                    iterations=sum(iters) / steps, fgmres=int(sum(fg)), particles=sum(nparts) / steps, kms=kms.value,
                    launches=kl.value)

    L = m.lib()
    for w in range(args.warmup):
        sp.step(s, dt, **solve_for(args.precision))
        log("warm-up step %d done" % (w + 1))
    # the state the timed steps start from, for the FP32 comparison from the same state
    import shutil
    import tempfile
    shm = os.path.isdir("/dev/shm") and shutil.disk_usage("/dev/shm").free > 16 << 30
    tmpdir = tempfile.mkdtemp(dir="/dev/shm" if shm else None)
    start = os.path.join(tmpdir, "start.ckpt")
    s.t = round(s.t * 24) / 24   # a whole number of frames (the steps are frames' sub-steps)
    m.save_checkpoint(s, start)
    kernel = "jacobi" if args.kernel in ("rbgs", "jacobi") else args.kernel
    kid = KERNEL_ID["rbgs" if kernel == "jacobi" else kernel]
    R = timed(s, args.precision, args.steps, kid)
    log("timed %d steps: %.1f ms/step" % (args.steps, R["ms"]))
    slot, where, ns, free = s.topology()
    live_slots = int((where[:ns, 0] >= 0).sum())
    ms_step = R["ms"]
    avg_launch_ms = R["kms"] / max(R["launches"], 1)
    n_avg = R["particles"]
    # a 16^3 block covers 4^3 coarse cells
    nc = s.nc
    abytes = sparse_bytes(kernel, live_slots * 4096, n_avg, args.precision, nc[0] * nc[1] * nc[2], live_slots * 64)
    achieved = abytes / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    traffic = (measured_traffic("traffic_per_algorithmic_byte_fp16_jacobi", abytes)
               if kernel == "jacobi" and args.precision == 1 else None)
    out = {
        "metric": METRIC, "value": round(ms_step, 3), "unit": "ms/step", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "blocks": list(int(x) for x in s.grid.nb), "coarse_cells": list(s.nc),
                   "live_blocks": live_slots, "slots": int(ns), "fine_cells": live_slots * 4096, "particles": int(n_avg),
                   "dt_s": round(dt, 5), "tol": args.tol, "policy": "fcg-then-fgmres(20)",
                   "precision": ["fp32", "fp16", "bf16"][args.precision], "refine": args.refine,
                   "stir": "surface stir, 32 seeded narutowaves-shaped waves, race eddy 3",
                   "phases": "two advection-reflection halves of refine, advect, p2g, redistance, classify, "
                             "trapped air, stir forces, project, kick, extrapolate, reflect, g2p, resample"},
        "pressure_ms": round(R["pressure_ms"], 3),
        "iterations": round(R["iterations"], 2),
        "fgmres_steps": R["fgmres"],
        "mparticles_per_s": round(n_avg / (ms_step * 1e-3) / 1e6, 2),
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "launches": R["launches"], "timed_every": KTIMER_STRIDE, "avg_launch_us": round(avg_launch_ms * 1e3, 3),
                     "algorithmic_bytes_per_launch": int(abytes)},
    }
    out["ms_per_step_median"] = round(R["median_ms"], 3)
    # the other hot kernels, each timed over one more step (their launches' mean duration)
    census = {}
    for name in ("search", "advance", "p2g", "g2p", "bin"):
        Q = timed(s, args.precision, 1, KERNEL_ID[name])
        if Q["launches"]:
            us = 1e3 * Q["kms"] / Q["launches"]
            ab = sparse_bytes(name, live_slots * 4096, s.n, args.precision, nc[0] * nc[1] * nc[2], live_slots * 64)
            gbs = ab / (us * 1e-6) / 1e9
            census[name] = {"avg_launch_us": round(us, 2), "algorithmic_bytes": int(ab), "achieved": round(gbs, 1),
                            "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["kernels"] = census
    log("kernel census done")
    # the blocks the marks ask for at the state reached, by kind (sflip.jl:518-526): the
    # surface's, then with the flow's variation, then within reach of a particle
    reach = float(np.float32((math.ceil(s.cfl) + 3) * np.float32(s.grid.dx)))
    out["config"]["marked_blocks"] = {
        "surface": int(sp.marks(s, vary=1e30, particles=False, rings=0).sum()),
        "surface_flow": int(sp.marks(s, particles=False, rings=0).sum()),
        "surface_flow_particles": int(sp.marks(s, rings=0).sum()),
        "reach_m": round(reach, 3)}
    if args.precision != 0 and not args.no_fp32:
        # the same steps with the FP32 V-cycle (the same FP64 refinement) from the same state
        # as the mixed-precision ones: the comparison the metric names
        del s
        f = m.load_checkpoint(start, blocks=12000)
        F = timed(f, 0, args.steps, kid)
        log("FP32 V-cycle steps: %.1f ms/step" % F["ms"])
        out["fp32"] = {"ms_per_step": round(F["ms"], 3), "ms_per_step_median": round(F["median_ms"], 3),
                       "pressure_ms": round(F["pressure_ms"], 3), "iterations": round(F["iterations"], 2),
                       "jacobi_avg_launch_us": round(1e3 * F["kms"] / max(F["launches"], 1), 3),
                       "from": "the same state as the mixed-precision steps (a checkpoint after the warm-up)"}
        out["speedup_vs_fp32"] = {"step": round(F["ms"] / ms_step, 3),
                                  "pressure": round(F["pressure_ms"] / max(R["pressure_ms"], 1e-9), 3)}
        # the reference's plain FP32 solve (no refinement) and the mixed-precision one on the
        # same system: the first half-step's projection from the saved state
        note = ("the reference's FP32 flexible CG then FGMRES(20), no FP64 refinement, at most 240 iterations, "
                "on the first half-step's system from the saved state")
        for label, prec, ref in (("fp32_plain", 0, 0), ("mp_same_system", args.precision, args.refine)):
            h = m.load_checkpoint(start, blocks=12000)
            if prec:
                # the reduced-precision hierarchy's storage is allocated by a basin's first use:
                # one V-cycle does it, then the saved state again (the V-cycle overwrote P)
                o = sp.solve_options(**dict(sp.pressure_defaults(h), precision=prec))
                h.ctx.check(L.mael_sparse_vcycle(h.h, C.byref(o), 0))
                h.ctx.check(L.mael_sparse_restore(h.h, start.encode()))
            sp.before_projection(h, dt / 2, reflect=True)
            h.sync()
            t0 = time.perf_counter()
            try:
                st = sp.project(h, dt / 2, tol=args.tol, maxiter=240, restart=20, policy=2, boost=1.5, precision=prec,
                                refine=ref)
                h.sync()
                out[label] = {"converged": True, "iterations": st.iterations, "fgmres": bool(st.used_fgmres),
                              "rel_residual": float("%.3e" % st.rel_residual),
                              "ms": round(1e3 * (time.perf_counter() - t0), 2)}
            except m.MaelstromError as e:
                # the reference raises here too (fgmres.jl:90): FP32 alone does not reach the tolerance
                h.sync()
                out[label] = {"converged": False, "error": str(e), "ms": round(1e3 * (time.perf_counter() - t0), 2)}
            del h
        out["fp32_plain"]["note"] = note
        log("plain FP32 solve: %s" % out["fp32_plain"])
        s = f
    os.remove(start)
    os.rmdir(tmpdir)
    if args.mp_error:
        out["mp_error"] = sparse_mp_error(s, dt, args.tol, args.refine)
    if not args.no_cpu_baseline:
        log("CPU baseline")
        out["cpu_baseline"] = sparse_cpu_baseline(s, out["iterations"], dt)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sparse5cm", choices=sorted(WORKLOADS))
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--precision", type=int, default=None,
                    help="V-cycle storage: 0 FP32, 1 FP16, 2 BF16 (default: 1 on the sparse workloads, 0 on the dense)")
    ap.add_argument("--refine", type=int, default=10,
                    help="sparse solve: FP64 residual recomputation every k iterations (0: the plain FP32 solve)")
    ap.add_argument("--kernel", default="rbgs", help="the kernel the roofline object reports")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="sparse: skip the FP32 V-cycle comparison steps")
    ap.add_argument("--mp-error", action="store_true",
                    help="sparse: report the solves' error against an FP64 oracle solve of the same system (minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build's hot path runs on a single MI355X (--gpus %d not supported)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    kw, dx, desc = WORKLOADS[args.workload]
    if args.precision is None:
        args.precision = 1 if args.workload.startswith("sparse") else 0
    if args.workload.startswith("sparse"):
        return sparse_main(args, desc, kw, dx)
    import maelstrom_amd as m
    from maelstrom_amd._lib import KERNEL_ID
    b = m.Basin(**kw)
    s = m.FlipBasin(b, dx=dx, band=4, nmin=8, cfl=4.0, capacity=30_000_000)
    m.initialize(s)
    # run.jl:44: dt = left / ceil(left / (cfl dx / max(maxspeed, 1e-3))) within a 1/24 s frame
    left = 1.0 / 24
    dt = left / math.ceil(left / (s.cfl * dx / max(m.maxspeed(s), 1e-3)))
    # the reference's warm start; FP64 refinement to reach 1e-6 (its FP32 PCG stalls near 1e-5)
    solve = dict(tol=args.tol, maxiter=400, precision=args.precision, warm=1, refine=1)
    for _ in range(args.warmup):
        m.step(s, dt, **solve)
    s.sync()
    L = m.lib()
    s.ctx.check(L.mael_dense_ktimer(s.h, KERNEL_ID[args.kernel], 64 * 1024))
    iters, pressure_ms, nparts = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = m.step(s, dt, **solve)
        iters.append(st.iterations)
        pressure_ms.append(st.ms)
        nparts.append(s.n)
    s.sync()
    t1 = time.perf_counter()
    import ctypes as C
    kms, kl = C.c_double(), C.c_int32()
    s.ctx.check(L.mael_dense_ktime(s.h, C.byref(kms), C.byref(kl)))
    ms_step = 1e3 * (t1 - t0) / args.steps
    avg_launch_ms = kms.value / max(kl.value, 1)
    n_avg = sum(nparts) / len(nparts)
    abytes = algorithmic_bytes(args.kernel, s, n_avg)
    achieved = abytes / (avg_launch_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(ms_step, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "cells": list(s.n_cells), "particles": int(n_avg), "dt_s": round(dt, 5),
                   "tol": args.tol, "precision": ["fp32", "fp16", "bf16"][args.precision], "solve_start": "warm",
                   "phases": "advect, p2g, redistance, classify, trapped air, stir forces, project, extrapolate, "
                             "g2p, resample"},
        "pressure_ms": round(sum(pressure_ms) / len(pressure_ms), 3),
        "cg_iterations": round(sum(iters) / len(iters), 2),
        "mparticles_per_s": round(n_avg / (ms_step * 1e-3) / 1e6, 2),
        "roofline": {"bound": "hbm", "kernel": args.kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "launches": kl.value, "timed_every": KTIMER_STRIDE, "avg_launch_us": round(avg_launch_ms * 1e3, 3),
                     "algorithmic_bytes_per_launch": int(abytes)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(s, out["cg_iterations"], dt)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
