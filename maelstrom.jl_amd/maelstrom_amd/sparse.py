"""`SparseBasin`: the two-resolution narrow-band FLIP basin on the HIP library
(sflip.jl:242-2318, fgmres.jl).

The Python mirror of the reference's Julia API for the hot path: `SparseBasin`,
`refine`, `resample`, `particles2grid`, `project` (with `pressure!`'s policies),
`grid2particles`, `initialize` and `step` (Julia's bang names without the `!`). Every
phase runs in the gfx950 library; this module keeps only what the reference keeps on
the host: time, step count, the reference volume, and the block table's bookkeeping
mirrored from the library.
"""
import ctypes as C
import math
import warnings

import numpy as np

from . import _lib
from ._lib import SPARSE_ID, PARTICLE_FIELDS, MaelstromError, SparseGrid, SparseOpts, SparseStats, StirC
from .basin import Stir, surface

BCELLS, BFACES = 4096, 4352
_INT32 = {"KEY", "RANK", "COUNTS", "TOTALS", "OFFSETS", "SLOT", "WHERE"}
_UINT8 = {"TYPES_F", "TYPES_C", "FU_F", "FU_C", "FV_F", "FV_C", "FW_F", "FW_C"}
_INT64 = {"MOM_U", "MOM_V", "MOM_W", "WT_U", "WT_V", "WT_W"}


def sparsegrid(basin, dx, top=None, pad=2):
    """sflip.jl:46-53 `SparseGrid(b; dx, top = b.rimwall, pad = 2)`."""
    g = SparseGrid()
    top = basin.rimwall if top is None else top
    st = _lib.lib().mael_sparsegrid(basin.radius, basin.depth, top, dx, pad, C.byref(g))
    if st != _lib.OK:
        raise MaelstromError(st, "bad sparse grid for dx=%r" % dx)
    return g


class SparseBasin:
    """sflip.jl:301 `SparseBasin(b, backend; dx, stir, band = 4, nmin = 8, cfl = 8, capacity = 30_000_000,
    blocks = 8000, rings = 0, vary = 0.5, reflect = true)`."""

    def __init__(self, basin, dx, stir=None, band=4, nmin=8, cfl=8.0, capacity=30_000_000, blocks=8000, rings=0,
                 vary=0.5, reflect=True, ctx=None, top=None, pad=2):
        self.basin = basin
        self.ctx = ctx or _lib.default_context()
        self.grid = sparsegrid(basin, dx, top, pad)
        self.top, self.pad = top, pad
        self.band, self.nmin, self.cfl, self.capacity = int(band), int(nmin), float(cfl), int(capacity)
        self.blocks, self.rings, self.vary, self.reflect = int(blocks), int(rings), float(vary), bool(reflect)
        self.h = C.c_void_p()
        self.ctx.check(_lib.lib().mael_sparse_create(self.ctx.h, C.byref(self.grid), self.blocks, self.capacity,
                                                     self.band, self.nmin, C.byref(self.h)))
        self.volume0 = 0.0
        self.t = 0.0
        self.steps = 0
        self.iterations = []   # each solve's iterations, as the reference's s.iterations
        self.solve_ms = []     # each solve's wall time (ms), set-up and projection of the faces included
        self.last_solve = None
        set_stir(self, Stir(basin) if stir is None else stir)

    # ── arrays ─────────────────────────────────────────────────────────────
    @property
    def nc(self):
        return self.grid.nc

    @property
    def nblocks(self):
        return int(np.prod([int(x) for x in self.grid.nb]))

    def shape(self, name):
        """numpy shape of an array: coarse parts (z, y, x); fine parts and particles flat."""
        n1, n2, n3 = self.nc
        coarse = {"PHI_C": (n3, n2, n1), "PHI0_C": (n3, n2, n1), "P_C": (n3, n2, n1), "TYPES_C": (n3, n2, n1),
                  "RHS_C": (n3, n2, n1), "U_C": (n3, n2, n1 + 1), "U0_C": (n3, n2, n1 + 1), "FU_C": (n3, n2, n1 + 1),
                  "V_C": (n3, n2 + 1, n1), "V0_C": (n3, n2 + 1, n1), "FV_C": (n3, n2 + 1, n1),
                  "W_C": (n3 + 1, n2, n1), "W0_C": (n3 + 1, n2, n1), "FW_C": (n3 + 1, n2, n1)}
        for ax, sh in (("U", (n3, n2, n1 + 1)), ("V", (n3, n2 + 1, n1)), ("W", (n3 + 1, n2, n1))):
            coarse["UB_%s_C" % ax] = coarse["UB0_%s_C" % ax] = sh
        if name in coarse:
            return coarse[name]
        return (self._info(name)[1],)

    def _info(self, name):
        ptr, cnt, es = C.c_void_p(), C.c_int64(), C.c_int32()
        self.ctx.check(_lib.lib().mael_sparse_array(self.h, SPARSE_ID[name], C.byref(ptr), C.byref(cnt), C.byref(es)))
        dt = np.int32 if name in _INT32 else (np.uint8 if name in _UINT8 else (np.int64 if name in _INT64 else np.float32))
        return ptr.value, cnt.value, dt

    def get(self, name, count=None):
        """Download array `name`; particle arrays are cut to `count` (default n), fine
        parts to the slots in use unless `count` is given."""
        _, cnt, dt = self._info(name)
        if name in PARTICLE_FIELDS or name in ("KEY", "RANK") or name.startswith("PB_"):
            cnt = self.n if count is None else count
        elif count is not None:
            cnt = count
        elif name.endswith("_F") or name in ("COUNTS", "TOTALS", "OFFSETS"):
            cnt = self.nslots * BCELLS if name in ("PHI_F", "PHI0_F", "P_F", "TYPES_F", "RHS_F", "COUNTS", "TOTALS",
                                                    "OFFSETS") else self.nslots * BFACES
        elif name in _INT64:
            cnt = self.nslots * BFACES
        out = np.empty(cnt, dtype=dt)
        if out.size:
            self.ctx.check(_lib.lib().mael_sparse_download(self.h, SPARSE_ID[name], out.ctypes.data, out.size))
        if name.endswith("_C"):
            out = out.reshape(self.shape(name))
        return out

    def set(self, name, arr):
        _, cnt, dt = self._info(name)
        a = np.ascontiguousarray(arr, dtype=dt).ravel()
        if a.size > cnt:
            raise ValueError("%s holds %d elements, got %d" % (name, cnt, a.size))
        if a.size:
            self.ctx.check(_lib.lib().mael_sparse_upload(self.h, SPARSE_ID[name], a.ctypes.data, a.size))

    @property
    def n(self):
        v = C.c_int64()
        self.ctx.check(_lib.lib().mael_sparse_count(self.h, C.byref(v)))
        return v.value

    @n.setter
    def n(self, value):
        self.ctx.check(_lib.lib().mael_sparse_set_count(self.h, int(value)))

    def particles(self):
        return {f: self.get(f) for f in PARTICLE_FIELDS}

    def set_particles(self, ps, n=None):
        n = len(ps["PX"]) if n is None else n
        for f in PARTICLE_FIELDS:
            self.set(f, np.asarray(ps[f][:n], dtype=np.float32))
        self.n = n

    # ── topology ───────────────────────────────────────────────────────────
    def topology(self):
        """(slot per block, where per slot (blocks x 3), nslots, free list)."""
        slot = np.zeros(self.nblocks, np.int32)
        where = np.zeros(3 * self.blocks, np.int32)
        free = np.zeros(self.blocks, np.int32)
        ns, nf = C.c_int32(), C.c_int32()
        self.ctx.check(_lib.lib().mael_sparse_get_topology(self.h, slot.ctypes.data, where.ctypes.data, C.byref(ns),
                                                           free.ctypes.data, C.byref(nf)))
        return slot, where.reshape(-1, 3), ns.value, free[:nf.value].copy()

    @property
    def nslots(self):
        ns = C.c_int32()
        self.ctx.check(_lib.lib().mael_sparse_get_topology(self.h, None, None, C.byref(ns), None, None))
        return ns.value

    def set_topology(self, slot, where, nslots, free=()):
        slot = np.ascontiguousarray(slot, np.int32).ravel()
        where = np.ascontiguousarray(where, np.int32).ravel()
        free = np.ascontiguousarray(free, np.int32).ravel()
        self.ctx.check(_lib.lib().mael_sparse_set_topology(self.h, slot.ctypes.data, where.ctypes.data, int(nslots),
                                                           free.ctypes.data if free.size else None, free.size))

    def sync(self):
        self.ctx.check(_lib.lib().mael_sparse_sync(self.h))

    def close(self):
        if self.h:
            _lib.lib().mael_sparse_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ── phases ───────────────────────────────────────────────────────────────────

def reach(s):
    """sflip.jl:522: the particles' reach, (ceil(cfl) + 3) dx."""
    return float(np.float32((math.ceil(s.cfl) + 3) * s.grid.dx))


def marks(s, particles=True, vary=None, rings=None):
    """The marks of `refine!` (sflip.jl:518-526), one byte a block (`vary` and `rings`
    default to the basin's)."""
    out = np.zeros(s.nblocks, np.uint8)
    vary = s.vary if vary is None else float(vary)
    rings = s.rings if rings is None else int(rings)
    s.ctx.check(_lib.lib().mael_sparse_marks(s.h, vary, reach(s), rings, 1 if particles else 0, out.ctypes.data))
    return out


def refine(s, want=None):
    """sflip.jl:516-576 `refine!(s)` -> (added, removed); with `want`, from those marks."""
    na, nr = C.c_int32(), C.c_int32()
    L = _lib.lib()
    if want is None:
        s.ctx.check(L.mael_sparse_refine(s.h, s.vary, reach(s), s.rings, C.byref(na), C.byref(nr)))
    else:
        w = np.ascontiguousarray(want, np.uint8)
        s.ctx.check(L.mael_sparse_refine_marks(s.h, w.ctypes.data, C.byref(na), C.byref(nr)))
    return na.value, nr.value


def resample(s, seeding=False):
    """sflip.jl:2227-2246 `resample!(s; seeding)`: returns the new particle count."""
    total = C.c_int64()
    s.ctx.check(_lib.lib().mael_sparse_resample(s.h, 1 if seeding else 0, (s.steps + 1) & 0xffffffff, C.byref(total)))
    return total.value


def particles2grid(s, dt):
    """sflip.jl:808-831 `particles2grid!(s, dt)`."""
    s.ctx.check(_lib.lib().mael_sparse_p2g(s.h, float(dt)))
    return s


def grid2particles(s, dt):
    """sflip.jl:2141-2144 `grid2particles!(s, dt)`."""
    s.ctx.check(_lib.lib().mael_sparse_g2p(s.h, float(dt), float(s.basin.g)))
    return s


def grid2particles_adjoint(s, dt):
    """The transpose of `grid2particles!` at the state it left (no counterpart in the
    reference): UB_*_F = G2P^T (PB_*), the coarse parts 0; PB_* becomes the cotangent of
    the input particles (v passes through where the grid was not reached)."""
    s.ctx.check(_lib.lib().mael_sparse_g2p_adjoint(s.h, float(dt)))
    return s


def particles2grid_adjoint(s, dt):
    """The transpose of `particles2grid!` at the state it left: PB_* += P2G^T (UB_*) for
    the particles' v and APIC rows, UB0_* = the part for u0 (fine pools and coarse grid)."""
    s.ctx.check(_lib.lib().mael_sparse_p2g_adjoint(s.h, float(dt)))
    return s


def project_adjoint(s, **solve):
    """The transpose of `project!`'s linear part at the state it left, in place on UB_*
    (the restriction's transpose, then UB += D^T A^-1 (-G^T UB)); `solve` as `project`'s.
    P is left holding the adjoint pressure."""
    o = pressure_defaults(s)
    o.update(solve)
    opts = solve_options(**o)
    st = SparseStats()
    s.ctx.check(_lib.lib().mael_sparse_project_adjoint(s.h, C.byref(opts), C.byref(st)))
    return st


def solve_options(tol=5e-4, maxiter=240, sweeps=1, coarse=40, restart=8, policy=2, precision=0, boost=1.5,
                  refine=None):
    """`pressure!(::SparseBasin{B,<:SurfaceStir})`'s defaults (fgmres.jl:96); policy 0 is the
    plain flexible CG of sflip.jl:1867 (its defaults: tol 1e-4, maxiter 200, boost 1),
    1 FGMRES alone (fgmres.jl:26). `precision` 1 / 2 keeps the V-cycle in FP16 / BF16;
    `refine` k > 0 keeps the solution in FP64 with its residual recomputed in FP64 every
    k iterations (default: 10 with a reduced-precision V-cycle, else 0, the reference's
    FP32 solve)."""
    if refine is None:
        refine = 10 if precision else 0
    return SparseOpts(float(tol), int(maxiter), int(sweeps), int(coarse), int(restart), int(policy), int(precision),
                      float(boost), int(refine))


def volume(s):
    """sflip.jl:1990-1996 `volume(s)`, m^3."""
    v = C.c_double()
    s.ctx.check(_lib.lib().mael_sparse_volume(s.h, C.byref(v)))
    return v.value


def gamma_for(s, dt, recover=0.5):
    """sflip.jl:2000-2001: the volume control's divergence γ."""
    loss = (s.volume0 - volume(s)) / s.volume0 if s.volume0 else 0.0
    return float(np.float32(min(max(loss / recover, -0.01 / dt), 0.01 / dt)))


def pressure_defaults(s):
    """The solve `pressure!(s)` dispatches to: with a surface stir, fgmres.jl:96 (tol 5e-4,
    maxiter 240, boost 1.5, flexible CG for at most 80 iterations, then FGMRES); else
    sflip.jl:1867 (tol 1e-4, maxiter 200, boost 1, flexible CG that warns when it stops
    short)."""
    if s.stir is not None and s.stir.surface:
        return dict(tol=5e-4, maxiter=240, boost=1.5, policy=2)
    return dict(tol=1e-4, maxiter=200, boost=1.0, policy=0)


def project(s, dt, recover=0.5, gamma=None, **solve):
    """sflip.jl:1998-2015 `project!(s, dt; recover)`; `solve` overrides the options of
    `pressure_defaults(s)`."""
    g = gamma_for(s, dt, recover) if gamma is None else float(gamma)
    o = pressure_defaults(s)
    o.update(solve)
    opts = solve_options(**o)
    st = SparseStats()
    s.ctx.check(_lib.lib().mael_sparse_project(s.h, g, C.byref(opts), C.byref(st)))
    if opts.policy == 0 and st.rel_residual > opts.tol:
        # sflip.jl:1902: the plain flexible CG warns and keeps its last iterate
        warnings.warn("pressure solve did not converge: residual %.3g after %d iterations"
                      % (st.rel_residual, st.iterations), RuntimeWarning, stacklevel=2)
    s.iterations.append(st.iterations)
    s.solve_ms.append(st.ms)
    s.last_solve = st
    return st


def classify(s):
    s.ctx.check(_lib.lib().mael_sparse_classify(s.h))
    return s


def redistance(s, iterations=None):
    """sflip.jl:910-919 `redistance!(s; iterations = band + 3)`."""
    it = s.band + 3 if iterations is None else int(iterations)
    s.ctx.check(_lib.lib().mael_sparse_redistance(s.h, it))
    return s


def keep(s):
    """φ0 = φ, u0 = u (sflip.jl:700-703)."""
    s.ctx.check(_lib.lib().mael_sparse_keep(s.h))


def surface_table(s):
    """The balanced surface at 2048 radii over [0, hypot(lo)] (sflip.jl:684-685)."""
    g = s.grid
    rmax = np.float32(math.hypot(g.lo[0], g.lo[1]))
    r = np.linspace(np.float32(0), rmax, 2048, dtype=np.float32)
    return np.ascontiguousarray(surface(s.basin, r), dtype=np.float32), rmax


def initialize(s):
    """sflip.jl:681-713 `initialize!(s)`: the balanced surface and swirl on the coarse grid,
    fine blocks round the surface with both on them too, the band seeded, their deposit."""
    L = _lib.lib()
    eta, rmax = surface_table(s)
    args = (eta.ctypes.data, eta.size, float(rmax), float(s.basin.Γ), float(s.basin.core))
    s.ctx.check(L.mael_sparse_init_coarse(s.h, *args))
    refine(s)
    s.ctx.check(L.mael_sparse_init_fine(s.h, *args))
    redistance(s, 2 * s.band + 4)
    classify(s)
    s.ctx.check(L.mael_sparse_solidify(s.h))
    s.ctx.check(L.mael_sparse_fill_flags(s.h, 1))
    s.n = 0
    resample(s, seeding=True)
    resample(s)
    keep(s)
    particles2grid(s, 1.0)
    redistance(s)
    classify(s)
    s.volume0 = volume(s)
    s.t = 0.0
    return s


def stir_struct(stir):
    """The C ABI's form of a stirring model: the parameters (basin.jl:13-30 Basin, :69-80
    Race, :92-103 `Stir`, 191-205 `SurfaceStir`), the waves at least `shortest` long scaled
    by the race (basin.jl:101), 7 floats each, and SurfaceStir's surface table."""
    b, race = stir.basin, stir.race
    c = StirC()
    c.radius, c.band, c.relax, c.core, c.gamma_circ = b.radius, b.band, b.relax, b.core, b.Γ
    c.current[0], c.current[1] = race.current
    c.eddy, c.eddyscale, c.eddytime, c.guide = race.eddy, race.eddyscale, race.eddytime, race.guide
    c.free, c.turbulence, c.boil, c.reach = race.free, race.turbulence, race.boil, race.reach
    c.origin[0], c.origin[1] = stir.origin
    c.offset = stir.offset
    c.surface = 1 if stir.surface else 0
    c.ramp = stir.ramp
    w = [(wv.d[0], wv.d[1], wv.k, wv.omega, race.waves * wv.a, wv.q, wv.phase) for wv in stir.waves
         if 2 * math.pi / wv.k >= b.shortest]
    waves = np.ascontiguousarray(np.array(w, np.float32).reshape(-1), np.float32)
    heights = np.zeros(0, np.float32)
    if stir.surface:   # basin.jl:197-198: the balanced surface over [0, sqrt(2) (radius + 1)]
        c.hradius = float(np.float32(math.sqrt(2.0) * (b.radius + 1.0)))
        r = np.linspace(np.float32(0), np.float32(c.hradius), 2048, dtype=np.float32)
        heights = np.ascontiguousarray(surface(b, r), dtype=np.float32)
    return c, waves, len(w), heights


def set_stir(s, stir):
    """The stirring model `forces!` applies (basin.jl:92-103 `Stir`, 191-205 `SurfaceStir`)."""
    c, waves, nw, heights = stir_struct(stir)
    s.ctx.check(_lib.lib().mael_sparse_set_stir(s.h, C.byref(c), waves.ctypes.data if waves.size else None, nw,
                                                heights.ctypes.data if heights.size else None, heights.size))
    s.stir = stir


def maxspeed(s):
    """sflip.jl:2264-2272 `maxspeed(s)`."""
    v = C.c_float()
    s.ctx.check(_lib.lib().mael_sparse_maxspeed(s.h, C.byref(v)))
    return v.value


def before_projection(s, dt, reflect=False, carried_reflected=False):
    """The phases of `substep!` before its projection (sflip.jl:2295-2306): refine, advect,
    particles to grid, redistance, classify, trapped air, forces, and with `reflect` the
    pre-projection velocity kept for the reflection."""
    L = _lib.lib()
    refine(s)
    s.ctx.check(L.mael_sparse_advect(s.h, float(dt), 1 if carried_reflected else 0, float(s.basin.g), int(s.steps)))
    particles2grid(s, dt)
    redistance(s)
    classify(s)
    s.ctx.check(L.mael_sparse_filltrapped(s.h, 32))
    s.ctx.check(L.mael_sparse_forces(s.h, float(dt), float(s.t), float(s.basin.g)))
    if reflect:
        s.ctx.check(L.mael_sparse_save_reflect(s.h))


def substep(s, dt, reflect=False, carried_reflected=False, **solve):
    """sflip.jl:2295-2318 `substep!(s, dt; reflect, carried)`: one step `dt` carrying the
    velocity (or the reflected one); with `reflect` the reflected velocity is left in
    ur, vr, wr for the particles and the next half-step. Returns the solve's stats."""
    L = _lib.lib()
    before_projection(s, dt, reflect, carried_reflected)
    st = project(s, dt, **solve)
    if reflect:
        s.ctx.check(L.mael_sparse_kick(s.h))
    s.ctx.check(L.mael_sparse_extrapolate(s.h, 6, 2, 1 if reflect else 0))
    if reflect:
        s.ctx.check(L.mael_sparse_reflected(s.h))
    s.ctx.check(L.mael_sparse_set_flow(s.h, 1 if reflect else 0))
    grid2particles(s, dt)
    s.steps += 1
    resample(s)
    s.ctx.check(L.mael_sparse_set_flow(s.h, 0))
    s.t += dt
    return st


def step(s, dt, **solve):
    """sflip.jl:2288-2292 `step!(s, dt)`: with `s.reflect`, advection-reflection's two
    halves (the first ending in the reflected velocity, the second carrying it), else one
    substep. Returns the last half's solve stats; `s.iterations` holds both."""
    if not s.reflect:
        return substep(s, dt, **solve)
    substep(s, dt / 2, reflect=True, **solve)
    return substep(s, dt / 2, carried_reflected=True, **solve)


advance = step
