"""`FlipBasin`: the dense narrow-band FLIP basin on the HIP library (flip.jl:156-1410).

The Python mirror of the reference's Julia API for the hot path: `FlipBasin`,
`initialize`, `resample`, `particles2grid`, `project`, `grid2particles`, `step`
(Julia's bang names without the `!`). Every phase runs in the gfx950 library; this
module only holds host-side bookkeeping (time, step count, the volume control's
reference volume) exactly where the reference keeps it on the host.
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import ARRAY_ID, PARTICLE_FIELDS, DenseGrid, MaelstromError, SolveOpts, SolveStats
from .basin import Stir, surface

_DTYPES = {1: np.uint8, 4: np.float32, 8: np.int64}
_INT_ARRAYS = {"KEY", "RANK", "COUNTS", "TOTALS", "OFFSETS"}


def flipgrid(basin, dx, top=None, pad=2):
    """flip.jl:48-62 `FlipGrid(b; dx, top = b.rimwall, pad = 2)`."""
    g = DenseGrid()
    top = basin.rimwall if top is None else top
    st = _lib.lib().mael_flipgrid(basin.radius, basin.depth, top, dx, pad, C.byref(g))
    if st != _lib.OK:
        raise MaelstromError(st, "bad grid for dx=%r" % dx)
    return g


class FlipBasin:
    """flip.jl:189 `FlipBasin(b, backend; dx, stir = Stir(b), band = 4, nmin = 8, cfl = 2, capacity = 12_000_000)`.

    `stir` is the stirring model `forces!` applies (basin.jl:92-181 `Stir`); `stir=False`
    leaves gravity alone.
    """

    def __init__(self, basin, dx, band=4, nmin=8, cfl=2.0, capacity=12_000_000, ctx=None, top=None, pad=2, stir=None):
        self.basin = basin
        self.ctx = ctx or _lib.default_context()
        self.grid = flipgrid(basin, dx, top, pad)
        self.top, self.pad = top, pad
        self.band, self.nmin, self.cfl, self.capacity = int(band), int(nmin), float(cfl), int(capacity)
        self.h = C.c_void_p()
        self.ctx.check(_lib.lib().mael_dense_create(self.ctx.h, C.byref(self.grid), self.capacity, self.band,
                                                    self.nmin, C.byref(self.h)))
        self.volume0 = 0.0
        self.t = 0.0
        self.steps = 0
        self.iterations = []
        self.last_solve = None
        self.stir = None
        if stir is not False:
            set_stir(self, Stir(basin) if stir is None else stir)

    # ── arrays ─────────────────────────────────────────────────────────────
    @property
    def n_cells(self):
        return tuple(int(x) for x in self.grid.n)

    def shape(self, name):
        """numpy shape (z, y, x) of a grid array, or (capacity,) of a particle array."""
        n1, n2, n3 = self.n_cells
        if name in ("U", "U0", "FU", "MOM_U", "WT_U", "UB_U", "UB0_U"):
            return (n3, n2, n1 + 1)
        if name in ("V", "V0", "FV", "MOM_V", "WT_V", "UB_V", "UB0_V"):
            return (n3, n2 + 1, n1)
        if name in ("W", "W0", "FW", "MOM_W", "WT_W", "UB_W", "UB0_W"):
            return (n3 + 1, n2, n1)
        if name in ("PHI", "PHI0", "TYPES", "P", "RHS", "COUNTS", "TOTALS", "OFFSETS"):
            return (n3, n2, n1)
        return (self.capacity,)

    def _info(self, name):
        ptr, cnt, es = C.c_void_p(), C.c_int64(), C.c_int32()
        self.ctx.check(_lib.lib().mael_dense_array(self.h, ARRAY_ID[name], C.byref(ptr), C.byref(cnt), C.byref(es)))
        dt = np.int32 if name in _INT_ARRAYS else _DTYPES[es.value]
        return ptr.value, cnt.value, dt

    def get(self, name, count=None):
        """Download array `name` (`Array(s.x)`); particle arrays are cut to `count` (default n)."""
        _, cnt, dt = self._info(name)
        if name in PARTICLE_FIELDS or name in ("KEY", "RANK") or name.startswith("PB_"):
            cnt = self.n if count is None else count
            out = np.empty(cnt, dtype=dt)
        else:
            out = np.empty(self.shape(name), dtype=dt)
            assert out.size == cnt
        if out.size:
            self.ctx.check(_lib.lib().mael_dense_download(self.h, ARRAY_ID[name], out.ctypes.data, out.size))
        return out

    def set(self, name, arr):
        """Upload into array `name` (`copyto!(s.x, arr)`)."""
        _, cnt, dt = self._info(name)
        a = np.ascontiguousarray(arr, dtype=dt).ravel()
        if a.size > cnt:
            raise ValueError("%s holds %d elements, got %d" % (name, cnt, a.size))
        if a.size:
            self.ctx.check(_lib.lib().mael_dense_upload(self.h, ARRAY_ID[name], a.ctypes.data, a.size))

    def copy(self, dst, src):
        self.ctx.check(_lib.lib().mael_dense_copy(self.h, ARRAY_ID[dst], ARRAY_ID[src]))

    @property
    def n(self):
        v = C.c_int64()
        self.ctx.check(_lib.lib().mael_dense_count(self.h, C.byref(v)))
        return v.value

    @n.setter
    def n(self, value):
        self.ctx.check(_lib.lib().mael_dense_set_count(self.h, int(value)))

    def particles(self):
        """The live particles as a dict of arrays (x, y, z, vx, vy, vz, c0..c8, lag)."""
        return {f: self.get(f) for f in PARTICLE_FIELDS}

    def set_particles(self, ps, n=None):
        n = len(ps["PX"]) if n is None else n
        for f in PARTICLE_FIELDS:
            self.set(f, np.asarray(ps[f][:n], dtype=np.float32))
        self.n = n

    def sync(self):
        self.ctx.check(_lib.lib().mael_dense_sync(self.h))

    def close(self):
        if self.h:
            _lib.lib().mael_dense_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ── phases ───────────────────────────────────────────────────────────────────

def resample(s, seeding=False):
    """flip.jl:1364-1382 `resample!(s; seeding)`: returns the new particle count."""
    total = C.c_int64()
    s.ctx.check(_lib.lib().mael_dense_resample(s.h, 1 if seeding else 0, (s.steps + 1) & 0xffffffff, C.byref(total)))
    return total.value


def particles2grid(s, dt):
    """flip.jl:587-606 `particles2grid!(s, dt)`."""
    s.ctx.check(_lib.lib().mael_dense_p2g(s.h, float(dt)))
    return s


def grid2particles(s, dt):
    """flip.jl:1246-1252 `grid2particles!(s, dt)`."""
    s.ctx.check(_lib.lib().mael_dense_g2p(s.h, float(dt), float(s.basin.g)))
    return s


def solve_options(tol=1e-4, maxiter=60, sweeps=2, coarse=40, precision=0, refine=0, warm=1):
    """`pressure!`'s options (flip.jl:1082 defaults) plus precision, FP64 refinement (off: the
    reference's FP32 PCG) and warm start."""
    return SolveOpts(float(tol), int(maxiter), int(sweeps), int(coarse), int(precision), int(refine), int(warm))


def volume(s):
    """flip.jl:293 `volume(s)`, m^3."""
    v = C.c_double()
    s.ctx.check(_lib.lib().mael_dense_volume(s.h, C.byref(v)))
    return v.value


def gamma_for(s, dt, recover=0.5):
    """flip.jl:1146-1147: the volume control's divergence γ."""
    loss = (s.volume0 - volume(s)) / s.volume0 if s.volume0 else 0.0
    return float(np.float32(min(max(loss / recover, -0.01 / dt), 0.01 / dt)))


def project(s, dt, recover=0.5, gamma=None, **solve):
    """flip.jl:1144-1156 `project!(s, dt; recover)`; `solve` are `pressure!`'s options
    (tol, maxiter, sweeps, coarse) plus precision and refine. Returns the solve's stats."""
    g = gamma_for(s, dt, recover) if gamma is None else float(gamma)
    opts = solve_options(**solve)
    st = SolveStats()
    s.ctx.check(_lib.lib().mael_dense_project(s.h, float(dt), g, C.byref(opts), C.byref(st)))
    s.iterations.append(st.iterations)
    s.last_solve = st
    return st


# ── adjoints (no counterpart in the reference, which has no AD) ──────────────
# The transposes of the hot path's linear maps at the state the forward phase left, for
# reverse-mode gradients: cotangents of the face velocities in UB_U/V/W (of u0 in
# UB0_U/V/W) and of the particles' v and APIC rows in PB_VX ... PB_C8.

def grid2particles_adjoint(s, dt):
    """Transpose of `grid2particles!` in u: UB = G2P^T(PB); PB becomes the cotangent of
    G2P's input particles (v passes through where the grid was not reached)."""
    s.ctx.check(_lib.lib().mael_dense_g2p_adjoint(s.h, float(dt)))
    return s


def project_adjoint(s, **solve):
    """Transpose of `project!`'s linear part, in place on UB (a transposed MG-PCG solve;
    P is left holding the adjoint pressure). Returns the solve's stats."""
    solve.setdefault("warm", 0)
    opts = solve_options(**solve)
    st = SolveStats()
    s.ctx.check(_lib.lib().mael_dense_project_adjoint(s.h, C.byref(opts), C.byref(st)))
    return st


def particles2grid_adjoint(s, dt):
    """Transpose of `particles2grid!` in the particles' v and APIC rows (PB += P2G^T(UB),
    added to what `grid2particles_adjoint` passed through) and in u0 (UB0 =), at the
    positions, lags, level set and weights of the last deposit."""
    s.ctx.check(_lib.lib().mael_dense_p2g_adjoint(s.h, float(dt)))
    return s


def classify(s):
    """flip.jl:674-677 `classify!(s)`."""
    s.ctx.check(_lib.lib().mael_dense_classify(s.h))
    return s


def redistance(s, iterations=None):
    """flip.jl:672-679 `redistance!(s; iterations = band + 3)`."""
    it = s.band + 3 if iterations is None else int(iterations)
    s.ctx.check(_lib.lib().mael_dense_redistance(s.h, it))
    return s


def initialize(s):
    """flip.jl:325-348 `initialize!(s)`: the balanced vortex, the band seeded with
    particles, their deposit."""
    g = s.grid
    rmax = np.float32(math.hypot(g.lo[0], g.lo[1]))
    r = np.linspace(np.float32(0), rmax, 2048, dtype=np.float32)
    eta = np.ascontiguousarray(surface(s.basin, r), dtype=np.float32)
    L = _lib.lib()
    s.ctx.check(L.mael_dense_init_vortex(s.h, eta.ctypes.data, eta.size, float(rmax), float(s.basin.Γ),
                                         float(s.basin.core)))
    redistance(s, 2 * s.band + 4)
    classify(s)
    s.ctx.check(L.mael_dense_solidify(s.h))
    for f in ("FU", "FV", "FW"):
        s.set(f, np.ones(s.shape(f), dtype=np.uint8))
    s.n = 0
    resample(s, seeding=True)
    resample(s)
    s.copy("PHI0", "PHI")
    for a, b in (("U0", "U"), ("V0", "V"), ("W0", "W")):
        s.copy(a, b)
    particles2grid(s, 1.0)
    redistance(s)
    classify(s)
    s.volume0 = volume(s)
    s.t = 0.0
    return s


def set_stir(s, stir):
    """The stirring model `forces!` applies (basin.jl:92-103 `Stir`)."""
    from .sparse import stir_struct
    if stir.surface:
        raise ValueError("the dense basin takes Stir, not SurfaceStir")
    c, waves, nw, _ = stir_struct(stir)
    s.ctx.check(_lib.lib().mael_dense_set_stir(s.h, C.byref(c), waves.ctypes.data if waves.size else None, nw))
    s.stir = stir


def maxspeed(s):
    """flip.jl:1387 `maxspeed(s)`: the largest face speed."""
    v = C.c_float()
    s.ctx.check(_lib.lib().mael_dense_maxspeed(s.h, C.byref(v)))
    return v.value


def advect(s, dt):
    """flip.jl:490-499 `advect!(s, dt)`: the particles (ST-FLIP, seed `s.steps`), then the
    grid's velocity and level set into u0, v0, w0 and φ0."""
    s.ctx.check(_lib.lib().mael_dense_advect(s.h, float(dt), float(s.basin.g), int(s.steps)))
    return s


def filltrapped(s, maxrounds=32):
    """flip.jl:754-760 `filltrapped!(s; maxrounds = 32)`."""
    s.ctx.check(_lib.lib().mael_dense_filltrapped(s.h, int(maxrounds)))
    return s


def forces(s, dt):
    """flip.jl:786-798 `forces!(s, dt)`: gravity and the stir at time `s.t`, then the walls."""
    s.ctx.check(_lib.lib().mael_dense_forces(s.h, float(dt), float(s.t), float(s.basin.g)))
    return s


def extrapolate(s, layers=4):
    """flip.jl:1194-1206 `extrapolate!(s; layers = 4)`."""
    s.ctx.check(_lib.lib().mael_dense_extrapolate(s.h, int(layers)))
    return s


def step(s, dt, **solve):
    """flip.jl:1396-1410 `step!(s, dt)`: advect the particles and the grid, particles to
    grid, redistance, classify, trapped air, forces, project, extrapolate six layers, grid
    to particles, resample. `solve` overrides `pressure!`'s options. Returns the solve's
    stats."""
    advect(s, dt)
    particles2grid(s, dt)
    redistance(s)
    classify(s)
    filltrapped(s)
    forces(s, dt)
    st = project(s, dt, **solve)
    extrapolate(s, layers=6)
    grid2particles(s, dt)
    s.steps += 1
    resample(s)
    s.t += dt
    return st


advance = step
