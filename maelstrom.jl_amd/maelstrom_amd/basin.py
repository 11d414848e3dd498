"""The simulated basin (basin.jl:13-30) and the balanced vortex surface (sea.jl:38-58).

Only what the hot path's inputs need: the basin's geometry and vortex, and the free
surface in balance with it, which `initialize` turns into the level set.
"""
from dataclasses import dataclass, field, replace

import numpy as np


@dataclass(frozen=True)
class Basin:
    """basin.jl:13-30 `Basin`: field units as there (m, m^2/s, s)."""
    radius: float = 14.0
    depth: float = 6.0
    spacing: float = 0.3
    Γ: float = 150.0
    core: float = 2.5
    drain: float = 1.2
    band: float = 3.0
    relax: float = 0.3
    shortest: float = 3.0
    headroom: float = 6.0
    g: float = 9.81
    ρ0: float = 1000.0
    sound: float = 100.0
    layers: int = 4
    rimwall: float = 3.0
    arc: float = 1.5


def narutobasin(**kw):
    """basin.jl:47-49 `narutobasin`: a Naruto-sized basin (note rimwall 1.5)."""
    d = dict(radius=10.0, depth=3.0, spacing=0.1, Γ=45.0, core=1.5, drain=0.0, band=2.5, shortest=1.0,
             headroom=3.0, sound=60.0, rimwall=1.5)
    d.update(kw)
    return Basin(**d)


def oseen(G, core, r):
    """sea.jl:39: azimuthal speed of the Lamb-Oseen vortex, FP32."""
    r = np.asarray(r, dtype=np.float32)
    G, core = np.float32(G), np.float32(core)
    q = r / core
    with np.errstate(over="ignore"):
        return G / (np.float32(2.0) * np.float32(np.pi) * np.maximum(r, np.float32(1e-4))) * (
            np.float32(1.0) - np.exp(-(q * q)))


def balanced(G, core, g, r, steps=2000):
    """sea.jl:48-58: the free surface over the vortex in balance at radius r (FP32, trapezoid in log r)."""
    r = np.asarray(r, dtype=np.float32)
    f32 = np.float32
    a = np.log(np.maximum(r, f32(1e-3)))
    z = np.log(f32(50.0) * f32(core) + f32(50.0) * np.maximum(r, f32(1.0)))
    h = (z - a) / f32(steps)
    s = np.zeros_like(r)
    for i in range(steps + 1):
        rho = np.exp(a + f32(i) * h)
        o = oseen(G, core, rho)
        s += (f32(0.5) if i in (0, steps) else f32(1.0)) * (o * o) / f32(g)
    return -s * h


def surface(b, r):
    """basin.jl:36 `surface`: the basin's balanced free surface at radius r."""
    return balanced(b.Γ, b.core, b.g, r)


__all__ = ["Basin", "narutobasin", "oseen", "balanced", "surface", "replace", "Race", "Wave", "Stir", "waves", "narutowaves"]


@dataclass
class Race:
    """basin.jl:69-80 `Race`: the tidal race the stirring reads."""
    current: tuple = (0.0, 0.0)
    eddy: float = 0.6
    eddyscale: float = 3.5
    eddytime: float = 4.0
    guide: float = 4.0
    free: float = 0.6
    turbulence: float = 0.0
    boil: float = 0.5
    reach: float = 2.5
    waves: float = 2.5


@dataclass
class Wave:
    """sea.jl:6-13 `Wave`: direction, wave number, angular frequency, amplitude, steepness, phase."""
    d: tuple
    k: float
    omega: float
    a: float
    q: float
    phase: float


def waves(n=48, wind=0.6, spread=1.1, hs=0.35, lmin=0.5, lmax=24.0, seed=3):
    """sea.jl:22-32 `waves(; n, wind, spread, hs, λmin, λmax, seed)`: `n` Gerstner waves,
    wavelengths log-spaced from `lmax` down to `lmin`, amplitudes as wavelength^1.25
    scaled so the sea's significant height is about `hs`, steepness 0.55, deep-water
    frequency sqrt(9.81 k), directions within `wind ± spread`. Each wave draws its
    direction, then its phase; the draws come from numpy's PCG64 seeded with `seed`, not
    from Julia's MersenneTwister, so the table has the reference's shape and spectrum but
    not its exact angles and phases."""
    rng = np.random.default_rng(seed)
    lams = np.exp(np.linspace(np.log(lmax), np.log(lmin), n))
    raw = lams ** 1.25
    scale = hs / (4.0 * np.sqrt(np.sum(raw * raw) / 2.0))
    out = []
    for lam, r in zip(lams, raw):
        th = wind + spread * (2.0 * rng.random() - 1.0)
        k = 2.0 * np.pi / lam
        ph = 2.0 * np.pi * rng.random()
        f = lambda v: float(np.float32(v))
        out.append(Wave((f(np.cos(th)), f(np.sin(th))), f(k), f(np.sqrt(9.81 * k)), f(scale * r), 0.55, f(ph)))
    return out


def narutowaves(hs=1.4, lmin=1.0, lmax=10.0, wind=0.4, spread=0.9, seed=5):
    """sea.jl:35-36 `narutowaves`: a wind sea of 1.4 m significant height, 1-10 m long, 32 waves."""
    return waves(n=32, wind=wind, spread=spread, hs=hs, lmin=lmin, lmax=lmax, seed=seed)


@dataclass
class Stir:
    """basin.jl:92-103 `Stir(b, waves; origin, offset, race)`; with `surface`, basin.jl:191-205
    `SurfaceStir(b, waves; race, ramp)`. The waves are a table given by the caller (the
    reference draws them from Julia's random stream)."""
    basin: Basin
    waves: list = field(default_factory=list)
    race: Race = field(default_factory=Race)
    origin: tuple = (0.0, 0.0)
    offset: float = 0.0
    surface: bool = False
    ramp: float = 2.0
