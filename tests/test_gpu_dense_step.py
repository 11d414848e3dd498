"""GPU parity of the dense basin's support phases (the rest of `step!(::FlipBasin)`,
flip.jl:1396-1410) with the CPU oracle, through the C ABI: ST-FLIP particle advection and
the grid's semi-Lagrangian advection bit for bit, extrapolation and redistancing bit for
bit, trapped air
against a connected-component labelling, the stirring forces to FP32 rounding of the
transcendental functions (the library's expf, sinf and cosf are not the host's), the
fastest speed exactly."""
import numpy as np
import pytest

from test_gpu_dense import PF, make_basin, rel_l2, scene  # noqa: F401  (scene is a fixture)

pytestmark = pytest.mark.gpu


def _seeded(mael, orc, scene, seed=4):
    s, sc = make_basin(mael, orc, scene=scene)
    s.n = 0
    mael.resample(s, seeding=True)
    mael.resample(s)
    rng = np.random.default_rng(seed)
    n = s.n
    ps = {f: s.get(f) for f in PF}
    ps["LAG"] = rng.uniform(-0.03, 0.03, n).astype(np.float32)
    for c in ("VX", "VY", "VZ"):
        ps[c] = (ps[c] + rng.normal(0, 0.5, n)).astype(np.float32)
    s.set_particles(ps, n)
    return s, sc, ps, rng


def test_advect_matches_oracle(mael, orc, ctx, scene):
    s, sc, ps, rng = _seeded(mael, orc, scene)
    g = sc["g"]
    n = s.n
    flags = []
    for name, f in (("FU", "u"), ("FV", "v"), ("FW", "w")):
        fl = (rng.random(sc[f].shape) > 0.05).astype(np.uint8)
        s.set(name, fl)
        flags.append(fl)
    dt = 0.12 if scene == "basin64" else 0.04
    s.steps = 7
    mael.advect(s, dt)
    out = orc.dense_advect_particles(g, ps, n, sc["u"], sc["v"], sc["w"], *flags, dt, 9.81, seed=7)
    for f in ("PX", "PY", "PZ", "LAG"):
        assert np.array_equal(s.get(f).view(np.uint32), out[f].view(np.uint32)), f
    for name, src, off in (("U0", "u", (0.0, 0.5, 0.5)), ("V0", "v", (0.5, 0.0, 0.5)), ("W0", "w", (0.5, 0.5, 0.0)),
                           ("PHI0", "phi", (0.5, 0.5, 0.5))):
        want = orc.dense_advect_grid(g, sc["u"], sc["v"], sc["w"], sc[src], off, dt)
        assert np.array_equal(s.get(name).view(np.uint32), want.view(np.uint32)), name


def test_extrapolate_matches_oracle(mael, orc, ctx, scene):
    s, sc = make_basin(mael, orc, scene=scene)
    g = sc["g"]
    rng = np.random.default_rng(12)
    want = []
    for a, (name, fl) in enumerate((("U", "FU"), ("V", "FV"), ("W", "FW"))):
        u = rng.normal(0, 1, sc[name.lower()].shape).astype(np.float32)
        f = (rng.random(u.shape) < 0.3).astype(np.uint8)
        s.set(name, u)
        s.set(fl, f)
        want.append(orc.dense_extrapolate(g, a, u, f, 6))
    mael.extrapolate(s, layers=6)
    for (name, fl), (u, f) in zip((("U", "FU"), ("V", "FV"), ("W", "FW")), want):
        assert np.array_equal(s.get(fl), f), fl
        assert np.array_equal(s.get(name).view(np.uint32), u.view(np.uint32)), name


def test_redistance_matches_oracle(mael, orc, ctx, scene):
    """`redistance!` (flip.jl:611-679) bit for bit: the scene's level set with noise (so
    cells of both signs meet away from the surface too), `band + 3` sweeps."""
    s, sc = make_basin(mael, orc, scene=scene)
    g = sc["g"]
    rng = np.random.default_rng(21)
    phi = (sc["phi"] + rng.normal(0, 0.3 * g.dx, sc["phi"].shape)).astype(np.float32)
    s.set("PHI", phi)
    it = s.band + 3
    print("case: dense redistance, %d sweeps, %s" % (it, scene), flush=True)
    mael.redistance(s, it)
    want = orc.dense_redistance(g, phi, it)
    assert np.array_equal(s.get("PHI").view(np.uint32), want.view(np.uint32))


def _stir_params(orc, b, race, stir):
    return orc.stir_params(radius=b.radius, band=b.band, relax=b.relax, core=b.core, gamma=b.Γ,
                           current_x=race.current[0], current_y=race.current[1], eddy=race.eddy, eddyscale=race.eddyscale,
                           eddytime=race.eddytime, guide=race.guide, free=race.free, turbulence=race.turbulence,
                           boil=race.boil, reach=race.reach, origin_x=stir.origin[0], origin_y=stir.origin[1],
                           offset=stir.offset)


@pytest.mark.parametrize("waves", [False, True])
def test_forces_match_oracle(mael, orc, ctx, scene, waves):
    s, sc = make_basin(mael, orc, scene=scene)
    g = sc["g"]
    b = s.basin
    race = mael.Race(eddy=1.5, turbulence=0.8, boil=0.7)
    stir = mael.Stir(b, waves=mael.narutowaves() if waves else [], race=race, origin=(3.0, -2.0), offset=0.5)
    mael.dense.set_stir(s, stir)
    _, table, nw, _ = mael.sparse.stir_struct(stir)
    rng = np.random.default_rng(3)
    u, v, w = ((sc[f] + rng.normal(0, 0.3, sc[f].shape)).astype(np.float32) for f in "uvw")
    for name, arr in (("U", u), ("V", v), ("W", w)):
        s.set(name, arr)
    dt, t = 0.05, 2.5
    s.t = t
    mael.forces(s, dt)
    want = orc.dense_forces(g, _stir_params(orc, b, race, stir), table, u, v, w, dt, t, 9.81)
    for name, ref, before in zip("UVW", want, (u, v, w)):
        got = s.get(name)
        change = np.abs(ref - before).max()
        assert change > 1e-3, name
        err = np.abs(got.astype(np.float64) - ref).max()
        assert err <= 2e-6 * max(1.0, np.abs(ref).max()), (name, err)


def test_filltrapped_matches_flood_fill(mael, orc, ctx):
    """`filltrapped!` (flip.jl:688-760): the AIR reached from the top layer's AIR through
    face-adjacent AIR stays; every other AIR cell becomes LIQUID at least 2 dx deep."""
    import scipy.ndimage as ndi
    SOLID, LIQUID, AIR = 0, 1, 2
    s, sc = make_basin(mael, orc, scene="basin64")
    shape = sc["types"].shape
    rng = np.random.default_rng(17)
    for trial in range(3):
        t = np.where(rng.random(shape) < 0.45 + 0.1 * trial, AIR, LIQUID).astype(np.uint8)
        t[rng.random(shape) < 0.05] = SOLID
        phi = rng.normal(0, 1, shape).astype(np.float32)
        s.set("TYPES", t)
        s.set("PHI", phi)
        mael.filltrapped(s, maxrounds=256)
        lab, _ = ndi.label(t == AIR, structure=ndi.generate_binary_structure(3, 1))
        top = np.unique(lab[-1][lab[-1] > 0])
        trapped = (t == AIR) & ~np.isin(lab, top)
        want = t.copy()
        want[trapped] = LIQUID
        assert np.array_equal(s.get("TYPES"), want), trial
        wphi = phi.copy()
        wphi[trapped] = np.minimum(phi[trapped], -np.float32(2) * np.float32(sc["g"].dx))
        assert np.array_equal(s.get("PHI"), wphi), trial


def test_maxspeed_exact(mael, orc, ctx, scene):
    s, sc = make_basin(mael, orc, scene=scene)
    want = max(np.abs(sc[f]).max() for f in "uvw")
    assert mael.maxspeed(s) == want


def test_steps_keep_water_basin64(mael, ctx):
    """Three whole steps of `step!(::FlipBasin)` on the 64x64x32 basin at CFL 4 with the
    reference's stir: finite, the volume within 5 % (test/runtests.jl:56)."""
    b = mael.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)
    s = mael.FlipBasin(b, dx=0.6, capacity=1_000_000, cfl=4.0)
    mael.initialize(s)
    for _ in range(3):
        dt = min(s.cfl * s.grid.dx / max(mael.maxspeed(s), 1e-3), 1 / 24 * 8)
        st = mael.step(s, dt)
        assert st.rel_residual <= 1e-4
    assert np.isfinite(mael.maxspeed(s)) and s.n > 0
    assert 0.95 < mael.volume(s) / s.volume0 < 1.05
