"""GPU parity of the bake writers' device side with the oracle, through the C ABI: the
band's kept particles packed as Float16 bit for bit and in particle order (dense and
sparse), the quantised level sets byte for byte (flip.jl:1859-1867, sflip.jl:2427-2432),
and the sinks' files read back by ParticleFrames and LevelSetFrames."""
import numpy as np
import pytest

from test_gpu_dense import PF, make_basin

pytestmark = pytest.mark.gpu


def _seeded_dense(mael, orc):
    s, sc = make_basin(mael, orc, scene="basin64")
    s.n = 0
    mael.resample(s, seeding=True)
    mael.resample(s)
    rng = np.random.default_rng(6)
    ps = {f: s.get(f) for f in PF}
    ps["LAG"] = rng.uniform(-0.05, 0.05, s.n).astype(np.float32)
    s.set_particles(ps, s.n)
    return s, sc, ps


def test_dense_pack_and_quantize(mael, orc, ctx):
    from maelstrom_amd import bake
    s, sc, ps = _seeded_dense(mael, orc)
    pos, vel = bake.pack_band(s, keep=3.5, throat=3.0)
    wpos, wvel = orc.pack_band(ps, s.n, 3.5, 3.0)
    assert 0 < len(wpos) < s.n
    assert np.array_equal(pos.view(np.uint16), wpos.view(np.uint16))
    assert np.array_equal(vel.view(np.uint16), wvel.view(np.uint16))
    rng_ = np.float32((s.band + 2) * np.float32(s.grid.dx))
    q = bake.quantize_levelset(s, rng_)
    assert np.array_equal(q, orc.dense_quantize(sc["g"], sc["phi"], rng_))


def test_sparse_pack_and_quantize(mael, orc, ctx):
    from maelstrom_amd import bake
    from test_gpu_sparse import PNAMES, _scene, _upload
    o = _scene()
    s = _upload(mael, o)
    rng = np.random.default_rng(9)
    lag = rng.uniform(-0.05, 0.05, o.n).astype(np.float32)
    s.set("LAG", lag)
    ps = {nm.upper(): o.array(nm)[:o.n].copy() for nm in PNAMES}
    ps["LAG"] = lag
    pos, vel = bake.pack_band(s, keep=3.5, throat=3.0)
    wpos, wvel = orc.pack_band(ps, o.n, 3.5, 3.0)
    assert 0 < len(wpos) < o.n
    assert np.array_equal(pos.view(np.uint16), wpos.view(np.uint16))
    assert np.array_equal(vel.view(np.uint16), wvel.view(np.uint16))
    cell, rng_ = np.float32(2) * np.float32(0.6), np.float32(6) * np.float32(0.6)
    q = bake.quantize_levelset(s, rng_)
    assert np.array_equal(q, o.quantize(cell, rng_))


def test_sinks_round_trip(mael, orc, ctx, tmp_path):
    from maelstrom_amd import bake
    b = mael.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)
    s = mael.FlipBasin(b, dx=0.6, capacity=500_000, cfl=4.0)
    mael.initialize(s)
    path = tmp_path / "bake"
    parts, levels = bake.BandParticles(path, b), bake.LevelSets(path, s)
    want = []
    for k in range(3):
        if k:
            mael.step(s, 1 / 24)
        parts.record(s)
        levels.record(s)
        want.append((bake.pack_band(s), bake.quantize_levelset(s, levels.range)))
    parts.close()
    levels.close()
    pf, lf = bake.ParticleFrames(path), bake.LevelSetFrames(path)
    assert len(pf) == len(lf) == 3
    for k, ((pos, vel), q) in enumerate(want, start=1):
        assert np.array_equal(pf.frame(k).T.view(np.uint16), pos.view(np.uint16))
        assert np.array_equal(pf.velocity(k).T.view(np.uint16), vel.view(np.uint16))
        assert np.array_equal(lf.data[k - 1], q)
