"""Checkpoint and restore (checkpoint.jl:7-66, test/runtests.jl:40-45) through the C ABI:
a restored basin holds the saved state bit for bit (arrays, topology, clocks), and the
next step from it gives the same bits as the next step of the basin that was saved (the
steps are deterministic). Also the reference's errors: a save between frames, a restore
with too small a particle capacity."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PF = ["PX", "PY", "PZ", "VX", "VY", "VZ", "C0", "C1", "C2", "C3", "C4", "C5", "C6", "C7", "C8", "LAG"]
DENSE = ["PHI", "TYPES", "U", "V", "W", "FU", "FV", "FW", "P"]
SPARSE = ["PHI", "TYPES", "U", "V", "W", "FU", "FV", "FW", "P"]


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def test_dense_round_trip_and_next_step(mael, tmp_path):
    print("case: dense checkpoint 64x64x32", flush=True)
    b = mael.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)
    s = mael.FlipBasin(b, dx=0.6, capacity=500_000, cfl=4.0)
    mael.initialize(s)
    for _ in range(2):
        mael.step(s, 1 / 24)
    with pytest.raises(ValueError, match="Save at a completed frame"):
        s.t += 0.01
        mael.save_checkpoint(s, tmp_path / "x.ckpt")
    s.t -= 0.01
    s.t = round(s.t * 24) / 24
    path = mael.save_checkpoint(s, tmp_path / "dense.ckpt")
    r = mael.load_checkpoint(path)
    assert (r.t, r.steps, r.n, r.volume0, r.iterations) == (s.t, s.steps, s.n, s.volume0, s.iterations)
    for k in DENSE + PF:
        assert _same(r.get(k), s.get(k)), k
    assert _same(r.get("PHI0"), s.get("PHI"))
    for x in (s, r):
        mael.step(x, 1 / 24)
    for k in DENSE + PF:
        assert _same(r.get(k), s.get(k)), k
    with pytest.raises(mael.MaelstromError, match="Particle capacity is too small"):
        mael.load_checkpoint(path, capacity=10)


def test_sparse_round_trip_into_larger_pool(mael, tmp_path):
    print("case: sparse checkpoint 60 cm", flush=True)
    b = mael.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)
    stir = mael.Stir(b, waves=mael.narutowaves(), race=mael.Race(eddy=3.0), surface=True)
    s = mael.SparseBasin(b, dx=0.6, blocks=64, capacity=300_000, cfl=4, stir=stir)
    mael.sparse.initialize(s)
    mael.sparse.step(s, 1 / 24)
    path = mael.save_checkpoint(s, tmp_path / "sparse.ckpt")
    r = mael.load_checkpoint(path, blocks=96)
    assert r.blocks == 96 and r.stir.surface and len(r.stir.waves) == len(stir.waves)
    assert (r.t, r.steps, r.n, r.volume0, r.iterations) == (s.t, s.steps, s.n, s.volume0, s.iterations)
    ts, tr = s.topology(), r.topology()
    assert np.array_equal(ts[0], tr[0]) and ts[2] == tr[2] and np.array_equal(ts[3], tr[3])
    assert np.array_equal(ts[1][:ts[2]], tr[1][:tr[2]])
    ns = s.nslots
    for k in SPARSE:
        cells = 4096 if k in ("PHI", "TYPES", "P") else 4352
        assert _same(r.get(k + "_F")[:ns * cells], s.get(k + "_F")[:ns * cells]), k
        assert _same(r.get(k + "_C"), s.get(k + "_C")), k
    for k in PF:
        assert _same(r.get(k), s.get(k)), k
    for x in (s, r):
        mael.sparse.step(x, 1 / 24)
    assert r.nslots == s.nslots
    ns = s.nslots
    for k in SPARSE:
        cells = 4096 if k in ("PHI", "TYPES", "P") else 4352
        assert _same(r.get(k + "_F")[:ns * cells], s.get(k + "_F")[:ns * cells]), k
    for k in PF:
        assert _same(r.get(k), s.get(k)), k
