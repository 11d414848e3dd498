"""GPU parity of the whitewater (flip.jl:1411-1511, sflip.jl:2320-2417, whitewater.jl) with
the oracle, through the C ABI: two frames of emission from the resampled particles and
diffuse steps, the ring buffer bit for bit (positions, velocities, lives, kinds, head),
on the dense and the sparse basin; and the bake's whitewater files read back.

The stream the waves ride (the basin's swirl, `carried`) goes through the library's expf,
which is not the host's; the bit-for-bit runs take no stream, and a run with it is
checked to emit and keep its particles finite."""
import numpy as np
import pytest

from test_gpu_dense import PF, make_basin

pytestmark = pytest.mark.gpu

KEYS = ("X", "Y", "Z", "VX", "VY", "VZ", "LIFE", "KIND")


def _rules(orc, b, white_rules, stream):
    r = white_rules
    return orc.white_rules(air=r.air, crest=r.crest, energy=r.energy, kair=r.kair, kcrest=r.kcrest, life=r.life,
                           buoyancy=r.buoyancy, drag=r.drag, kick=r.kick, tossed=r.tossed, within=r.within,
                           floor=-b.depth, gravity=b.g, stream=1 if stream else 0, gamma=b.Γ, core=b.core)


def _compare(w, white, head):
    assert w.head == head
    for k in KEYS:
        got = w.get(k)
        assert np.array_equal(got.view(np.uint8), white[k].view(np.uint8)), k


def test_dense_whitewater_matches_oracle(mael, orc, ctx):
    from maelstrom_amd import whitewater as W
    s, sc = make_basin(mael, orc, scene="basin64")
    s.n = 0
    mael.resample(s, seeding=True)
    mael.resample(s)
    n = s.n
    rng = np.random.default_rng(2)
    ps = {f: s.get(f) for f in PF}
    for k, sd in (("VX", 4.0), ("VY", 4.0), ("VZ", 2.0)):
        ps[k] = (ps[k] + rng.normal(0, sd, n)).astype(np.float32)
    s.set_particles(ps, n)
    b = s.basin
    rules = W.WhitewaterRules(kair=60.0, kcrest=60.0, kick=2.0, tossed=0.5)
    w = W.Whitewater(b, capacity=400_000, rules=rules, stream=False)
    R = _rules(orc, b, rules, False)
    white = orc.empty_white(400_000)
    head, frames = 0, 0
    offsets, totals = s.get("OFFSETS").ravel(), s.get("TOTALS").ravel()
    for frame in range(2):
        print("case: dense whitewater frame %d" % frame, flush=True)
        W.whitewater(w, s, 1 / 24)
        head, frames = orc.dense_whitewater(sc["g"], ps, n, offsets, totals, sc["phi"], sc["u"], sc["v"], sc["w"], R, white,
                                            head, frames, 1 / 24)
        assert w.frames == frames == frame + 1
        _compare(w, white, head)
    assert head > 100
    kinds = w.get("KIND")[w.get("LIFE") > 0]
    assert set(np.unique(kinds)) >= {1, 2}


def test_ring_buffer_wraps_as_in_order(mael, orc, ctx):
    """A buffer smaller than a frame's emission keeps the last `capacity` particles made,
    as the sequential order does."""
    from maelstrom_amd import whitewater as W
    s, sc = make_basin(mael, orc, scene="basin64")
    s.n = 0
    mael.resample(s, seeding=True)
    mael.resample(s)
    n = s.n
    rng = np.random.default_rng(8)
    ps = {f: s.get(f) for f in PF}
    for k, sd in (("VX", 4.0), ("VY", 4.0), ("VZ", 2.0)):
        ps[k] = (ps[k] + rng.normal(0, sd, n)).astype(np.float32)
    s.set_particles(ps, n)
    b = s.basin
    rules = W.WhitewaterRules(kair=300.0, kcrest=300.0)
    w = W.Whitewater(b, capacity=5000, rules=rules, stream=False)
    R = _rules(orc, b, rules, False)
    white = orc.empty_white(5000)
    offsets, totals = s.get("OFFSETS").ravel(), s.get("TOTALS").ravel()
    head, frames = 0, 0
    for frame in range(2):
        W.whitewater(w, s, 1 / 24)
        head, frames = orc.dense_whitewater(sc["g"], ps, n, offsets, totals, sc["phi"], sc["u"], sc["v"], sc["w"], R, white,
                                            head, frames, 1 / 24)
        assert head > 5000 * (frame + 1)
        _compare(w, white, head)


def test_sparse_whitewater_matches_oracle(mael, orc, ctx):
    from maelstrom_amd import whitewater as W
    from test_gpu_sparse import PNAMES, _scene, _upload
    o = _scene()
    rng = np.random.default_rng(4)
    for nm in ("vx", "vy", "vz"):
        arr = o.array(nm)
        arr[:o.n] += rng.normal(0, 3.0, o.n).astype(np.float32)
    s = _upload(mael, o)
    s.steps = 2
    mael.sparse.resample(s)
    r = o.resample(seed=3)
    assert s.n == r["total"]
    b = s.basin
    rules = W.WhitewaterRules(kair=60.0, kcrest=60.0)
    w = W.Whitewater(b, capacity=400_000, rules=rules, stream=False)
    R = _rules(orc, b, rules, False)
    white = orc.empty_white(400_000)
    head, frames = 0, 0
    C = o.nslots * 4096
    offsets, totals = s.get("OFFSETS")[:C], s.get("TOTALS")[:C]
    assert np.array_equal(offsets, r["offsets"][:C])
    for frame in range(2):
        print("case: sparse whitewater frame %d" % frame, flush=True)
        W.whitewater(w, s, 1 / 24)
        head, frames = o.whitewater(offsets, totals, R, white, head, frames, 1 / 24)
        _compare(w, white, head)
    assert head > 100


def test_whitewater_frames_round_trip(mael, ctx, tmp_path):
    """Whitewater riding the vortex over a few steps of the 64x64x32 basin, written by the
    sink and read back; every live particle finite."""
    from maelstrom_amd import whitewater as W
    b = mael.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)
    s = mael.FlipBasin(b, dx=0.6, capacity=500_000, cfl=4.0)
    mael.initialize(s)
    # no stream: the vortex's smooth swirl is the water's whole speed, and the emission
    # counts speed relative to the stream
    frames = W.WhitewaterFrames(tmp_path / "bake", b, capacity=200_000, stream=False,
                                rules=W.WhitewaterRules(kair=200.0, kcrest=200.0, kick=1.0))
    packed = []
    for _ in range(3):
        mael.step(s, 1 / 24)
        frames.record(s)
        packed.append(W.pack_white(frames.w))
    frames.close()
    data = W.WhitewaterData(tmp_path / "bake")
    assert len(data) == 3
    for k, (out, vel) in enumerate(packed, start=1):
        assert np.array_equal(data.frame(k).T.view(np.uint16), out.view(np.uint16))
        assert np.array_equal(data.velocity(k).T.view(np.uint16), vel.view(np.uint16))
        assert np.isfinite(out.astype(np.float32)).all()
    assert sum(data.counts) > 0
