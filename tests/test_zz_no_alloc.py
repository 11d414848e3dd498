"""mael_dense_resample_noalloc (the resampling scan's tile sums in the library's static
device array, no allocation) against mael_dense_resample, through the C ABI: bit-identical
particles, keys, ranks, counts, totals and offsets, on two calls in a row, the second on a
smaller grid; and the size limit's error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PF = ["PX", "PY", "PZ", "VX", "VY", "VZ", "C0", "C1", "C2", "C3", "C4", "C5", "C6", "C7", "C8", "LAG"]
OUT = ("KEY", "RANK", "COUNTS", "TOTALS", "OFFSETS")


def _basin(mael, orc, big):
    if big:   # Basin(radius 18, depth 12, Γ 225, core 2.5) at 0.6 m: 64x64x32
        b = mael.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)
        args = (18.0, 12.0, 3.0, 0.6, 225.0, 2.5)
    else:     # the reference's dense test basin at 0.2 m: 44x44x20
        b = mael.narutobasin(radius=4.0, depth=2.0, Γ=15.0, core=1.0, band=1.0, headroom=1.5)
        args = (4.0, 2.0, 1.5, 0.2, 15.0, 1.0)
    s = mael.FlipBasin(b, dx=args[3], capacity=1_000_000)
    g, phi, u, v, w, types = orc.dense_scene(*args)
    for name, arr in (("PHI", phi), ("U", u), ("V", v), ("W", w), ("TYPES", types)):
        s.set(name, arr)
    return s


def _resample(mael, s, noalloc, seeding, seed):
    L = mael.lib()
    fn = L.mael_dense_resample_noalloc if noalloc else L.mael_dense_resample
    total = C.c_int64()
    s.ctx.check(fn(s.h, 1 if seeding else 0, seed, C.byref(total)))
    return total.value


def test_noalloc_resample_matches(mael, orc, ctx):
    for big in (True, False):   # the second call on a smaller grid
        print("case: resample_noalloc, %s grid" % ("64x64x32" if big else "44x44x20"), flush=True)
        a, b = _basin(mael, orc, big), _basin(mael, orc, big)
        for s in (a, b):
            s.n = 0
        for seeding, seed in ((True, 1), (False, 1), (False, 7)):
            na = _resample(mael, a, False, seeding, seed)
            nb = _resample(mael, b, True, seeding, seed)
            assert na == nb > 1000
            for f in PF:
                assert np.array_equal(a.get(f).view(np.uint32), b.get(f).view(np.uint32)), f
            for f in OUT[2:]:
                assert np.array_equal(a.get(f), b.get(f)), f
        ps = a.particles()
        n = a.n
        rng = np.random.default_rng(3)
        perm = rng.permutation(n)
        for s in (a, b):   # an unsorted input: ranks and the survivors of crowded cells
            s.set_particles({f: ps[f][perm] for f in PF}, n)
        na = _resample(mael, a, False, False, 9)
        nb = _resample(mael, b, True, False, 9)
        assert na == nb
        for f in OUT:
            assert np.array_equal(a.get(f, count=n) if f in ("KEY", "RANK") else a.get(f),
                                  b.get(f, count=n) if f in ("KEY", "RANK") else b.get(f)), f
        for f in PF:
            assert np.array_equal(a.get(f).view(np.uint32), b.get(f).view(np.uint32)), f


def test_noalloc_resample_rejects_large_grids(mael, ctx):
    """A grid of more than 16,384 scan tiles (here 256x256x288, 18.9 M cells) is refused with
    the limit in the message, and the basin keeps its particles."""
    b = mael.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)
    s = mael.FlipBasin(b, dx=0.15, capacity=1000, top=27.0)
    assert int(np.prod(s.n_cells)) > 16384 * 1024
    s.n = 0
    with pytest.raises(mael.MaelstromError, match="more than the 16386 of its static scratch"):
        _resample(mael, s, True, True, 1)
    assert s.n == 0
