"""The dense basin at full size: `Basin(radius 18, depth 12, Γ 225, core 2.5)` at 0.15 m,
256x256x128 cells, 6 multigrid levels, the band seeded by `initialize`. The FP64 oracle
does not finish a solve of this size in a test's time, so the checks here are the ones
that hold at any size (the smaller scenes of test_gpu_dense.py hold the oracle parity):

- P2G is order-independent: a permutation of the particles gives the same bits;
- P2G conserves momentum: the deposited momentum and weight sum to the particles'
  temporally weighted velocities and weights (the B-spline weights are a partition of
  unity with zero first moment, flip.jl:99-105, 513-537);
- the projection leaves the divergence the solve's residual allows: the FP64 residual of
  the returned pressure, and the divergence of the projected faces, within the solve's
  tolerance plus the FP32 rounding of P and of the faces, bounded cell by cell;
- a few whole steps keep the water's volume (test/runtests.jl:56).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PF = ["PX", "PY", "PZ", "VX", "VY", "VZ", "C0", "C1", "C2", "C3", "C4", "C5", "C6", "C7", "C8", "LAG"]
EPS32 = 2.0 ** -24


@pytest.fixture(scope="module")
def full(mael, orc):
    print("case: dense 256x256x128 initialize", flush=True)
    b = mael.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)
    s = mael.FlipBasin(b, dx=0.15, capacity=6_000_000, cfl=4.0)
    assert tuple(s.grid.n) == (256, 256, 128) and s.grid.levels == 6
    mael.initialize(s)
    assert s.n > 1_000_000
    g = orc.flipgrid(18.0, 12.0, 3.0, 0.15)
    return s, g


def _temporal(lag, dt):
    """flip.jl:503-511 in FP64."""
    r = np.clip(-lag.astype(np.float64) / dt, -0.5, 0.5) - 0.5
    return (35.0 / 16.0) * (1.0 - r * r) ** 3


def test_p2g_order_independent_and_conserving(mael, full):
    s, g = full
    n = s.n
    dt = np.float32(0.06)
    ps = {f: s.get(f) for f in PF}
    rng = np.random.default_rng(7)
    ps["LAG"] = rng.uniform(-0.03, 0.03, n).astype(np.float32)
    s.set_particles(ps, n)
    mael.particles2grid(s, dt)
    keys = ("MOM_U", "MOM_V", "MOM_W", "WT_U", "WT_V", "WT_W", "U", "V", "W", "PHI")
    first = [s.get(k) for k in keys]
    perm = rng.permutation(n)
    s.set_particles({f: ps[f][perm] for f in PF}, n)
    mael.particles2grid(s, dt)
    for k, x in zip(keys, first):
        assert np.array_equal(x.view(np.uint8), s.get(k).view(np.uint8)), k
    # every particle's 27 deposit nodes lie on the grid (the basin is padded), so the sums
    # are the particles' own
    wT = _temporal(ps["LAG"], dt)
    n1, n2, n3 = (int(x) for x in s.grid.n)
    for a, ax in enumerate("UVW"):
        off = [0.5, 0.5, 0.5]
        off[a] = 0.0
        dims = [n1 + (a == 0), n2 + (a == 1), n3 + (a == 2)]
        for c, key in enumerate(("PX", "PY", "PZ")):
            q = (ps[key].astype(np.float64) - g.lo[c]) / g.dx - off[c]
            base = np.floor(q - 0.5)
            assert base.min() >= 0 and base.max() + 2 < dims[c], (ax, key)
        mom = s.get("MOM_" + ax).astype(np.float64).sum() * 2.0 ** -32
        wt = s.get("WT_" + ax).astype(np.float64).sum() * 2.0 ** -32
        want_m = float((wT * ps["V" + "XYZ"[a]].astype(np.float64)).sum())
        want_w = float(wT.sum())
        # each of the 27 n contributions is rounded once to 2^-32
        slack = 27 * n * 2.0 ** -32
        assert abs(wt - want_w) <= slack + 1e-9 * want_w, (ax, wt, want_w)
        scale = float((wT * np.abs(ps["V" + "XYZ"[a]].astype(np.float64))).sum())
        assert abs(mom - want_m) <= slack + 1e-6 * scale, (ax, mom, want_m)


def test_projection_divergence_bounded(mael, orc, full):
    s, g = full
    dt, gamma = 0.06, 0.001
    s.ctx.check(mael.lib().mael_dense_gravity(s.h, dt, 9.81))
    u, v, w = s.get("U"), s.get("V"), s.get("W")
    s.set("P", np.zeros(orc.cells_shape(g), np.float32))
    st = mael.project(s, dt, gamma=gamma, tol=1e-6, maxiter=200, refine=1)
    print("full-size projection: %d iterations, rel residual %.3e" % (st.iterations, st.rel_residual), flush=True)
    assert st.rel_residual <= 1e-6
    types, phi, P = s.get("TYPES"), s.get("PHI"), s.get("P").astype(np.float64)
    b = orc.dense_rhs(g, u, v, w, types, gamma).astype(np.float64)
    bn = np.linalg.norm(b)
    liquid = types == 1
    # the returned P is the FP64 solution rounded to FP32: its residual is the solve's
    # plus A (P64 - P), bounded cell by cell by 2^-24 |A| |P|, and |A| |P| = 2 D |P| - A |P|
    # (A has diagonal D and off-diagonal entries -1)
    r = b - orc.dense_apply(g, types, phi, P)
    ones = orc.dense_apply(g, types, phi, np.ones_like(P))
    nliq = np.zeros_like(P)
    L = liquid.astype(np.float64)
    nliq[:, :, 1:] += L[:, :, :-1]; nliq[:, :, :-1] += L[:, :, 1:]
    nliq[:, 1:, :] += L[:, :-1, :]; nliq[:, :-1, :] += L[:, 1:, :]
    nliq[1:] += L[:-1]; nliq[:-1] += L[1:]
    D = np.where(liquid, ones + nliq, 0.0)
    absAP = 2.0 * D * np.abs(P) - orc.dense_apply(g, types, phi, np.abs(P))
    bound_p = EPS32 * np.linalg.norm(np.where(liquid, absAP, 0.0))
    print("residual of P %.3e, rounding bound %.3e (relative to |b|)" % (np.linalg.norm(r) / bn, bound_p / bn))
    assert np.linalg.norm(r) <= 1.01e-6 * bn + bound_p
    # the faces: D(u') = r up to the rounding of the face update (k_project: a difference,
    # a division by the ghost-fluid fraction and a subtraction, each rounded once)
    U = [s.get(k).astype(np.float64) for k in ("U", "V", "W")]
    b2 = orc.dense_rhs(g, *(x.astype(np.float32) for x in U), types, gamma).astype(np.float64)
    au, av, aw = (np.abs(x) for x in U)
    ou, ov, ow = (np.abs(x.astype(np.float64)) for x in (u, v, w))
    per = (au[:, :, 1:] + au[:, :, :-1] + av[:, 1:, :] + av[:, :-1, :] + aw[1:] + aw[:-1]
           + ou[:, :, 1:] + ou[:, :, :-1] + ov[:, 1:, :] + ov[:, :-1, :] + ow[1:] + ow[:-1])
    bound_f = 4.0 * g.dx * EPS32 * np.linalg.norm(np.where(liquid, per, 0.0))
    print("divergence left %.3e, face rounding bound %.3e (relative to |b|)"
          % (np.linalg.norm(b2) / bn, bound_f / bn))
    assert np.linalg.norm(b2 - r) <= bound_f
    assert np.linalg.norm(b2) <= 1.01e-6 * bn + bound_p + bound_f


def test_steps_keep_volume(mael, full):
    s, g = full
    v0 = s.volume0
    for k in range(3):
        dt = min(s.cfl * s.grid.dx / max(mael.maxspeed(s), 1e-3), 1 / 24)
        st = mael.step(s, dt)
        print("full-size step %d: dt %.4f, %d iterations, rel residual %.2e" % (k, dt, st.iterations,
                                                                               st.rel_residual), flush=True)
        assert st.rel_residual <= 1e-4
    assert np.isfinite(mael.maxspeed(s))
    assert 0.95 < mael.volume(s) / v0 < 1.05
