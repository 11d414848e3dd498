"""Pins of the oracle's whitewater (oracle/whitewater.hpp; flip.jl:1411-1511,
whitewater.jl:131-240):

- emission: against an independent NumPy restatement that finds each fluid particle's
  neighbours by brute force over all particles (not through the cell runs of the resampled
  order), on a few hundred particles under a planar surface: the same number of diffuse
  particles from each fluid particle (up to draws within rounding of their threshold), at
  the positions the reference's jitter puts them, in particle order in the ring buffer;
- the diffuse step's closed forms in a uniform flow under a planar surface: spray falls
  ballistically, foam takes the flow's velocity, a bubble rises and is dragged towards
  the flow, and a particle that leaves the basin dies.
"""
import math

import numpy as np

import oracle as O

f32 = np.float32


def _setup(n=400, seed=1, z0=-1.0):
    g = O.flipgrid(18.0, 12.0, 3.0, 0.6)
    rng = np.random.default_rng(seed)
    ps = O.empty_particles(n)
    ps["PX"] = rng.uniform(2.0, 4.0, n).astype(f32)
    ps["PY"] = rng.uniform(-1.0, 1.0, n).astype(f32)
    ps["PZ"] = rng.uniform(z0 - 0.8, z0, n).astype(f32)
    for k, s in (("VX", 6.0), ("VY", 6.0), ("VZ", 3.0)):
        ps[k] = rng.normal(0, s, n).astype(f32)
    # the resampled order: by cell (x fastest), stable
    cell = [np.clip(np.floor((ps[k].astype(f32) - f32(g.lo[a])) / f32(g.dx)).astype(int), 0, g.dims[a] - 1)
            for a, k in enumerate(("PX", "PY", "PZ"))]
    key = cell[0] + g.dims[0] * (cell[1] + g.dims[1] * cell[2])
    order = np.argsort(key, kind="stable")
    ps = {k: v[order].copy() for k, v in ps.items()}
    key = key[order]
    totals = np.bincount(key, minlength=int(np.prod(g.dims))).astype(np.int32)
    offsets = np.cumsum(totals).astype(np.int32)
    n1, n2, n3 = g.dims
    zc = g.lo[2] + (np.arange(n3) + 0.5) * g.dx
    phi = np.broadcast_to((zc - z0)[:, None, None], (n3, n2, n1)).astype(f32).copy()
    return g, ps, offsets, totals, phi


def _hash(i, j):
    return O.hash01(np.atleast_1d(np.int32(i)), np.atleast_1d(np.int32(j)))[0]


def _brute(g, ps, n, rules, dt, frames, near=1.5):
    """The emission counts by brute force (whitewater.jl:131-175), the normal (0, 0, 1) of
    the planar level set within `near` cells of it. Returns (whole, margin) per particle:
    `margin` how close the draw was to deciding the fraction the other way."""
    h = f32(g.dx)
    X = np.stack([ps[k][:n] for k in ("PX", "PY", "PZ")], 1).astype(np.float64)
    V = np.stack([ps[k][:n] for k in ("VX", "VY", "VZ")], 1).astype(np.float64)
    air_r, energy_r, kair = rules[0:2], rules[4:6], rules[6]
    whole = np.zeros(n, int)
    margin = np.full(n, np.inf)
    for i in range(n):
        x, v = X[i], V[i]
        if x[2] < rules[15] + 2 * h:
            continue
        r = math.hypot(x[0], x[1])
        G, core = rules[18], rules[19]
        turn = G / (2 * math.pi * r) * (1 - math.exp(-(r / core) ** 2)) / r
        vr = v - np.array([-turn * x[1], turn * x[0], 0.0])
        rs = np.linalg.norm(vr)
        cl = lambda I, rr: (min(I, rr[1]) - min(I, rr[0])) / (rr[1] - rr[0])
        energy = cl(0.5 * rs * rs, energy_r)
        draw = float(_hash(i + 1, frames))
        d = np.linalg.norm(X - x, axis=1)
        nb = (d > 0) & (d < h)
        air = 0.0
        for q in np.flatnonzero(nb):
            w = 1 - d[q] / h
            vij = v - V[q]
            lv = np.linalg.norm(vij)
            if lv > 0:
                air += lv * (1 - (vij / lv) @ ((x - X[q]) / d[q])) * w
        tot = energy * kair * cl(air, air_r) * dt   # the planar normals make no crest
        whole[i] = int(math.floor(tot)) + (1 if draw < tot - math.floor(tot) else 0)
        margin[i] = abs(draw - (tot - math.floor(tot)))
    return whole, margin


def test_emission_matches_brute_force():
    g, ps, offsets, totals, phi = _setup()
    n = len(ps["PX"])
    rules = O.white_rules(kair=400.0)
    cap = 20000
    white = O.empty_white(cap)
    u, v, w = (np.zeros(O.face_shape(g, a), f32) for a in range(3))
    dt = 1 / 24
    # the emission alone (no diffuse sub-step after it)
    head, frames = O.dense_whitewater(g, ps, n, offsets, totals, phi, u, v, w, rules, white, 0, 0, dt, substeps=0,
                                      near=1.5)
    assert frames == 1
    whole, margin = _brute(g, ps, n, rules, f32(dt), 1)
    sure = margin > 1e-4
    assert whole.sum() > 50 and sure.mean() > 0.95
    # the positions: particle i's m-th at slot head0 + its offset + m, x_i + jitter
    got = np.stack([white[k] for k in ("X", "Y", "Z")], 1)
    slot, made = 0, 0
    jh = f32(0.5) * f32(g.dx)
    counts = []
    for i in range(n):
        # the number this particle made is the run of slots whose position is within the
        # jitter box of x_i (the boxes of different particles may overlap, so use the
        # brute-force count where it is sure)
        k = whole[i]
        if not sure[i]:
            return   # the rest of the ring buffer's layout depends on this uncertain count
        for m in range(1, k + 1):
            slot += 1
            a = _hash(np.int32(i + 1 + m * 7919), 1 + 31)
            bb = _hash(np.int32(i + 1 - m * 104729), 1 + 77)
            cc = _hash(m, i + 1 + 1)
            jit = np.array([(a - f32(0.5)) * jh, (bb - f32(0.5)) * jh, (cc - f32(0.5)) * jh], f32)
            want = np.array([ps["PX"][i], ps["PY"][i], ps["PZ"][i]], f32) + jit
            np.testing.assert_array_equal(got[slot % cap], want)
            assert white["LIFE"][slot % cap] == f32(2) + f32(6) * _hash(1 - m, i + 1)
            np.testing.assert_array_equal([white[k][slot % cap] for k in ("VX", "VY", "VZ")],
                                          [ps[k][i] for k in ("VX", "VY", "VZ")])
        made += k
    assert head == made


def test_diffuse_closed_forms():
    g, ps, offsets, totals, phi = _setup(n=1, z0=-1.0)
    dx = f32(g.dx)
    rules = O.white_rules(kair=0.0, kcrest=0.0, buoyancy=2.0, drag=0.5, stream=0)
    U = (f32(1.0), f32(-0.5), f32(0.25))
    u, v, w = (np.full(O.face_shape(g, a), U[a], f32) for a in range(3))
    white = O.empty_white(4)
    # spray well above the surface, foam on it, a bubble well below, one below the floor
    pos = [(3.0, 0.0, 0.0), (3.0, 0.0, -1.0), (3.0, 0.0, -4.0), (3.0, 0.0, -11.95)]
    vel = [(0.5, 0.2, 1.0), (0.0, 0.0, 0.0), (0.3, 0.1, -0.2), (0.0, 0.0, -50.0)]
    for k, (p, q) in enumerate(zip(pos, vel)):
        for a, key in enumerate(("X", "Y", "Z")):
            white[key][k] = p[a]
            white["V" + key][k] = q[a]
        white["LIFE"][k] = 5.0
    dt = f32(0.1)
    O.dense_whitewater(g, ps, 0, offsets, totals, phi, u, v, w, rules, white, 0, 0, float(dt), substeps=1)
    G = f32(9.81)
    # spray: v_z - g dt, then x + dt v
    vs = np.array([0.5, 0.2, f32(1.0) + (-G * dt)], f32)
    np.testing.assert_array_equal([white["VX"][0], white["VY"][0], white["VZ"][0]], vs)
    np.testing.assert_array_equal(white["Z"][0], f32(0.0) + dt * vs[2])
    assert white["KIND"][0] == 1 and white["LIFE"][0] == f32(5.0)
    # foam: the flow's velocity, aged
    np.testing.assert_allclose([white["VX"][1], white["VY"][1], white["VZ"][1]], U, rtol=1e-6)
    assert white["KIND"][1] == 2 and white["LIFE"][1] == f32(5.0) - dt
    # bubble: buoyancy and drag towards the flow
    vb = [f32(0.3) + f32(0.5) * (U[0] - f32(0.3)), f32(0.1) + f32(0.5) * (U[1] - f32(0.1)),
          (f32(-0.2) + dt * (f32(2.0) * G)) + f32(0.5) * (U[2] - f32(-0.2))]
    np.testing.assert_allclose([white["VX"][2], white["VY"][2], white["VZ"][2]], vb, rtol=1e-6)
    assert white["KIND"][2] == 3
    # below the floor: dead
    assert white["LIFE"][3] == 0
