"""Pins of the oracle's dense support phases (oracle/dense_step.hpp): closed forms the
mathematics fixes, and an independent NumPy restatement where a rule, not a formula, is
restated.

- grid advection (flip.jl:356-369): in a uniform flow the semi-Lagrangian trace of an
  affine field is the field at x - u dt;
- particle advection (flip.jl:438-464): in a uniform, reached flow a particle moves by its
  own time times the flow, its new lag is what it did not advance; where the flow was
  not reached it flies ballistically, x + act v - g act^2 / 2 along z;
- extrapolation (flip.jl:1160-1206): the faces flagged after L layers are a breadth-first
  search L steps out from the valid ones through faces off the walls, a constant field
  stays constant on them, and the rest is zero;
- the stir (basin.jl:105-181): its closed forms at the wall (relaxation to the swirl and
  the waves' orbital velocity), round the core (the guide), the boils where the noise is
  sampled at lattice points (the hash itself), and the eddies' discrete divergence (they
  are the curl of a stream function).
"""
import math

import numpy as np
import pytest

import oracle as O

f32 = np.float32
G = 9.81


@pytest.fixture(scope="module")
def grid():
    return O.flipgrid(18.0, 12.0, 3.0, 0.6)


def _uniform(g, U, V, W):
    return (np.full(O.face_shape(g, 0), U, f32), np.full(O.face_shape(g, 1), V, f32),
            np.full(O.face_shape(g, 2), W, f32))


def test_grid_advection_affine(grid):
    g = grid
    n1, n2, n3 = g.dims
    U, V, W = 1.3, -0.7, 0.4
    u, v, w = _uniform(g, U, V, W)
    k, j, i = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    x, y, z = (g.lo[0] + (i + 0.5) * g.dx, g.lo[1] + (j + 0.5) * g.dx, g.lo[2] + (k + 0.5) * g.dx)
    phi = (0.3 * x - 0.2 * y + 0.5 * z + 1.0).astype(f32)
    dt = 0.9   # more than a cell: several RK2 sub-steps
    out = O.dense_advect_grid(g, u, v, w, phi, (0.5, 0.5, 0.5), dt)
    exp = 0.3 * (x - U * dt) - 0.2 * (y - V * dt) + 0.5 * (z - W * dt) + 1.0
    inner = (i > 3) & (i < n1 - 4) & (j > 3) & (j < n2 - 4) & (k > 3) & (k < n3 - 4)
    assert np.abs(out - exp)[inner].max() < 1e-5 * np.abs(exp).max()
    # a face field (u at its faces) the same way
    k, j, i = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1 + 1), indexing="ij")
    x, y, z = (g.lo[0] + i * g.dx, g.lo[1] + (j + 0.5) * g.dx, g.lo[2] + (k + 0.5) * g.dx)
    fu = (0.1 * x + 0.4 * y - 0.3 * z).astype(f32)
    out = O.dense_advect_grid(g, u, v, w, fu, (0.0, 0.5, 0.5), dt)
    exp = 0.1 * (x - U * dt) + 0.4 * (y - V * dt) - 0.3 * (z - W * dt)
    inner = (i > 3) & (i < n1 - 3) & (j > 3) & (j < n2 - 4) & (k > 3) & (k < n3 - 4)
    assert np.abs(out - exp)[inner].max() < 1e-5 * np.abs(exp).max()


def _interior_particles(g, n, rng):
    r = rng.uniform(0, 10.0, n)
    th = rng.uniform(0, 2 * math.pi, n)
    ps = O.empty_particles(n)
    ps["PX"] = (r * np.cos(th)).astype(f32)
    ps["PY"] = (r * np.sin(th)).astype(f32)
    ps["PZ"] = rng.uniform(-8.0, -2.0, n).astype(f32)
    for c, s in (("VX", 2.0), ("VY", 2.0), ("VZ", 1.0)):
        ps[c] = rng.normal(0, s, n).astype(f32)
    ps["LAG"] = rng.uniform(-0.02, 0.02, n).astype(f32)
    return ps


def _act(ps, dt, h, seed):
    """The particle's own time and new lag (flip.jl:446-450) in the reference's FP32 order."""
    n = len(ps["PX"])
    dt = f32(dt)
    vx, vy, vz = ps["VX"], ps["VY"], ps["VZ"]
    speed = np.sqrt(vx * vx + vy * vy + vz * vz)
    c = np.clip(speed * dt / f32(h), f32(0), f32(1))
    gm = c * c * (f32(3) - f32(2) * c)
    xi = O.hash01(np.arange(1, n + 1, dtype=np.int32), np.full(n, seed, np.int32)) - f32(0.5)
    tau = ps["LAG"]
    act = np.clip(dt + tau + gm * xi * dt, f32(0), f32(2) * dt)
    return act, dt + tau - act


def test_particle_advection_uniform_and_ballistic(grid):
    g = grid
    rng = np.random.default_rng(3)
    ps = _interior_particles(g, 4000, rng)
    dt, seed = 0.3, 11
    act, lag = _act(ps, dt, g.dx, seed)
    U, V, W = 0.8, -1.1, 0.3
    u, v, w = _uniform(g, U, V, W)
    ones = [np.ones(O.face_shape(g, a), np.uint8) for a in range(3)]
    out = O.dense_advect_particles(g, ps, 4000, u, v, w, *ones, dt, G, seed)
    assert np.array_equal(out["LAG"].view(np.uint32), lag.view(np.uint32))
    a64 = act.astype(np.float64)
    for key, vel in (("PX", U), ("PY", V), ("PZ", W)):
        np.testing.assert_allclose(out[key], ps[key].astype(np.float64) + a64 * vel, rtol=0, atol=2e-5)
    # no face reached: ballistic along the particle's own velocity
    zeros = [np.zeros(O.face_shape(g, a), np.uint8) for a in range(3)]
    out = O.dense_advect_particles(g, ps, 4000, u, v, w, *zeros, dt, G, seed)
    for key, vk in (("PX", "VX"), ("PY", "VY")):
        np.testing.assert_allclose(out[key], ps[key] + a64 * ps[vk], rtol=0, atol=2e-5)
    np.testing.assert_allclose(out["PZ"], ps["PZ"] + a64 * ps["VZ"] - 0.5 * G * a64 ** 2, rtol=0, atol=2e-5)


def _walled(g, a):
    """flip.jl:779-784 walled, per face of axis a (NumPy)."""
    n1, n2, n3 = g.dims
    shape = O.face_shape(g, a)
    k, j, i = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
    cx = lambda ii: g.lo[0] + (ii + 0.5) * g.dx
    cy = lambda jj: g.lo[1] + (jj + 0.5) * g.dx
    cz = lambda kk: g.lo[2] + (kk + 0.5) * g.dx
    solid = lambda ii, jj, kk: (cx(ii).astype(f32) ** 2 + cy(jj).astype(f32) ** 2 > f32(g.radius) ** 2) | (
        cz(kk).astype(f32) < f32(g.floor))
    bi, bj, bk = i - (a == 0), j - (a == 1), k - (a == 2)
    edge = ~((bi >= 0) & (bj >= 0) & (bk >= 0) & (i < n1) & (j < n2) & (k < n3))
    return edge | solid(bi, bj, bk) | solid(i, j, k)


def _bfs_flags(valid, walled, layers):
    f = valid.copy()
    for _ in range(layers):
        nb = np.zeros_like(f)
        nb[:, :, 1:] |= f[:, :, :-1]; nb[:, :, :-1] |= f[:, :, 1:]
        nb[:, 1:, :] |= f[:, :-1, :]; nb[:, :-1, :] |= f[:, 1:, :]
        nb[1:] |= f[:-1]; nb[:-1] |= f[1:]
        f = f | (nb & ~walled)
    return f


@pytest.mark.parametrize("a", [0, 1, 2])
def test_extrapolation_reach_and_constant(grid, a):
    g = grid
    rng = np.random.default_rng(20 + a)
    shape = O.face_shape(g, a)
    valid = rng.random(shape) < 0.02
    u = rng.normal(0, 1, shape).astype(f32)
    u[valid] = f32(1.5)
    walled = _walled(g, a)
    for layers in (1, 4, 6):
        out, flags = O.dense_extrapolate(g, a, u, valid.astype(np.uint8), layers)
        want = _bfs_flags(valid, walled, layers)
        assert np.array_equal(flags.astype(bool), want), layers
        reached = want & ~walled
        assert np.all(out[reached] == f32(1.5))
        assert np.all(out[~reached] == 0)


def _acc(params, x, y, z, vel, t=0.0, waves=()):
    return O.stir(params, np.asarray(waves, np.float32), x, y, z, vel, t).astype(np.float64)


def test_stir_closed_forms():
    P = O.stir_params()
    R, core, Gm, relax, guide = 18.0, 2.5, 225.0, 0.3, 4.0
    oseen = lambda r: Gm / (2 * math.pi * r) * (1 - math.exp(-(r / core) ** 2))
    vel = np.array([0.3, -0.2, 0.1], np.float32)
    # at the wall (k = 1): relaxed to the swirl over `relax` seconds, nothing else
    for (x, y, z) in ((18.0, 0.0, -3.0), (12.0, 13.5, -0.5), (-17.0, -6.0, -9.0)):
        r = math.hypot(x, y)
        v = oseen(r)
        want = np.array([-v * y / r - vel[0], v * x / r - vel[1], -vel[2]]) / relax
        np.testing.assert_allclose(_acc(P, x, y, z, vel), want, rtol=1e-5, atol=1e-6)
    # round the core, deep (k = 0, churn = 0, hold = 1): the guide to the swirl alone
    for (x, y, z) in ((1.0, 0.5, -3.0), (-0.3, 1.7, -5.0), (0.0, 0.0, -2.0)):
        r = math.hypot(x, y)
        turn = oseen(r) / r if r > 1e-4 else 0.0
        want = np.array([-turn * y - vel[0], turn * x - vel[1], 0.0]) / guide
        np.testing.assert_allclose(_acc(P, x, y, z, vel), want, rtol=1e-5, atol=1e-6)
    # a wave at the wall: its orbital velocity is the target (basin.jl:139-151)
    wave = (0.6, 0.8, 0.7, 2.6, 0.25, 0.55, 1.1)   # d, k, omega, a, q, phase
    x, y, z, t = 18.5, 2.0, -1.2, 3.0
    th = wave[2] * (wave[0] * x + wave[1] * y) - wave[3] * t + wave[6]
    e = wave[4] * wave[3] * math.exp(wave[2] * z)
    r = math.hypot(x, y)
    v = oseen(r)
    want = np.array([-v * y / r + e * math.sin(th) * wave[0] - vel[0], v * x / r + e * math.sin(th) * wave[1] - vel[1],
                     -e * math.cos(th) - vel[2]]) / relax
    np.testing.assert_allclose(_acc(P, x, y, z, vel, t, [wave]), want, rtol=2e-5, atol=2e-6)


def test_stir_boils_at_lattice_points():
    """With eddyscale 0.6 and t = 0 the boils' noise is sampled at integer points, where
    it is the hash itself (sea.jl:68-76): acc_z = boil (max(h - 0.6, 0) / 0.4 - 0.15)."""
    P = O.stir_params(eddyscale=0.6)
    vel = np.zeros(3, np.float32)
    seen = set()
    for x in range(5, 12):
        for y in range(-11, -4):
            if math.hypot(x, y) >= 14.9:   # inside the band's inner edge (k = 0)
                continue
            acc = _acc(P, float(x), float(y), 0.0, vel)
            h = float(O.hash01(x + 41, y)[0])
            want = 0.5 * (max(h - 0.6, 0.0) / 0.4 - 0.15)
            assert abs(acc[2] - want) < 1e-6, (x, y)
            seen.add(h > 0.6)
    assert seen == {True, False}


def test_stir_eddies_are_divergence_free():
    """In the free layer away from the core and the wall, with turbulence = eddy, the
    horizontal push is the eddies' velocity itself: the curl of the stream function by
    central differences (basin.jl:109-119), whose central-difference divergence with the
    same step vanishes."""
    P = O.stir_params(eddyscale=1.0, turbulence=0.6)
    vel = np.zeros(3, np.float32)
    d = 0.25
    rng = np.random.default_rng(5)
    worst, scale = 0.0, 0.0
    for _ in range(40):
        x, y = rng.uniform(6.0, 9.0), rng.uniform(-9.0, -6.0)
        ex = lambda xx, yy: _acc(P, xx, yy, 0.0, vel)[0]
        ey = lambda xx, yy: _acc(P, xx, yy, 0.0, vel)[1]
        div = (ex(x + d, y) - ex(x - d, y) + ey(x, y + d) - ey(x, y - d)) / (2 * d)
        worst = max(worst, abs(div))
        scale = max(scale, abs(ex(x, y)), abs(ey(x, y)))
    assert scale > 0.05
    assert worst < 1e-4 * scale / d


def _cell_centres(g):
    n1, n2, n3 = g.dims
    k, j, i = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    x = (f32(g.lo[0]) + (i.astype(f32) + f32(0.5)) * f32(g.dx)).astype(f32)
    y = (f32(g.lo[1]) + (j.astype(f32) + f32(0.5)) * f32(g.dx)).astype(f32)
    z = (f32(g.lo[2]) + (k.astype(f32) + f32(0.5)) * f32(g.dx)).astype(f32)
    return i, j, k, x, y, z


def _solid(g, x, y, z):   # flip.jl:64-66
    return (x * x + y * y > f32(g.radius) * f32(g.radius)) | (z < f32(g.floor))


def test_dense_redistance_planar_and_rules(grid):
    """The redistancing sweep (flip.jl:611-679): a signed distance to a tilted plane is a
    fixed point of the Godunov update (its upwind neighbours are exact, and the three-axis
    root reproduces a linear distance), so it is kept wherever it is below the reach
    (iterations + 1) dx and the cell is clear of the walls by more than the sweeps carry;
    above the reach it is clamped to exactly the reach, sign kept. One sweep's solid cells
    take the lowest value of their non-solid neighbours (or their own, clamped), as a NumPy
    restatement of the rule has them."""
    g = grid
    n1, n2, n3 = g.dims
    i, j, k, x, y, z = _cell_centres(g)
    nrm = np.array([0.48, -0.36, 0.8])
    phi = (nrm[0] * x + nrm[1] * y + nrm[2] * z + 3.0).astype(f32)
    it = 4
    h = f32(g.dx)
    reach = f32(it + 1) * h
    out = O.dense_redistance(g, phi, it)
    solid = _solid(g, x, y, z)
    # distance to the walls, the floor and the grid's edges, in cells
    r = np.hypot(x, y)
    m = it + 1
    clear = ((r < g.radius - m * g.dx) & (z > g.floor + m * g.dx) & (i >= m) & (i < n1 - m) & (j >= m) &
             (j < n2 - m) & (k >= m) & (k < n3 - m))
    near = clear & (np.abs(phi) < reach - h)
    assert near.sum() > 2000
    np.testing.assert_allclose(out[near], phi[near], rtol=0, atol=3e-5)
    far = clear & (np.abs(phi) > reach + h)
    assert far.sum() > 2000
    assert np.all(np.abs(out[far]) == reach) and np.all(np.sign(out[far]) == np.sign(phi[far]))
    # one sweep's solid cells, restated
    rng = np.random.default_rng(3)
    phi = rng.uniform(-4, 4, phi.shape).astype(f32)
    one = O.dense_redistance(g, phi, 1)
    reach1 = f32(2) * h
    pad = np.pad(phi, 1, constant_values=np.nan)
    opn = np.pad(~solid, 1, constant_values=False)
    best = np.full(phi.shape, np.inf, f32)
    for dk, dj, di in ((0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)):
        sl = (slice(1 + dk, 1 + dk + n3), slice(1 + dj, 1 + dj + n2), slice(1 + di, 1 + di + n1))
        best = np.where(opn[sl], np.minimum(best, pad[sl]), best)
    exp = np.where(np.isfinite(best), best, np.clip(phi, -reach1, reach1))
    assert solid.sum() > 1000 and np.isfinite(best[solid]).sum() > 500
    np.testing.assert_array_equal(one[solid], exp[solid])
