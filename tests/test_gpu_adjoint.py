"""Adjoints of the dense hot path (no counterpart in the reference, which has no AD).

With the particles' positions and lags, the level set, the cell types and the valid-face
flags held at the state the forward phase left, P2G is linear in the particles' v and
APIC rows (and in u0), the projection is affine in u, and G2P is linear in u; the
adjoint entry points are their transposes. Checks:

- dot-product identities <L x, y> = <x, L^T y> for G2P, P2G and the projection, on
  random x and y (relative 1e-5: FP32 arithmetic, 2^-32 fixed-point deposits, a
  projection solved to 1e-10);
- the gradient of a scalar loss through P2G -> projection -> G2P, chained through the
  three adjoints, against central finite differences of the GPU forward chain, with
  respect to the particles' APIC rows, to u0, to their velocities, and to the vortex
  circulation Gamma (the particles' and the grid's velocities are proportional to Gamma
  at fixed geometry): relative 1e-4. The splat moves each particle by its own velocity,
  so P2G's level set depends on v; the linearization holds it fixed, and so does the
  differenced chain (the level set of the base state is restored before the
  projection). The loss is then affine in the rows and in u0 (exact differences, unit
  step); in v, the small step keeps P2G's band test away from its edges.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VN = ["VX", "VY", "VZ"]
CN = ["C%d" % k for k in range(9)]
FACES = ["U", "V", "W"]
TIGHT = dict(tol=1e-10, maxiter=400, refine=1, warm=0)


def _basin(mael):
    """The reference's dense test basin (test/runtests.jl:49) at 0.2 m, initialized."""
    b = mael.narutobasin(radius=4.0, depth=2.0, Γ=15.0, core=1.0, band=1.0, headroom=1.5)
    s = mael.FlipBasin(b, dx=0.2, capacity=400_000)
    mael.initialize(s)
    mael.project(s, 0.05, gamma=0.0, **TIGHT)   # the valid-face flags of a projection
    return s


def _dot(a, b):
    return float(np.dot(np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()))


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_g2p_adjoint_dot(mael):
    print("case: dense g2p adjoint", flush=True)
    s = _basin(mael)
    rng = np.random.default_rng(1)
    n = s.n
    assert n > 1000
    u = {f: rng.normal(0, 1, s.shape(f)).astype(np.float32) for f in FACES}
    for f in FACES:
        s.set(f, u[f])
    fw = s.get("FW")
    fw[rng.random(fw.shape) < 0.005] = 0   # some particles miss the grid
    s.set("FW", fw)
    for f in VN + CN:
        s.set(f, np.zeros(n, np.float32))
    mael.grid2particles(s, 0.05)   # gravity acts only where the grid is missed
    y = {f: s.get(f) for f in VN + CN}
    ybar = {f: rng.normal(0, 1, n).astype(np.float32) for f in VN + CN}
    for f in VN + CN:
        s.set("PB_" + f, ybar[f])
    mael.grid2particles_adjoint(s, 0.05)
    lhs = sum(_dot(y[f], ybar[f]) for f in VN + CN)
    # the gravity constant of the missed particles is not part of the linear map
    miss = s.get("PB_VX") != 0
    lhs -= _dot(y["VZ"][miss], ybar["VZ"][miss])
    rhs = sum(_dot(u[f], s.get("UB_" + f)) for f in FACES)
    assert miss.any() and (~miss).sum() > n // 10
    assert _rel(lhs, rhs) < 1e-5, (lhs, rhs)
    # the input cotangent: v passes through where the grid was missed, rows never do
    assert np.array_equal(s.get("PB_VX")[miss], ybar["VX"][miss])
    assert not s.get("PB_VX")[~miss].any() and not s.get("PB_C4").any()


def test_p2g_adjoint_dot(mael):
    print("case: dense p2g adjoint", flush=True)
    s = _basin(mael)
    rng = np.random.default_rng(2)
    n = s.n
    x = {f: rng.normal(0, 1, n).astype(np.float32) for f in VN + CN}
    for f in VN + CN:
        s.set(f, x[f])
    for f in ("U0", "V0", "W0"):
        s.set(f, np.zeros(s.shape(f), np.float32))
    mael.particles2grid(s, 0.05)
    u = {f: s.get(f) for f in FACES}
    ubar = {f: rng.normal(0, 1, s.shape(f)).astype(np.float32) for f in FACES}
    for f in FACES:
        s.set("UB_" + f, ubar[f])
    for f in VN + CN:   # the adjoint adds into PB
        s.set("PB_" + f, np.zeros(n, np.float32))
    mael.particles2grid_adjoint(s, 0.05)
    lhs = sum(_dot(u[f], ubar[f]) for f in FACES)
    rhs = sum(_dot(x[f], s.get("PB_" + f)) for f in VN + CN)
    assert _rel(lhs, rhs) < 1e-5, (lhs, rhs)
    # and in u0, with the particles' velocities at zero
    u0 = {f: rng.normal(0, 1, s.shape(f)).astype(np.float32) for f in ("U0", "V0", "W0")}
    for f in VN + CN:
        s.set(f, np.zeros(n, np.float32))
    for f in u0:
        s.set(f, u0[f])
    mael.particles2grid(s, 0.05)
    u = {f: s.get(f) for f in FACES}
    mael.particles2grid_adjoint(s, 0.05)
    lhs = sum(_dot(u[f], ubar[f]) for f in FACES)
    rhs = sum(_dot(u0[f + "0"], s.get("UB0_" + f)) for f in FACES)
    assert _rel(lhs, rhs) < 1e-5, (lhs, rhs)


def test_project_adjoint_dot(mael):
    print("case: dense project adjoint", flush=True)
    s = _basin(mael)
    rng = np.random.default_rng(3)
    u = {f: rng.normal(0, 1, s.shape(f)).astype(np.float32) for f in FACES}
    for f in FACES:
        s.set(f, u[f])
    st = mael.project(s, 0.05, gamma=0.0, **TIGHT)
    assert st.rel_residual <= 1e-10
    un = {f: s.get(f) for f in FACES}
    ubar = {f: rng.normal(0, 1, s.shape(f)).astype(np.float32) for f in FACES}
    for f in FACES:
        s.set("UB_" + f, ubar[f])
    sa = mael.project_adjoint(s, **TIGHT)
    assert sa.rel_residual <= 1e-10
    lhs = sum(_dot(un[f], ubar[f]) for f in FACES)
    rhs = sum(_dot(u[f], s.get("UB_" + f)) for f in FACES)
    assert _rel(lhs, rhs) < 1e-5, (lhs, rhs)


def test_chain_gradient_finite_differences(mael):
    """d/dv, d/du0 and d/dGamma of L = sum_p w_p . v_p after P2G -> projection -> G2P."""
    print("case: dense adjoint chain vs finite differences", flush=True)
    s = _basin(mael)
    rng = np.random.default_rng(4)
    n = s.n
    dt = 0.05
    base = {f: s.get(f) for f in VN + CN}
    base0 = {f: s.get(f) for f in ("U0", "V0", "W0")}
    wv = {f: rng.normal(0, 1, n).astype(np.float32) for f in VN}

    mael.particles2grid(s, dt)
    phi = s.get("PHI")   # the base state's level set, held fixed

    def loss(parts, grid0):
        for f in VN + CN:
            s.set(f, parts[f])
        for f in grid0:
            s.set(f, grid0[f])
        mael.particles2grid(s, dt)
        s.set("PHI", phi)
        mael.project(s, dt, gamma=0.0, **TIGHT)
        mael.grid2particles(s, dt)
        return sum(_dot(wv[f], s.get(f)) for f in VN)

    loss(base, base0)   # the forward state the adjoints transpose
    for f in VN:
        s.set("PB_" + f, wv[f])
    for f in CN:
        s.set("PB_" + f, np.zeros(n, np.float32))
    mael.grid2particles_adjoint(s, dt)
    mael.project_adjoint(s, **TIGHT)
    mael.particles2grid_adjoint(s, dt)
    gv = {f: s.get("PB_" + f) for f in VN + CN}
    g0 = {f: s.get("UB0_" + f[0]) for f in ("U0", "V0", "W0")}

    def fd(dp, d0, h):
        plus = loss({f: base[f] + h * dp[f] for f in base}, {f: base0[f] + h * d0[f] for f in base0})
        minus = loss({f: base[f] - h * dp[f] for f in base}, {f: base0[f] - h * d0[f] for f in base0})
        return (plus - minus) / (2 * h)

    zero_p = {f: np.zeros(n, np.float32) for f in base}
    zero_0 = {f: np.zeros_like(base0[f]) for f in base0}
    # the particles' APIC rows
    dp = {f: (rng.normal(0, 1, n) if f in CN else np.zeros(n)).astype(np.float32) for f in base}
    want = sum(_dot(gv[f], dp[f]) for f in base)
    assert _rel(fd(dp, zero_0, 1.0), want) < 1e-4
    # the particles' velocities
    dp = {f: (rng.normal(0, 1, n) if f in VN else np.zeros(n)).astype(np.float32) for f in base}
    want = sum(_dot(gv[f], dp[f]) for f in base)
    assert _rel(fd(dp, zero_0, 1e-2), want) < 1e-4
    # u0
    d0 = {f: rng.normal(0, 1, base0[f].shape).astype(np.float32) for f in base0}
    want = sum(_dot(g0[f], d0[f]) for f in base0)
    assert _rel(fd(zero_p, d0, 1.0), want) < 1e-4
    # Gamma: at fixed geometry the velocities scale with it, dv/dGamma = v / Gamma
    G = s.basin.Γ
    dpg = {f: (base[f] / G).astype(np.float32) for f in base}
    d0g = {f: (base0[f] / G).astype(np.float32) for f in base0}
    want = sum(_dot(gv[f], dpg[f]) for f in base) + sum(_dot(g0[f], d0g[f]) for f in base0)
    # a relative step of 1e-2: the chain is FP32, and at 1e-3 the difference of two losses
    # of order 100 is 0.2, within FP32's rounding noise of 1e-4 relative
    assert _rel(fd(dpg, d0g, 1e-2 * G), want) < 1e-4
