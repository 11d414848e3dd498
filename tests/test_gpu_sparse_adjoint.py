"""Adjoints of the sparse hot path (no counterpart in the reference, which has no AD).

With the particles' positions and lags, the level set, the topology, the cell types, the
deposit weights and the valid-face flags held at the state the forward phase left, P2G
is linear in the particles' v and APIC rows and in u0 (fine pools and coarse grid), the
projection is affine in the faces (linear at gamma = 0), and G2P is linear in the fine
faces; the adjoint entry points are their transposes over every stored value, both
copies of a face shared by two fine blocks included. Checks:

- dot-product identities <L x, y> = <x, L^T y> for G2P, P2G (in the particles and in u0)
  and the projection, on random x and y over the slots in use (relative 1e-5: FP32
  arithmetic, 2^-32 fixed-point deposits, projections solved to 1e-9 with FP64
  refinement);
- the gradient of a scalar loss through P2G -> projection -> G2P, chained through the
  three adjoints, against central differences of the GPU forward chain in the APIC rows
  and in u0, where the chain is affine (relative 1e-4).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VN = ["VX", "VY", "VZ"]
CN = ["C%d" % k for k in range(9)]
AX = ["U", "V", "W"]
BFACES = 4352
TIGHT = dict(tol=1e-9, maxiter=600, refine=10)


@pytest.fixture(scope="module")
def M():
    import maelstrom_amd
    return maelstrom_amd


def _basin(M):
    """The sparse test basin at 0.6 m (a few dozen fine blocks round the surface over a
    coarse grid), initialized on the GPU."""
    b = M.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)
    s = M.SparseBasin(b, dx=0.6, blocks=64, capacity=300_000)
    M.sparse.initialize(s)
    return s


def _live(s):
    """1 on the fine face copies of slots in use, 0 on free slots."""
    where = s.get("WHERE").reshape(-1, 4)[:s.nslots]
    return np.repeat((where[:, 0] >= 0).astype(np.float32), BFACES)


def _dot(a, b):
    return float(np.dot(np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()))


def _rel(a, b):
    """Relative difference of two inner products; neither may vanish (a vanishing pair
    would pass any bar)."""
    print("  %.9g vs %.9g" % (a, b), flush=True)
    assert a != 0.0 and b != 0.0, (a, b)
    return abs(a - b) / max(abs(a), abs(b))


def _faces(s, rng, prefix=""):
    """Random face values over the fine copies in use and the coarse grid, by upload name."""
    live = _live(s)
    out = {}
    for ax in AX:
        out[prefix + ax + "_F"] = (rng.normal(0, 1, live.size) * live).astype(np.float32)
        out[prefix + ax + "_C"] = rng.normal(0, 1, s.shape(prefix + ax + "_C")).astype(np.float32)
    return out


def test_sparse_g2p_adjoint_dot(M):
    print("case: sparse g2p adjoint", flush=True)
    s = _basin(M)
    rng = np.random.default_rng(1)
    n = s.n
    assert n > 1000
    u = _faces(s, rng)
    for k, v in u.items():
        s.set(k, v)
    fw = s.get("FW_F")
    fw[rng.random(fw.shape) < 0.01] = 0   # some particles miss the grid
    s.set("FW_F", fw)
    for f in VN + CN:
        s.set(f, np.zeros(n, np.float32))
    M.sparse.grid2particles(s, 0.05)   # gravity acts only where the grid is missed
    y = {f: s.get(f) for f in VN + CN}
    ybar = {f: rng.normal(0, 1, n).astype(np.float32) for f in VN + CN}
    for f in VN + CN:
        s.set("PB_" + f, ybar[f])
    M.sparse.grid2particles_adjoint(s, 0.05)
    lhs = sum(_dot(y[f], ybar[f]) for f in VN + CN)
    miss = s.get("PB_VX") != 0   # the gravity constant of the missed particles is no part of the map
    lhs -= _dot(y["VZ"][miss], ybar["VZ"][miss])
    rhs = sum(_dot(u[ax + "_F"], s.get("UB_" + ax + "_F")) for ax in AX)
    assert miss.any() and (~miss).sum() > n // 10
    assert _rel(lhs, rhs) < 1e-5, (lhs, rhs)
    for ax in AX:   # G2P reads no coarse face
        assert not s.get("UB_" + ax + "_C").any()
    np.testing.assert_array_equal(s.get("PB_VX")[miss], ybar["VX"][miss])
    assert not s.get("PB_VX")[~miss].any() and not s.get("PB_C4").any()


def test_sparse_p2g_adjoint_dot(M):
    print("case: sparse p2g adjoint", flush=True)
    s = _basin(M)
    rng = np.random.default_rng(2)
    n = s.n
    zero0 = {k: np.zeros_like(v) for k, v in _faces(s, rng, "").items()}
    x = {f: rng.normal(0, 1, n).astype(np.float32) for f in VN + CN}
    for f in VN + CN:
        s.set(f, x[f])
    for k, v in zero0.items():
        s.set(k.replace("_", "0_", 1), v)
    M.sparse.particles2grid(s, 0.05)
    u = {k: s.get(k) for k in zero0}
    ubar = _faces(s, rng)
    for k, v in ubar.items():
        s.set("UB_" + k, v)
    for f in VN + CN:   # the adjoint adds into PB
        s.set("PB_" + f, np.zeros(n, np.float32))
    M.sparse.particles2grid_adjoint(s, 0.05)
    lhs = sum(_dot(u[k], ubar[k]) for k in u)
    rhs = sum(_dot(x[f], s.get("PB_" + f)) for f in VN + CN)
    assert _rel(lhs, rhs) < 1e-5, (lhs, rhs)
    # and in u0 (fine copies and coarse faces), with the particles' v and rows at zero
    u0 = _faces(s, rng)
    for f in VN + CN:
        s.set(f, np.zeros(n, np.float32))
    for k, v in u0.items():
        s.set(k.replace("_", "0_", 1), v)
    M.sparse.particles2grid(s, 0.05)
    u = {k: s.get(k) for k in u0}
    M.sparse.particles2grid_adjoint(s, 0.05)
    lhs = sum(_dot(u[k], ubar[k]) for k in u)
    rhs = sum(_dot(u0[k], s.get("UB0_" + k)) for k in u0)
    assert _rel(lhs, rhs) < 1e-5, (lhs, rhs)


def test_sparse_project_adjoint_dot(M):
    print("case: sparse project adjoint", flush=True)
    s = _basin(M)
    rng = np.random.default_rng(3)
    u = _faces(s, rng)
    for k, v in u.items():
        s.set(k, v)
    st = M.sparse.project(s, 0.05, gamma=0.0, **TIGHT)
    assert st.rel_residual <= 1e-9, st.rel_residual
    un = {k: s.get(k) for k in u}
    ubar = _faces(s, rng)
    for k, v in ubar.items():
        s.set("UB_" + k, v)
    sa = M.sparse.project_adjoint(s, **TIGHT)
    assert sa.rel_residual <= 1e-9, sa.rel_residual
    lhs = sum(_dot(un[k], ubar[k]) for k in u)
    rhs = sum(_dot(u[k], s.get("UB_" + k)) for k in u)
    assert _rel(lhs, rhs) < 1e-5, (lhs, rhs)


def test_sparse_chain_gradient_finite_differences(M):
    """d/d(rows) and d/du0 of L = sum_p w_p . v_p after P2G -> projection -> G2P: the
    loss is affine in both at fixed geometry, so unit central differences are exact up
    to rounding."""
    print("case: sparse adjoint chain vs finite differences", flush=True)
    s = _basin(M)
    rng = np.random.default_rng(4)
    n = s.n
    dt = 0.05
    base = {f: s.get(f) for f in VN + CN}
    base0 = {k.replace("_", "0_", 1): s.get(k.replace("_", "0_", 1)) for k in _faces(s, rng)}
    wv = {f: rng.normal(0, 1, n).astype(np.float32) for f in VN}

    def loss(parts, grid0):
        for f in VN + CN:
            s.set(f, parts[f])
        for k in grid0:
            s.set(k, grid0[k])
        M.sparse.particles2grid(s, dt)
        M.sparse.project(s, dt, gamma=0.0, **TIGHT)
        M.sparse.grid2particles(s, dt)
        return sum(_dot(wv[f], s.get(f)) for f in VN)

    loss(base, base0)   # the forward state the adjoints transpose
    for f in VN:
        s.set("PB_" + f, wv[f])
    for f in CN:
        s.set("PB_" + f, np.zeros(n, np.float32))
    M.sparse.grid2particles_adjoint(s, dt)
    M.sparse.project_adjoint(s, **TIGHT)
    M.sparse.particles2grid_adjoint(s, dt)
    gv = {f: s.get("PB_" + f) for f in VN + CN}
    g0 = {k: s.get("UB0_" + k.replace("0_", "_", 1)) for k in base0}

    def fd(dp, d0, h):
        plus = loss({f: base[f] + h * dp[f] for f in base}, {k: base0[k] + h * d0[k] for k in base0})
        minus = loss({f: base[f] - h * dp[f] for f in base}, {k: base0[k] - h * d0[k] for k in base0})
        return (plus - minus) / (2 * h)

    live = _live(s)
    zero_p = {f: np.zeros(n, np.float32) for f in base}
    zero_0 = {k: np.zeros_like(base0[k]) for k in base0}
    dp = {f: (rng.normal(0, 1, n) if f in CN else np.zeros(n)).astype(np.float32) for f in base}
    want = sum(_dot(gv[f], dp[f]) for f in base)
    assert _rel(fd(dp, zero_0, 1.0), want) < 1e-4
    d0 = {k: (rng.normal(0, 1, base0[k].shape) * (live if k.endswith("_F") else 1)).astype(np.float32) for k in base0}
    want = sum(_dot(g0[k], d0[k]) for k in base0)
    assert _rel(fd(zero_p, d0, 1.0), want) < 1e-4
