"""CPU tests that pin the dense oracle: known answers, closed forms, invariants, an
independent numpy restatement of the integer pipeline, and scipy's direct solver.

Each check is chosen so a plausible slip (a dropped term, a wrong sign or index, an
off-by-one in the 1-based to 0-based translation) fails it.
"""
import numpy as np
import pytest

# ── shapes the reference's sizing gives (known answers) ──────────────────────


@pytest.mark.parametrize("radius,depth,top,dx,dims,levels", [
    (18.0, 12.0, 3.0, 0.6, (64, 64, 32), 4),    # the 64x64x32 config (Basin defaults, rimwall 3)
    (18.0, 12.0, 1.5, 0.6, (64, 64, 24), 4),    # the same with narutobasin's rimwall 1.5
    (18.0, 12.0, 3.0, 0.3, (128, 128, 64), 5),
    (18.0, 12.0, 3.0, 0.15, (256, 256, 128), 6),
    (4.0, 2.0, 1.5, 0.2, (44, 44, 20), 3),      # the reference's dense test basin
])
def test_flipgrid_sizes(orc, radius, depth, top, dx, dims, levels):
    g = orc.flipgrid(radius, depth, top, dx)
    assert g.dims == dims and g.levels == levels
    # lo = (-nx dx / 2, -nx dx / 2, -depth - dx)
    assert g.lo[0] == np.float32(-dims[0]) * np.float32(dx) / np.float32(2)
    assert g.lo[2] == np.float32(-depth) - np.float32(dx)


def test_flipgrid_matches_library(orc):
    """The C ABI's sizing (host code, no GPU needed) agrees with the oracle's."""
    import maelstrom_amd as m
    from maelstrom_amd import Basin
    for dx in (0.6, 0.3, 0.15, 0.2, 0.45):
        b = Basin(radius=18.0, depth=12.0)
        gm = m.flipgrid(b, dx)
        go = orc.flipgrid(18.0, 12.0, 3.0, dx)
        assert tuple(gm.n) == go.dims and gm.levels == go.levels and tuple(gm.lo) == tuple(go.lo)


# ── hash01 (sea.jl:61-66) ────────────────────────────────────────────────────

def _hash01_numpy(i, j):
    i = np.asarray(i, np.int64).astype(np.uint32)
    j = np.asarray(j, np.int64).astype(np.uint32)
    with np.errstate(over="ignore"):
        k = (i * np.uint32(0x27d4eb2d)) ^ (j * np.uint32(0x165667b1))
        k = (k ^ (k >> np.uint32(15))) * np.uint32(0x85ebca6b)
        k = (k ^ (k >> np.uint32(13))) * np.uint32(0xc2b2ae35)
    k = k ^ (k >> np.uint32(16))
    return k.astype(np.float32) / np.float32(4.2949673e9)


def test_hash01_two_restatements_agree(orc):
    rng = np.random.default_rng(1)
    i = rng.integers(-2**31, 2**31, 200000, dtype=np.int64).astype(np.int32)
    j = rng.integers(-2**31, 2**31, 200000, dtype=np.int64).astype(np.int32)
    assert np.array_equal(orc.hash01(i, j), _hash01_numpy(i, j))


def test_hash01_range_closed_at_one():
    # 4.2949673f9 is 2^32 in Float32, and Float32(k) rounds the top 128 k up to 2^32.
    assert np.float32(4.2949673e9) == np.float32(2.0**32)
    k = np.arange(0xFFFFFF00, 0x100000000, dtype=np.uint64).astype(np.uint32)
    v = k.astype(np.float32) / np.float32(2.0**32)
    assert np.all(v[k >= 0xFFFFFF80] == 1.0) and np.all(v[k < 0xFFFFFF80] < 1.0)
    assert int(np.sum(v == 1.0)) == 128


# ── B-spline, temporal weight, ahead, order keys ─────────────────────────────

@pytest.mark.parametrize("q", [0.5, 0.73, 1.2, 3.99, 7.0, 10.25])
def test_bspline_moments(orc, q):
    b, w = orc.bspline(q)
    d = np.arange(b, b + 3, dtype=np.float64) - q
    assert abs(w.sum() - 1) < 1e-6
    assert abs((w * d).sum()) < 1e-6
    assert abs((w * d * d).sum() - 0.25) < 1e-6


def test_temporal_mean_is_one(orc):
    dt = 0.05
    lags = np.linspace(-dt / 2, dt / 2, 20001)
    vals = np.array([orc.lib().orc_temporal(l, dt) for l in lags])
    assert abs(np.trapezoid(vals, lags) / dt - 1) < 1e-4
    # r = clamp(-lag/dt, ±1/2) - 1/2: the peak 35/16 at lag = -dt/2, (35/16)(3/4)^3 at lag = 0
    assert orc.lib().orc_temporal(-dt / 2, dt) == np.float32(35 / 16)
    assert orc.lib().orc_temporal(0.0, dt) == np.float32(35 / 16 * 0.421875)


def test_ahead_closed_forms(orc):
    assert abs(orc.lib().orc_ahead(1.0) - 0.2265625) < 1e-7
    for gm in (1e-2, 3e-2):
        assert abs(orc.lib().orc_ahead(gm) / (gm * gm / 3) - 1) < 1e-3
    # strength: the Courant number eased to 1 (flip.jl:466-470)
    assert orc.lib().orc_strength(0.0, 0.1, 0.2) == 0.0
    assert orc.lib().orc_strength(10.0, 0.1, 0.2) == 1.0


def test_orderkey_is_monotone_bijection(orc):
    rng = np.random.default_rng(2)
    f = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000),
                        np.array([0.0, -0.0, 1e30, -1e30, np.inf, -np.inf, 1e-45, -1e-45], np.float32)]).astype(np.float32)
    keys = np.array([orc.lib().orc_orderkey(float(x)) for x in f])
    back = np.array([orc.lib().orc_orderfloat(int(k)) for k in keys], np.float32)
    assert np.array_equal(back.view(np.uint32), f.view(np.uint32))
    o = np.argsort(f, kind="stable")
    fs, ks = f[o], keys[o]
    strict = fs[1:] > fs[:-1]
    assert np.all(ks[1:][strict] > ks[:-1][strict])


def test_theta_clamps(orc):
    assert orc.lib().orc_theta64(-1.0, 1.0) == 0.5
    assert orc.lib().orc_theta64(-1e-6, 1.0) == pytest.approx(0.01)
    assert orc.lib().orc_theta64(-1.0, 1e-9) == pytest.approx(1.0)


# ── trilinear and the APIC gather reproduce affine fields ────────────────────

def test_trilinear_affine_and_clamped(orc):
    n1, n2, n3 = 7, 6, 5
    k, j, i = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    A = (0.5 + 0.25 * i - 0.5 * j + 0.125 * k).astype(np.float32)
    for q in [(1.3, 2.6, 3.2), (0.0, 0.0, 0.0), (5.99, 4.5, 3.9)]:
        assert orc.trilinear(A, q) == pytest.approx(0.5 + 0.25 * q[0] - 0.5 * q[1] + 0.125 * q[2], abs=1e-5)
    # clamped outside: the value at the nearest face of the box
    assert orc.trilinear(A, (-3.0, -1.0, 10.0)) == pytest.approx(0.5 + 0.125 * 4, abs=1e-6)


def test_gather_reproduces_affine_face_field(orc):
    g = orc.flipgrid(4.0, 2.0, 1.5, 0.2)
    lo, dx = np.array(g.lo, np.float64), float(g.dx)
    grad = np.array([0.3, -0.7, 1.1])
    for a in range(3):
        sh = orc.face_shape(g, a)
        kk, jj, ii = np.meshgrid(np.arange(sh[0]), np.arange(sh[1]), np.arange(sh[2]), indexing="ij")
        off = [0.5, 0.5, 0.5]
        off[a] = 0.0
        X = lo[0] + (ii + off[0]) * dx
        Y = lo[1] + (jj + off[1]) * dx
        Z = lo[2] + (kk + off[2]) * dx
        u = (2.0 + grad[0] * X + grad[1] * Y + grad[2] * Z).astype(np.float32)
        for p in [(0.13, -0.41, -0.27), (1.0, 1.2, -1.5)]:
            vel, aff = orc.gather(u, g, a, p)
            assert vel == pytest.approx(2.0 + grad @ np.array(p), abs=2e-5)
            assert np.allclose(aff, grad, atol=2e-3)


# ── the dense scene used by the phase tests ──────────────────────────────────

@pytest.fixture(scope="module")
def scene(orc):
    # the reference's dense test basin: narutobasin(radius 4, depth 2, Γ 15, core 1), dx 0.2
    g, phi, u, v, w, types = orc.dense_scene(4.0, 2.0, 1.5, 0.2, 15.0, 1.0)
    ps, n, *_ = orc.dense_resample(g, orc.empty_particles(1), 0, 400000, phi, u, v, w, seeding=True, seed=1)
    ps, n, *_ = orc.dense_resample(g, ps, n, 400000, phi, u, v, w, seed=1)
    return dict(g=g, phi=phi, u=u, v=v, w=w, types=types, ps=ps, n=n)


def _bin_numpy(g, ps, n, phi, deep, above):
    """An independent vectorised restatement of flip.jl:1258-1287's keep test and key."""
    lo = np.array(g.lo, np.float32)
    dx = np.float32(g.dx)
    n1, n2, n3 = g.dims
    x, y, z = ps["PX"][:n], ps["PY"][:n], ps["PZ"][:n]
    c = [(x - lo[0]) / dx, (y - lo[1]) / dx, (z - lo[2]) / dx]
    with np.errstate(invalid="ignore"):
        keep = np.isfinite(c[0]) & np.isfinite(c[1]) & np.isfinite(c[2])
        for a, m in enumerate((n1, n2, n3)):
            keep &= (c[a] >= 0) & (c[a] < m)
        keep &= ~((x * x + y * y > np.float32(g.radius) * np.float32(g.radius)) | (z < np.float32(g.floor)))
    idx = np.where(keep)[0]
    q = [np.clip(c[a][idx] - np.float32(0.5), np.float32(0), np.float32(m - 1)) for a, m in enumerate((n1, n2, n3))]
    ii = [np.minimum(q[a].astype(np.int64), m - 2) for a, m in enumerate((n1, n2, n3))]
    f = [q[a] - ii[a].astype(np.float32) for a in range(3)]
    P = phi
    def at(di, dj, dk):
        return P[ii[2] + dk, ii[1] + dj, ii[0] + di]
    gx, gy, gz = (np.float32(1) - f[0]), (np.float32(1) - f[1]), (np.float32(1) - f[2])
    c00 = at(0, 0, 0) * gx + at(1, 0, 0) * f[0]
    c10 = at(0, 1, 0) * gx + at(1, 1, 0) * f[0]
    c01 = at(0, 0, 1) * gx + at(1, 0, 1) * f[0]
    c11 = at(0, 1, 1) * gx + at(1, 1, 1) * f[0]
    val = (c00 * gy + c10 * f[1]) * gz + (c01 * gy + c11 * f[1]) * f[2]
    ok = (-np.float32(deep) <= val) & (val < np.float32(above))
    key = np.full(n, -1, np.int64)
    cells = [np.floor(c[a][idx]).astype(np.int64) for a in range(3)]
    lin = cells[0] + n1 * (cells[1] + n2 * cells[2])
    key[idx[ok]] = lin[ok]
    return key


def test_resample_keys_match_numpy_restatement(orc, scene):
    g, n = scene["g"], scene["n"]
    assert n > 1000
    rng = np.random.default_rng(3)
    ps = {k: v.copy() for k, v in scene["ps"].items()}
    # move particles about so some leave the band, the grid or the walls, and some cells crowd
    ps["PX"] += rng.normal(0, 0.15, n).astype(np.float32)
    ps["PZ"] += rng.normal(0, 0.15, n).astype(np.float32)
    ps["PX"][:50] = np.nan
    dx = np.float32(g.dx)
    out, nn, key, rank, counts, totals, offsets = orc.dense_resample(g, ps, n, 400000, scene["phi"], scene["u"],
                                                                     scene["v"], scene["w"], seed=7)
    ref = _bin_numpy(g, ps, n, scene["phi"], np.float32(4) * dx, np.float32(3) * dx)
    assert np.array_equal(key, ref)
    # counts are the key histogram, ranks the stable order within each key
    hist = np.bincount(key[key >= 0], minlength=counts.size)
    assert np.array_equal(hist, counts)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    starts = np.searchsorted(ks, ks, side="left")
    exp_rank = np.empty(n, np.int64)
    exp_rank[order] = np.arange(n) - starts
    assert np.array_equal(rank[key >= 0], exp_rank[key >= 0]) and np.all(rank[key < 0] == -1)
    assert np.array_equal(offsets, np.cumsum(totals)) and nn == offsets[-1]
    # survivors: the first min(count, 2 nmin) of each cell, in order, at the cell's place
    keep = (key >= 0) & (rank < 16)
    src = np.where(keep)[0]
    dst = offsets[key[src]] - totals[key[src]] + rank[src]
    assert np.array_equal(out["PX"][dst], ps["PX"][src]) and np.array_equal(out["C4"][dst], ps["C4"][src])
    # everything written lies in its own cell (spawned particles included)
    lo = np.array(g.lo, np.float32)
    n1, n2, n3 = g.dims
    cell = np.repeat(np.arange(totals.size), totals)
    ci = cell % n1
    cj = (cell // n1) % n2
    ck = cell // (n1 * n2)
    fx = np.floor((out["PX"] - lo[0]) / dx).astype(np.int64)
    fy = np.floor((out["PY"] - lo[1]) / dx).astype(np.int64)
    fz = np.floor((out["PZ"] - lo[2]) / dx).astype(np.int64)
    assert np.array_equal(fx, ci) and np.array_equal(fy, cj) and np.array_equal(fz, ck)


def test_resample_cellcounts(orc, scene):
    """kept = min(count, 2 nmin); add = nmin - kept in the band [-deep, -dx] when short (flip.jl:1289-1297)."""
    g = scene["g"]
    _, nn, key, rank, counts, totals, offsets = orc.dense_resample(g, scene["ps"], scene["n"], 400000, scene["phi"],
                                                                   scene["u"], scene["v"], scene["w"], seed=2)
    dx = np.float32(g.dx)
    phi = scene["phi"].ravel()
    solid = scene["types"].ravel() == 0
    kept = np.minimum(counts, 16)
    inband = ~solid & (-np.float32(4) * dx <= phi) & (phi <= -dx)
    add = np.where(inband & (kept < 8), 8 - kept, 0)
    assert np.array_equal(totals, kept + add)


def test_spawn_positions_follow_hash(orc, scene):
    """A spawned particle's position is corner + (octant + hash01 jitter) dx/2 (flip.jl:1329-1335)."""
    g = scene["g"]
    phi = scene["phi"]
    out, nn, key, rank, counts, totals, offsets = orc.dense_resample(g, orc.empty_particles(1), 0, 400000, phi,
                                                                     scene["u"], scene["v"], scene["w"], seed=5)
    l = int(np.argmax(totals > 0))
    m = np.arange(totals[l])
    h = ((l + 1) * 31 + m).astype(np.int64)
    h = ((h + 2**31) % 2**32 - 2**31).astype(np.int32)
    jx = _hash01_numpy(h, 5 + 1)
    o = m % 8
    n1 = g.dims[0]
    corner = np.float32(g.lo[0]) + np.float32(l % n1) * np.float32(g.dx)
    xp = corner + ((o & 1).astype(np.float32) + jx) * (np.float32(0.5) * np.float32(g.dx))
    base = offsets[l] - totals[l]
    assert np.array_equal(out["PX"][base:base + totals[l]], xp.astype(np.float32))


def test_capacity_error_text(orc, scene):
    with pytest.raises(RuntimeError, match=r"^the narrow band needs \d+ particles, more than the 10 allocated$"):
        orc.dense_resample(scene["g"], scene["ps"], scene["n"], 10, scene["phi"], scene["u"], scene["v"], scene["w"])


# ── P2G / G2P ────────────────────────────────────────────────────────────────

def _scatter_numpy(g, ps, idx, a, dt):
    """Brute-force FP64 scatter of flip.jl:513-537 for a few particles."""
    sh = g.dims
    n1, n2, n3 = sh[0] + (a == 0), sh[1] + (a == 1), sh[2] + (a == 2)
    mom = {}
    wt = {}
    off = np.array([0.5, 0.5, 0.5])
    off[a] = 0.0
    for p in idx:
        x = np.array([ps["PX"][p], ps["PY"][p], ps["PZ"][p]], np.float64)
        lag = float(ps["LAG"][p])
        r = min(max(-lag / dt, -0.5), 0.5) - 0.5
        wT = 35 / 16 * (1 - r * r) ** 3
        q = (x - np.array(g.lo, np.float64)) / g.dx - off
        base = np.floor(q - 0.5).astype(int)
        f = q - base
        W = np.stack([0.5 * (1.5 - f) ** 2, 0.75 - (f - 1) ** 2, 0.5 * (f - 0.5) ** 2])
        v = float(ps[["VX", "VY", "VZ"][a]][p])
        c = np.array([ps["C%d" % (3 * a + b)][p] for b in range(3)], np.float64)
        for mz in range(3):
            for my in range(3):
                for mx in range(3):
                    I, J, K = base + np.array([mx, my, mz])
                    if not (0 <= I < n1 and 0 <= J < n2 and 0 <= K < n3):
                        continue
                    wgt = wT * W[mx, 0] * W[my, 1] * W[mz, 2]
                    d = (np.array([I, J, K]) - q) * g.dx
                    mom[(K, J, I)] = mom.get((K, J, I), 0.0) + wgt * (v + c @ d)
                    wt[(K, J, I)] = wt.get((K, J, I), 0.0) + wgt
    return mom, wt


def test_p2g_matches_brute_force(orc, scene):
    g = scene["g"]
    rng = np.random.default_rng(4)
    idx = rng.choice(scene["n"], 120, replace=False)
    ps = {k: v[idx].copy() for k, v in scene["ps"].items()}
    ps["LAG"] = rng.uniform(-0.02, 0.02, 120).astype(np.float32)
    for b in range(9):
        ps["C%d" % b] = rng.normal(0, 0.5, 120).astype(np.float32)
    dt = 0.04
    _, _, _, _, M, W = orc.dense_p2g(g, ps, 120, dt, scene["phi"], scene["u"], scene["v"], scene["w"])
    for a in range(3):
        mom, wt = _scatter_numpy(g, ps, range(120), a, dt)
        for key, val in mom.items():
            assert M[a][key] == pytest.approx(val, rel=1e-5, abs=5e-6)
            assert W[a][key] == pytest.approx(wt[key], rel=1e-5, abs=5e-6)
        assert np.count_nonzero(W[a]) == len(wt)


def test_p2g_conserves_momentum(orc, scene):
    """Partition of unity and zero first moment: sum mom = sum wT v, sum wt = sum wT."""
    g = scene["g"]
    ps = scene["ps"]
    n = scene["n"]
    dt = 0.04
    _, _, _, _, M, W = orc.dense_p2g(g, ps, n, dt, scene["phi"], scene["u"], scene["v"], scene["w"])
    r = np.clip(-ps["LAG"][:n].astype(np.float64) / dt, -0.5, 0.5) - 0.5
    wT = 35 / 16 * (1 - r * r) ** 3
    for a, f in enumerate(("VX", "VY", "VZ")):
        # the weights are FP32 (as the reference's), so the sums hold to FP32 rounding per particle
        scale = np.sum(np.abs(wT * ps[f][:n]))
        assert abs(M[a].sum() - np.sum(wT * ps[f][:n])) <= 1e-6 * scale
        assert W[a].sum() == pytest.approx(np.sum(wT), rel=1e-6)


def test_g2p_reproduces_affine_field(orc, scene):
    """G2P of an affine velocity returns it exactly, with c = its gradient (APIC)."""
    g = scene["g"]
    lo, dx = np.array(g.lo, np.float64), float(g.dx)
    A = np.array([[0.2, -0.1, 0.05], [0.3, 0.0, -0.2], [-0.1, 0.25, 0.1]])
    fields, flags = [], []
    for a in range(3):
        sh = orc.face_shape(g, a)
        kk, jj, ii = np.meshgrid(np.arange(sh[0]), np.arange(sh[1]), np.arange(sh[2]), indexing="ij")
        off = [0.5, 0.5, 0.5]
        off[a] = 0.0
        X = np.stack([lo[0] + (ii + off[0]) * dx, lo[1] + (jj + off[1]) * dx, lo[2] + (kk + off[2]) * dx])
        fields.append((1.0 + np.tensordot(A[a], X, 1)).astype(np.float32))
        flags.append(np.ones(sh, np.uint8))
    n = scene["n"]
    out = orc.dense_g2p(g, scene["ps"], n, *fields, *flags, 0.04)
    X = np.stack([scene["ps"][f][:n].astype(np.float64) for f in ("PX", "PY", "PZ")])
    for a, f in enumerate(("VX", "VY", "VZ")):
        assert np.allclose(out[f], 1.0 + A[a] @ X, atol=2e-5)
        for b in range(3):
            assert np.allclose(out["C%d" % (3 * a + b)], A[a, b], atol=3e-3)
    # no valid flags: ballistic (vz -= g dt, c = 0, vx unchanged)
    out2 = orc.dense_g2p(g, scene["ps"], n, *fields, *[np.zeros_like(f) for f in flags], 0.04)
    assert np.array_equal(out2["VX"], scene["ps"]["VX"][:n])
    assert np.allclose(out2["VZ"], scene["ps"]["VZ"][:n] - np.float32(9.81 * 0.04))
    assert not np.any(out2["C4"])


# ── the pressure ─────────────────────────────────────────────────────────────

def test_operator_symmetric_and_matches_apply(orc, scene):
    g = scene["g"]
    A = orc.dense_matrix(g, scene["types"], scene["phi"])
    assert abs(A - A.T).max() == 0
    x = np.random.default_rng(5).standard_normal(orc.cells_shape(g))
    q = orc.dense_apply(g, scene["types"], scene["phi"], x)
    assert np.allclose(q.ravel(), A @ x.ravel(), atol=1e-12)
    # positive definite on the water: x.Ax > 0 for x supported on LIQUID cells
    liq = (scene["types"].ravel() == 1)
    xl = np.where(liq, x.ravel(), 0)
    assert xl @ (A @ xl) > 0


def test_vcycle_is_symmetric_linear(orc, scene):
    g = scene["g"]
    rng = np.random.default_rng(6)
    liq = scene["types"] == 1
    x = np.where(liq, rng.standard_normal(orc.cells_shape(g)), 0)
    y = np.where(liq, rng.standard_normal(orc.cells_shape(g)), 0)
    Mx = orc.dense_vcycle(g, scene["types"], scene["phi"], x)
    My = orc.dense_vcycle(g, scene["types"], scene["phi"], y)
    assert abs((Mx * y).sum() - (x * My).sum()) <= 1e-10 * abs((Mx * y).sum())
    Mxy = orc.dense_vcycle(g, scene["types"], scene["phi"], 2 * x + 3 * y)
    assert np.allclose(Mxy, 2 * Mx + 3 * My, atol=1e-10)
    assert (Mx * x).sum() > 0
    # the FP32 V-cycle is the same algorithm in single precision
    Mx32 = orc.dense_vcycle(g, scene["types"], scene["phi"], x.astype(np.float32), fp32=True)
    assert Mx32.dtype == np.float32
    assert np.linalg.norm(Mx32 - Mx) <= 1e-5 * np.linalg.norm(Mx)


def test_pcg_matches_scipy_direct_solve(orc, scene):
    import scipy.sparse.linalg as spla
    g = scene["g"]
    b = orc.dense_rhs(g, scene["u"], scene["v"], scene["w"] - np.float32(0.3), scene["types"], 0.001)
    P, it, rr = orc.dense_pressure(g, scene["types"], scene["phi"], b, tol=1e-12)
    assert rr <= 1e-12 and 0 < it < 100
    liq = np.where(scene["types"].ravel() == 1)[0]
    A = orc.dense_matrix(g, scene["types"], scene["phi"])[liq][:, liq]
    Pd = spla.spsolve(A.tocsc(), b.ravel()[liq])
    assert np.linalg.norm(P.ravel()[liq] - Pd) <= 1e-9 * np.linalg.norm(Pd)


def test_projection_leaves_target_divergence(orc, scene):
    """After the projection, -dx div u + dx^2 gamma = b - A P on water cells (flip.jl:811-822, 1113-1130)."""
    g = scene["g"]
    gamma = 0.002
    w = scene["w"] - np.float32(0.3)
    b = orc.dense_rhs(g, scene["u"], scene["v"], w, scene["types"], gamma)
    P, it, rr = orc.dense_pressure(g, scene["types"], scene["phi"], b, tol=1e-10)
    U, F = orc.dense_project_faces(g, P, scene["types"], scene["phi"], scene["u"], scene["v"], w)
    b2 = orc.dense_rhs(g, *U, scene["types"], gamma)
    # the projected velocities are stored in FP32: a few roundings of |u| per face
    fp32 = 8 * float(g.dx) * max(np.abs(x).max() for x in U) * 2.0 ** -23
    assert np.abs(b2).max() <= 1e-6 * np.abs(b).max() + fp32
    assert np.abs(b2).max() < 0.01 * np.abs(b).max()
