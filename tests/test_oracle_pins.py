"""Pins of the oracle functions that no closed form covers: the refinement marks, the
particles' level set (splat and union), the combine, the fine-to-coarse restriction and
new blocks' prolongation. Each is checked bit for bit against an independent NumPy
restatement of the reference text (tests/pins_np.py) on random and edge-case inputs, and
by a property the mathematics fixes (brute-force minimum distance, the combine's closed
forms, affine reproduction, means of constant fields). Each test also shows that a
planted slip in the restatement (a sign, an index, a comparison) disagrees with the
oracle on the same inputs, so the pin can tell a right answer from a plausible wrong one.
"""
import numpy as np
import pytest

import oracle as O
import pins_np as N

f32 = np.float32
DX = 0.6


def _scene(**kw):
    kw.setdefault("blocks", 64)
    kw.setdefault("capacity", 300000)
    return O.sparse_scene(dx=DX, **kw)


def _topo(o):
    return N.Topo(o.grid, o.array("slot"), o.array("where"))


def _coarse(o, name):
    g = o.grid
    n1, n2, n3 = g.nc
    shp = {"phi_c": (n3, n2, n1), "P_c": (n3, n2, n1), "types_c": (n3, n2, n1), "u_c": (n3, n2, n1 + 1),
           "v_c": (n3, n2 + 1, n1), "w_c": (n3 + 1, n2, n1), "fu_c": (n3, n2, n1 + 1), "fv_c": (n3, n2 + 1, n1),
           "fw_c": (n3 + 1, n2, n1)}[name]
    return o.array(name).reshape(shp)


def _np_marks(o, vary, reach, rings, particles):
    t = _topo(o)
    parts = None
    if particles:
        n = o.n
        parts = (o.array("px")[:n], o.array("py")[:n], o.array("pz")[:n])
    return N.marks(t, _coarse(o, "phi_c"), _coarse(o, "u_c"), _coarse(o, "v_c"), _coarse(o, "w_c"), vary, parts,
                   reach, rings)


# ── marks (sflip.jl:384-455) ─────────────────────────────────────────────────

def test_marks_pinned_random():
    """The oracle's marks equal the NumPy restatement on the start state and on random
    level sets, velocities and particles, for every kind, with and without dilation."""
    o = _scene()
    reach = f32((8 + 3) * f32(DX))
    rng = np.random.default_rng(3)
    kinds = {"surface": 0, "flow": 0, "particles": 0}
    for trial in range(4):
        if trial:
            phi = o.array("phi_c")
            phi[:] += rng.normal(0, 1.5, phi.size).astype(f32)
            for nm in ("u_c", "v_c", "w_c"):
                a = o.array(nm)
                a[:] += rng.normal(0, 0.4, a.size).astype(f32)
            n = o.n
            for nm in ("px", "py", "pz"):
                a = o.array(nm)
                a[:n] += rng.normal(0, 1.0, n).astype(f32)
        for particles in (False, True):
            for rings in (0, 1, 2):
                got = o.marks(vary=0.5, reach=reach, rings=rings, particles=particles)
                exp = _np_marks(o, 0.5, reach, rings, particles)
                np.testing.assert_array_equal(got, exp, err_msg="trial %d particles %s rings %d" % (trial, particles,
                                                                                                     rings))
        t = _topo(o)
        kinds["surface"] += int(N.marksurface(t, _coarse(o, "phi_c")).sum())
        kinds["flow"] += int(N.markflow(t, _coarse(o, "u_c"), _coarse(o, "v_c"), _coarse(o, "w_c"),
                                        _coarse(o, "phi_c"), 0.5).sum())
    assert all(v > 0 for k, v in kinds.items() if k != "particles"), kinds


def _blank(o):
    """Everything air and still, no particles: nothing is marked."""
    o.array("phi_c")[:] = f32(100.0)
    for nm in ("u_c", "v_c", "w_c"):
        o.array(nm)[:] = 0
    o.n = 0


def test_marks_edge_cases_and_planted_slips():
    o = _scene(particles=False)
    t = _topo(o)
    g = o.grid
    n1, n2, n3 = g.nc
    reach = f32(11 * f32(DX))
    (I, J, K), (px, py, pz) = t.coarsecentres()
    solid = t.issolid(px, py, pz)

    # (1) a sign change towards a solid neighbour does not mark; towards a fluid one it does
    _blank(o)
    phi = _coarse(o, "phi_c")
    k0 = n3 // 2
    row = np.flatnonzero(~solid[k0, n2 // 2])
    i_edge = row[0]                         # the first non-solid cell along x: its -x neighbour is solid
    phi[k0, n2 // 2, i_edge - 1] = f32(-50.0)
    assert o.marks(particles=False).sum() == 0
    assert N.marksurface(t, phi, slip="solid").sum() > 0        # the planted slip marks it
    phi[k0, n2 // 2, i_edge + 1] = f32(-50.0)                  # now a fluid neighbour changes sign
    got = o.marks(particles=False)
    assert got.sum() > 0 and np.array_equal(got, N.marksurface(t, phi))

    # (2) |φ| exactly 1.5 coarse cells is not the surface; just under is
    _blank(o)
    phi = _coarse(o, "phi_c")
    c = f32(1.5) * t.cdx
    phi[k0, n2 // 2, n1 // 2] = c
    assert o.marks(particles=False).sum() == 0
    assert N.marksurface(t, phi, slip="radius").sum() > 0
    phi[k0, n2 // 2, n1 // 2] = np.nextafter(c, f32(0))
    assert o.marks(particles=False).sum() == 1

    # (3) a velocity jump of exactly `vary` does not mark; one ulp more does
    _blank(o)
    phi = _coarse(o, "phi_c")
    phi[:] = f32(-5.0)
    phi[solid] = f32(-5.0)
    u = _coarse(o, "u_c")
    i0, j0 = n1 // 2, n2 // 2
    u[k0, j0, i0 + 1:] = f32(1.0)            # cell i0 has mean 0.5, i0+1 mean 1.0: a jump of 0.5 ...
    vary = f32(0.5)
    m = o.marks(vary=vary, particles=False)
    exp = N.markflow(t, u, _coarse(o, "v_c"), _coarse(o, "w_c"), phi, vary) | N.marksurface(t, phi)
    np.testing.assert_array_equal(m, exp)
    slipped = N.markflow(t, u, _coarse(o, "v_c"), _coarse(o, "w_c"), phi, vary, slip="geq") | N.marksurface(t, phi)
    assert not np.array_equal(m, slipped)
    m2 = o.marks(vary=np.nextafter(vary, f32(0)), particles=False)
    assert m2.sum() > m.sum()

    # (4) particles exactly on block faces and at the grid's ends
    _blank(o)
    S = f32(16) * f32(DX)
    lo = np.array(list(g.lo), f32)
    pts = []
    for k in range(3):
        for jj in range(3):
            pts.append([lo[0] + f32(8 + k) * S - reach, lo[1] + f32(10 + jj) * S + reach, lo[2] + f32(3) * S])
    pts += [[lo[0], lo[1], lo[2]], [lo[0] + S * f32(g.nb[0]), lo[1] + S * f32(g.nb[1]), lo[2] + S * f32(g.nb[2])],
            [np.nan, 0, 0], [0, np.inf, 0]]
    pts = np.array(pts, f32)
    n = len(pts)
    for a, nm in enumerate(("px", "py", "pz")):
        o.array(nm)[:n] = pts[:, a]
    o.n = n
    got = o.marks(reach=reach, particles=True)
    x, y, z = (o.array(nm)[:n] for nm in ("px", "py", "pz"))
    np.testing.assert_array_equal(got, N.markparticles(t, x, y, z, reach))
    assert not np.array_equal(got, N.markparticles(t, x, y, z, reach, slip="onecorner"))
    o.n = 0


# ── restriction (sflip.jl:580-612) ───────────────────────────────────────────

@pytest.mark.parametrize("name,axis,flag", [("phi", None, False), ("P", None, False), ("types", None, True),
                                            ("u", 0, False), ("v", 1, False), ("w", 2, False),
                                            ("fu", 0, True), ("fv", 1, True), ("fw", 2, True)])
def test_restriction_pinned(name, axis, flag):
    o = _scene(particles=False)
    t = _topo(o)
    rng = np.random.default_rng(7)
    fine = o.array(name + "_f")
    if flag:
        fine[:] = (rng.random(fine.size) < 0.97).astype(np.uint8)
    else:
        fine[:] = rng.normal(0, 2, fine.size).astype(f32)
    coarse = _coarse(o, name + "_c")
    coarse[:] = (rng.random(coarse.shape) < 0.5) if flag else rng.normal(0, 1, coarse.shape).astype(f32)
    before = coarse.copy()
    if axis is None:
        exp = N.restrictcells(t, before, fine, flag=flag)
    else:
        exp = N.restrictfaces(t, before, fine, axis, flag=flag)
    o.restrict_field(name)
    np.testing.assert_array_equal(_coarse(o, name + "_c"), exp)
    assert (exp != before).any()
    if axis is None and not flag:
        # a planted index slip (b and c swapped) reads other cells and disagrees
        assert not np.array_equal(N.restrictcells(t, before, fine, slip="index"), exp)
        # a constant fine field restricts to that constant on every covered cell
        fine[:] = f32(0.375)
        o.restrict_field(name)
        (I, J, K), _ = t.coarsecentres()
        cov = t.covered(I, J, K)
        assert cov.sum() > 0 and (_coarse(o, name + "_c")[cov] == f32(0.375)).all()


def test_restriction_flag_rule():
    """A covered coarse cell's flag is set only when all 64 fine flags are; a coarse face's
    only when all 16 fine faces exist and are set (sflip.jl:580-581)."""
    o = _scene(particles=False)
    t = _topo(o)
    (I, J, K), _ = t.coarsecentres()
    cov = np.argwhere(t.covered(I, J, K))
    kc, jc, ic = cov[len(cov) // 2]
    types = o.array("types_f")
    types[:] = 1
    f0 = t.fineat(np.array([4 * ic + 1]), np.array([4 * jc + 2]), np.array([4 * kc + 3]))[0]
    types[f0] = 0
    o.restrict_field("types")
    tc = _coarse(o, "types_c")
    assert tc[kc, jc, ic] == 0 and tc[cov[:, 0], cov[:, 1], cov[:, 2]].sum() == len(cov) - 1


# ── new blocks from the coarse grid (sflip.jl:465-496) ───────────────────────

def test_prolongation_pinned_and_affine():
    o = _scene(particles=False)
    g = o.grid
    rng = np.random.default_rng(9)
    t0 = _topo(o)
    (I, J, K), (px, py, pz) = t0.coarsecentres()
    # an affine level set and pressure on the coarse grid; random velocities and flags
    A = (f32(0.25) * px - f32(0.5) * py + f32(0.125) * pz + f32(1.0)).astype(f32)
    _coarse(o, "phi_c")[:] = A
    _coarse(o, "P_c")[:] = (f32(2.0) * pz).astype(f32)
    for nm in ("u_c", "v_c", "w_c"):
        c = _coarse(o, nm)
        c[:] = rng.normal(0, 1, c.shape).astype(f32)
    for nm in ("fu_c", "fv_c", "fw_c"):
        c = _coarse(o, nm)
        c[:] = (rng.random(c.shape) < 0.9).astype(np.uint8)
    # release half the blocks and add new ones: the arrivals are prolonged
    want = np.zeros(int(np.prod(list(g.nb))), np.uint8)
    slot = o.array("slot")
    live = np.flatnonzero(slot)
    want[live[::2]] = 1
    free_blocks = np.flatnonzero(slot == 0)
    want[rng.choice(free_blocks, 20, replace=False)] = 1
    added, removed = o.refine(want)
    assert len(added) > 0 and len(removed) > 0
    t = _topo(o)
    for s in added:
        cs = slice((s - 1) * N.BCELLS, s * N.BCELLS)
        fs = slice((s - 1) * N.BFACES, s * N.BFACES)
        np.testing.assert_array_equal(o.array("phi_f")[cs], N.prolongcells(t, _coarse(o, "phi_c"), s))
        np.testing.assert_array_equal(o.array("P_f")[cs], N.prolongcells(t, _coarse(o, "P_c"), s))
        for a, (u, f) in enumerate((("u", "fu"), ("v", "fv"), ("w", "fw"))):
            np.testing.assert_array_equal(o.array(u + "_f")[fs], N.prolongfaces(t, _coarse(o, u + "_c"), s, a))
            np.testing.assert_array_equal(o.array(f + "_f")[fs], N.prolongfaces(t, _coarse(o, f + "_c"), s, a,
                                                                                 flag=True))
    # affine reproduction: away from the grid's clamped edges the prolonged level set is
    # the affine field at the fine centres (FP32 rounding only)
    errs = []
    for s in added:
        w = t.where[s - 1]
        i, j, k = N._cell_local()
        P = [t.lo[a] + ((X + w[a] * 16).astype(f32) + f32(0.5)) * t.dx for a, X in enumerate((i, j, k))]
        exact = 0.25 * P[0].astype(np.float64) - 0.5 * P[1] + 0.125 * P[2] + 1.0
        q = [(P[a] - t.lo[a]) / t.cdx - f32(0.5) for a in range(3)]
        inside = np.all([(q[a] >= 0) & (q[a] <= t.nc[a] - 1) for a in range(3)], axis=0)
        got = o.array("phi_f")[(s - 1) * N.BCELLS:s * N.BCELLS]
        errs.append(np.abs(got[inside] - exact[inside]).max() if inside.any() else 0.0)
    assert max(errs) < 2e-5 * 40, max(errs)
    # removed slots are zeroed (sflip.jl:457-463)
    for s in removed:
        if s not in added:
            assert not o.array("phi_f")[(s - 1) * N.BCELLS:s * N.BCELLS].any()


# ── the particles' level set (flip.jl:539-565, sflip.jl:790-794) ─────────────

def _dense_small():
    """A small dense basin with a few hundred moving particles near the surface."""
    g, phi, u, v, w, types = O.dense_scene(4.0, 2.0, 1.5, 0.2, 15.0, 1.0)
    rng = np.random.default_rng(4)
    n = 700
    n1, n2, n3 = g.dims
    lo = np.array(list(g.lo), f32)
    ps = {f: np.zeros(n, f32) for f in O.PFIELDS}
    ps["PX"][:] = lo[0] + rng.uniform(2, n1 - 2, n).astype(f32) * f32(g.dx)
    ps["PY"][:] = lo[1] + rng.uniform(2, n2 - 2, n).astype(f32) * f32(g.dx)
    ps["PZ"][:] = lo[2] + rng.uniform(n3 * 0.4, n3 - 1, n).astype(f32) * f32(g.dx)
    for f in ("VX", "VY", "VZ"):
        ps[f][:] = rng.normal(0, 2.0, n).astype(f32)
    ps["VX"][:5] = 0                         # some still particles (strength 0)
    ps["VY"][:5] = 0
    ps["VZ"][:5] = 0
    for k in range(9):
        ps["C%d" % k][:] = rng.normal(0, 0.5, n).astype(f32)
    ps["LAG"][:] = rng.uniform(-0.05, 0, n).astype(f32)
    ps["PX"][5] = lo[0] + f32(7) * f32(g.dx)          # exactly on a cell face
    return g, phi, u, v, w, ps, n


def test_dense_splat_union_bruteforce():
    g, phi0, u0, v0, w0, ps, n = _dense_small()
    dt = f32(0.05)
    phi, *_ = O.dense_p2g(g, ps, n, dt, phi0, u0, v0, w0)
    dx = f32(g.dx)
    lo = np.array(list(g.lo), f32)
    xp = N.splat_positions(ps["PX"], ps["PY"], ps["PZ"], ps["VX"], ps["VY"], ps["VZ"], ps["LAG"], dt, dx)
    n1, n2, n3 = g.dims
    lin = lambda I, J, K: I + n1 * (J + n2 * K)
    best = N.splat_bruteforce(lo, dx, (n1, n2, n3), lin, xp)
    exp = np.minimum(phi0.ravel() + dx, np.where(np.isinf(best), f32(1e30), best)).astype(f32)
    np.testing.assert_array_equal(phi.ravel(), exp)
    touched = np.isfinite(best)
    assert touched.sum() > 1000
    # the splat value is a distance: never below the nearest particle's distance less dx/2
    cen = np.stack(np.meshgrid(*(lo[a] + (np.arange(m) + 0.5) * dx for a, m in enumerate((n1, n2, n3))),
                               indexing="ij"), -1).reshape(-1, 3)
    cen = cen.reshape(n1, n2, n3, 3).transpose(2, 1, 0, 3).reshape(-1, 3)
    P = np.stack(xp, -1).astype(np.float64)
    dmin = np.full(cen.shape[0], np.inf)
    for p0 in range(0, n, 256):
        d = np.sqrt(((cen[:, None, :] - P[None, p0:p0 + 256]) ** 2).sum(-1)).min(1)
        dmin = np.minimum(dmin, d)
    assert (best[touched] >= dmin[touched] - 0.5 * dx - 1e-5).all()
    # a planted reach slip (two cells instead of one) changes the level set
    slipped = N.splat_bruteforce(lo, dx, (n1, n2, n3), lin, xp, slip="reach")
    assert not np.array_equal(np.minimum(phi0.ravel() + dx, np.where(np.isinf(slipped), f32(1e30), slipped)), exp)


def test_sparse_splat_union_bruteforce():
    o = _scene()
    t = _topo(o)
    dt = f32(0.1)
    phi0_f = o.array("phi0_f").copy()
    n = o.n
    parts = {nm: o.array(nm)[:n].copy() for nm in ("px", "py", "pz", "vx", "vy", "vz", "lag")}
    o.p2g(dt)
    ns = o.nslots
    got = o.array("phi_f")[:ns * N.BCELLS]
    xp = N.splat_positions(*(parts[k] for k in ("px", "py", "pz", "vx", "vy", "vz", "lag")), dt, t.dx)
    # every particle's 27 cells through the restated fine index map, minimised per cell
    best = np.full(ns * N.BCELLS, np.inf, f32)
    ci = [np.floor((xp[a] - t.lo[a]) / t.dx).astype(np.int64) for a in range(3)]
    radius = f32(0.5) * t.dx
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                I, J, K = ci[0] + di, ci[1] + dj, ci[2] + dk
                idx = t.fineat(I, J, K)
                ok = idx >= 0
                cen = [t.lo[a] + (X.astype(f32) + f32(0.5)) * t.dx for a, X in enumerate((I, J, K))]
                e = [cen[a] - xp[a] for a in range(3)]
                d = (np.sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]) - radius).astype(f32)
                np.minimum.at(best, idx[ok], d[ok])
    live = np.repeat(t.where[:ns, 0] >= 0, N.BCELLS)
    exp = np.where(live, np.minimum(phi0_f[:ns * N.BCELLS] + t.dx, np.where(np.isinf(best), f32(1e30), best)),
                   got).astype(f32)
    np.testing.assert_array_equal(got, exp)
    assert np.isfinite(best).sum() > 10000


# ── the combine (flip.jl:567-585, sflip.jl:796-806) ──────────────────────────

def test_dense_combine_pinned_and_closed_forms():
    g, phi0, u0, v0, w0, ps, n = _dense_small()
    dt = f32(0.05)
    band, nmin = 4, 8
    phi, U, V, W, M, Wt = O.dense_p2g(g, ps, n, dt, phi0, u0, v0, w0, band=band, nmin=nmin)
    r = f32((band - 1) * f32(g.dx))
    full = f32(nmin)
    for a, (out, base) in enumerate(zip((U, V, W), (u0, v0, w0))):
        pf = N.dense_face_phi(phi, a)
        m, wt = M[a].astype(f32), Wt[a].astype(f32)
        exp = N.combine(pf, base, m, wt, r, full)
        np.testing.assert_array_equal(out, exp, err_msg="axis %d" % a)
        near = pf >= -r
        # closed forms: no weight -> u0 exactly; full weight -> mom / wt; far below -> u0
        z = near & (wt == 0)
        assert z.any() and np.array_equal(out[z], base[z])
        f = near & (wt >= full)
        np.testing.assert_array_equal(out[f], (m[f] / wt[f]).astype(f32))
        far = ~near
        assert far.any() and np.array_equal(out[far], base[far])
        part = near & (wt > 0) & (wt < full)
        assert part.any()
        np.testing.assert_allclose(out[part], (m[part] + (full - wt[part]) * base[part]) / full, rtol=1e-6, atol=1e-6)
        # a planted sign slip (the band tested as φ <= r) disagrees
        slipped = np.where(pf <= r, (m + np.maximum(full - wt, 0) * base) / np.maximum(full, wt), base).astype(f32)
        assert not np.array_equal(slipped, exp)


def test_sparse_combine_pinned():
    o = _scene()
    t = _topo(o)
    dt = f32(0.1)
    u0 = [o.array(nm + "0_f").copy() for nm in "uvw"]
    momwt = o.p2g(dt, momwt=True)
    ns = o.nslots
    F = ns * N.BFACES
    r = f32(3 * f32(DX))
    phi_f, phi_c = o.array("phi_f"), _coarse(o, "phi_c")
    for a, nm in enumerate("uvw"):
        slots = np.repeat(np.arange(1, ns + 1), N.BFACES)
        loc = np.tile(np.arange(N.BFACES), ns)
        i, j, k = N._face_local(a)
        order = np.argsort(N.Topo.faceindex(1, a, i, j, k))       # local layout -> (i, j, k)
        i, j, k = i[order], j[order], k[order]
        w = t.where[slots - 1]
        live = w[:, 0] >= 0
        I, J, K = i[loc] + 16 * w[:, 0], j[loc] + 16 * w[:, 1], k[loc] + 16 * w[:, 2]
        off = N.FACEOFF[a]
        P = [t.lo[b] + (X.astype(f32) + off[b]) * t.dx for b, X in enumerate((I, J, K))]
        pf = N.cellsample(t, phi_f, phi_c, *P)
        exp = N.combine(pf, u0[a][:F], momwt[a].astype(f32), momwt[3 + a].astype(f32), r, f32(8))
        got = o.array(nm + "_f")[:F]
        np.testing.assert_array_equal(got[live], exp[live], err_msg=nm)


def _plane_phi(o, n, c):
    """The signed distance to the plane n.x = c at every fine and coarse cell centre."""
    g = o.grid
    ns = o.nslots
    where = o.array("where").reshape(-1, 3)[:ns]
    k, j, i = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    phif = o.array("phi_f")
    for s in range(ns):
        if where[s, 0] < 0:
            continue
        x = g.lo[0] + (where[s, 0] * 16 + i + 0.5) * g.dx
        y = g.lo[1] + (where[s, 1] * 16 + j + 0.5) * g.dx
        z = g.lo[2] + (where[s, 2] * 16 + k + 0.5) * g.dx
        phif[s * 4096:(s + 1) * 4096] = (n[0] * x + n[1] * y + n[2] * z - c).astype(f32).ravel()
    n1, n2, n3 = g.nc
    K, J, I = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    h = 4 * g.dx
    x, y, z = g.lo[0] + (I + 0.5) * h, g.lo[1] + (J + 0.5) * h, g.lo[2] + (K + 0.5) * h
    o.array("phi_c")[:] = (n[0] * x + n[1] * y + n[2] * z - c).astype(f32).ravel()


def test_redistance_keeps_a_planar_distance():
    """The redistancing sweep (sflip.jl:910-919 with flip.jl:610-664's eikonal update) leaves
    a signed distance to a tilted plane unchanged where it is not clamped: the three-axis
    Godunov update reproduces a linear distance function exactly. A plane far from the
    basin's walls and floor, cells away from the clamp at (iterations + 1) h."""
    o = _scene()
    nrm = np.array([0.48, -0.36, 0.8])   # unit length
    _plane_phi(o, nrm, -5.0)
    before_f, before_c = o.array("phi_f").copy(), o.array("phi_c").copy()
    it = 3
    o.redistance(it)
    ns = o.nslots
    after_f = o.array("phi_f")[:ns * 4096]
    bf = before_f[:ns * 4096]
    types = o.array("types_f")[:ns * 4096]
    # inside the blocks: a cell on a block's side reads the coarse grid across it, and each
    # Jacobi sweep carries that a cell further in
    loc = np.arange(4096)
    li, lj, lk = loc % 16, (loc // 16) % 16, loc // 256
    m = it
    interior = np.tile((li >= m) & (li < 16 - m) & (lj >= m) & (lj < 16 - m) & (lk >= m) & (lk < 16 - m), ns)
    # and away from the walls and the floor (solid cells take their water neighbours' value)
    # by more than the sweeps' reach
    where = o.array("where").reshape(-1, 3)[:ns]
    g = o.grid
    cx = g.lo[0] + (np.repeat(where[:, 0], 4096) * 16 + np.tile(li, ns) + 0.5) * g.dx
    cy = g.lo[1] + (np.repeat(where[:, 1], 4096) * 16 + np.tile(lj, ns) + 0.5) * g.dx
    cz = g.lo[2] + (np.repeat(where[:, 2], 4096) * 16 + np.tile(lk, ns) + 0.5) * g.dx
    clear = (np.hypot(cx, cy) < g.radius - (it + 2) * g.dx) & (cz > g.floor + (it + 2) * g.dx)
    inner = (np.abs(bf) < (it - 1) * 0.6) & (np.abs(bf) > 0.6) & (types != 0) & interior & clear
    assert inner.sum() > 300
    np.testing.assert_allclose(after_f[inner], bf[inner], rtol=0, atol=2e-5)
    # far from the plane the value is clamped at the sweep's reach, its sign kept
    far = np.abs(bf) > (it + 2) * 0.6 * 4
    assert np.all(np.sign(after_f[far]) == np.sign(bf[far]))
