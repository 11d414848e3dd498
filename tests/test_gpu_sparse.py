"""Parity of the sparse basin's HIP path (SparseBasin) with the CPU oracle, through the
C ABI: the block allocator and marks bit-exact, resampling and G2P bit-exact, P2G to
FP32 rounding of the deposit, the composite-multigrid projection against the FP64
oracle solve, and the reference's error texts."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

PNAMES = ["px", "py", "pz", "vx", "vy", "vz"] + ["c%d" % k for k in range(9)] + ["lag"]
BCELLS, BFACES = 4096, 4352
DX = 0.6


@pytest.fixture(scope="module")
def M():
    import maelstrom_amd
    return maelstrom_amd


def _basin(M):
    return M.Basin(radius=18.0, depth=12.0, Γ=225.0, core=2.5)


def _scene(dx=DX, **kw):
    kw.setdefault("blocks", 64)
    kw.setdefault("capacity", 300000)
    return O.sparse_scene(dx=dx, **kw)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _upload(M, o, cfl=8.0):
    """A GPU SparseBasin holding the oracle's state (topology, fields, particles)."""
    s = M.SparseBasin(_basin(M), dx=float(o.grid.dx), blocks=o.blocks, capacity=o.capacity, band=o.band, nmin=o.nmin,
                      cfl=cfl)
    g = o.grid
    assert list(s.grid.nb) == list(g.nb) and np.float32(s.grid.dx) == np.float32(g.dx)
    s.set_topology(o.array("slot"), o.array("where"), o.nslots, o.free_list())
    for gname, oname in [("PHI", "phi"), ("PHI0", "phi0"), ("P", "P"), ("TYPES", "types"), ("U", "u"), ("V", "v"),
                         ("W", "w"), ("U0", "u0"), ("V0", "v0"), ("W0", "w0"), ("FU", "fu"), ("FV", "fv"),
                         ("FW", "fw")]:
        s.set(gname + "_F", o.array(oname + "_f"))
        s.set(gname + "_C", o.array(oname + "_c"))
    n = o.n
    for nm in PNAMES:
        s.set(nm.upper(), o.array(nm)[:n])
    s.n = n
    return s


def _fine(o, name, per):
    return o.array(name)[:o.nslots * per]


# ── topology ─────────────────────────────────────────────────────────────────

def test_refine_allocator_bitexact(M):
    """Random mark sequences: slots, wheres, free list and the prolonged fields match."""
    print("case: refine allocator, 15 random mark sets", flush=True)
    o = _scene(particles=False)
    s = _upload(M, o)
    rng = np.random.default_rng(5)
    nblk = s.nblocks
    for it in range(15):
        want = (rng.random(nblk) < rng.uniform(0.1, 0.6)).astype(np.uint8)
        if want.sum() > 64:
            want[np.flatnonzero(want)[64:]] = 0
        oa, orem = o.refine(want)
        na, nr = M.sparse.refine(s, want)
        assert (na, nr) == (len(oa), len(orem))
        slot, where, ns, free = s.topology()
        assert ns == o.nslots
        np.testing.assert_array_equal(slot, o.array("slot"))
        np.testing.assert_array_equal(where, o.array("where").reshape(-1, 3))
        np.testing.assert_array_equal(free, o.free_list())
    # the fields of the last arrivals came from the coarse grid exactly as the oracle's
    np.testing.assert_array_equal(s.get("PHI_F"), _fine(o, "phi_f", BCELLS))
    np.testing.assert_array_equal(s.get("U_F"), _fine(o, "u_f", BFACES))
    np.testing.assert_array_equal(s.get("FW_F"), _fine(o, "fw_f", BFACES))


def test_refine_pool_error(M):
    o = _scene(particles=False)
    s = _upload(M, o)
    want = np.ones(s.nblocks, np.uint8)
    s2 = M.SparseBasin(_basin(M), dx=DX, blocks=8, capacity=10)
    with pytest.raises(M.MaelstromError, match="the sparse grid needs 64 fine blocks, more than its 8 allocated"):
        M.sparse.refine(s2, want)
    del s


def test_marks_bitexact(M):
    o = _scene()
    s = _upload(M, o, cfl=8.0)
    for particles in (False, True):
        for rings in (0, 1):
            s.rings = rings
            got = M.sparse.marks(s, particles=particles)
            exp = o.marks(vary=0.5, reach=np.float32((8 + 3) * np.float32(DX)), rings=rings, particles=particles)
            np.testing.assert_array_equal(got, exp)


def test_initialize_matches_oracle(M):
    """initialize! on the GPU from the same surface table gives the oracle's blocks,
    particles and level set."""
    print("case: sparse initialize", flush=True)
    o = _scene()
    s = M.SparseBasin(_basin(M), dx=DX, blocks=64, capacity=300000)
    M.sparse.initialize(s)
    slot, where, ns, free = s.topology()
    assert ns == o.nslots
    np.testing.assert_array_equal(slot, o.array("slot"))
    assert s.n == o.n
    # the positions are bit-exact; the start's velocities come from the library's own
    # Lamb-Oseen (expf), so the level set the first deposit splats (positions moved along
    # those velocities) agrees to a few ulps
    np.testing.assert_array_equal(s.get("TYPES_F"), _fine(o, "types_f", BCELLS))
    for nm in ("px", "py", "pz"):
        np.testing.assert_array_equal(s.get(nm.upper()), o.array(nm)[:o.n])
    np.testing.assert_allclose(s.get("PHI_F"), _fine(o, "phi_f", BCELLS), rtol=0, atol=2e-5)
    assert abs(M.sparse.volume(s) / o.volume() - 1) < 1e-6


# ── particles ────────────────────────────────────────────────────────────────

def test_resample_bitexact(M):
    print("case: sparse resample", flush=True)
    o = _scene()
    s = _upload(M, o)
    n = o.n
    s.steps = 4
    total = M.sparse.resample(s)
    r = o.resample(seed=5)
    assert total == r["total"] == s.n
    C = o.nslots * BCELLS
    np.testing.assert_array_equal(s.get("KEY", count=n), r["key"])
    np.testing.assert_array_equal(s.get("RANK", count=n), r["rank"])
    np.testing.assert_array_equal(s.get("COUNTS"), r["counts"][:C])
    np.testing.assert_array_equal(s.get("TOTALS"), r["totals"][:C])
    np.testing.assert_array_equal(s.get("OFFSETS"), r["offsets"][:C])
    for nm in PNAMES:
        np.testing.assert_array_equal(s.get(nm.upper()), o.array(nm)[:total], err_msg=nm)


def test_resample_seeding_and_capacity(M):
    o = _scene(particles=False)
    s = _upload(M, o)
    s.n = 0
    total = M.sparse.resample(s, seeding=True)
    r = o.resample(seeding=True, seed=1)
    assert total == r["total"]
    for nm in PNAMES:
        np.testing.assert_array_equal(s.get(nm.upper()), o.array(nm)[:total], err_msg=nm)
    s2 = M.SparseBasin(_basin(M), dx=DX, blocks=64, capacity=1000)
    s2.set_topology(o.array("slot"), o.array("where"), o.nslots, o.free_list())
    s2.set("PHI_F", o.array("phi_f"))
    s2.set("PHI_C", o.array("phi_c"))
    with pytest.raises(M.MaelstromError, match="the narrow band needs [0-9]+ particles, more than the 1000 allocated"):
        M.sparse.resample(s2, seeding=True)


def test_p2g_parity(M):
    print("case: sparse p2g", flush=True)
    o = _scene()
    s = _upload(M, o)
    dt = np.float32(0.1)
    M.sparse.particles2grid(s, dt)
    momwt = o.p2g(dt, momwt=True)
    F = o.nslots * BFACES
    for a, nm in enumerate("UVW"):
        got_m = s.get("MOM_" + nm).astype(np.float64) / 2.0 ** 32
        got_w = s.get("WT_" + nm).astype(np.float64) / 2.0 ** 32
        scale = max(1.0, np.abs(momwt[a]).max())
        np.testing.assert_allclose(got_m, momwt[a][:F], rtol=0, atol=1e-6 * scale)
        np.testing.assert_allclose(got_w, momwt[3 + a][:F], rtol=0, atol=1e-6 * max(1.0, momwt[3 + a].max()))
    np.testing.assert_array_equal(s.get("PHI_F"), _fine(o, "phi_f", BCELLS))
    np.testing.assert_array_equal(s.get("PHI_C"), o.array("phi_c").reshape(s.shape("PHI_C")))
    for nm in "uvw":
        exp = _fine(o, nm + "_f", BFACES)
        got = s.get(nm.upper() + "_F")
        err = np.linalg.norm(got - exp) / max(np.linalg.norm(exp), 1e-30)
        assert err < 1e-6, (nm, err)
        np.testing.assert_allclose(s.get(nm.upper() + "_C").ravel(), o.array(nm + "_c"), rtol=0,
                                   atol=1e-5 * max(1.0, np.abs(exp).max()))


def _orders(n, seed=5):
    """Particle orders that drive the tiled transfers down each of their paths: cell order
    (a box per workgroup), 128-particle runs from the two halves interleaved (every
    workgroup straddles two regions, so a box per wavefront), and a random order (mostly
    the direct path)."""
    half = n // 2 // 128 * 128
    runs = []
    for k in range(0, half, 128):
        runs += [np.arange(k, k + 128), np.arange(half + k, half + k + 128)]
    inter = np.concatenate(runs + [np.arange(2 * half, n)])
    rnd = np.random.default_rng(seed).permutation(n)
    return {"sorted": np.arange(n), "interleaved": inter, "random": rnd}


def _permute(M, s, o, order):
    for nm in PNAMES:
        s.set(nm.upper(), o.array(nm)[:o.n][order])


def test_p2g_deterministic(M):
    """The deposit and the splat give the same bits whatever the particle order (every
    path of the tiled kernels), and on repeated runs."""
    print("case: sparse p2g orders", flush=True)
    o = _scene()
    ref = None
    for name, order in _orders(o.n).items():
        for rep in range(2 if name == "sorted" else 1):
            s = _upload(M, o)
            _permute(M, s, o, order)
            M.sparse.particles2grid(s, 0.1)
            got = [s.get(x) for x in ("U_F", "V_F", "W_F", "PHI_F")]
            if ref is None:
                ref = got
                continue
            for nm, x, y in zip(("U_F", "V_F", "W_F", "PHI_F"), ref, got):
                np.testing.assert_array_equal(x, y, err_msg="%s order %s" % (nm, name))


def test_g2p_bitexact(M):
    print("case: sparse g2p", flush=True)
    o = _scene()
    rng = np.random.default_rng(2)
    for nm in ("u_f", "v_f", "w_f"):
        arr = o.array(nm)
        arr[:] += rng.normal(0, 0.2, arr.size).astype(np.float32)
    fw = o.array("fw_f")
    fw[rng.random(fw.size) < 0.01] = 0
    s = _upload(M, o)
    dt = np.float32(0.08)
    orders = _orders(o.n)
    M.sparse.grid2particles(s, dt)
    before = {nm: o.array(nm)[:o.n].copy() for nm in PNAMES}
    o.g2p(dt, 9.81)
    for nm in PNAMES:
        np.testing.assert_array_equal(s.get(nm.upper()), o.array(nm)[:o.n], err_msg=nm)
    for name in ("interleaved", "random"):   # the wavefront and direct paths of the tiled gather
        order = orders[name]
        s = _upload(M, o)
        for nm in PNAMES:
            s.set(nm.upper(), before[nm][order])
        M.sparse.grid2particles(s, dt)
        for nm in PNAMES:
            np.testing.assert_array_equal(s.get(nm.upper()), o.array(nm)[:o.n][order], err_msg="%s order %s" % (nm, name))


# ── projection ───────────────────────────────────────────────────────────────

def _projection_inputs(o, seed=11):
    rng = np.random.default_rng(seed)
    for nm in ("u_f", "v_f", "w_f", "u_c", "v_c", "w_c"):
        arr = o.array(nm)
        arr[:] += rng.normal(0, 0.3, arr.size).astype(np.float32)
    o.solidify()
    o.restrict_velocity()
    o.array("P_f")[:] = 0
    o.array("P_c")[:] = 0


def _pressure(o, s):
    ns = o.nslots
    Pe = np.concatenate([o.array("P_f")[:ns * BCELLS], o.array("P_c")])
    Pg = np.concatenate([s.get("P_F"), s.get("P_C").ravel()])
    return Pg, Pe


def _check_projected(o, s, bar=1e-5):
    """P, u, v, w (fine pool and coarse grid) within `bar` relative L2 of the oracle's
    converged FP64 projection, P also in max-abs; valid-face flags exact."""
    Pg, Pe = _pressure(o, s)
    err = rel_l2(Pg, Pe)
    print("P rel L2 %.3e" % err, flush=True)
    assert err < bar, err
    np.testing.assert_allclose(Pg, Pe, rtol=0, atol=2e-5 * np.abs(Pe).max())
    for nm in "uvw":
        gf, ef = s.get(nm.upper() + "_F"), _fine(o, nm + "_f", BFACES)
        gc, ec = s.get(nm.upper() + "_C").ravel(), o.array(nm + "_c")
        e = rel_l2(np.concatenate([gf, gc]), np.concatenate([ef, ec]))
        assert e < bar, (nm, e)
        np.testing.assert_array_equal(s.get("F" + nm.upper() + "_F"), _fine(o, "f" + nm + "_f", BFACES))
        np.testing.assert_array_equal(s.get("F" + nm.upper() + "_C").ravel(), o.array("f" + nm + "_c"))


@pytest.mark.parametrize("dx,policy,tol,refine,bar", [(0.6, 0, 1e-6, 0, 1e-5), (0.6, 2, 1e-6, 0, 1e-5),
                                                       (0.6, 1, 1e-6, 0, 5e-5), (0.6, 1, 1e-7, 10, 1e-5),
                                                       (0.2, 2, 1e-6, 0, 1e-5)])
def test_project_parity(M, dx, policy, tol, refine, bar):
    """The pressure to `tol` against the FP64 oracle's solve of the same system, then the
    faces and their flags: within `bar` relative L2. The flexible CG (policies 0, 2)
    minimises the error's energy norm and is within 1e-5 at a 1e-6 residual; FGMRES alone
    (policy 1) minimises the residual, and at the same 1e-6 residual its error is about
    2e-5, so it meets 1e-5 at 1e-7 (with FP64 residual replacement). At 0.2 m the grid
    holds a few hundred blocks (48x48x32 coarse cells)."""
    kw = dict(blocks=2000, capacity=3_000_000) if dx < 0.5 else {}
    o = _scene(dx=dx, **kw)
    print("case: sparse project dx %.2f policy %d tol %g, %d blocks" % (dx, policy, tol, o.nslots), flush=True)
    _projection_inputs(o)
    s = _upload(M, o)
    gamma = 0.01
    st = M.sparse.project(s, 0.1, gamma=gamma, tol=tol, maxiter=400, policy=policy, boost=1.0 if policy == 0 else 1.5,
                          refine=refine)
    assert st.rel_residual <= tol * 1.01, st.rel_residual
    o.project(gamma=gamma, tol=1e-12, maxiter=600, policy=0)
    _check_projected(o, s, bar)


def test_fcg_iterations_track_oracle(M):
    """The flexible CG with the V-cycle preconditioner (sflip.jl:1867-1904) takes the
    oracle's FP32 iteration count on the same system, within one: a wrong preconditioner
    still converges, only in more iterations."""
    o = _scene()
    _projection_inputs(o, seed=17)
    s = _upload(M, o)
    st = M.sparse.project(s, 0.1, gamma=0.01, tol=1e-6, maxiter=400, policy=0, boost=1.0)
    r = o.project(gamma=0.01, tol=1e-6, maxiter=400, policy=0, fp32=True)
    print("fcg iterations gpu %d oracle %d" % (st.fcg_iterations, r["iterations"]), flush=True)
    assert abs(st.fcg_iterations - r["iterations"]) <= 1


# tolerance of z = M r against the FP64 oracle's V-cycle, by V-cycle storage precision
VCYCLE_BAR = {0: 1e-5, 1: 3e-3, 2: 2e-2}


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_vcycle_matches_oracle(M, precision):
    """z = M r, the composite V-cycle alone (sflip.jl:1218-1311, damped Jacobi, Galerkin
    levels, coarsest solve), against the oracle's FP64 V-cycle on the same r, for FP32,
    FP16 and BF16 storage; and the path with the first launches queued ahead of the
    Krylov loop's host read gives the plain path's bits."""
    from maelstrom_amd import _lib as L
    o = _scene()
    s = _upload(M, o)
    ns = o.nslots
    rng = np.random.default_rng(21)
    tf, tc = _fine(o, "types_f", BCELLS), o.array("types_c")
    rf = np.where(tf == 1, rng.normal(0, 1, tf.size), 0).astype(np.float32)
    rc = np.where(tc == 1, rng.normal(0, 1, tc.size), 0).astype(np.float32)
    s.set("RHS_F", rf)
    s.set("RHS_C", rc)
    opts = M.sparse.solve_options(precision=precision, boost=1.0)
    zs = []
    for queued in (0, 1):
        s.ctx.check(L.lib().mael_sparse_vcycle(s.h, C.byref(opts), queued))
        zs.append(np.concatenate([s.get("P_F"), s.get("P_C").ravel()]))
    assert np.array_equal(zs[0].view(np.uint32), zs[1].view(np.uint32))
    zf, zc = o.vcycle(np.concatenate([rf, np.zeros(o.array("P_f").size - rf.size, np.float32)]), rc, sweeps=opts.sweeps,
                      coarse=opts.coarse, boost=1.0)
    ze = np.concatenate([zf[:ns * BCELLS], zc])
    err = rel_l2(zs[0], ze)
    print("V-cycle precision %d rel L2 vs FP64 oracle %.3e" % (precision, err), flush=True)
    assert err < VCYCLE_BAR[precision], err


@pytest.mark.parametrize("precision,refine", [(1, 0), (2, 0), (0, 10), (1, 10), (2, 10)])
def test_project_mixed_precision(M, precision, refine):
    """The FP16 / BF16 hierarchy (FP32 CG, FP64 sums and coarsest solve), with and without
    FP64 iterative refinement, reaches the same 1e-6 true residual and the same pressure
    and faces as the FP64 oracle, within 1e-5 relative L2."""
    print("case: sparse project precision %d refine %d" % (precision, refine), flush=True)
    o = _scene()
    _projection_inputs(o, seed=13)
    s = _upload(M, o)
    st = M.sparse.project(s, 0.1, gamma=0.005, tol=1e-6, maxiter=400, policy=2, boost=1.5, precision=precision,
                          refine=refine)
    assert st.rel_residual <= 1e-6 * 1.01, st.rel_residual
    o.project(gamma=0.005, tol=1e-12, maxiter=600, policy=0)
    _check_projected(o, s)


def test_project_errors(M):
    o = _scene()
    s = _upload(M, o)
    nan = np.full(o.nslots * BFACES, np.nan, np.float32)
    s.set("U_F", nan)
    with pytest.raises(M.MaelstromError, match="Nonfinite pressure right-hand side"):
        M.sparse.project(s, 0.1, gamma=0.0, policy=0)
    with pytest.raises(M.MaelstromError, match="Nonfinite pressure RHS"):
        M.sparse.project(s, 0.1, gamma=0.0, policy=1)


@pytest.mark.parametrize("surface", [False, True])
def test_steps_keep_volume(M, surface):
    """initialize! and three whole steps (advection-reflection halves, the stir, trapped
    air, extrapolation): finite, the volume kept to a few percent."""
    print("case: sparse steps, surface stir %s" % surface, flush=True)
    b = _basin(M)
    stir = M.Stir(b, waves=M.narutowaves(), race=M.Race(eddy=3.0), surface=surface)
    s = M.SparseBasin(b, dx=DX, blocks=64, capacity=300000, cfl=4, stir=stir)
    M.sparse.initialize(s)
    v0 = s.volume0
    tol = M.sparse.pressure_defaults(s)["tol"]   # 5e-4 with the surface stir, else 1e-4
    for _ in range(3):
        st = M.sparse.step(s, 0.1)
        assert st.rel_residual <= tol * 1.01, st.rel_residual
    v = M.sparse.volume(s)
    assert 0.95 < v / v0 < 1.05
    assert np.isfinite(s.get("U_F")).all() and s.n > 0


def test_filltrapped_matches_flood_fill(M):
    """`filltrapped!` (sflip.jl:950-990): the line sweeps, repeated until nothing changes,
    reach exactly the AIR cells joined to the top layer's AIR cells through face-adjacent
    AIR cells; every other uncovered AIR coarse cell becomes LIQUID. Checked against a
    connected-component labelling on random coarse types, with no fine blocks."""
    import scipy.ndimage as ndi
    from maelstrom_amd import _lib as L
    print("case: sparse filltrapped vs flood fill", flush=True)
    SOLID, LIQUID, AIR = 0, 1, 2
    s = M.SparseBasin(_basin(M), dx=DX, blocks=64, capacity=10000)   # no fine blocks yet
    shape = s.shape("TYPES_C")
    rng = np.random.default_rng(11)
    for trial in range(3):
        # AIR pockets in liquid, with solid bits and open shafts to the top
        t = np.where(rng.random(shape) < 0.45 + 0.1 * trial, AIR, LIQUID).astype(np.uint8)
        t[rng.random(shape) < 0.05] = SOLID
        s.set("TYPES_C", t)
        s.set("PHI_C", np.zeros(shape, np.float32))
        s.ctx.check(L.lib().mael_sparse_filltrapped(s.h, 256))
        got = s.get("TYPES_C").reshape(shape)
        lab, _ = ndi.label(t == AIR, structure=ndi.generate_binary_structure(3, 1))
        top = np.unique(lab[-1][lab[-1] > 0])
        reached = np.isin(lab, top) & (t == AIR)
        want = t.copy()
        want[(t == AIR) & ~reached] = LIQUID
        assert np.array_equal(got, want), (trial, int((got != want).sum()))


def test_redistance_bitexact(M):
    """`redistance!` (sflip.jl:833-919) on the GPU gives the oracle's level set bit for bit,
    from a tilted plane's distance (which the oracle keeps where the sweeps are exact)."""
    from test_oracle_pins import _plane_phi
    o = _scene()
    _plane_phi(o, np.array([0.48, -0.36, 0.8]), -5.0)
    s = _upload(M, o)
    M.sparse.redistance(s, 3)
    o.redistance(3)
    np.testing.assert_array_equal(s.get("PHI_F").view(np.uint32), _fine(o, "phi_f", BCELLS).view(np.uint32))
    np.testing.assert_array_equal(s.get("PHI_C").ravel().view(np.uint32), o.array("phi_c").view(np.uint32))
