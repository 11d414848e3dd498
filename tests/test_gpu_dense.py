"""GPU parity of the dense hot path: each phase through the C ABI against the oracle on
the same inputs.

Bars: bit-exact for binning, ranks, counts, totals, offsets and the resampled
particle arrays (spawned records included: positions and gathered velocities follow
the reference's unfused FP32 operation order on both sides); bit-exact G2P; the level
set of P2G bit-exact and its velocities within 1e-5 relative L2 of the FP64 oracle;
pressure and projected velocities within 1e-5 relative L2 of the FP64 oracle's
converged solve.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PF = ["PX", "PY", "PZ", "VX", "VY", "VZ", "C0", "C1", "C2", "C3", "C4", "C5", "C6", "C7", "C8", "LAG"]


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# The scenes the dense parity tests run on: the reference's dense test basin
# (test/runtests.jl:49, narutobasin(radius 4, depth 2, Γ 15, core 1) at 0.2 m, 44x44x20)
# and the one-step workload (`Basin(radius 18, depth 12, Γ 225, core 2.5)` at 0.6 m, default
# rim wall 3 m: 64x64x32 cells, 4 multigrid levels).
SCENES = {
    "testbasin": dict(basin=dict(radius=4.0, depth=2.0, Γ=15.0, core=1.0, band=1.0, headroom=1.5), naruto=True,
                      top=1.5, dx=0.2),
    "basin64": dict(basin=dict(radius=18.0, depth=12.0, Γ=225.0, core=2.5), naruto=False, top=3.0, dx=0.6),
}


@pytest.fixture(params=sorted(SCENES))
def scene(request):
    return request.param


def make_basin(mael, orc, capacity=1_000_000, scene="testbasin"):
    """A GPU FlipBasin holding the oracle's balanced-vortex state of `scene`."""
    sc = SCENES[scene]
    kw = sc["basin"]
    b = mael.narutobasin(**kw) if sc["naruto"] else mael.Basin(**kw)
    s = mael.FlipBasin(b, dx=sc["dx"], capacity=capacity)
    g, phi, u, v, w, types = orc.dense_scene(kw["radius"], kw["depth"], sc["top"], sc["dx"], kw["Γ"], kw["core"])
    assert tuple(s.grid.n) == g.dims and tuple(s.grid.lo) == tuple(g.lo)
    if scene == "basin64":
        assert g.dims == (64, 64, 32) and s.grid.levels == 4
    print("case: dense scene %s %s" % (scene, g.dims), flush=True)
    for name, arr in (("PHI", phi), ("U", u), ("V", v), ("W", w), ("TYPES", types)):
        s.set(name, arr)
    return s, dict(g=g, phi=phi, u=u, v=v, w=w, types=types)


def gpu_particles(s):
    return {f: s.get(f) for f in PF}


def test_context_opens(mael, ctx):
    assert ctx.h


def test_resample_bit_exact(mael, orc, ctx, scene):
    s, sc = make_basin(mael, orc, scene=scene)
    g = sc["g"]
    args = (sc["phi"], sc["u"], sc["v"], sc["w"])
    # seeding from empty, then the second pass that drops the NaN seeds (flip.jl:336-339)
    s.n = 0
    n1 = mael.resample(s, seeding=True)
    ops, on1, *_ = orc.dense_resample(g, orc.empty_particles(1), 0, 1_000_000, *args, seeding=True, seed=1)
    assert n1 == on1 > 1000
    gp = gpu_particles(s)
    for f in PF:
        assert np.array_equal(gp[f].view(np.uint32), ops[f].view(np.uint32)), f
    n2 = mael.resample(s)
    ops2, on2, key, rank, counts, totals, offsets = orc.dense_resample(g, ops, on1, 1_000_000, *args, seed=1)
    assert n2 == on2
    assert np.array_equal(s.get("KEY", on1), key) and np.array_equal(s.get("RANK", on1), rank)
    for name, ref in (("COUNTS", counts), ("TOTALS", totals), ("OFFSETS", offsets)):
        assert np.array_equal(s.get(name).ravel(), ref), name
    gp = gpu_particles(s)
    for f in PF:
        assert np.array_equal(gp[f].view(np.uint32), ops2[f].view(np.uint32)), f
    # unsorted, crowded input: shuffle, jitter, pile a cell up past 2 nmin, drop some out of the band
    rng = np.random.default_rng(11)
    perm = rng.permutation(n2)
    ps = {f: ops2[f][perm].copy() for f in PF}
    ps["PX"] += rng.normal(0, 0.1, n2).astype(np.float32)
    ps["PZ"] += rng.normal(0, 0.1, n2).astype(np.float32)
    for f in ("PX", "PY", "PZ"):   # 40 on one kept particle's unjittered position
        ps[f][:40] = ops2[f][0]
    ps["PY"][40:60] = np.nan
    s.set_particles(ps, n2)
    s.steps = 6
    n3 = mael.resample(s)
    ops3, on3, key, rank, counts, totals, offsets = orc.dense_resample(g, ps, n2, 1_000_000, *args, seed=7)
    assert n3 == on3
    assert np.array_equal(s.get("KEY", n2), key) and np.array_equal(s.get("RANK", n2), rank)
    assert counts.max() > 16
    gp = gpu_particles(s)
    for f in PF:
        assert np.array_equal(gp[f].view(np.uint32), ops3[f].view(np.uint32)), f


def test_resample_deterministic(mael, orc, ctx):
    s, sc = make_basin(mael, orc)
    s.n = 0
    mael.resample(s, seeding=True)
    mael.resample(s)
    ps = gpu_particles(s)
    n = s.n
    rng = np.random.default_rng(3)
    perm = rng.permutation(n)
    ps = {f: ps[f][perm] for f in PF}
    outs = []
    for _ in range(2):
        s.set_particles(ps, n)
        mael.resample(s)
        outs.append((s.get("RANK", n), gpu_particles(s)))
    assert np.array_equal(outs[0][0], outs[1][0])
    for f in PF:
        assert np.array_equal(outs[0][1][f].view(np.uint32), outs[1][1][f].view(np.uint32))


def test_capacity_error(mael, orc, ctx):
    s, sc = make_basin(mael, orc, capacity=100)
    s.n = 0
    with pytest.raises(mael.MaelstromError, match=r"^the narrow band needs \d+ particles, more than the 100 allocated$") as e:
        mael.resample(s, seeding=True)
    assert e.value.code == -2
    assert s.n == 0


def seeded(mael, orc, scene="testbasin"):
    s, sc = make_basin(mael, orc, scene=scene)
    s.n = 0
    mael.resample(s, seeding=True)
    mael.resample(s)
    return s, sc


def test_p2g_parity_and_determinism(mael, orc, ctx, scene):
    s, sc = seeded(mael, orc, scene)
    n = s.n
    rng = np.random.default_rng(5)
    ps = gpu_particles(s)
    ps["LAG"] = rng.uniform(-0.03, 0.03, n).astype(np.float32)
    for b in range(9):
        ps["C%d" % b] = rng.normal(0, 0.4, n).astype(np.float32)
    ps["VZ"] = ps["VZ"] + rng.normal(0, 0.2, n).astype(np.float32)
    s.set_particles(ps, n)
    phi0 = sc["phi"] + np.float32(0.05)
    s.set("PHI0", phi0)
    for a, f in (("U0", "u"), ("V0", "v"), ("W0", "w")):
        s.set(a, sc[f])
    dt = 0.04
    mael.particles2grid(s, dt)
    phi, U, V, W, M, Wt = orc.dense_p2g(sc["g"], ps, n, dt, phi0, sc["u"], sc["v"], sc["w"])
    assert np.array_equal(s.get("PHI").view(np.uint32), phi.view(np.uint32))
    for name, ref in (("U", U), ("V", V), ("W", W)):
        assert rel_l2(s.get(name), ref) < 1e-5, name
    for a, ax in enumerate("UVW"):
        mom = s.get("MOM_" + ax).astype(np.float64) * 2.0 ** -32
        wt = s.get("WT_" + ax).astype(np.float64) * 2.0 ** -32
        assert rel_l2(mom, M[a]) < 1e-6 and rel_l2(wt, Wt[a]) < 1e-6
    first = [s.get(k) for k in ("MOM_U", "MOM_V", "MOM_W", "WT_U", "WT_V", "WT_W", "U", "V", "W", "PHI")]
    mael.particles2grid(s, dt)
    second = [s.get(k) for k in ("MOM_U", "MOM_V", "MOM_W", "WT_U", "WT_V", "WT_W", "U", "V", "W", "PHI")]
    for x, y in zip(first, second):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_g2p_bit_exact(mael, orc, ctx, scene):
    s, sc = seeded(mael, orc, scene)
    n = s.n
    rng = np.random.default_rng(8)
    fields, flags = [], []
    for name, f, fl in (("U", "u", "FU"), ("V", "v", "FV"), ("W", "w", "FW")):
        arr = (sc[f] + rng.normal(0, 0.3, sc[f].shape)).astype(np.float32)
        flag = (rng.random(sc[f].shape) > 0.002).astype(np.uint8)
        s.set(name, arr)
        s.set(fl, flag)
        fields.append(arr)
        flags.append(flag)
    ps = gpu_particles(s)
    mael.grid2particles(s, 0.05)
    ref = orc.dense_g2p(sc["g"], ps, n, *fields, *flags, 0.05, 9.81)
    gp = gpu_particles(s)
    for f in PF:
        assert np.array_equal(gp[f].view(np.uint32), ref[f].view(np.uint32)), f


def _projection_case(mael, orc, scene, dt=0.05, gamma=0.001):
    s, sc = make_basin(mael, orc, scene=scene)
    # gravity, then the walls (flip.jl:775-777): the projection's input velocities
    s.ctx.check(mael.lib().mael_dense_gravity(s.h, dt, 9.81))
    u, v, w = s.get("U"), s.get("V"), s.get("W")
    b = orc.dense_rhs(sc["g"], u, v, w, sc["types"], gamma)
    return s, sc, (u, v, w), b


@pytest.mark.parametrize("tol,refine", [(1e-6, 1), (1e-4, 0)])
def test_projection_parity(mael, orc, ctx, scene, tol, refine):
    """The pressure and the projected faces against the FP64 oracle's converged solve of
    the same system: within 1e-5 relative L2 at 1e-6 (FP64 refinement), within the
    residual's reach at the reference's FP32 PCG and its 1e-4 default; flags exact."""
    dt, gamma = 0.05, 0.001
    s, sc, (u, v, w), b = _projection_case(mael, orc, scene, dt, gamma)
    g = sc["g"]
    s.set("P", np.zeros(orc.cells_shape(g), np.float32))
    st = mael.project(s, dt, gamma=gamma, tol=tol, maxiter=200, refine=refine)
    assert st.rel_residual <= tol and st.iterations > 0
    assert rel_l2(s.get("RHS"), b) < 1e-6
    P, it, rr = orc.dense_pressure(g, sc["types"], sc["phi"], b, tol=1e-13, maxiter=500)
    # the solution error is bounded by the residual times the conditioning
    err = rel_l2(s.get("P"), P)
    assert err < (1e-5 if tol <= 1e-6 else 5e-3), err
    U, F = orc.dense_project_faces(g, P, sc["types"], sc["phi"], u, v, w)
    for name, ref in (("U", U[0]), ("V", U[1]), ("W", U[2])):
        assert rel_l2(s.get(name), ref) < (1e-5 if tol <= 1e-6 else 1e-3), name
    for name, ref in (("FU", F[0]), ("FV", F[1]), ("FW", F[2])):
        assert np.array_equal(s.get(name), ref), name


def test_fp32_pcg_tracks_oracle(mael, orc, ctx, scene):
    """The reference's plain FP32 PCG (no refinement, flip.jl:1075-1111), warm-started as
    in a step: the same iteration count as the oracle's FP32 PCG on the same system (a
    wrong preconditioner still converges, only in more iterations), and its recursive
    residual at the tolerance while the true one sits at FP32's floor."""
    dt, gamma = 0.05, 0.001
    s, sc, (u, v, w), b = _projection_case(mael, orc, scene, dt, gamma)
    g = sc["g"]
    rng = np.random.default_rng(4)
    P0 = (rng.normal(0, 1e-3, orc.cells_shape(g))).astype(np.float32)
    s.set("P", P0)
    st = mael.project(s, dt, gamma=gamma, tol=1e-6, maxiter=200)   # refine 0, warm start
    _, it, rr = orc.dense_pressure(g, sc["types"], sc["phi"], b, P0=P0, tol=1e-6, maxiter=200, fp32=True)
    assert abs(st.iterations - it) <= 1, (st.iterations, it)
    assert st.rel_residual_rec <= 1e-6
    assert st.rel_residual < 1e-4, st.rel_residual
    # the drop-in default is the reference's: no refinement, warm start
    opts = mael.dense.solve_options()
    assert opts.refine == 0 and opts.warm == 1 and opts.tol == 1e-4 and opts.sweeps == 2


@pytest.mark.parametrize("queued", [0, 1])
def test_vcycle_matches_oracle(mael, orc, ctx, scene, queued):
    """z = M r: the V-cycle preconditioner alone (flip.jl:986-1014) against the oracle's
    FP64 V-cycle on the same r; the path with the first launches queued ahead of a host
    read (as the CG loop runs it) gives the plain path's bits."""
    s, sc = make_basin(mael, orc, scene=scene)
    g = sc["g"]
    rng = np.random.default_rng(9)
    r = rng.normal(0, 1, orc.cells_shape(g)).astype(np.float32)
    r[sc["types"] != 1] = 0
    s.set("RHS", r)
    s.ctx.check(mael.lib().mael_dense_vcycle(s.h, None, queued))
    z = s.get("P")
    zo = orc.dense_vcycle(g, sc["types"], sc["phi"], r)
    print("V-cycle rel L2 vs FP64 oracle %.3e" % rel_l2(z, zo))
    assert rel_l2(z, zo) < 1e-5, rel_l2(z, zo)
    # the oracle's FP32 V-cycle takes every sum and division in the same order: the same
    # bits (the fused two-colour smoothing passes included)
    z32 = orc.dense_vcycle(g, sc["types"], sc["phi"], r, fp32=True)
    print("V-cycle vs FP32 oracle: %d of %d differ" % ((z.view(np.uint32) != z32.view(np.uint32)).sum(), z.size))
    assert np.array_equal(z.view(np.uint32), z32.view(np.uint32))
    s.ctx.check(mael.lib().mael_dense_vcycle(s.h, None, 1 - queued))
    assert np.array_equal(s.get("P").view(np.uint32), z.view(np.uint32))


def test_initialize_and_steps_keep_water(mael, ctx):
    """test/runtests.jl:48-57: a dense basin steps, keeps a finite speed and its volume within 5 %."""
    b = mael.narutobasin(radius=4.0, depth=2.0, Γ=15.0, core=1.0, band=1.0, headroom=1.5)
    s = mael.FlipBasin(b, dx=0.2, capacity=1_000_000)
    mael.initialize(s)
    assert s.n > 0
    for _ in range(3):
        st = mael.step(s, 1 / 48)
        assert st.rel_residual <= 1e-4
    assert np.isfinite(mael.maxspeed(s))
    assert 0.95 < mael.volume(s) / s.volume0 < 1.05
