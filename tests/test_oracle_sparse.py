"""The sparse basin's CPU oracle, pinned by what fixes it without the reference:
known grid sizes, an independent restatement of the block allocator, the all-fine grid
against the (already pinned) dense oracle, affine reproduction, the operator's symmetry
and Galerkin structure, and direct solves."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O

BCELLS, BFACES = 4096, 4352


@pytest.fixture(scope="module")
def scene():
    return O.sparse_scene(dx=0.6, blocks=64, capacity=300000)


# ── the grid (sflip.jl:46-53) ────────────────────────────────────────────────

@pytest.mark.parametrize("dx,nb", [(0.05, (48, 48, 20)), (0.6, (4, 4, 4)), (0.1, (24, 24, 12)), (0.2, (12, 12, 8))])
def test_sparsegrid_sizes(dx, nb):
    g = O.sparsegrid(18.0, 12.0, 3.0, dx)
    assert tuple(g.nb) == nb
    S = np.float32(16) * np.float32(dx)
    assert np.isclose(g.lo[2], -12.0 - 4 * dx, atol=1e-5)
    assert np.isclose(g.lo[0], -nb[0] * S / 2, atol=1e-4)
    # x, y cover the cylinder and two fine cells; z from under the floor to the rim
    assert -g.lo[0] >= 18.0 + 2 * dx and g.lo[2] + nb[2] * S >= 3.0


def test_s3_dense_levels():
    """5 cm: coarse 192x192x80 halving to 12x12x5 (at most 1024 cells or an odd side)."""
    o = O.SparseOracle(dx=0.05, blocks=1, capacity=1)
    assert o.levels() == [(192, 192, 80), (96, 96, 40), (48, 48, 20), (24, 24, 10), (12, 12, 5)]


# ── the block allocator (sflip.jl:528-551), against an independent restatement ──

class Allocator:
    """A second restatement of refine!'s host allocator, written from its description:
    release every unwanted block first (in block order), then give wanted blocks a slot,
    taking the most recently freed one, or a new one past the high-water mark."""

    def __init__(self, nblocks, pool):
        self.slot = [0] * nblocks
        self.where = {}
        self.free = []
        self.nslots = 0
        self.pool = pool

    def refine(self, want, nb):
        need = sum(1 for m in want if m)
        if need > self.pool:
            raise ValueError(need)
        removed = [b for b in range(len(want)) if not want[b] and self.slot[b]]
        for b in removed:
            s = self.slot[b]
            self.free.append(s)
            self.where.pop(s)
            self.slot[b] = 0
        added = []
        for b in range(len(want)):
            if want[b] and not self.slot[b]:
                if self.free:
                    s = self.free.pop()
                else:
                    self.nslots += 1
                    s = self.nslots
                self.slot[b] = s
                self.where[s] = (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))
                added.append(s)
        return added


def _check_same(o, ref, nb):
    slot = o.array("slot")
    where = o.array("where").reshape(-1, 3)
    assert list(slot) == ref.slot
    assert o.nslots == ref.nslots
    for s in range(1, o.blocks + 1):
        exp = ref.where.get(s, (-1, -1, -1))
        assert tuple(where[s - 1]) == exp


def test_refine_allocator_matches_restatement():
    rng = np.random.default_rng(7)
    o = O.SparseOracle(dx=0.6, blocks=40, capacity=1)
    nb = tuple(o.grid.nb)
    nblk = int(np.prod(nb))
    ref = Allocator(nblk, 40)
    for it in range(60):
        want = (rng.random(nblk) < rng.uniform(0.05, 0.55)).astype(np.uint8)
        if want.sum() > 40:
            want[np.flatnonzero(want)[40:]] = 0
        added, removed = o.refine(want)
        rem_blocks = [b for b in range(nblk) if not want[b] and ref.slot[b]]
        rem_slots = [ref.slot[b] for b in rem_blocks]
        assert list(added) == ref.refine(want, nb)
        assert list(removed) == rem_slots
        _check_same(o, ref, nb)
        assert list(o.free_list()) == ref.free


def test_refine_lifo_and_release_first():
    o = O.SparseOracle(dx=0.6, blocks=3, capacity=1)
    nblk = int(np.prod(list(o.grid.nb)))
    w = np.zeros(nblk, np.uint8)
    w[[0, 5, 9]] = 1
    a, r = o.refine(w)
    assert list(a) == [1, 2, 3] and list(r) == []
    # a full pool: three leave and three arrive in one call, which fits only because
    # departures are released before arrivals are placed (sflip.jl:529-533)
    w2 = np.zeros(nblk, np.uint8)
    w2[[1, 2, 3]] = 1
    a, r = o.refine(w2)
    assert list(r) == [1, 2, 3]
    assert list(a) == [3, 2, 1]          # LIFO: the last freed slot is taken first
    slot = o.array("slot")
    assert (slot[1], slot[2], slot[3]) == (3, 2, 1)
    assert o.nslots == 3


def test_refine_pool_error_text():
    o = O.SparseOracle(dx=0.6, blocks=2, capacity=1)
    nblk = int(np.prod(list(o.grid.nb)))
    w = np.zeros(nblk, np.uint8)
    w[:3] = 1
    with pytest.raises(O.OracleError, match="the sparse grid needs 3 fine blocks, more than its 2 allocated"):
        o.refine(w)


# ── the all-fine grid equals the dense grid ──────────────────────────────────

def _allfine(dx=0.6, capacity=20000):
    o = O.SparseOracle(dx=dx, blocks=64, capacity=capacity)
    g = o.grid
    nblk = int(np.prod(list(g.nb)))
    o.refine(np.ones(nblk, np.uint8))
    dg = O.Grid()
    for a in range(3):
        dg.lo[a] = g.lo[a]
        dg.n[a] = 16 * g.nb[a]
    dg.dx, dg.radius, dg.floor, dg.levels = g.dx, g.radius, g.floor, 1
    return o, g, dg


def _pool_face_to_dense(o, g, a):
    """For every pool face along a: its global (I, J, K) and dense linear index."""
    where = o.array("where").reshape(-1, 3)[:o.nslots]
    f = np.arange(BFACES)
    if a == 0:
        i, r = f % 17, f // 17
        j, k = r & 15, r >> 4
    elif a == 1:
        i, r = f & 15, f >> 4
        j, k = r % 17, r // 17
    else:
        i, j, k = f & 15, (f >> 4) & 15, f >> 8
    n = [16 * int(x) for x in g.nb]
    n[a] += 1
    I = (where[:, 0:1] * 16 + i[None]).ravel()
    J = (where[:, 1:2] * 16 + j[None]).ravel()
    K = (where[:, 2:3] * 16 + k[None]).ravel()
    return I + n[0] * (J + n[1] * K)


def _random_particles(o, g, n, rng, margin=2.5):
    ps = {f: np.zeros(n, np.float32) for f in O.PFIELDS}
    S = 16 * np.float32(g.dx)
    for a, f in enumerate(("PX", "PY", "PZ")):
        lo = np.float32(g.lo[a]) + np.float32(margin * g.dx)
        hi = np.float32(g.lo[a]) + np.float32(g.nb[a]) * S - np.float32(margin * g.dx)
        ps[f] = rng.uniform(lo, hi, n).astype(np.float32)
    for f in ("VX", "VY", "VZ"):
        ps[f] = rng.normal(0, 2, n).astype(np.float32)
    for k in range(9):
        ps["C%d" % k] = rng.normal(0, 0.5, n).astype(np.float32)
    ps["LAG"] = rng.uniform(-0.05, 0.05, n).astype(np.float32)
    names = ["px", "py", "pz", "vx", "vy", "vz"] + ["c%d" % k for k in range(9)] + ["lag"]
    for nm, f in zip(names, O.PFIELDS):
        o.array(nm)[:n] = ps[f]
    o.n = n
    return ps


def test_allfine_p2g_equals_dense():
    """Deposits through the topology (finefaceat's canonical upper-block copy, then
    syncplanes) equal the dense grid's at every face copy."""
    rng = np.random.default_rng(3)
    o, g, dg = _allfine()
    ps = _random_particles(o, g, 3000, rng)
    dt = np.float32(0.1)
    momwt = o.p2g(dt, momwt=True)
    nd = [16 * int(x) for x in g.nb]
    phi0 = np.zeros(nd[::-1], np.float32)
    u0 = [np.zeros(O.face_shape(dg, a), np.float32) for a in range(3)]
    _, _, _, _, M, W = O.dense_p2g(dg, ps, 3000, dt, phi0, *u0)
    for a in range(3):
        idx = _pool_face_to_dense(o, g, a)
        m_pool = momwt[a][:idx.size]
        w_pool = momwt[3 + a][:idx.size]
        np.testing.assert_allclose(m_pool, M[a].ravel()[idx], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(w_pool, W[a].ravel()[idx], rtol=1e-12, atol=1e-12)


def test_allfine_g2p_equals_dense():
    rng = np.random.default_rng(4)
    o, g, dg = _allfine()
    ps = _random_particles(o, g, 2000, rng)
    vel = []
    for a, nm in enumerate("uvw"):
        shape = O.face_shape(dg, a)
        u = rng.normal(0, 1, shape).astype(np.float32)
        vel.append(u)
        idx = _pool_face_to_dense(o, g, a)
        o.array(nm + "_f")[:idx.size] = u.ravel()[idx]
        o.array("f" + nm + "_f")[:] = 1
    dt = np.float32(0.05)
    o.g2p(dt)
    flags = [np.ones(O.face_shape(dg, a), np.uint8) for a in range(3)]
    out = O.dense_g2p(dg, ps, 2000, *vel, *flags, dt)
    names = ["vx", "vy", "vz"] + ["c%d" % k for k in range(9)]
    for nm, f in zip(names, ["VX", "VY", "VZ"] + ["C%d" % k for k in range(9)]):
        np.testing.assert_array_equal(o.array(nm)[:2000], out[f])


def test_cellsample_and_finegather_affine(scene):
    """Through the topology, G2P of an affine face field returns it exactly (v and c = grad)."""
    o = O.sparse_scene(dx=0.6, blocks=64, capacity=300000, particles=False)
    g = o.grid
    where = o.array("where").reshape(-1, 3)[:o.nslots]
    A = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [-0.1, 0.2, 0.25]], np.float64)
    c0 = np.array([1.0, -2.0, 0.5])
    for a, nm in enumerate("uvw"):
        f = np.arange(BFACES)
        if a == 0:
            i, r = f % 17, f // 17
            j, k = r & 15, r >> 4
        elif a == 1:
            i, r = f & 15, f >> 4
            j, k = r % 17, r // 17
        else:
            i, j, k = f & 15, (f >> 4) & 15, f >> 8
        off = [0.5, 0.5, 0.5]
        off[a] = 0.0
        px = g.lo[0] + (where[:, 0:1] * 16 + i[None] + off[0]) * g.dx
        py = g.lo[1] + (where[:, 1:2] * 16 + j[None] + off[1]) * g.dx
        pz = g.lo[2] + (where[:, 2:3] * 16 + k[None] + off[2]) * g.dx
        val = c0[a] + A[a, 0] * px + A[a, 1] * py + A[a, 2] * pz
        o.array(nm + "_f")[:val.size] = val.ravel().astype(np.float32)
        o.array("f" + nm + "_f")[:] = 1
    # particles at the centres of fine cells deep inside live blocks (all 27 nodes fine)
    slot = o.array("slot")
    nb = list(g.nb)
    pts = []
    for s in range(o.nslots):
        bi, bj, bk = where[s]
        ok = True
        for d in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
            b = (bi + d[0], bj + d[1], bk + d[2])
            if not all(0 <= b[m] < nb[m] for m in range(3)) or slot[b[0] + nb[0] * (b[1] + nb[1] * b[2])] == 0:
                ok = False
        if ok or len(pts) < 50:
            for q in range(20):
                pts.append(g.lo + (np.array([bi, bj, bk]) * 16 + 3 + q % 10 * 1.0 + 0.37) * g.dx)
    pts = np.array(pts[:200], np.float32)
    n = len(pts)
    for m, nm in enumerate(("px", "py", "pz")):
        o.array(nm)[:n] = pts[:, m]
    o.n = n
    o.g2p(np.float32(0.1))
    for a, nm in enumerate(("vx", "vy", "vz")):
        exp = c0[a] + pts.astype(np.float64) @ A[a]
        np.testing.assert_allclose(o.array(nm)[:n], exp, rtol=0, atol=2e-4 * (1 + np.abs(exp).max()))
        for b in range(3):
            np.testing.assert_allclose(o.array("c%d" % (3 * a + b))[:n], A[a, b], atol=3e-3)


# ── the scene ────────────────────────────────────────────────────────────────

def test_scene_state(scene):
    g, ns, nfree, n = scene.info()
    assert ns > 0 and nfree == 0 and n > 0
    types_f = scene.array("types_f")[:ns * BCELLS]
    assert (types_f == 1).any() and (types_f == 2).any()
    # every particle is finite and at most a cell's jitter past the wall (spawns are
    # placed by their cell's centre, which is inside)
    px, py, pz = (scene.array(x)[:n] for x in ("px", "py", "pz"))
    assert np.isfinite(px).all()
    assert (np.sqrt(px * px + py * py) <= 18.0 + np.sqrt(2) * g.dx).all()
    # the volume is the basin's water to a few percent (cylinder of radius 18 m, depth 12 m,
    # less the funnel)
    assert 0.8 * np.pi * 18 ** 2 * 12 < scene.volume() < 1.02 * np.pi * 18 ** 2 * 12


def test_sparse_resample_invariants(scene):
    o = O.sparse_scene(dx=0.6, blocks=64, capacity=300000)
    n = o.n
    px = o.array("px")[:n].copy()
    r = o.resample(seed=5)
    key, rank, counts, totals, offsets = r["key"], r["rank"], r["counts"], r["totals"], r["offsets"]
    kept = key >= 0
    assert (np.bincount(key[kept], minlength=counts.size) == counts).all()
    # stable ranks: within a cell, by particle index
    for c in np.unique(key[kept])[:200]:
        members = np.flatnonzero(key == c)
        assert list(rank[members]) == list(range(members.size))
    assert (offsets == np.cumsum(totals)).all() and r["total"] == offsets[-1] == o.n
    assert (totals >= np.minimum(counts, 2 * o.nmin)).all()
    del px


# ── the composite operator (sflip.jl:1139-1216, 1323-1448) ───────────────────

def test_operator_symmetric_positive(scene):
    A = scene.matrix(16)
    rows = np.flatnonzero(np.diff(A.indptr))
    As = A[rows][:, rows]
    assert abs(As - As.T).max() < 1e-9
    vals = set(np.round(-As.data[As.data < 0], 6))
    assert vals <= {1.0, 0.4, 4.0}
    # SPD: Dirichlet faces at the surface; the smallest eigenvalue is positive
    lo = spla.eigsh(As.tocsc(), k=1, sigma=0, which="LM", return_eigenvectors=False)[0]
    assert lo > 0


def test_galerkin_level1(scene):
    """Level 1 is Pt A0 P exactly, P piecewise constant (each 8^3 cell its eight 16^3 children,
    coarse cells themselves): Braun et al.'s aggregation."""
    A0 = scene.matrix(16).tocsr()
    A1 = scene.matrix(8).tocsr()
    ns = scene.nslots
    nf0, nf1 = ns * 4096, ns * 512
    nc = A0.shape[0] - nf0
    c = np.arange(4096)
    i, j, k = c & 15, (c >> 4) & 15, c >> 8
    parent = (i >> 1) + 8 * ((j >> 1) + 8 * (k >> 1))
    rows = np.concatenate([np.arange(nf0), nf0 + np.arange(nc)])
    s = np.repeat(np.arange(ns), 4096)
    cols = np.concatenate([s * 512 + np.tile(parent, ns), nf1 + np.arange(nc)])
    P = sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(nf0 + nc, nf1 + nc))
    G = (P.T @ A0 @ P).tocsr()
    act = np.flatnonzero(np.diff(A1.indptr))
    D = (G[act][:, act] - A1[act][:, act])
    assert abs(D).max() < 1e-9
    # and nothing outside level 1's active rows
    other = np.setdiff1d(np.arange(G.shape[0]), act)
    assert abs(G[other]).max() < 1e-12 if other.size else True


def test_vcycle_linear_and_symmetric(scene):
    A = scene.matrix(16)
    act = np.flatnonzero(np.diff(A.indptr))
    ns = scene.nslots
    nf = ns * 4096
    rng = np.random.default_rng(1)

    def M(v):
        zf, zc = scene.vcycle(v[:nf], v[nf:])
        return np.concatenate([zf, zc])

    x = np.zeros(A.shape[0]); y = np.zeros(A.shape[0])
    x[act] = rng.normal(size=act.size); y[act] = rng.normal(size=act.size)
    Mx, My = M(x), M(y)
    np.testing.assert_allclose(M(2 * x - 3 * y), 2 * Mx - 3 * My, rtol=1e-9, atol=1e-9 * abs(Mx).max())
    assert abs(x @ My - y @ Mx) < 1e-8 * abs(x @ My)
    assert x @ Mx > 0


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_solvers_match_direct(policy):
    """Flexible CG, FGMRES and the CG-then-FGMRES policy converge to spsolve's solution."""
    o = O.sparse_scene(dx=0.6, blocks=64, capacity=300000)
    rng = np.random.default_rng(11)
    for nm in ("u_f", "v_f", "w_f", "u_c", "v_c", "w_c"):
        arr = o.array(nm)
        arr[:] += rng.normal(0, 0.3, arr.size).astype(np.float32)
    o.solidify()
    o.restrict_velocity()
    o.array("P_f")[:] = 0
    o.array("P_c")[:] = 0
    A = o.matrix(16)
    bf, bc = o.rhs(0.01)
    b = np.concatenate([bf, bc])
    act = np.flatnonzero(np.diff(A.indptr))
    x = np.zeros_like(b)
    x[act] = spla.spsolve(A[act][:, act].tocsc(), b[act])
    st = o.project(gamma=0.01, tol=1e-10, maxiter=400, policy=policy, restart=8)
    assert st["rel_residual"] < 1e-9
    ns = o.nslots
    P = np.concatenate([o.array("P_f")[:ns * 4096], o.array("P_c")]).astype(np.float64)
    np.testing.assert_allclose(P[act], x[act], rtol=0, atol=1e-5 * abs(x).max())
    # after the projection the fine and coarse cells' divergence is gamma h^2 (the rhs at
    # the new velocity vanishes), to the solve's FP32 velocity rounding
    bf2, bc2 = o.rhs(0.01)
    b2 = np.concatenate([bf2, bc2])
    assert abs(b2[act]).max() < 2e-5 * abs(b[act]).max() + 1e-6


def test_chunked_fp64_solve_matches_direct():
    """The continuable FP64 solve (solve64) in chunks of 4 iterations, each warm-started from
    the last chunk's FP64 pressure, converges to spsolve's solution to 1e-9 relative; a
    chunk that restarted from the FP32-rounded pressure instead could not go below the
    rounding of P."""
    o = O.sparse_scene(dx=0.6, blocks=64, capacity=300000)
    rng = np.random.default_rng(12)
    for nm in ("u_f", "v_f", "w_f", "u_c", "v_c", "w_c"):
        arr = o.array(nm)
        arr[:] += rng.normal(0, 0.3, arr.size).astype(np.float32)
    o.solidify()
    o.restrict_velocity()
    o.array("P_f")[:] = 0
    o.array("P_c")[:] = 0
    A = o.matrix(16)
    bf, bc = o.rhs(0.0)
    b = np.concatenate([bf, bc])
    act = np.flatnonzero(np.diff(A.indptr))
    x = np.zeros_like(b)
    x[act] = spla.spsolve(A[act][:, act].tocsc(), b[act])
    total, st = 0, None
    for k in range(60):
        st = o.solve64(gamma=0.0, tol=1e-12, maxiter=4, cont=k > 0)
        total += st["iterations"]
        if st["rel_residual"] <= 1e-12:
            break
    assert st["rel_residual"] <= 1e-12 and total > 8   # several chunks were needed
    pf, pc = o.p64()
    ns = o.nslots
    P = np.concatenate([pf[:ns * 4096], pc])
    assert np.linalg.norm(P[act] - x[act]) <= 1e-9 * np.linalg.norm(x[act])


def test_policy_error_texts():
    """A zero right-hand side returns at once; a nonfinite one raises the reference's text."""
    o = O.sparse_scene(dx=0.6, blocks=64, capacity=300000)
    for nm in ("u_f", "v_f", "w_f", "u_c", "v_c", "w_c"):
        o.array(nm)[:] = 0
    o.array("P_f")[:] = 0
    o.array("P_c")[:] = 0
    st = o.project(gamma=0.0, policy=0)
    assert st["iterations"] == 0
    o.array("u_f")[:o.nslots * 4352] = np.nan
    with pytest.raises(O.OracleError, match="Nonfinite pressure right-hand side"):
        o.project(gamma=0.0, policy=0)
    with pytest.raises(O.OracleError, match="Nonfinite pressure RHS"):
        o.project(gamma=0.0, policy=1)
