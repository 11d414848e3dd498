"""Independent NumPy restatements of reference functions, used to pin the C++ oracle.

Each function is written from the reference's Julia text (file:line cited), not from the
oracle, in FP32 with the reference's operation order, so that agreement with the oracle
is bitwise. Arrays follow numpy's (z, y, x) order for grids; pool arrays are flat.
`slip` arguments plant a named mistake, for the tests that show a pin would catch it.
"""
import numpy as np

f32 = np.float32
BLOCK, RATIO, BCELLS, BFACES = 16, 4, 4096, 4352
NB6 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]   # flip.jl:84-92 neighbour order
FACEOFF = [(f32(0), f32(0.5), f32(0.5)), (f32(0.5), f32(0), f32(0.5)), (f32(0.5), f32(0.5), f32(0))]  # flip.jl:71-73
COFF = (f32(0.5), f32(0.5), f32(0.5))                                                                  # flip.jl:74


# ── the block table (sflip.jl:92-128) ────────────────────────────────────────

class Topo:
    """`Topology` (sflip.jl:73-78) with its index maps, vectorised over integer arrays."""

    def __init__(self, grid, slot, where):
        self.lo = np.array([grid.lo[0], grid.lo[1], grid.lo[2]], f32)
        self.dx = f32(grid.dx)
        self.nb = tuple(int(x) for x in grid.nb)
        self.radius, self.floor = f32(grid.radius), f32(grid.floor)
        self.slot = np.asarray(slot, np.int64).ravel()
        self.where = np.asarray(where, np.int64).reshape(-1, 3)
        self.nc = tuple(4 * b for b in self.nb)

    @property
    def cdx(self):
        return f32(RATIO) * self.dx                                    # sflip.jl:58 coarsedx

    def cellslot(self, I, J, K):                                       # sflip.jl:92-97
        bi, bj, bk = I >> 4, J >> 4, K >> 4
        ok = (bi >= 0) & (bi < self.nb[0]) & (bj >= 0) & (bj < self.nb[1]) & (bk >= 0) & (bk < self.nb[2])
        lin = np.where(ok, bi + self.nb[0] * (bj + self.nb[1] * bk), 0)
        return np.where(ok, self.slot[lin], 0)

    def fineat(self, I, J, K):                                         # sflip.jl:98-104 (0-based, -1 none)
        s = self.cellslot(I, J, K)
        return np.where(s > 0, (s - 1) * BCELLS + (I & 15) + 16 * ((J & 15) + 16 * (K & 15)), -1)

    @staticmethod
    def faceindex(s, a, i, j, k):                                      # sflip.jl:105-109 (0-based)
        if a == 0:
            f = i + 17 * (j + 16 * k)
        elif a == 1:
            f = i + 16 * (j + 17 * k)
        else:
            f = i + 16 * (j + 16 * k)
        return (s - 1) * BFACES + f

    def finefaceat(self, a, I, J, K):                                  # sflip.jl:110-126 (0-based, -1 none)
        s = self.cellslot(I, J, K)
        e = [0, 0, 0]
        e[a] = 1
        Im, Jm, Km = I - e[0], J - e[1], K - e[2]
        sm = self.cellslot(Im, Jm, Km)
        apart = ((Im >> 4) != (I >> 4)) | ((Jm >> 4) != (J >> 4)) | ((Km >> 4) != (K >> 4))
        li = np.full_like(Im, 16) if a == 0 else Im & 15
        lj = np.full_like(Jm, 16) if a == 1 else Jm & 15
        lk = np.full_like(Km, 16) if a == 2 else Km & 15
        own = self.faceindex(s, a, I & 15, J & 15, K & 15)
        low = self.faceindex(sm, a, li, lj, lk)
        return np.where(s > 0, own, np.where((sm > 0) & apart, low, -1))

    def covered(self, I, J, K):                                        # sflip.jl:127-128
        return self.cellslot(I * RATIO, J * RATIO, K * RATIO) != 0

    def issolid(self, px, py, pz):                                     # sflip.jl:60
        return (px * px + py * py > self.radius * self.radius) | (pz < self.floor)

    def coarsecentres(self):                                           # sflip.jl:66 coarsecentre, (z, y, x)
        n1, n2, n3 = self.nc
        K, J, I = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
        c = self.cdx
        return (I, J, K), tuple(self.lo[a] + ((X.astype(f32) + COFF[a]) * c) for a, X in enumerate((I, J, K)))


# ── refinement marks (sflip.jl:384-455) ──────────────────────────────────────

def _shift(A, d, fill):
    """B[z, y, x] = A at the neighbour (x + d0, y + d1, z + d2), `fill` where out of range."""
    out = np.full_like(A, fill)
    n3, n2, n1 = A.shape
    sx, sy, sz = d
    dst = (slice(max(0, -sz), n3 - max(0, sz)), slice(max(0, -sy), n2 - max(0, sy)), slice(max(0, -sx), n1 - max(0, sx)))
    src = (slice(max(0, sz), n3 - max(0, -sz)), slice(max(0, sy), n2 - max(0, -sy)), slice(max(0, sx), n1 - max(0, -sx)))
    out[dst] = A[src]
    return out


def _markblocks(t, hit, IJK):
    I, J, K = IJK
    m = np.zeros(np.prod(t.nb), np.uint8)
    lin = (I[hit] >> 2) + t.nb[0] * ((J[hit] >> 2) + t.nb[1] * (K[hit] >> 2))
    m[lin] = 1
    return m


def marksurface(t, phic, slip=None):
    """sflip.jl:384-399: a non-solid coarse cell within 1.5 coarse cells of zero, or whose
    sign differs from a non-solid 6-neighbour's, marks its block."""
    IJK, (px, py, pz) = t.coarsecentres()
    solid = t.issolid(px, py, pz)
    c = phic
    surf = np.abs(c) < f32(1.5) * t.cdx
    if slip == "radius":
        surf = np.abs(c) <= f32(1.5) * t.cdx + t.cdx
    for d in NB6:
        inr = _shift(np.ones_like(c, bool), d, False)
        nsolid = _shift(solid, d, True)
        nphi = _shift(c, d, f32(0))
        ok = inr & (True if slip == "solid" else ~nsolid)
        surf |= ok & ((nphi < 0) != (c < 0))
    return _markblocks(t, surf & ~solid, IJK)


def markflow(t, uc, vc, wc, phic, vary, slip=None):
    """sflip.jl:401-424: a water cell whose cell velocity (face means) jumps by more than
    `vary` to a water neighbour marks its block."""
    IJK, (px, py, pz) = t.coarsecentres()
    solid = t.issolid(px, py, pz)
    h = f32(0.5)
    cvel = (h * (uc[:, :, :-1] + uc[:, :, 1:]), h * (vc[:, :-1, :] + vc[:, 1:, :]), h * (wc[:-1] + wc[1:]))
    water = (phic < 0) & ~solid
    jump = np.zeros_like(phic)
    for d in NB6:
        ok = _shift(np.ones_like(water), d, False) & _shift(water, d, False)
        e = [_shift(cv, d, f32(0)) - cv for cv in cvel]
        j = np.sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2])
        jump = np.where(ok, np.maximum(jump, j), jump)
    hit = (jump >= f32(vary)) if slip == "geq" else (jump > f32(vary))
    return _markblocks(t, hit & water, IJK)


def markparticles(t, x, y, z, reach, slip=None):
    """sflip.jl:426-445: each finite particle marks the blocks of the eight corners of the
    box `reach` round it, clamped to the grid."""
    ok = np.isfinite(x) & np.isfinite(y) & np.isfinite(z)
    x, y, z = x[ok], y[ok], z[ok]
    S = f32(BLOCK) * t.dx
    m = np.zeros(np.prod(t.nb), np.uint8)
    reach = f32(reach)
    for c in (0, 1):
        for b in (0, 1):
            for a in (0, 1):
                sg = [(s if slip == "onecorner" else 2 * s - 1) for s in (a, b, c)]
                q = [((p - t.lo[ax]) + f32(sg[ax]) * reach) / S for ax, p in enumerate((x, y, z))]
                bi = [np.clip(np.floor(qq).astype(np.int64), 0, t.nb[ax] - 1) for ax, qq in enumerate(q)]
                m[bi[0] + t.nb[0] * (bi[1] + t.nb[1] * bi[2])] = 1
    return m


def dilate(t, m):
    """sflip.jl:447-455: one ring of the 26 neighbours."""
    A = m.reshape(t.nb[2], t.nb[1], t.nb[0])
    out = np.zeros_like(A)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                out |= _shift(A, (di, dj, dk), 0)
    return out.ravel()


def marks(t, phic, uc, vc, wc, vary, parts=None, reach=None, rings=0):
    """sflip.jl:516-526: surface, flow, particles, then `rings` dilations."""
    m = marksurface(t, phic) | markflow(t, uc, vc, wc, phic, vary)
    if parts is not None and len(parts[0]):
        m |= markparticles(t, *parts, reach)
    for _ in range(rings):
        m = dilate(t, m)
    return m


# ── sampling (interpolation.jl:3-17, sflip.jl:132-155) ───────────────────────

def trilinear(A, qx, qy, qz):
    """interpolation.jl:3-17 over arrays of index-space points (A in (z, y, x))."""
    A = A.astype(f32) if A.dtype != f32 else A
    n3, n2, n1 = A.shape
    x = np.clip(qx, f32(0), f32(n1 - 1)).astype(f32)
    y = np.clip(qy, f32(0), f32(n2 - 1)).astype(f32)
    z = np.clip(qz, f32(0), f32(n3 - 1)).astype(f32)
    i = np.minimum(x.astype(np.int64), n1 - 2)
    j = np.minimum(y.astype(np.int64), n2 - 2)
    k = np.minimum(z.astype(np.int64), n3 - 2)
    fx, fy, fz = x - i.astype(f32), y - j.astype(f32), z - k.astype(f32)
    one = f32(1)
    c00 = A[k, j, i] * (one - fx) + A[k, j, i + 1] * fx
    c10 = A[k, j + 1, i] * (one - fx) + A[k, j + 1, i + 1] * fx
    c01 = A[k + 1, j, i] * (one - fx) + A[k + 1, j, i + 1] * fx
    c11 = A[k + 1, j + 1, i] * (one - fx) + A[k + 1, j + 1, i + 1] * fx
    return (c00 * (one - fy) + c10 * fy) * (one - fz) + (c01 * (one - fy) + c11 * fy) * fz


def lerp8(c, fx, fy, fz):
    """sflip.jl:132-134 (c in the order 000, 100, 010, 110, 001, 101, 011, 111)."""
    one = f32(1)
    return (((c[0] * (one - fx) + c[1] * fx) * (one - fy) + (c[2] * (one - fx) + c[3] * fx) * fy) * (one - fz)
            + ((c[4] * (one - fx) + c[5] * fx) * (one - fy) + (c[6] * (one - fx) + c[7] * fx) * fy) * fz)


def cellsample(t, fine, coarse, px, py, pz):
    """sflip.jl:136-155: fine trilinear where all eight taps are fine, else the coarse grid's."""
    q = [((p - t.lo[a]) / t.dx) - COFF[a] for a, p in enumerate((px, py, pz))]
    fl = [np.floor(qq) for qq in q]
    I, J, K = (v.astype(np.int64) for v in fl)
    taps = [t.fineat(I + (m & 1), J + ((m >> 1) & 1), K + ((m >> 2) & 1)) for m in range(8)]
    allfine = np.all([tp >= 0 for tp in taps], axis=0)
    F = fine.astype(f32)
    vals = [np.where(tp >= 0, F[np.maximum(tp, 0)], f32(0)) for tp in taps]
    fv = lerp8(vals, q[0] - fl[0], q[1] - fl[1], q[2] - fl[2])
    cq = [((p - t.lo[a]) / t.cdx) - COFF[a] for a, p in enumerate((px, py, pz))]
    cv = trilinear(coarse, *cq)
    return np.where(allfine, fv, cv).astype(f32)


# ── new blocks from the coarse grid (sflip.jl:465-496) ───────────────────────

def _cell_local():
    c = np.arange(BCELLS)
    return c & 15, (c >> 4) & 15, c >> 8


def _face_local(a):
    f = np.arange(BFACES)
    if a == 0:
        r = f // 17
        return f % 17, r & 15, r >> 4
    if a == 1:
        r = f >> 4
        return f & 15, r % 17, r // 17
    return f & 15, (f >> 4) & 15, f >> 8


def prolongcells(t, coarse, s, flag=False):
    """sflip.jl:468-478: slot `s` (1-based) cells at the coarse grid's trilinear sample."""
    i, j, k = _cell_local()
    w = t.where[s - 1]
    P = [t.lo[a] + ((X + w[a] * 16).astype(f32) + COFF[a]) * t.dx for a, X in enumerate((i, j, k))]
    v = trilinear(coarse, *[((P[a] - t.lo[a]) / t.cdx) - COFF[a] for a in range(3)])
    return (v > f32(0.999)).astype(np.uint8) if flag else v


def prolongfaces(t, coarse, s, a, flag=False):
    """sflip.jl:480-496: slot `s` faces along `a` at the coarse faces' trilinear sample."""
    i, j, k = _face_local(a)
    w = t.where[s - 1]
    off = FACEOFF[a]
    P = [t.lo[b] + ((X + w[b] * 16).astype(f32) + off[b]) * t.dx for b, X in enumerate((i, j, k))]
    v = trilinear(coarse, *[((P[b] - t.lo[b]) / t.cdx) - off[b] for b in range(3)])
    out = (v > f32(0.999)).astype(np.uint8) if flag else v
    idx = Topo.faceindex(s, a, i, j, k) - (s - 1) * BFACES
    res = np.empty(BFACES, out.dtype)
    res[idx] = out
    return res


# ── fine to coarse (sflip.jl:580-612) ────────────────────────────────────────

def restrictcells(t, coarse, fine, flag=False, slip=None):
    """sflip.jl:583-594: covered coarse cells take the mean of their 64 fine cells
    (flags: 1 only if all 64 are set)."""
    out = coarse.copy()
    (I, J, K), _ = t.coarsecentres()
    cov = t.covered(I, J, K)
    I, J, K = I[cov], J[cov], K[cov]
    acc = np.zeros(I.shape, f32)
    F = fine.astype(f32)
    for c in range(4):
        for b in range(4):
            for a in range(4):
                if slip == "index":
                    idx = t.fineat(RATIO * I + a, RATIO * J + c, RATIO * K + b)
                else:
                    idx = t.fineat(RATIO * I + a, RATIO * J + b, RATIO * K + c)
                acc = acc + F[idx]
    out[cov] = (acc >= f32(63.5)).astype(out.dtype) if flag else acc / f32(64)
    return out


def restrictfaces(t, coarse, fine, a, flag=False):
    """sflip.jl:596-612: a coarse face takes the mean of its 16 fine faces where all 16 exist
    (flags: 1 only if all are set)."""
    out = coarse.copy()
    n3, n2, n1 = coarse.shape
    K, J, I = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    acc = np.zeros(I.shape, f32)
    found = np.zeros(I.shape, np.int64)
    F = fine.astype(f32)
    for q in range(4):
        for p in range(4):
            FI = RATIO * I + (0 if a == 0 else p)
            FJ = RATIO * J + (p if a == 0 else (0 if a == 1 else q))
            FK = RATIO * K + (0 if a == 2 else q)
            idx = t.finefaceat(a, FI, FJ, FK)
            ok = idx >= 0
            acc = np.where(ok, acc + F[np.maximum(idx, 0)], acc)
            found += ok
    full = found == 16
    out[full] = (acc[full] >= f32(15.5)).astype(out.dtype) if flag else acc[full] / f32(16)
    return out


# ── the particles' level set and the combine (flip.jl:466-585, sflip.jl:790-806) ──

def strength(speed, dt, h):
    """flip.jl:466-470."""
    c = np.clip((speed * f32(dt)) / f32(h), f32(0), f32(1)).astype(f32)
    return (c * c) * (f32(3) - f32(2) * c)


def ahead(g):
    """flip.jl:482-486."""
    h2 = (f32(0.25) * g) * g
    num = h2 * (f32(9) / f32(16) - h2 * (f32(7) / f32(10) - (h2 * f32(3)) / f32(7)))
    den = f32(27) / f32(64) + h2 * (f32(3) / f32(16) - h2 * (f32(3) / f32(20) + h2 / f32(7)))
    return num / den


def splat_positions(x, y, z, vx, vy, vz, lag, dt, dx):
    """flip.jl:546-547: where each particle counts for the splat."""
    sp = np.sqrt((vx * vx + vy * vy) + vz * vz)
    s = lag + ahead(strength(sp, dt, dx)) * f32(dt)
    return x + s * vx, y + s * vy, z + s * vz


def splat_bruteforce(lo, dx, ncells, cellidx, xp, slip=None):
    """flip.jl:539-560 by brute force: every cell against every particle, keeping those whose
    cell (floor of the displaced position) is within one cell on every axis; the distance
    to the cell centre less dx/2, minimised. `cellidx(I, J, K)` maps global cells to the
    output's flat index (-1 for none). Returns (value per output cell, +inf where none)."""
    n1, n2, n3 = ncells
    K, J, I = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    out_idx = cellidx(I, J, K)
    keep = out_idx >= 0
    I, J, K, out_idx = I[keep], J[keep], K[keep], out_idx[keep]
    ci = [np.floor((xp[a] - lo[a]) / dx).astype(np.int64) for a in range(3)]
    cen = [lo[a] + (X.astype(f32) + f32(0.5)) * dx for a, X in enumerate((I, J, K))]
    best = np.full(out_idx.max() + 1 if out_idx.size else 0, np.inf, f32)
    radius = f32(0.5) * dx
    for p0 in range(0, xp[0].size, 512):
        sl = slice(p0, p0 + 512)
        near = np.ones((I.size, xp[0][sl].size), bool)
        for a, X in enumerate((I, J, K)):
            lim = 2 if slip == "reach" else 1
            near &= np.abs(X[:, None] - ci[a][None, sl]) <= lim
        e = [cen[a][:, None] - xp[a][None, sl] for a in range(3)]
        d = np.sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]) - radius
        d = np.where(near, d, np.inf).min(axis=1).astype(f32)
        np.minimum.at(best, out_idx, d)
    return best


def combine(phif, u0, mom, wt, r, full):
    """flip.jl:567-585 / sflip.jl:796-806 per face, from the face's level set value."""
    r, full = f32(r), f32(full)
    return np.where(phif >= -r, (mom + np.maximum(full - wt, f32(0)) * u0) / np.maximum(full, wt), u0).astype(f32)


def dense_face_phi(phi, a):
    """flip.jl:575-578: the face's level set, the mean of the cells either side (the one
    cell at the grid's edges). phi in (z, y, x); the result has the face array's shape."""
    n3, n2, n1 = phi.shape
    shp = [n3, n2, n1]
    shp[2 - a] += 1
    K, J, Ii = np.meshgrid(np.arange(shp[0]), np.arange(shp[1]), np.arange(shp[2]), indexing="ij")
    b = [Ii, J, K]
    bb = [Ii.copy(), J.copy(), K.copy()]
    bb[a] = bb[a] - 1
    inb = (bb[0] >= 0) & (bb[0] < n1) & (bb[1] >= 0) & (bb[1] < n2) & (bb[2] >= 0) & (bb[2] < n3)
    ina = (b[0] < n1) & (b[1] < n2) & (b[2] < n3)
    fa = np.where(inb, phi[np.clip(bb[2], 0, n3 - 1), np.clip(bb[1], 0, n2 - 1), np.clip(bb[0], 0, n1 - 1)],
                  phi[np.minimum(K, n3 - 1), np.minimum(J, n2 - 1), np.minimum(Ii, n1 - 1)])
    fb = np.where(ina, phi[np.minimum(K, n3 - 1), np.minimum(J, n2 - 1), np.minimum(Ii, n1 - 1)], fa)
    return f32(0.5) * (fa + fb)
