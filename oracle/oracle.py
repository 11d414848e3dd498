"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (oracle/liboracle.so).

Only tests/, the smoke check and bench.py's CPU-baseline leg use this module, as the
checker. It builds the oracle with `make -C oracle` on first use if needed.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

PFIELDS = ["PX", "PY", "PZ", "VX", "VY", "VZ", "C0", "C1", "C2", "C3", "C4", "C5", "C6", "C7", "C8", "LAG"]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIBPATH


class Grid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("dx", C.c_float), ("n", C.c_int32 * 3), ("radius", C.c_float),
                ("floor", C.c_float), ("levels", C.c_int32)]

    @property
    def dims(self):
        return tuple(int(x) for x in self.n)


class CParticles(C.Structure):
    _fields_ = [("x", C.c_void_p * 3), ("v", C.c_void_p * 3), ("c", C.c_void_p * 9), ("lag", C.c_void_p),
                ("n", C.c_int64), ("capacity", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        p, i32, i64, f32, f64, u32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double, C.c_uint32
        G, PP = C.POINTER(Grid), C.POINTER(CParticles)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_flipgrid": (C.c_int, [f32, f32, f32, f32, i32, G]),
            "orc_hash01": (f32, [i32, i32]),
            "orc_hash01_many": (None, [p, p, p, i64]),
            "orc_bspline": (None, [f32, C.POINTER(i32), p]),
            "orc_orderkey": (i32, [f32]),
            "orc_orderfloat": (f32, [i32]),
            "orc_temporal": (f32, [f32, f32]),
            "orc_strength": (f32, [f32, f32, f32]),
            "orc_ahead": (f32, [f32]),
            "orc_oseen": (f32, [f32, f32, f32]),
            "orc_balanced": (f32, [f32, f32, f32, f32]),
            "orc_theta64": (f64, [f64, f64]),
            "orc_trilinear": (f32, [p, i32, i32, i32, f32, f32, f32]),
            "orc_gather": (None, [p, G, i32, f32, f32, f32, p, p]),
            "orc_dense_resample": (C.c_int, [G, PP, PP, p, p, p, p, i32, i32, i32, u32, p, p, p, p, p, C.POINTER(i64)]),
            "orc_dense_p2g": (C.c_int, [G, PP, f32, i32, i32] + [p] * 14),
            "orc_dense_p2g_f32": (C.c_int, [G, PP, f32, i32, i32, p, p, p, p, p, p, p, p]),
            "orc_dense_g2p": (C.c_int, [G, PP, p, p, p, p, p, p, f32, f32]),
            "orc_dense_advect_particles": (C.c_int, [G, PP, p, p, p, p, p, p, f32, f32, i32]),
            "orc_dense_advect_grid": (None, [G, p, p, p, p, p, i32, i32, i32, p, f32]),
            "orc_dense_extrapolate": (None, [G, i32, p, p, i32]),
            "orc_dense_redistance": (None, [G, p, i32]),
            "orc_dense_forces": (None, [G, p, p, i32, p, p, p, f32, f32, f32]),
            "orc_stir": (None, [p, p, i32, f32, f32, f32, p, f32, p]),
            "orc_dense_whitewater": (None, [G, PP, p, p, p, p, p, p, p] + [p] * 8 + [i64, p, f64, i32, f32, f32, f32]),
            "orc_dense_classify": (None, [G, p, p]),
            "orc_dense_volume": (f64, [G, p]),
            "orc_dense_rhs": (None, [G, p, p, p, p, f32, p]),
            "orc_dense_apply": (None, [G, p, p, p, p]),
            "orc_dense_vcycle_f32": (None, [G, p, p, p, p, i32, i32]),
            "orc_dense_vcycle": (None, [G, p, p, p, p, i32, i32]),
            "orc_dense_matrix": (i64, [G, p, p, p, p, p]),
            "orc_dense_pressure": (i32, [G, p, p, p, p, f64, i32, i32, i32, C.POINTER(f64)]),
            "orc_dense_pressure_f32": (i32, [G, p, p, p, p, f64, i32, i32, i32, C.POINTER(f64)]),
            "orc_dense_project_faces": (None, [G, p, p, p, p, p, p, p, p, p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ── grid ─────────────────────────────────────────────────────────────────────

def flipgrid(radius, depth, top, dx, pad=2):
    g = Grid()
    lib().orc_flipgrid(radius, depth, top, dx, pad, C.byref(g))
    return g


def cells_shape(g):
    n1, n2, n3 = g.dims
    return (n3, n2, n1)


def face_shape(g, a):
    n1, n2, n3 = g.dims
    return (n3 + (a == 2), n2 + (a == 1), n1 + (a == 0))


# ── particles ────────────────────────────────────────────────────────────────

def empty_particles(capacity):
    return {f: np.zeros(capacity, dtype=np.float32) for f in PFIELDS}


def _cparts(ps, n, capacity):
    arrs = {f: np.array(ps[f], dtype=np.float32, copy=True) for f in PFIELDS}   # never alias the caller's
    for f in PFIELDS:
        if arrs[f].size < capacity:
            a = np.zeros(capacity, np.float32)
            a[:arrs[f].size] = arrs[f]
            arrs[f] = a
    cp = CParticles()
    for a in range(3):
        cp.x[a] = arrs[PFIELDS[a]].ctypes.data
        cp.v[a] = arrs[PFIELDS[3 + a]].ctypes.data
    for a in range(9):
        cp.c[a] = arrs[PFIELDS[6 + a]].ctypes.data
    cp.lag = arrs["LAG"].ctypes.data
    cp.n = n
    cp.capacity = capacity
    return cp, arrs


# ── scalar helpers ───────────────────────────────────────────────────────────

def hash01(i, j):
    i = _c(np.atleast_1d(i), np.int32)
    j = _c(np.atleast_1d(j), np.int32)
    i, j = np.broadcast_arrays(i, j)
    i, j = _c(i, np.int32), _c(j, np.int32)
    out = np.empty(i.shape, np.float32)
    lib().orc_hash01_many(i.ctypes.data, j.ctypes.data, out.ctypes.data, i.size)
    return out


def bspline(q):
    b = C.c_int32()
    w = np.zeros(3, np.float32)
    lib().orc_bspline(q, C.byref(b), w.ctypes.data)
    return b.value, w


def trilinear(A, q):
    A = _c(A, np.float32)
    n3, n2, n1 = A.shape
    return lib().orc_trilinear(A.ctypes.data, n1, n2, n3, q[0], q[1], q[2])


def gather(u, g, a, p):
    u = _c(u, np.float32)
    vel = np.zeros(1, np.float32)
    aff = np.zeros(3, np.float32)
    lib().orc_gather(u.ctypes.data, C.byref(g), a, p[0], p[1], p[2], vel.ctypes.data, aff.ctypes.data)
    return float(vel[0]), aff


# ── the dense phases ─────────────────────────────────────────────────────────

def dense_resample(g, ps, n, capacity, phi, u, v, w, band=4, nmin=8, seeding=False, seed=1):
    """flip.jl:1364-1382; returns (particles, n_new, key, rank, counts, totals, offsets)."""
    src, keep_src = _cparts(ps, n, max(n, 1))
    dst_arrs = empty_particles(capacity)
    dst, keep_dst = _cparts(dst_arrs, 0, capacity)
    C_ = int(np.prod(g.dims))
    key = np.zeros(max(n, 1), np.int32)
    rank = np.zeros(max(n, 1), np.int32)
    counts = np.zeros(C_, np.int32)
    totals = np.zeros(C_, np.int32)
    offsets = np.zeros(C_, np.int32)
    total = C.c_int64()
    phi, u, v, w = (_c(x, np.float32) for x in (phi, u, v, w))
    st = lib().orc_dense_resample(C.byref(g), C.byref(src), C.byref(dst), phi.ctypes.data, u.ctypes.data,
                                  v.ctypes.data, w.ctypes.data, band, nmin, 1 if seeding else 0, seed & 0xffffffff,
                                  key.ctypes.data, rank.ctypes.data, counts.ctypes.data, totals.ctypes.data,
                                  offsets.ctypes.data, C.byref(total))
    if st != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    nn = total.value
    out = {f: keep_dst[f][:nn].copy() for f in PFIELDS}
    return out, nn, key[:n], rank[:n], counts, totals, offsets


def dense_p2g(g, ps, n, dt, phi0, u0, v0, w0, band=4, nmin=8, fp32=False):
    """flip.jl:587-606 in FP64 (or FP32 with fp32=True); returns (phi, u, v, w, mom, wt)."""
    cp, keep = _cparts(ps, n, max(n, 1))
    phi0, u0, v0, w0 = (_c(x, np.float32) for x in (phi0, u0, v0, w0))
    phi = np.zeros(cells_shape(g), np.float32)
    U = [np.zeros(face_shape(g, a), np.float32) for a in range(3)]
    if fp32:
        lib().orc_dense_p2g_f32(C.byref(g), C.byref(cp), dt, band, nmin, phi0.ctypes.data, u0.ctypes.data,
                                v0.ctypes.data, w0.ctypes.data, phi.ctypes.data, *(x.ctypes.data for x in U))
        return phi, U[0], U[1], U[2], None, None
    M = [np.zeros(face_shape(g, a), np.float64) for a in range(3)]
    W = [np.zeros(face_shape(g, a), np.float64) for a in range(3)]
    lib().orc_dense_p2g(C.byref(g), C.byref(cp), dt, band, nmin, phi0.ctypes.data, u0.ctypes.data, v0.ctypes.data,
                        w0.ctypes.data, phi.ctypes.data, *(x.ctypes.data for x in U), *(x.ctypes.data for x in M),
                        *(x.ctypes.data for x in W))
    return phi, U[0], U[1], U[2], M, W


def dense_g2p(g, ps, n, u, v, w, fu, fv, fw, dt, gravity=9.81):
    """flip.jl:1246-1252; returns the updated particle arrays."""
    cp, keep = _cparts(ps, n, max(n, 1))
    u, v, w = (_c(x, np.float32) for x in (u, v, w))
    fu, fv, fw = (_c(x, np.uint8) for x in (fu, fv, fw))
    lib().orc_dense_g2p(C.byref(g), C.byref(cp), u.ctypes.data, v.ctypes.data, w.ctypes.data, fu.ctypes.data,
                        fv.ctypes.data, fw.ctypes.data, dt, gravity)
    return {f: keep[f][:n].copy() for f in PFIELDS}


def dense_advect_particles(g, ps, n, u, v, w, fu, fv, fw, dt, gravity=9.81, seed=0):
    """flip.jl:438-464 advectparticles_kernel!; returns the updated particle arrays."""
    cp, keep = _cparts(ps, n, max(n, 1))
    u, v, w = (_c(x, np.float32) for x in (u, v, w))
    fu, fv, fw = (_c(x, np.uint8) for x in (fu, fv, fw))
    lib().orc_dense_advect_particles(C.byref(g), C.byref(cp), u.ctypes.data, v.ctypes.data, w.ctypes.data, fu.ctypes.data,
                                     fv.ctypes.data, fw.ctypes.data, C.c_float(dt), C.c_float(gravity), C.c_int32(seed))
    return {f: keep[f][:n].copy() for f in PFIELDS}


def dense_advect_grid(g, u, v, w, src, off, dt):
    """flip.jl:356-369 advect_kernel!: `src` (samples at `off` in cells) traced back along (u, v, w)."""
    u, v, w, src = (_c(x, np.float32) for x in (u, v, w, src))
    out = np.zeros_like(src)
    n3, n2, n1 = src.shape
    o = np.asarray(off, np.float32)
    lib().orc_dense_advect_grid(C.byref(g), u.ctypes.data, v.ctypes.data, w.ctypes.data, src.ctypes.data, out.ctypes.data,
                                n1, n2, n3, o.ctypes.data, C.c_float(dt))
    return out


def dense_redistance(g, phi, iterations):
    """flip.jl:672-679 redistance!(s; iterations): returns the level set after `iterations`
    Jacobi sweeps."""
    phi = _c(phi, np.float32).copy()
    lib().orc_dense_redistance(C.byref(g), phi.ctypes.data, iterations)
    return phi


def dense_extrapolate(g, a, u, f, layers):
    """flip.jl:1194-1206 extrapolate! of component `a`: returns (u, flags)."""
    u, f = _c(u, np.float32).copy(), _c(f, np.uint8).copy()
    lib().orc_dense_extrapolate(C.byref(g), a, u.ctypes.data, f.ctypes.data, layers)
    return u, f


STIR_KEYS = ("radius", "band", "relax", "core", "gamma", "current_x", "current_y", "eddy", "eddyscale", "eddytime",
             "guide", "free", "turbulence", "boil", "reach", "origin_x", "origin_y", "offset")


def stir_params(**kw):
    """The Stir model's parameters in the order the oracle takes them (basin.jl:13-30 Basin
    and :69-80 Race defaults, for the Basin(radius 18, depth 12, Γ 225, core 2.5) scene)."""
    d = dict(radius=18.0, band=3.0, relax=0.3, core=2.5, gamma=225.0, current_x=0.0, current_y=0.0, eddy=0.6,
             eddyscale=3.5, eddytime=4.0, guide=4.0, free=0.6, turbulence=0.0, boil=0.5, reach=2.5, origin_x=0.0,
             origin_y=0.0, offset=0.0)
    d.update(kw)
    return np.array([d[k] for k in STIR_KEYS], np.float32)


def dense_forces(g, params, waves, u, v, w, dt, t, gravity=9.81):
    """flip.jl:786-798 forces! with `Stir`; returns (u, v, w)."""
    u, v, w = (_c(x, np.float32).copy() for x in (u, v, w))
    wv = _c(np.asarray(waves, np.float32).reshape(-1), np.float32)
    lib().orc_dense_forces(C.byref(g), _c(params, np.float32).ctypes.data, wv.ctypes.data if wv.size else None,
                           wv.size // 7, u.ctypes.data, v.ctypes.data, w.ctypes.data, C.c_float(dt), C.c_float(t),
                           C.c_float(gravity))
    return u, v, w


def stir(params, waves, x, y, z, vel, t):
    """basin.jl:153-181 at one point: the acceleration for velocity `vel`."""
    wv = _c(np.asarray(waves, np.float32).reshape(-1), np.float32)
    vel = _c(np.asarray(vel, np.float32), np.float32)
    acc = np.zeros(3, np.float32)
    lib().orc_stir(_c(params, np.float32).ctypes.data, wv.ctypes.data if wv.size else None, wv.size // 7, C.c_float(x),
                   C.c_float(y), C.c_float(z), vel.ctypes.data, C.c_float(t), acc.ctypes.data)
    return acc


def dense_quantize(g, phi, range_):
    """flip.jl:1859-1867 quantize_kernel!: the level set averaged over cells twice as wide
    (the 8 summed in the reference's order, x fastest), clamped to +-range, as bytes
    (round to nearest even). `phi` is (z, y, x); returns (z/2, y/2, x/2)."""
    phi = np.asarray(phi, np.float32)
    n3, n2, n1 = phi.shape
    s = np.zeros((n3 // 2, n2 // 2, n1 // 2), np.float32)
    for c in range(2):
        for b in range(2):
            for a in range(2):
                s = s + phi[c:2 * (n3 // 2):2, b:2 * (n2 // 2):2, a:2 * (n1 // 2):2]
    f = np.clip(np.float32(0.125) * s / np.float32(range_), np.float32(-1), np.float32(1))
    return np.rint(np.float32(127.5) * (f + np.float32(1))).astype(np.uint8)


def pack_band(ps, n, keep, throat):
    """flip.jl:1880-1893 packband_kernel!: each particle at the frame's time (x + lag v) kept
    if above -keep or within `throat` of the axis; the kept ones' positions and velocities
    as Float16 (round to nearest even), in particle order, each (m, 3)."""
    f = {k: np.asarray(ps[k][:n], np.float32) for k in PFIELDS}
    lag = f["LAG"]
    xp = [f["PX"] + lag * f["VX"], f["PY"] + lag * f["VY"], f["PZ"] + lag * f["VZ"]]
    kept = (xp[2] > -np.float32(keep)) | (xp[0] * xp[0] + xp[1] * xp[1] < np.float32(throat) * np.float32(throat))
    pos = np.stack([x[kept] for x in xp], axis=1).astype(np.float16)
    vel = np.stack([f[k][kept] for k in ("VX", "VY", "VZ")], axis=1).astype(np.float16)
    return pos, vel


WHITE_KEYS = ("X", "Y", "Z", "VX", "VY", "VZ", "LIFE", "KIND")


def white_rules(air=(5.0, 20.0), crest=(2.0, 8.0), energy=(2.0, 20.0), kair=40.0, kcrest=80.0, life=(2.0, 8.0),
                buoyancy=2.0, drag=0.5, kick=0.0, tossed=1.0, within=np.inf, floor=-12.0, gravity=9.81, stream=1,
                gamma=225.0, core=2.5):
    """whitewater.jl:30-44 WhitewaterRules with the whitewater's floor, gravity and stream,
    as the float array the oracle takes."""
    return np.array([*air, *crest, *energy, kair, kcrest, *life, buoyancy, drag, kick, tossed, within, floor, gravity,
                     stream, gamma, core], np.float32)


def empty_white(capacity):
    w = {k: np.zeros(capacity, np.float32) for k in WHITE_KEYS[:-1]}
    w["KIND"] = np.zeros(capacity, np.uint8)
    return w


def _white_args(w):
    return [w[k].ctypes.data for k in WHITE_KEYS]


def dense_whitewater(g, ps, n, offsets, totals, phi, u, v, w, rules, white, head, frames, dt, substeps=4, near=1.5,
                     above=0.25, below=0.75):
    """flip.jl:1496-1511 whitewater! on the dense basin; `white` (dict of arrays) is updated
    in place; returns (head, frames)."""
    cp, keep = _cparts(ps, n, max(n, 1))
    offsets, totals = _c(offsets, np.int32), _c(totals, np.int32)
    phi, u, v, w = (_c(x, np.float32) for x in (phi, u, v, w))
    cap = len(white["LIFE"])
    hf = np.array([head, frames], np.int64)
    lib().orc_dense_whitewater(C.byref(g), C.byref(cp), offsets.ctypes.data, totals.ctypes.data, phi.ctypes.data,
                               u.ctypes.data, v.ctypes.data, w.ctypes.data, _c(rules, np.float32).ctypes.data,
                               *_white_args(white), cap, hf.ctypes.data, float(dt), substeps, near, above, below)
    return int(hf[0]), int(hf[1])


def dense_classify(g, phi):
    phi = _c(phi, np.float32)
    t = np.zeros(cells_shape(g), np.uint8)
    lib().orc_dense_classify(C.byref(g), phi.ctypes.data, t.ctypes.data)
    return t


def dense_volume(g, phi):
    phi = _c(phi, np.float32)
    return lib().orc_dense_volume(C.byref(g), phi.ctypes.data)


def dense_rhs(g, u, v, w, types, gamma):
    u, v, w = (_c(x, np.float32) for x in (u, v, w))
    types = _c(types, np.uint8)
    b = np.zeros(cells_shape(g), np.float64)
    lib().orc_dense_rhs(C.byref(g), u.ctypes.data, v.ctypes.data, w.ctypes.data, types.ctypes.data, gamma,
                        b.ctypes.data)
    return b


def dense_apply(g, types, phi, x):
    types, phi, x = _c(types, np.uint8), _c(phi, np.float32), _c(x, np.float64)
    q = np.zeros(cells_shape(g), np.float64)
    lib().orc_dense_apply(C.byref(g), types.ctypes.data, phi.ctypes.data, x.ctypes.data, q.ctypes.data)
    return q


def dense_vcycle(g, types, phi, r, sweeps=2, coarse=40, fp32=False):
    """flip.jl:986-1014 vcycle!: z = M r, in FP64 (or FP32 with `fp32`)."""
    dt = np.float32 if fp32 else np.float64
    types, phi, r = _c(types, np.uint8), _c(phi, np.float32), _c(r, dt)
    z = np.zeros(cells_shape(g), dt)
    fn = lib().orc_dense_vcycle_f32 if fp32 else lib().orc_dense_vcycle
    fn(C.byref(g), types.ctypes.data, phi.ctypes.data, r.ctypes.data, z.ctypes.data, sweeps, coarse)
    return z


def dense_matrix(g, types, phi):
    """The fine level's operator as a scipy CSR matrix over all cells."""
    import scipy.sparse as sp
    types, phi = _c(types, np.uint8), _c(phi, np.float32)
    nnz = lib().orc_dense_matrix(C.byref(g), types.ctypes.data, phi.ctypes.data, None, None, None)
    rows = np.zeros(nnz, np.int64)
    cols = np.zeros(nnz, np.int64)
    vals = np.zeros(nnz, np.float64)
    lib().orc_dense_matrix(C.byref(g), types.ctypes.data, phi.ctypes.data, rows.ctypes.data, cols.ctypes.data,
                           vals.ctypes.data)
    N = int(np.prod(g.dims))
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def dense_pressure(g, types, phi, b, P0=None, tol=1e-12, maxiter=500, sweeps=2, coarse=40, fp32=False):
    """flip.jl:1075-1111 in FP64 (or FP32); returns (P, iterations, relres)."""
    types, phi = _c(types, np.uint8), _c(phi, np.float32)
    dt = np.float32 if fp32 else np.float64
    b = _c(b, dt)
    P = np.zeros(cells_shape(g), dt) if P0 is None else _c(P0, dt).copy()
    rr = C.c_double()
    fn = lib().orc_dense_pressure_f32 if fp32 else lib().orc_dense_pressure
    it = fn(C.byref(g), types.ctypes.data, phi.ctypes.data, b.ctypes.data, P.ctypes.data, tol, maxiter, sweeps, coarse,
            C.byref(rr))
    return P, it, rr.value


def dense_project_faces(g, P, types, phi, u, v, w):
    P, types, phi = _c(P, np.float64), _c(types, np.uint8), _c(phi, np.float32)
    U = [_c(x, np.float32).copy() for x in (u, v, w)]
    F = [np.zeros(face_shape(g, a), np.uint8) for a in range(3)]
    lib().orc_dense_project_faces(C.byref(g), P.ctypes.data, types.ctypes.data, phi.ctypes.data,
                                  *(x.ctypes.data for x in U), *(x.ctypes.data for x in F))
    return U, F


# ── a scene for CPU tests (the balanced vortex of initialize!, restated in numpy) ──

def _oseen(G, core, r):
    r = np.asarray(r, np.float32)
    q = r / np.float32(core)
    return np.float32(G) / (np.float32(2) * np.float32(np.pi) * np.maximum(r, np.float32(1e-4))) * (
        np.float32(1) - np.exp(-(q * q)))


def surface_table(G, core, gravity, rmax, n=2048, steps=2000):
    """sea.jl:48-58 `balanced` at n radii in [0, rmax] (FP32)."""
    r = np.linspace(np.float32(0), np.float32(rmax), n, dtype=np.float32)
    a = np.log(np.maximum(r, np.float32(1e-3)))
    z = np.log(np.float32(50) * np.float32(core) + np.float32(50) * np.maximum(r, np.float32(1)))
    h = (z - a) / np.float32(steps)
    s = np.zeros_like(r)
    for i in range(steps + 1):
        o = _oseen(G, core, np.exp(a + np.float32(i) * h))
        s += (np.float32(0.5) if i in (0, steps) else np.float32(1)) * (o * o) / np.float32(gravity)
    return -s * h


def dense_scene(radius, depth, top, dx, G, core, gravity=9.81):
    """The balanced vortex on a FlipGrid (flip.jl:298-330): phi, u, v, w, types."""
    g = flipgrid(radius, depth, top, dx)
    n1, n2, n3 = g.dims
    lo = np.array(g.lo, np.float32)
    dxf = np.float32(g.dx)
    rmax = np.float32(np.hypot(lo[0], lo[1]))
    eta = surface_table(G, core, gravity, rmax)
    k, j, i = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    px = lo[0] + (i.astype(np.float32) + np.float32(0.5)) * dxf
    py = lo[1] + (j.astype(np.float32) + np.float32(0.5)) * dxf
    pz = lo[2] + (k.astype(np.float32) + np.float32(0.5)) * dxf
    r = np.sqrt(px * px + py * py)
    t = np.clip(r / rmax, 0, 1).astype(np.float32) * np.float32(eta.size - 1)
    ia = np.minimum(t.astype(np.int64), eta.size - 2)
    f = t - ia.astype(np.float32)
    phi = (pz - (eta[ia] * (np.float32(1) - f) + eta[ia + 1] * f)).astype(np.float32)
    vel = []
    for a in range(3):
        sh = face_shape(g, a)
        kk, jj, ii = np.meshgrid(np.arange(sh[0]), np.arange(sh[1]), np.arange(sh[2]), indexing="ij")
        fx = lo[0] + (ii.astype(np.float32) + np.float32(0 if a == 0 else 0.5)) * dxf
        fy = lo[1] + (jj.astype(np.float32) + np.float32(0 if a == 1 else 0.5)) * dxf
        rr = np.sqrt(fx * fx + fy * fy)
        turn = np.where(rr > 1e-4, _oseen(G, core, rr) / np.maximum(rr, np.float32(1e-4)), 0).astype(np.float32)
        vel.append((-turn * fy if a == 0 else (turn * fx if a == 1 else np.zeros_like(fx))).astype(np.float32))
    types = dense_classify(g, phi)
    return g, phi, vel[0], vel[1], vel[2], types


# ── the sparse basin (sflip.jl) ──────────────────────────────────────────────

class SGrid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("dx", C.c_float), ("nb", C.c_int32 * 3), ("radius", C.c_float),
                ("floor", C.c_float)]

    @property
    def nc(self):
        return tuple(int(x) * 4 for x in self.nb)


_SPARSE_SIG = None


def _slib():
    global _SPARSE_SIG
    L = lib()
    if _SPARSE_SIG is None:
        p, i32, i64, f32, f64, u32, u8 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double, C.c_uint32, C.c_uint8
        P32, P64, PF64 = C.POINTER(i32), C.POINTER(i64), C.POINTER(f64)
        _SPARSE_SIG = {
            "orc_sparsegrid": (C.c_int, [f32, f32, f32, f32, i32, C.POINTER(SGrid)]),
            "orc_sparse_create": (p, [f32, f32, f32, f32, i32, i32, i64, i32, i32]),
            "orc_sparse_destroy": (None, [p]),
            "orc_sparse_info": (None, [p, C.POINTER(SGrid), P32, P32, P64, P32]),
            "orc_sparse_free_list": (None, [p, p]),
            "orc_sparse_set_n": (None, [p, i64]),
            "orc_sparse_set_topology_counts": (None, [p, i32, p, i32]),
            "orc_sparse_array": (C.c_int, [p, C.c_char_p, C.POINTER(p), P64]),
            "orc_sparse_marks": (None, [p, f32, f32, i32, i32, p]),
            "orc_sparse_refine": (C.c_int, [p, p, p, P32, p, P32]),
            "orc_sparse_init_coarse": (None, [p, p, i32, f32, f32, f32]),
            "orc_sparse_init_fine": (None, [p, p, i32, f32, f32, f32]),
            "orc_sparse_quantize": (None, [p, f32, f32, i32, i32, i32, p]),
            "orc_sparse_whitewater": (None, [p, p, p, p] + [p] * 8 + [i64, p, f64, i32, f32, f32, f32]),
            "orc_sparse_classify": (None, [p]),
            "orc_sparse_solidify": (None, [p]),
            "orc_sparse_restrict_velocity": (None, [p]),
            "orc_sparse_restrict_field": (C.c_int, [p, C.c_char_p]),
            "orc_sparse_redistance": (None, [p, i32]),
            "orc_sparse_volume": (f64, [p]),
            "orc_sparse_fill_flags": (None, [p, u8]),
            "orc_sparse_keep": (None, [p]),
            "orc_sparse_resample": (C.c_int, [p, i32, u32, P64, p, p, p, p, p]),
            "orc_sparse_p2g": (None, [p, f32, i32, p]),
            "orc_sparse_g2p": (None, [p, f32, f32]),
            "orc_sparse_rhs": (None, [p, f32, p, p]),
            "orc_sparse_matrix": (i64, [p, i32, p, p, p]),
            "orc_sparse_apply": (None, [p, p, p, p, p]),
            "orc_sparse_vcycle": (None, [p, p, p, p, p, i32, i32, f64]),
            "orc_sparse_levels": (i32, [p, p]),
            "orc_sparse_project": (C.c_int, [p, f32, f64, i32, i32, i32, f64, i32, i32, i32, P32, PF64, P32]),
            "orc_sparse_solve64": (C.c_int, [p, f32, f64, i32, i32, i32, f64, i32, P32, PF64]),
            "orc_sparse_p64": (C.c_int, [p, p, p]),
        }
        for name, (res, args) in _SPARSE_SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
    return L


def sparsegrid(radius, depth, top, dx, pad=2):
    g = SGrid()
    _slib().orc_sparsegrid(radius, depth, top, dx, pad, C.byref(g))
    return g


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_U8 = ("types_f", "types_c", "fu_f", "fu_c", "fv_f", "fv_c", "fw_f", "fw_c")
_I32 = ("slot", "where")


class SparseOracle:
    """One sparse basin's state in the CPU oracle; arrays are numpy views into it."""

    def __init__(self, radius=18.0, depth=12.0, top=3.0, dx=0.6, pad=2, blocks=64, capacity=400000, band=4, nmin=8):
        L = _slib()
        self.h = L.orc_sparse_create(radius, depth, top, dx, pad, blocks, capacity, band, nmin)
        self.band, self.nmin, self.blocks, self.capacity = band, nmin, blocks, capacity
        self.params = dict(radius=radius, depth=depth, top=top, dx=dx, pad=pad)

    def __del__(self):
        try:
            if self.h:
                _slib().orc_sparse_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def info(self):
        g, ns, nf, n, b = SGrid(), C.c_int32(), C.c_int32(), C.c_int64(), C.c_int32()
        _slib().orc_sparse_info(self.h, C.byref(g), C.byref(ns), C.byref(nf), C.byref(n), C.byref(b))
        return g, ns.value, nf.value, n.value

    @property
    def grid(self):
        return self.info()[0]

    @property
    def nslots(self):
        return self.info()[1]

    @property
    def n(self):
        return self.info()[3]

    @n.setter
    def n(self, v):
        _slib().orc_sparse_set_n(self.h, int(v))

    def free_list(self):
        k = self.info()[2]
        out = np.zeros(k, np.int32)
        _slib().orc_sparse_free_list(self.h, _ptr(out))
        return out

    def array(self, name):
        """A numpy view (no copy) of a state array; particle fields are the current buffer."""
        ptr, n = C.c_void_p(), C.c_int64()
        if _slib().orc_sparse_array(self.h, name.encode(), C.byref(ptr), C.byref(n)) != 0:
            raise KeyError(name)
        dt = np.uint8 if name in _U8 else (np.int32 if name in _I32 else np.float32)
        buf = (C.c_char * (n.value * np.dtype(dt).itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dt)

    def marks(self, vary=0.5, reach=None, rings=0, particles=True):
        g = self.grid
        if reach is None:
            reach = np.float32((np.ceil(8.0) + 3) * g.dx)
        out = np.zeros(int(np.prod(list(g.nb))), np.uint8)
        _slib().orc_sparse_marks(self.h, vary, reach, rings, 1 if particles else 0, _ptr(out))
        return out

    def refine(self, want):
        want = _c(want, np.uint8)
        added = np.zeros(self.blocks, np.int32)
        removed = np.zeros(self.blocks, np.int32)
        na, nr = C.c_int32(), C.c_int32()
        st = _slib().orc_sparse_refine(self.h, _ptr(want), _ptr(added), C.byref(na), _ptr(removed), C.byref(nr))
        if st != 0:
            raise OracleError(st, lib().orc_last_error().decode())
        return added[:na.value].copy(), removed[:nr.value].copy()

    def init_coarse(self, eta, rmax, G, core):
        eta = _c(eta, np.float32)
        _slib().orc_sparse_init_coarse(self.h, _ptr(eta), eta.size, rmax, G, core)

    def init_fine(self, eta, rmax, G, core):
        eta = _c(eta, np.float32)
        _slib().orc_sparse_init_fine(self.h, _ptr(eta), eta.size, rmax, G, core)

    def quantize(self, cell, range_):
        """sflip.jl:2427-2432 sparsequantize_kernel!, (8 nb2, 8 nb1, 8 nb0) bytes."""
        nb = self.grid.nb
        out = np.zeros((8 * nb[2], 8 * nb[1], 8 * nb[0]), np.uint8)
        _slib().orc_sparse_quantize(self.h, C.c_float(cell), C.c_float(range_), 8 * nb[0], 8 * nb[1], 8 * nb[2],
                                    out.ctypes.data)
        return out

    def whitewater(self, offsets, totals, rules, white, head, frames, dt, substeps=4, near=1.5, above=0.25, below=0.75):
        """sflip.jl:2401-2417 whitewater! on this basin's state; returns (head, frames)."""
        offsets, totals = _c(offsets, np.int32), _c(totals, np.int32)
        hf = np.array([head, frames], np.int64)
        _slib().orc_sparse_whitewater(self.h, offsets.ctypes.data, totals.ctypes.data, _c(rules, np.float32).ctypes.data,
                                      *_white_args(white), len(white["LIFE"]), hf.ctypes.data, float(dt), substeps, near,
                                      above, below)
        return int(hf[0]), int(hf[1])

    def classify(self):
        _slib().orc_sparse_classify(self.h)

    def solidify(self):
        _slib().orc_sparse_solidify(self.h)

    def restrict_velocity(self):
        _slib().orc_sparse_restrict_velocity(self.h)

    def restrict_field(self, name):
        """restrictcells (phi, P, types) or restrictfaces (u, v, w, fu, fv, fw) of one field."""
        if _slib().orc_sparse_restrict_field(self.h, name.encode()) != 0:
            raise KeyError(name)

    def redistance(self, iterations=None):
        _slib().orc_sparse_redistance(self.h, self.band + 3 if iterations is None else iterations)

    def volume(self):
        return _slib().orc_sparse_volume(self.h)

    def fill_flags(self, v=1):
        _slib().orc_sparse_fill_flags(self.h, v)

    def keep(self):
        _slib().orc_sparse_keep(self.h)

    def resample(self, seeding=False, seed=1):
        n, ns = self.n, self.nslots
        key, rank = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        cells = ns * 4096
        counts, totals, offsets = (np.zeros(max(cells, 1), np.int32) for _ in range(3))
        total = C.c_int64()
        st = _slib().orc_sparse_resample(self.h, 1 if seeding else 0, seed, C.byref(total), _ptr(key), _ptr(rank),
                                         _ptr(counts), _ptr(totals), _ptr(offsets))
        if st != 0:
            raise OracleError(st, lib().orc_last_error().decode())
        return dict(total=total.value, key=key[:n], rank=rank[:n], counts=counts[:cells], totals=totals[:cells],
                    offsets=offsets[:cells])

    def p2g(self, dt, fp32=False, momwt=False):
        out = np.zeros(6 * self.nslots * 4352, np.float64) if momwt else None
        _slib().orc_sparse_p2g(self.h, dt, 1 if fp32 else 0, _ptr(out))
        return None if out is None else out.reshape(6, -1)

    def g2p(self, dt, gravity=9.81):
        _slib().orc_sparse_g2p(self.h, dt, gravity)

    def rhs(self, gamma=0.0):
        nc = self.grid.nc
        bf = np.zeros(self.nslots * 4096, np.float64)
        bc = np.zeros(nc[0] * nc[1] * nc[2], np.float64)
        _slib().orc_sparse_rhs(self.h, gamma, _ptr(bf), _ptr(bc))
        return bf, bc

    def matrix(self, N=16):
        """The level's composite operator as scipy CSR over (fine pool, then coarse) unknowns."""
        import scipy.sparse as sp
        nnz = _slib().orc_sparse_matrix(self.h, N, None, None, None)
        r, c, v = np.zeros(nnz, np.int64), np.zeros(nnz, np.int64), np.zeros(nnz, np.float64)
        _slib().orc_sparse_matrix(self.h, N, _ptr(r), _ptr(c), _ptr(v))
        nc = self.grid.nc
        size = self.nslots * N ** 3 + nc[0] * nc[1] * nc[2]
        return sp.csr_matrix((v, (r, c)), shape=(size, size))

    def apply(self, xf, xc):
        xf, xc = _c(xf, np.float64), _c(xc, np.float64)
        qf, qc = np.zeros_like(xf), np.zeros_like(xc)
        _slib().orc_sparse_apply(self.h, _ptr(xf), _ptr(xc), _ptr(qf), _ptr(qc))
        return qf, qc

    def vcycle(self, rf, rc, sweeps=1, coarse=40, boost=1.0):
        rf, rc = _c(rf, np.float64), _c(rc, np.float64)
        zf, zc = np.zeros_like(rf), np.zeros_like(rc)
        _slib().orc_sparse_vcycle(self.h, _ptr(rf), _ptr(rc), _ptr(zf), _ptr(zc), sweeps, coarse, boost)
        return zf, zc

    def levels(self):
        n = _slib().orc_sparse_levels(self.h, None)
        dims = np.zeros(3 * n, np.int32)
        _slib().orc_sparse_levels(self.h, _ptr(dims))
        return [tuple(int(x) for x in dims[3 * l:3 * l + 3]) for l in range(n)]

    def project(self, gamma=0.0, tol=1e-6, maxiter=200, sweeps=1, coarse=40, boost=1.0, restart=8, policy=0, fp32=False):
        it, rr, fg = C.c_int32(), C.c_double(), C.c_int32()
        st = _slib().orc_sparse_project(self.h, gamma, tol, maxiter, sweeps, coarse, boost, restart, policy,
                                        1 if fp32 else 0, C.byref(it), C.byref(rr), C.byref(fg))
        if st != 0:
            raise OracleError(st, lib().orc_last_error().decode())
        return dict(iterations=it.value, rel_residual=rr.value, fgmres=bool(fg.value))

    def solve64(self, gamma=0.0, tol=1e-10, maxiter=8, sweeps=1, coarse=40, boost=1.0, cont=False):
        """The FP64 flexible-CG pressure solve alone (no face projection), continuable: with
        `cont` it starts from the FP64 pressure the previous call left. Returns the
        iterations and the true relative residual."""
        it, rr = C.c_int32(), C.c_double()
        st = _slib().orc_sparse_solve64(self.h, gamma, tol, maxiter, sweeps, coarse, boost, 1 if cont else 0,
                                        C.byref(it), C.byref(rr))
        if st != 0:
            raise OracleError(st, lib().orc_last_error().decode())
        return dict(iterations=it.value, rel_residual=rr.value)

    def p64(self):
        """The FP64 pressure of the last solve64: (fine pools, coarse grid)."""
        nf, nc = self.array("P_f").size, self.array("P_c").size
        fine, coarse = np.zeros(nf, np.float64), np.zeros(nc, np.float64)
        if _slib().orc_sparse_p64(self.h, _ptr(fine), _ptr(coarse)) != 0:
            raise OracleError(-1, "no FP64 pressure yet")
        return fine, coarse


def sparse_scene(radius=18.0, depth=12.0, top=3.0, dx=0.6, G=225.0, core=2.5, gravity=9.81, blocks=64,
                 capacity=400000, band=4, nmin=8, vary=0.5, particles=True):
    """The sparse basin's start (sflip.jl:681-713) in the oracle: the balanced surface and
    the swirl on the coarse grid, fine blocks round the surface, the band seeded (two
    resamples with seed 1), then a first particles-to-grid and redistance."""
    o = SparseOracle(radius, depth, top, dx, 2, blocks, capacity, band, nmin)
    g = o.grid
    rmax = np.float32(np.hypot(np.float32(g.lo[0]), np.float32(g.lo[1])))
    eta = surface_table(G, core, gravity, rmax)
    o.init_coarse(eta, rmax, G, core)
    o.refine(o.marks(vary=vary, particles=False))
    o.init_fine(eta, rmax, G, core)
    o.redistance(2 * band + 4)
    o.classify()
    o.solidify()
    o.fill_flags(1)
    o.n = 0
    if particles:
        o.resample(seeding=True, seed=1)
        o.resample(seed=1)
        o.keep()
        o.p2g(1.0, fp32=True)
        o.redistance()
        o.classify()
    o.eta, o.rmax, o.G, o.core = eta, rmax, G, core
    return o
