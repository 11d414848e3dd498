"""Whitewater (whitewater.jl, flip.jl:1411-1511, sflip.jl:2320-2417): spray, foam and
bubbles from the narrow band, after Ihmsen, Akinci, Akinci & Teschner 2012.

`Whitewater(basin)` is the reference's `Whitewater(b, backend; ...)` (a ring buffer of
diffuse particles on the device), `whitewater(w, s, dt)` its `whitewater!(w, s, u, v, dt)`
for a dense or sparse basin (the emission reads the basin's resampled, cell-sorted
particles directly), and `WhitewaterFrames` / `WhitewaterData` the bake's whitewater
files. All the work runs in the library (`mael_white_*`, `mael_*_whitewater`).
"""
import ctypes as C
import dataclasses
import json
import math

import numpy as np

from . import _lib
from ._lib import WhiteRulesC

SPRAY, FOAM, BUBBLE = 1, 2, 3   # bakefiles.jl:71-73
FIELDS = {"X": 0, "Y": 1, "Z": 2, "VX": 3, "VY": 4, "VZ": 5, "LIFE": 6, "KIND": 7}


@dataclasses.dataclass(frozen=True)
class WhitewaterRules:
    """whitewater.jl:30-44 `WhitewaterRules`."""
    air: tuple = (5.0, 20.0)
    crest: tuple = (2.0, 8.0)
    energy: tuple = (2.0, 20.0)
    kair: float = 40.0
    kcrest: float = 80.0
    spray: int = 6
    bubble: int = 20
    life: tuple = (2.0, 8.0)
    buoyancy: float = 2.0
    drag: float = 0.5
    kick: float = 0.0
    tossed: float = 1.0
    within: float = math.inf


class Whitewater:
    """whitewater.jl:255-300 `Whitewater(b, backend; floor = -b.depth, g = b.g, capacity,
    rules, stream = b)`: the basin's whitewater, riding the vortex (`stream=False` for
    none). Its box and reach (lo, hi, h) are kept for the bake's metadata."""

    def __init__(self, basin, capacity=4_000_000, rules=None, stream=True, ctx=None):
        self.basin, self.capacity = basin, int(capacity)
        self.rules = rules or WhitewaterRules()
        self.ctx = ctx or _lib.default_context()
        b = basin
        self.h = 2.0 * b.spacing
        L = b.layers * b.spacing
        self.lo = (-b.radius - L - self.h, -b.radius - L - self.h, -b.depth - L - self.h)
        self.hi = (b.radius + L + self.h, b.radius + L + self.h, b.headroom)
        self.floor, self.g, self.stream = -b.depth, b.g, bool(stream)
        r = self.rules
        c = WhiteRulesC()
        c.air[:], c.crest[:], c.energy[:], c.life[:] = r.air, r.crest, r.energy, r.life
        c.kair, c.kcrest, c.spray, c.bubble = r.kair, r.kcrest, r.spray, r.bubble
        c.buoyancy, c.drag, c.kick, c.tossed, c.within = r.buoyancy, r.drag, r.kick, r.tossed, r.within
        c.floor, c.gravity, c.stream = self.floor, self.g, 1 if self.stream else 0
        c.gamma_circ, c.core = b.Γ, b.core
        self.h_ = C.c_void_p()
        self.ctx.check(_lib.lib().mael_white_create(self.ctx.h, self.capacity, C.byref(c), C.byref(self.h_)))

    @property
    def head(self):
        h, f = C.c_int64(), C.c_int64()
        self.ctx.check(_lib.lib().mael_white_state(self.h_, C.byref(h), C.byref(f)))
        return h.value

    @property
    def frames(self):
        h, f = C.c_int64(), C.c_int64()
        self.ctx.check(_lib.lib().mael_white_state(self.h_, C.byref(h), C.byref(f)))
        return f.value

    def set_state(self, head, frames):
        self.ctx.check(_lib.lib().mael_white_set_state(self.h_, int(head), int(frames)))

    def get(self, name):
        dt = np.uint8 if name == "KIND" else np.float32
        out = np.empty(self.capacity, dt)
        self.ctx.check(_lib.lib().mael_white_download(self.h_, FIELDS[name], out.ctypes.data, out.size))
        return out

    def set(self, name, arr):
        dt = np.uint8 if name == "KIND" else np.float32
        a = np.ascontiguousarray(arr, dtype=dt).ravel()
        self.ctx.check(_lib.lib().mael_white_upload(self.h_, FIELDS[name], a.ctypes.data, a.size))

    def close(self):
        if self.h_:
            _lib.lib().mael_white_destroy(self.h_)
            self.h_ = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def whitewater(w, s, dt, substeps=4, near=1.5, above=0.25, below=0.75):
    """flip.jl:1496-1511 / sflip.jl:2401-2417 `whitewater!(w, s, u, v, dt; substeps, near,
    above, below)`: the frame's emission from the basin's particles and `substeps` steps
    of the diffuse particles."""
    from .sparse import SparseBasin
    fn = _lib.lib().mael_sparse_whitewater if isinstance(s, SparseBasin) else _lib.lib().mael_dense_whitewater
    s.ctx.check(fn(s.h, w.h_, float(dt), int(substeps), float(near), float(above), float(below)))
    return w


def pack_white(w):
    """whitewater.jl:366-379 `packwhite_kernel!`: the live diffuse particles, `(m, 4)`
    Float16 (x, y, z, kind) and `(m, 3)` Float16 velocities, in slot order."""
    out = np.empty((w.capacity, 4), np.float16)
    vel = np.empty((w.capacity, 3), np.float16)
    m = C.c_int64()
    w.ctx.check(_lib.lib().mael_white_pack(w.h_, out.ctypes.data, vel.ctypes.data, w.capacity, C.byref(m)))
    return out[:m.value], vel[:m.value]


class WhitewaterFrames:
    """whitewater.jl:315-350, 381-390 `WhitewaterFrames(path, b, backend; fps)`: each
    `record` runs the whitewater a frame on and appends its live diffuse particles to
    `path.whitewater` (Float16 `4 x N`) and their velocities to `path.whitewater.vel`
    (`3 x N`); `close` writes the counts (JSON, as `.whitewater.json`)."""

    def __init__(self, path, basin, fps=24, **kw):
        self.w = Whitewater(basin, **kw)
        self.path, self.fps = str(path), float(fps)
        self.io = open(self.path + ".whitewater", "wb")
        self.vio = open(self.path + ".whitewater.vel", "wb")
        self.counts = []

    def record(self, s):
        whitewater(self.w, s, 1.0 / self.fps)
        out, vel = pack_white(self.w)
        self.io.write(out.tobytes())
        self.vio.write(vel.tobytes())
        self.counts.append(int(len(out)))
        return self

    def close(self):
        self.io.close()
        self.vio.close()
        with open(self.path + ".whitewater.json", "w") as f:
            json.dump(dict(counts=self.counts, velocities=True), f)


class WhitewaterData:
    """bakefiles.jl:75-100 `WhitewaterData(path)`: the frames memory-mapped; `frame(k)`
    (1-based) is `(4, N)` Float16 (x, y, z, kind), `velocity(k)` `(3, N)`."""

    def __init__(self, path):
        path = str(path)
        meta = json.load(open(path + ".whitewater.json"))
        self.counts = list(meta["counts"])
        self.offsets = np.concatenate([[0], np.cumsum(np.asarray(self.counts, np.int64))])
        total = int(self.offsets[-1])
        self.data = np.memmap(path + ".whitewater", np.float16, "r", shape=(4 * total,)) if total else np.zeros(0, np.float16)
        self.velocities = None
        if meta.get("velocities", False):
            self.velocities = (np.memmap(path + ".whitewater.vel", np.float16, "r", shape=(3 * total,)) if total
                               else np.zeros(0, np.float16))

    def __len__(self):
        return len(self.counts)

    def frame(self, k):
        return self.data[4 * self.offsets[k - 1]:4 * self.offsets[k]].reshape(-1, 4).T

    def velocity(self, k):
        if self.velocities is None:
            raise ValueError("this bake's whitewater kept no velocities")
        return self.velocities[3 * self.offsets[k - 1]:3 * self.offsets[k]].reshape(-1, 3).T
