"""The bake's files (bakefiles.jl, flip.jl:1846-1951, sflip.jl:2419-2440): the band's
particles kept for drawing (`BandParticles`, read back as `ParticleFrames`) and each
frame's level set quantised to bytes (`LevelSets`, read back as `LevelSetFrames`).

The data files are the reference's: `path.particles` and `path.velocities` hold each
frame's kept particles as Float16 `3 x N` one frame after the other, `path.levelset`
each frame's bytes (x fastest). The packing and the quantisation run on the device
(`mael_*_pack_band`, `mael_*_quantize_levelset`); only what is kept crosses to the host.
The metadata the reference writes with Julia's `Serialization` (`.particles.jls`,
`.levelset.jls`) is JSON here (`.particles.json`, `.levelset.json`): the same fields, in a
form any reader can open.
"""
import ctypes as C
import json

import numpy as np

from . import _lib


def _is_sparse(s):
    from .sparse import SparseBasin
    return isinstance(s, SparseBasin)


def pack_band(s, keep=3.5, throat=3.0):
    """`packband_kernel!` over the basin's particles (flip.jl:1880-1893): the kept ones'
    positions at the frame's time and velocities, each `(m, 3)` Float16."""
    n = s.n
    pos = np.empty((max(n, 1), 3), np.float16)
    vel = np.empty((max(n, 1), 3), np.float16)
    m = C.c_int64()
    fn = _lib.lib().mael_sparse_pack_band if _is_sparse(s) else _lib.lib().mael_dense_pack_band
    s.ctx.check(fn(s.h, float(keep), float(throat), pos.ctypes.data, vel.ctypes.data, max(n, 1), C.byref(m)))
    return pos[:m.value], vel[:m.value]


def quantize_levelset(s, range_):
    """`record!(::LevelSets, s)`'s quantised level set (flip.jl:1859-1867, sflip.jl:2427-2432)
    as a (z, y, x) byte array."""
    if _is_sparse(s):
        nb = [int(x) for x in s.grid.nb]
        shape = (8 * nb[2], 8 * nb[1], 8 * nb[0])
        fn = _lib.lib().mael_sparse_quantize_levelset
    else:
        n = [int(x) for x in s.grid.n]
        shape = (n[2] // 2, n[1] // 2, n[0] // 2)
        fn = _lib.lib().mael_dense_quantize_levelset
    out = np.empty(shape, np.uint8)
    s.ctx.check(fn(s.h, float(range_), out.ctypes.data))
    return out


class BandParticles:
    """flip.jl:1876-1904 `BandParticles(path, b, backend; keep, throat)` with bakefiles.jl:16-45
    `Particles`: every `record` appends the frame's kept particles (and their velocities)
    to `path.particles` (`path.velocities`); `close` writes the counts."""

    def __init__(self, path, basin=None, keep=3.5, throat=3.0, velocities=True):
        self.path, self.keep, self.throat = str(path), float(keep), float(throat)
        self.counts = []
        self.io = open(self.path + ".particles", "wb")
        self.vio = open(self.path + ".velocities", "wb") if velocities else None

    def record(self, s):
        pos, vel = pack_band(s, self.keep, self.throat)
        self.io.write(np.ascontiguousarray(pos).tobytes())
        if self.vio is not None:
            self.vio.write(np.ascontiguousarray(vel).tobytes())
        self.counts.append(int(len(pos)))
        return self

    def close(self):
        self.io.close()
        if self.vio is not None:
            self.vio.close()
        with open(self.path + ".particles.json", "w") as f:
            json.dump(dict(counts=self.counts, keep=self.keep, throat=self.throat), f)


class ParticleFrames:
    """bakefiles.jl:47-66 `ParticleFrames(path)`: the frames memory-mapped; `frame(k)`
    (1-based, as the reference) is `(3, N)` Float16, `velocity(k)` likewise."""

    def __init__(self, path):
        path = str(path)
        meta = json.load(open(path + ".particles.json"))
        self.counts = list(meta["counts"])
        self.keep, self.throat = meta["keep"], meta["throat"]
        self.offsets = np.concatenate([[0], np.cumsum(3 * np.asarray(self.counts, np.int64))])
        total = int(self.offsets[-1])
        self.data = np.memmap(path + ".particles", np.float16, "r", shape=(total,)) if total else np.zeros(0, np.float16)
        try:
            self.velocities = (np.memmap(path + ".velocities", np.float16, "r", shape=(total,)) if total
                               else np.zeros(0, np.float16))
        except FileNotFoundError:
            self.velocities = None

    def __len__(self):
        return len(self.counts)

    def frame(self, k):
        return self.data[self.offsets[k - 1]:self.offsets[k]].reshape(-1, 3).T

    def velocity(self, k):
        return self.velocities[self.offsets[k - 1]:self.offsets[k]].reshape(-1, 3).T


class LevelSets:
    """flip.jl:1846-1874 / sflip.jl:2421-2440 `LevelSets(path, s)`: every `record` appends
    the frame's level set on cells twice as wide, clamped to +-(band + 2) dx, as bytes to
    `path.levelset`; `close` writes the grid."""

    def __init__(self, path, s):
        self.path = str(path)
        dx = float(s.grid.dx)
        self.lo = [float(x) for x in s.grid.lo]
        self.cell = float(np.float32(2.0) * np.float32(dx))
        self.range = float(np.float32((s.band + 2) * np.float32(dx)))
        self.io = open(self.path + ".levelset", "wb")
        self.dims = None
        self.frames = 0

    def record(self, s):
        q = quantize_levelset(s, self.range)
        self.dims = [q.shape[2], q.shape[1], q.shape[0]]
        self.io.write(q.tobytes())
        self.frames += 1
        return self

    def close(self):
        self.io.close()
        with open(self.path + ".levelset.json", "w") as f:
            json.dump(dict(dims=self.dims, lo=self.lo, cell=self.cell, range=self.range, frames=self.frames), f)


class LevelSetFrames:
    """bakefiles.jl:140-160 `LevelSetFrames(path)`: the frames memory-mapped as
    `(frames, z, y, x)` bytes, the grid's corner and cell, and the range."""

    def __init__(self, path):
        path = str(path)
        m = json.load(open(path + ".levelset.json"))
        n1, n2, n3 = m["dims"]
        self.data = np.memmap(path + ".levelset", np.uint8, "r", shape=(m["frames"], n3, n2, n1))
        self.lo, self.cell, self.range = m["lo"], m["cell"], m["range"]

    def __len__(self):
        return self.data.shape[0]
