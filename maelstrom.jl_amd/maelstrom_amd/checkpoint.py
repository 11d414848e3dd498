"""Checkpoints: `save_checkpoint` / `load_checkpoint` (checkpoint.jl:7-66).

The device state goes through the library's versioned raw format
(`mael_dense_checkpoint` / `mael_sparse_checkpoint`); this module writes the host state
the reference keeps beside it (checkpoint.jl:16-20: the basin, grid, stirring model,
particle capacity, the sparse basin's rings, vary and reflect, time, steps, the reference
volume and the solver iterations) as the file's JSON metadata, and rebuilds the basin from
it on load. As in the reference a checkpoint is taken at a completed frame, and a sparse
one restores into a pool of at least the saved slots (`blocks`) and a capacity of at
least the saved particles.
"""
import ctypes as C
import dataclasses
import json

from . import _lib
from .basin import Basin, Race, Stir, Wave

FORMAT = 1


def _stir_dict(stir):
    if stir is None:
        return None
    return dict(waves=[dataclasses.asdict(w) for w in stir.waves], race=dataclasses.asdict(stir.race),
                origin=list(stir.origin), offset=stir.offset, surface=stir.surface, ramp=stir.ramp)


def _stir_from(d, basin):
    if d is None:
        return None
    race = Race(**{k: (tuple(v) if isinstance(v, list) else v) for k, v in d["race"].items()})
    waves = [Wave(tuple(w["d"]), w["k"], w["omega"], w["a"], w["q"], w["phase"]) for w in d["waves"]]
    return Stir(basin, waves=waves, race=race, origin=tuple(d["origin"]), offset=d["offset"], surface=d["surface"],
                ramp=d["ramp"])


def _meta(s, kind, fps):
    m = dict(format=FORMAT, kind=kind, fps=fps, basin=dataclasses.asdict(s.basin), dx=float(s.grid.dx), top=s.top,
             pad=s.pad,
             grid=dict(lo=[float(x) for x in s.grid.lo], dx=float(s.grid.dx)), band=s.band, nmin=s.nmin, cfl=s.cfl,
             capacity=s.capacity, stir=_stir_dict(s.stir), time=s.t, steps=s.steps, volume0=s.volume0,
             iterations=[int(x) for x in s.iterations])
    if kind == "sparse":
        m.update(rings=s.rings, vary=s.vary, reflect=s.reflect, blocks=s.blocks, nb=[int(x) for x in s.grid.nb])
    else:
        m.update(n=[int(x) for x in s.grid.n])
    return m


def save_checkpoint(s, path, fps=24):
    """checkpoint.jl:7-35 `save_sparse_checkpoint(s, w, path; fps)` (dense basins too):
    the state at a completed frame, to `path`."""
    if abs(s.t * fps - round(s.t * fps)) > 1e-6:
        raise ValueError("Save at a completed frame")
    from .sparse import SparseBasin
    kind = "sparse" if isinstance(s, SparseBasin) else "dense"
    meta = json.dumps(_meta(s, kind, fps)).encode()
    fn = _lib.lib().mael_sparse_checkpoint if kind == "sparse" else _lib.lib().mael_dense_checkpoint
    s.ctx.check(fn(s.h, str(path).encode(), meta))
    return path


def read_meta(path):
    """The checkpoint's kind (0 dense, 1 sparse) and host metadata, without a device."""
    kind, n = C.c_int32(), C.c_int64()
    st = _lib.lib().mael_checkpoint_meta(str(path).encode(), C.byref(kind), None, 0, C.byref(n))
    if st != _lib.OK:
        raise _lib.MaelstromError(st, "%s is not a readable checkpoint" % path)
    buf = C.create_string_buffer(n.value + 1)
    _lib.lib().mael_checkpoint_meta(str(path).encode(), C.byref(kind), buf, n.value + 1, C.byref(n))
    return kind.value, json.loads(buf.value.decode())


def load_checkpoint(path, ctx=None, blocks=None, capacity=None):
    """checkpoint.jl:37-66 `load_sparse_checkpoint(path, backend; blocks, capacity_override)`
    (dense basins too): a new basin in the saved state."""
    from .dense import FlipBasin
    from .sparse import SparseBasin
    kind, m = read_meta(path)
    if m.get("format") != FORMAT:
        raise ValueError("Checkpoint format differs")
    basin = Basin(**m["basin"])
    stir = _stir_from(m["stir"], basin)
    cap = m["capacity"] if capacity is None else int(capacity)
    if kind == 1:
        nb = m["blocks"] if blocks is None else max(int(blocks), m["blocks"])
        s = SparseBasin(basin, dx=m["dx"], stir=stir, band=m["band"], nmin=m["nmin"], cfl=m["cfl"], capacity=cap,
                        blocks=nb, rings=m["rings"], vary=m["vary"], reflect=m["reflect"], ctx=ctx, top=m["top"],
                        pad=m["pad"])
        if [int(x) for x in s.grid.nb] != m["nb"]:
            raise ValueError("Checkpoint grid differs")
        s.ctx.check(_lib.lib().mael_sparse_restore(s.h, str(path).encode()))
    else:
        s = FlipBasin(basin, dx=m["dx"], band=m["band"], nmin=m["nmin"], cfl=m["cfl"], capacity=cap, ctx=ctx,
                      stir=stir if stir is not None else False, top=m["top"], pad=m["pad"])
        if [int(x) for x in s.grid.n] != m["n"]:
            raise ValueError("Checkpoint grid differs")
        s.ctx.check(_lib.lib().mael_dense_restore(s.h, str(path).encode()))
    s.volume0, s.t, s.steps = m["volume0"], m["time"], m["steps"]
    s.iterations = list(m["iterations"])
    return s
