"""Pins of the bake writers' restatements (oracle.dense_quantize, pack_band, the sparse
oracle's quantize) and the host readers of the bake files.

- quantisation (flip.jl:1859-1867): a constant level set gives the closed-form byte,
  127.5 (f + 1) rounded to nearest even (f = 0 gives 128), clamped at 0 and 255;
- the sparse quantisation (sflip.jl:2427-2432) samples the level set trilinearly, which
  reproduces an affine level set: the bytes are the closed form's at the sample points;
- band packing (flip.jl:1880-1893): the keep rule's closed cases, and Float16 rounding to
  nearest even at halfway values;
- ParticleFrames and LevelSetFrames read the documented layout back.
"""
import json

import numpy as np
import pytest

import oracle as O

f32 = np.float32


def test_dense_quantize_closed_forms():
    g = O.flipgrid(18.0, 12.0, 3.0, 0.6)
    shape = O.cells_shape(g)
    rng_ = f32(6 * 0.6)
    for c, want in ((0.0, 128), (float(rng_), 255), (-float(rng_), 0), (100.0, 255), (-100.0, 0)):
        q = O.dense_quantize(g, np.full(shape, c, f32), rng_)
        assert q.shape == (shape[0] // 2, shape[1] // 2, shape[2] // 2)
        assert np.all(q == want), c
    c = f32(0.3) * rng_
    q = O.dense_quantize(g, np.full(shape, c, f32), rng_)
    f = np.clip(f32(0.125) * (f32(8) * c) / rng_, -1, 1)   # eight equal terms sum exactly
    assert np.all(q == int(np.rint(f32(127.5) * (f + f32(1)))))
    # averaging: a level set rising along z gives bytes non-decreasing along z
    z = (np.arange(shape[0], dtype=f32) - shape[0] / 2)[:, None, None] * f32(0.3)
    q = O.dense_quantize(g, np.broadcast_to(z, shape).astype(f32), rng_)
    assert np.all(np.diff(q.astype(int), axis=0) >= 0) and q.min() == 0 and q.max() == 255


def test_sparse_quantize_affine():
    o = O.sparse_scene(dx=0.6, blocks=64, capacity=300000)
    g = o.grid
    a, b = f32(0.7), f32(-1.5)
    # the affine level set a z + b at every fine and coarse cell centre
    ns = o.nslots
    where = o.array("where").reshape(-1, 3)[:ns]
    k, j, i = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    phif = o.array("phi_f")
    for s in range(ns):
        zc = g.lo[2] + (where[s, 2] * 16 + k + f32(0.5)) * g.dx
        phif[s * 4096:(s + 1) * 4096] = (a * zc + b).astype(f32).ravel()
    n1, n2, n3 = g.nc
    kc = np.arange(n3)[:, None, None] * np.ones((1, n2, n1))
    o.array("phi_c")[:] = (a * (g.lo[2] + (kc + f32(0.5)) * (4 * g.dx)) + b).astype(f32).ravel()
    cell, rng_ = f32(2) * f32(g.dx), f32(6) * f32(g.dx)
    q = o.quantize(cell, rng_)
    zs = g.lo[2] + (np.arange(q.shape[0]) + 1 - 0.5) * cell
    want = np.rint(127.5 * (np.clip((a * zs + b) / rng_, -1, 1) + 1))
    diff = np.abs(q.astype(int) - want[:, None, None].astype(int))
    assert diff.max() <= 1 and (diff == 0).mean() > 0.99


def test_pack_band_rule_and_rounding():
    n = 6
    ps = O.empty_particles(n)
    ps["PX"][:] = [0.0, 10.0, 10.0, 1.0, 10.0, 10.0]
    ps["PZ"][:] = [-9.0, -3.0, -4.0, -9.0, -3.6, -3.0]
    ps["VZ"][:] = [0.0, 0.0, 0.0, 0.0, 2.0, 0.0]
    ps["LAG"][:] = [0.0, 0.0, 0.0, 0.0, 0.1, 0.0]
    ps["PY"][:] = [0.0, 0.0, 0.0, 0.0, 0.0, 1 + 2.0 ** -11]   # halfway between two Float16s
    pos, vel = O.pack_band(ps, n, keep=3.5, throat=3.0)
    # kept: on the axis (0), the top 3.5 m (1), not 2 (z -4), within the throat (3), particle 4
    # at its frame-time position -3.6 + 0.1 * 2 = -3.4, particle 5
    assert pos.shape == (5, 3)
    np.testing.assert_array_equal(pos[:, 2].astype(f32), np.float16([-9.0, -3.0, -9.0, -3.4, -3.0]).astype(f32))
    assert pos[4, 1] == np.float16(1.0)   # ties to even
    assert np.float16(f32(1 + 3 * 2.0 ** -11)) == np.float16(1 + 2.0 ** -9)
    assert vel[3, 2] == np.float16(2.0)


def test_bake_readers(tmp_path):
    import sys
    from maelstrom_amd import bake
    p = str(tmp_path / "b")
    frames = [np.arange(12, dtype=np.float16).reshape(4, 3), np.arange(6, dtype=np.float16).reshape(2, 3) + 100]
    with open(p + ".particles", "wb") as f:
        for fr in frames:
            f.write(fr.tobytes())
    with open(p + ".velocities", "wb") as f:
        for fr in frames:
            f.write((-fr).tobytes())
    json.dump(dict(counts=[4, 2], keep=3.5, throat=3.0), open(p + ".particles.json", "w"))
    pf = bake.ParticleFrames(p)
    assert len(pf) == 2
    np.testing.assert_array_equal(pf.frame(1), frames[0].T)
    np.testing.assert_array_equal(pf.velocity(2), -frames[1].T)
    q = np.arange(2 * 3 * 4 * 5, dtype=np.uint8).reshape(2, 3, 4, 5)
    q.tofile(p + ".levelset")
    json.dump(dict(dims=[5, 4, 3], lo=[0, 0, 0], cell=1.2, range=3.6, frames=2), open(p + ".levelset.json", "w"))
    ls = bake.LevelSetFrames(p)
    assert len(ls) == 2 and np.array_equal(ls.data, q)
