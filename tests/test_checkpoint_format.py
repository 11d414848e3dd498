"""The checkpoint file's header and metadata (mael_checkpoint_meta, host only): a file in
the documented layout is read, and files with another magic or format version are
refused."""
import json
import struct

import pytest


def _write(path, magic=b"MAELCKPT", version=1, kind=1, meta=b'{"format": 1}', arrays=()):
    with open(path, "wb") as f:
        f.write(magic)
        f.write(struct.pack("<II", version, kind))
        f.write(struct.pack("<Q", len(meta)))
        f.write(meta)
        f.write(struct.pack("<II", len(arrays), 0))
        for name, es, count, off in arrays:
            f.write(name.encode().ljust(32, b"\0") + struct.pack("<IIQQQ", es, 0, count, off, 0))


def test_meta_round_trip(tmp_path, mael):
    from maelstrom_amd import checkpoint
    p = tmp_path / "a.ckpt"
    meta = {"format": 1, "kind": "sparse", "time": 0.25, "basin": {"radius": 18.0}}
    _write(p, meta=json.dumps(meta).encode(), arrays=[("phi_f", 4, 0, 256)])
    kind, m = checkpoint.read_meta(p)
    assert kind == 1 and m == meta


@pytest.mark.parametrize("magic,version", [(b"NOTACKPT", 1), (b"MAELCKPT", 2)])
def test_bad_header_refused(tmp_path, mael, magic, version):
    from maelstrom_amd import checkpoint
    p = tmp_path / "b.ckpt"
    _write(p, magic=magic, version=version)
    with pytest.raises(mael.MaelstromError):
        checkpoint.read_meta(p)


def test_truncated_refused(tmp_path, mael):
    from maelstrom_amd import checkpoint
    p = tmp_path / "c.ckpt"
    _write(p, meta=b'{"format": 1, "long": "' + b"x" * 100 + b'"}')
    data = p.read_bytes()
    p.write_bytes(data[:40])
    with pytest.raises(mael.MaelstromError):
        checkpoint.read_meta(p)
