"""The C-ABI library loads without a GPU and exports every function include/*.h declares."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b(mael_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("mael_create", "mael_dense_resample", "mael_dense_p2g", "mael_dense_project", "mael_dense_g2p"):
        assert n in names


def test_library_exports_every_declared_symbol(mael):
    L = ctypes.CDLL(mael.LIBPATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, missing


def test_host_only_entry_points_work_without_gpu(mael):
    import maelstrom_amd._lib as _l
    L = _l.lib()
    o = _l.SolveOpts()
    L.mael_solve_defaults(ctypes.byref(o))
    assert (o.tol, o.maxiter, o.sweeps, o.coarse) == (1e-4, 60, 2, 40)   # flip.jl:1082
    assert (o.precision, o.refine, o.warm) == (0, 0, 1)   # the reference's FP32 PCG, warm-started
    assert L.mael_version().startswith(b"maelstrom-hip")
    g = _l.DenseGrid()
    assert L.mael_flipgrid(18.0, 12.0, 3.0, 0.0, 2, ctypes.byref(g)) == _l.E_ARG
