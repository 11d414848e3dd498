"""ctypes binding of the HIP library's C ABI (include/maelstrom_hip.h).

The library is loaded from `maelstrom.jl_amd/lib/libmaelstrom_hip.so`. There is no
fallback: if it is missing or cannot open a GPU, the calls raise.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(os.path.dirname(HERE), "lib", "libmaelstrom_hip.so")

# status codes (include/maelstrom_hip.h)
OK, E_ARG, E_CAPACITY, E_BLOCKS, E_NONFINITE, E_CG_INDEFINITE, E_NOCONVERGE, E_BREAKDOWN, E_HIP, E_GRID, E_TIMEOUT = (
    0, -1, -2, -3, -4, -5, -6, -7, -8, -9, -10)

# array ids (include/maelstrom_hip.h)
ARRAYS = ["U", "V", "W", "U0", "V0", "W0", "PHI", "PHI0", "TYPES", "P", "FU", "FV", "FW", "RHS",
          "PX", "PY", "PZ", "VX", "VY", "VZ", "C0", "C1", "C2", "C3", "C4", "C5", "C6", "C7", "C8", "LAG",
          "KEY", "RANK", "COUNTS", "TOTALS", "OFFSETS",
          "MOM_U", "MOM_V", "MOM_W", "WT_U", "WT_V", "WT_W",
          "UB_U", "UB_V", "UB_W", "UB0_U", "UB0_V", "UB0_W",
          "PB_VX", "PB_VY", "PB_VZ"] + ["PB_C%d" % k for k in range(9)]
ARRAY_ID = {n: i for i, n in enumerate(ARRAYS)}
PARTICLE_FIELDS = ARRAYS[ARRAY_ID["PX"]:ARRAY_ID["LAG"] + 1]

# kernel ids of the launch timer (include/maelstrom_hip.h MAEL_K_*)
KERNELS = ["p2g", "splat", "combine", "g2p", "bin", "fixrank", "scatter", "spawn", "search", "advance", "rbgs",
           "residual", "restrict", "prolong", "coarsest", "resid64"]
KERNEL_ID = {n: i for i, n in enumerate(KERNELS)}

# sparse array ids (include/maelstrom_hip.h MAEL_S_*)
SPARSE_ARRAYS = ["PHI_F", "PHI_C", "PHI0_F", "PHI0_C", "P_F", "P_C", "TYPES_F", "TYPES_C",
                 "U_F", "U_C", "V_F", "V_C", "W_F", "W_C", "U0_F", "U0_C", "V0_F", "V0_C", "W0_F", "W0_C",
                 "FU_F", "FU_C", "FV_F", "FV_C", "FW_F", "FW_C", "RHS_F", "RHS_C"] + PARTICLE_FIELDS + [
                 "KEY", "RANK", "COUNTS", "TOTALS", "OFFSETS",
                 "MOM_U", "MOM_V", "MOM_W", "WT_U", "WT_V", "WT_W", "SLOT", "WHERE",
                 "UB_U_F", "UB_U_C", "UB_V_F", "UB_V_C", "UB_W_F", "UB_W_C",
                 "UB0_U_F", "UB0_U_C", "UB0_V_F", "UB0_V_C", "UB0_W_F", "UB0_W_C"] + [
                 "PB_" + f for f in ["VX", "VY", "VZ"] + ["C%d" % k for k in range(9)]]
SPARSE_ID = {n: i for i, n in enumerate(SPARSE_ARRAYS)}


class MaelstromError(RuntimeError):
    """A failed call: `code` is the status, the message the library's (the reference's text where it has one)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class DenseGrid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("dx", C.c_float), ("n", C.c_int32 * 3), ("radius", C.c_float),
                ("floor", C.c_float), ("levels", C.c_int32)]


class SparseGrid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("dx", C.c_float), ("nb", C.c_int32 * 3), ("radius", C.c_float),
                ("floor", C.c_float)]

    @property
    def nc(self):
        return tuple(4 * int(x) for x in self.nb)


class SparseOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxiter", C.c_int32), ("sweeps", C.c_int32), ("coarse", C.c_int32),
                ("restart", C.c_int32), ("policy", C.c_int32), ("precision", C.c_int32), ("boost", C.c_float),
                ("refine", C.c_int32)]


class SparseStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("fcg_iterations", C.c_int32), ("fgmres_iterations", C.c_int32),
                ("used_fgmres", C.c_int32), ("rel_residual", C.c_double), ("ms", C.c_double)]


class WhiteRulesC(C.Structure):
    _fields_ = [("air", C.c_float * 2), ("crest", C.c_float * 2), ("energy", C.c_float * 2), ("kair", C.c_float),
                ("kcrest", C.c_float), ("spray", C.c_int32), ("bubble", C.c_int32), ("life", C.c_float * 2),
                ("buoyancy", C.c_float), ("drag", C.c_float), ("kick", C.c_float), ("tossed", C.c_float),
                ("within", C.c_float), ("floor", C.c_float), ("gravity", C.c_float), ("stream", C.c_int32),
                ("gamma_circ", C.c_float), ("core", C.c_float)]


class StirC(C.Structure):
    _fields_ = [("radius", C.c_float), ("band", C.c_float), ("relax", C.c_float), ("core", C.c_float),
                ("gamma_circ", C.c_float), ("current", C.c_float * 2), ("eddy", C.c_float), ("eddyscale", C.c_float),
                ("eddytime", C.c_float), ("guide", C.c_float), ("free", C.c_float), ("turbulence", C.c_float),
                ("boil", C.c_float), ("reach", C.c_float), ("origin", C.c_float * 2), ("offset", C.c_float),
                ("surface", C.c_int32), ("hradius", C.c_float), ("ramp", C.c_float)]


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxiter", C.c_int32), ("sweeps", C.c_int32), ("coarse", C.c_int32),
                ("precision", C.c_int32), ("refine", C.c_int32), ("warm", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("restarts", C.c_int32), ("rel_residual", C.c_double),
                ("rel_residual_rec", C.c_double), ("ms", C.c_double)]


_lib = None


def lib():
    """The loaded library (loaded on first use; raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise MaelstromError(E_ARG, "the HIP library is not built: %s (run __graft_entry__.build())" % LIBPATH)
    L = C.CDLL(LIBPATH)
    p, i32, i64, f32, f64, u32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double, C.c_uint32
    sig = {
        "mael_create": (C.c_int, [C.c_int, C.POINTER(p)]),
        "mael_destroy": (None, [p]),
        "mael_last_error": (C.c_char_p, [p]),
        "mael_version": (C.c_char_p, []),
        "mael_flipgrid": (C.c_int, [f32, f32, f32, f32, i32, C.POINTER(DenseGrid)]),
        "mael_solve_defaults": (None, [C.POINTER(SolveOpts)]),
        "mael_dense_create": (C.c_int, [p, C.POINTER(DenseGrid), i64, i32, i32, C.POINTER(p)]),
        "mael_dense_destroy": (None, [p]),
        "mael_dense_array": (C.c_int, [p, i32, C.POINTER(p), C.POINTER(i64), C.POINTER(i32)]),
        "mael_dense_upload": (C.c_int, [p, i32, p, i64]),
        "mael_dense_download": (C.c_int, [p, i32, p, i64]),
        "mael_dense_count": (C.c_int, [p, C.POINTER(i64)]),
        "mael_dense_set_count": (C.c_int, [p, i64]),
        "mael_dense_resample": (C.c_int, [p, i32, u32, C.POINTER(i64)]),
        "mael_dense_resample_noalloc": (C.c_int, [p, i32, u32, C.POINTER(i64)]),
        "mael_dense_p2g": (C.c_int, [p, f32]),
        "mael_dense_g2p": (C.c_int, [p, f32, f32]),
        "mael_dense_project": (C.c_int, [p, f32, f32, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
        "mael_dense_vcycle": (C.c_int, [p, C.POINTER(SolveOpts), i32]),
        "mael_dense_g2p_adjoint": (C.c_int, [p, f32]),
        "mael_dense_project_adjoint": (C.c_int, [p, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
        "mael_dense_p2g_adjoint": (C.c_int, [p, f32]),
        "mael_dense_classify": (C.c_int, [p]),
        "mael_dense_volume": (C.c_int, [p, C.POINTER(f64)]),
        "mael_dense_redistance": (C.c_int, [p, i32]),
        "mael_dense_init_vortex": (C.c_int, [p, p, i32, f32, f32, f32]),
        "mael_dense_solidify": (C.c_int, [p]),
        "mael_dense_gravity": (C.c_int, [p, f32, f32]),
        "mael_dense_copy": (C.c_int, [p, i32, i32]),
        "mael_dense_advect": (C.c_int, [p, f32, f32, i32]),
        "mael_dense_filltrapped": (C.c_int, [p, i32]),
        "mael_dense_set_stir": (C.c_int, [p, C.POINTER(StirC), p, i32]),
        "mael_dense_forces": (C.c_int, [p, f32, f32, f32]),
        "mael_dense_extrapolate": (C.c_int, [p, i32]),
        "mael_dense_maxspeed": (C.c_int, [p, C.POINTER(f32)]),
        "mael_dense_ktimer": (C.c_int, [p, i32, i32]),
        "mael_dense_ktime": (C.c_int, [p, C.POINTER(f64), C.POINTER(i32)]),
        "mael_dense_sync": (C.c_int, [p]),
        "mael_sparsegrid": (C.c_int, [f32, f32, f32, f32, i32, C.POINTER(SparseGrid)]),
        "mael_sparse_create": (C.c_int, [p, C.POINTER(SparseGrid), i32, i64, i32, i32, C.POINTER(p)]),
        "mael_sparse_destroy": (None, [p]),
        "mael_sparse_array": (C.c_int, [p, i32, C.POINTER(p), C.POINTER(i64), C.POINTER(i32)]),
        "mael_sparse_upload": (C.c_int, [p, i32, p, i64]),
        "mael_sparse_download": (C.c_int, [p, i32, p, i64]),
        "mael_sparse_count": (C.c_int, [p, C.POINTER(i64)]),
        "mael_sparse_set_count": (C.c_int, [p, i64]),
        "mael_sparse_get_topology": (C.c_int, [p, p, p, C.POINTER(i32), p, C.POINTER(i32)]),
        "mael_sparse_set_topology": (C.c_int, [p, p, p, i32, p, i32]),
        "mael_sparse_marks": (C.c_int, [p, f32, f32, i32, i32, p]),
        "mael_sparse_refine_marks": (C.c_int, [p, p, C.POINTER(i32), C.POINTER(i32)]),
        "mael_sparse_refine": (C.c_int, [p, f32, f32, i32, C.POINTER(i32), C.POINTER(i32)]),
        "mael_sparse_resample": (C.c_int, [p, i32, u32, C.POINTER(i64)]),
        "mael_sparse_p2g": (C.c_int, [p, f32]),
        "mael_sparse_g2p": (C.c_int, [p, f32, f32]),
        "mael_sparse_defaults": (None, [C.POINTER(SparseOpts)]),
        "mael_sparse_project": (C.c_int, [p, f32, C.POINTER(SparseOpts), C.POINTER(SparseStats)]),
        "mael_sparse_vcycle": (C.c_int, [p, C.POINTER(SparseOpts), i32]),
        "mael_sparse_g2p_adjoint": (C.c_int, [p, f32]),
        "mael_sparse_p2g_adjoint": (C.c_int, [p, f32]),
        "mael_sparse_project_adjoint": (C.c_int, [p, C.POINTER(SparseOpts), C.POINTER(SparseStats)]),
        "mael_sparse_init_coarse": (C.c_int, [p, p, i32, f32, f32, f32]),
        "mael_sparse_init_fine": (C.c_int, [p, p, i32, f32, f32, f32]),
        "mael_sparse_classify": (C.c_int, [p]),
        "mael_sparse_solidify": (C.c_int, [p]),
        "mael_sparse_redistance": (C.c_int, [p, i32]),
        "mael_sparse_restrict_velocity": (C.c_int, [p]),
        "mael_sparse_keep": (C.c_int, [p]),
        "mael_sparse_fill_flags": (C.c_int, [p, C.c_uint8]),
        "mael_sparse_volume": (C.c_int, [p, C.POINTER(f64)]),
        "mael_sparse_gravity": (C.c_int, [p, f32, f32]),
        "mael_sparse_sync": (C.c_int, [p]),
        "mael_sparse_advect": (C.c_int, [p, f32, i32, f32, i32]),
        "mael_sparse_filltrapped": (C.c_int, [p, i32]),
        "mael_sparse_set_stir": (C.c_int, [p, C.POINTER(StirC), p, i32, p, i32]),
        "mael_sparse_forces": (C.c_int, [p, f32, f32, f32]),
        "mael_sparse_save_reflect": (C.c_int, [p]),
        "mael_sparse_kick": (C.c_int, [p]),
        "mael_sparse_reflected": (C.c_int, [p]),
        "mael_sparse_extrapolate": (C.c_int, [p, i32, i32, i32]),
        "mael_sparse_set_flow": (C.c_int, [p, i32]),
        "mael_sparse_maxspeed": (C.c_int, [p, C.POINTER(f32)]),
        "mael_dense_pack_band": (C.c_int, [p, f32, f32, p, p, i64, C.POINTER(i64)]),
        "mael_sparse_pack_band": (C.c_int, [p, f32, f32, p, p, i64, C.POINTER(i64)]),
        "mael_dense_quantize_levelset": (C.c_int, [p, f32, p]),
        "mael_sparse_quantize_levelset": (C.c_int, [p, f32, p]),
        "mael_white_create": (C.c_int, [p, i64, C.POINTER(WhiteRulesC), C.POINTER(p)]),
        "mael_white_destroy": (None, [p]),
        "mael_white_array": (C.c_int, [p, i32, C.POINTER(p), C.POINTER(i64), C.POINTER(i32)]),
        "mael_white_upload": (C.c_int, [p, i32, p, i64]),
        "mael_white_download": (C.c_int, [p, i32, p, i64]),
        "mael_white_state": (C.c_int, [p, C.POINTER(i64), C.POINTER(i64)]),
        "mael_white_set_state": (C.c_int, [p, i64, i64]),
        "mael_white_pack": (C.c_int, [p, p, p, i64, C.POINTER(i64)]),
        "mael_dense_whitewater": (C.c_int, [p, p, f64, i32, f32, f32, f32]),
        "mael_sparse_whitewater": (C.c_int, [p, p, f64, i32, f32, f32, f32]),
        "mael_dense_checkpoint": (C.c_int, [p, C.c_char_p, C.c_char_p]),
        "mael_sparse_checkpoint": (C.c_int, [p, C.c_char_p, C.c_char_p]),
        "mael_checkpoint_meta": (C.c_int, [C.c_char_p, C.POINTER(i32), p, i64, C.POINTER(i64)]),
        "mael_dense_restore": (C.c_int, [p, C.c_char_p]),
        "mael_sparse_restore": (C.c_int, [p, C.c_char_p]),
        "mael_sparse_ktimer": (C.c_int, [p, i32, i32]),
        "mael_sparse_ktime": (C.c_int, [p, C.POINTER(f64), C.POINTER(i32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class Context:
    """A device context (`mael_create`): the backend basins are built on."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        st = lib().mael_create(device, C.byref(self.h))
        if st != OK:
            raise MaelstromError(st, "cannot open HIP device %d (status %d)" % (device, st))

    def check(self, status):
        if status != OK:
            raise MaelstromError(status, lib().mael_last_error(self.h).decode())
        return status

    def close(self):
        if self.h:
            lib().mael_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    """The process's context on device 0 (`Mantle.defaultbackend()`)."""
    global _default
    if _default is None:
        _default = Context(0)
    return _default
