"""ctypes binding of libstab_hip.so (the C ABI of include/stab_hip.h).

The library is built in-tree by ``stab_amd.build``.  Loading fails loudly when
it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstab_hip.so")

STAB_OK = 0
STAB_MAX_NEV = 16
NOALLOC_SCRATCH_BYTES = 64 << 20
MODE_PLAIN_F64 = 0
MODE_MIXED = 1
MAT_A0, MAT_B0, MAT_C0, MAT_C1, MAT_C2, MAT_SHIFTED = range(6)
INFO_OK, INFO_SINGULAR, INFO_NOCONV, INFO_FALLBACK = 0, 1, 2, 3
LEAD_KRYLOV, LEAD_INVIT, TRACK = 0, 1, 2
FARFIELD_MAX, CONTINUUM_BAND = 0.1, 0.01

EXPORTS = ("stab_ctx_create", "stab_ctx_destroy", "stab_last_error", "stab_profile_set", "stab_grid",
           "stab_pencil", "stab_solve", "stab_evec_bytes", "stab_shifted_solve", "stab_lu_factors", "stab_copy_probe", "stab_solve_leading",
           "stab_solve_leading_noalloc",
           "stab_mlp_create", "stab_mlp_destroy", "stab_mlp_nparams", "stab_mlp_get_params", "stab_mlp_set_params",
           "stab_mlp_set_scaling", "stab_mlp_predict", "stab_mlp_loss_grad", "stab_mlp_train")


class Z(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class Point(C.Structure):
    """Mirror of ``stab_point``."""
    _fields_ = [("itype", C.c_int32), ("ny", C.c_int32), ("mattyp", C.c_int32), ("wallt", C.c_int32),
                ("top", C.c_int32), ("profile", C.c_int32), ("nev", C.c_int32), ("ider", C.c_int32),
                ("Ma", C.c_double), ("Re", C.c_double), ("Pr", C.c_double), ("Te", C.c_double),
                ("yi", C.c_double), ("ymax", C.c_double),
                ("alpha", Z), ("beta", Z), ("omega", Z), ("shift", Z)]


class Stats(C.Structure):
    _fields_ = [("ms_prep", C.c_double), ("ms_assemble", C.c_double), ("ms_lu", C.c_double),
                ("ms_eig", C.c_double), ("ms_total", C.c_double), ("lu_flops", C.c_int64),
                ("assemble_bytes", C.c_int64), ("npts", C.c_int32), ("nconv", C.c_int32),
                ("restarts_max", C.c_int32), ("nfallback", C.c_int32), ("ms_arnoldi", C.c_double),
                ("op_applies", C.c_int64), ("solve_bytes", C.c_double), ("ms_lu_kernel", C.c_double),
                ("graph", C.c_int32), ("lu_launches", C.c_int32)]


_lib = None


def load():
    """Load the in-tree library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with stab_amd.build.build()")
    lib = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)
    lib.stab_ctx_create.argtypes = [C.POINTER(vp), C.c_int]
    lib.stab_ctx_destroy.argtypes = [vp]
    lib.stab_last_error.argtypes = [vp]
    lib.stab_last_error.restype = C.c_char_p
    lib.stab_profile_set.argtypes = [vp, C.c_int, C.c_int, dp, dp]
    lib.stab_grid.argtypes = [vp, C.POINTER(Point), dp, dp, dp, dp, dp, dp, dp]
    lib.stab_pencil.argtypes = [vp, C.POINTER(Point), C.c_int, C.POINTER(Z)]
    lib.stab_solve.argtypes = [vp, C.c_int, C.POINTER(Point), C.c_int, C.POINTER(Z), C.POINTER(Z), dp, ip,
                               C.POINTER(Stats)]
    lib.stab_evec_bytes.argtypes = [C.c_int, C.POINTER(Point)]
    lib.stab_shifted_solve.argtypes = [vp, C.POINTER(Point), C.c_int, C.c_int, C.POINTER(Z), C.POINTER(Z), ip]
    lib.stab_lu_factors.argtypes = [vp, C.POINTER(Point), C.c_int, C.POINTER(Z), ip, dp, ip]
    lib.stab_copy_probe.argtypes = [vp, C.c_int64, C.c_int, dp]
    lib.stab_solve_leading.argtypes = [vp, C.c_int, C.POINTER(Point), C.c_int, C.c_int, C.c_int, C.POINTER(Z),
                                       C.POINTER(Z), C.POINTER(Z), dp, dp, ip, ip, C.POINTER(Stats)]
    lib.stab_solve_leading_noalloc.argtypes = lib.stab_solve_leading.argtypes
    fp = C.POINTER(C.c_float)
    lib.stab_mlp_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]
    lib.stab_mlp_destroy.argtypes = [vp]
    lib.stab_mlp_nparams.argtypes = [vp]
    lib.stab_mlp_get_params.argtypes = [vp, fp]
    lib.stab_mlp_set_params.argtypes = [vp, fp]
    lib.stab_mlp_set_scaling.argtypes = [vp, fp, fp, fp, fp]
    lib.stab_mlp_predict.argtypes = [vp, C.c_int, fp, fp]
    lib.stab_mlp_loss_grad.argtypes = [vp, C.c_int, fp, fp, fp, fp]
    lib.stab_mlp_train.argtypes = [vp, C.c_int, fp, fp, C.c_int, C.c_int, C.c_float, C.c_uint64, fp]
    lib.stab_evec_bytes.restype = C.c_size_t
    for name in EXPORTS:
        lib[name].restype = lib[name].restype if name in ("stab_last_error", "stab_evec_bytes") else C.c_int
    _lib = lib
    return lib
