"""ctypes binding of libwt2.so (the C ABI declared in include/wt2.h).

The library is loaded from this directory; if it is missing the import fails
loudly -- there is no CPU fallback in the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwt2.so")

WT2_OK = 0
WT2_ERR_ARG = -1
WT2_ERR_SHAPE = -2
WT2_ERR_TAU = -3
WT2_ERR_UNSUPPORTED = -4
WT2_ERR_HIP = -5
WT2_ERR_OOM = -6
WT2_ERR_NONFINITE = -7
WT2_ERR_STATE = -8

WT2_FP64 = 0
WT2_MIXED = 1

WT2_HOST_PTRS = 1
WT2_ALLOW_NONCOERCIVE = 2
WT2_NO_GRAPH = 4

#: every entry point include/wt2.h declares
EXPORTS = (
    "wt2_gll", "wt2_ctx_create", "wt2_ctx_create_mesh", "wt2_ctx_footprint", "wt2_ctx_destroy", "wt2_ctx_info", "wt2_ctx_set_background",
    "wt2_coords", "wt2_rhs3d", "wt2_rhs_full", "wt2_rk4_steps", "wt2_rk4_steps_noalloc", "wt2_energy", "wt2_set_stream",
    "wt2_batch_create", "wt2_batch_destroy", "wt2_batch_add", "wt2_batch_remove", "wt2_batch_step",
    "wt2_batch_step_async", "wt2_batch_sync",
    "wt2_batch_slot", "wt2_batch_info", "wt2_ctx_autotune", "wt2_ctx_tiles", "wt2_ctx_set_tiles", "wt2_launch_pass", "wt2_last_error", "wt2_version",
    "wt2_fo_coef", "wt2_fo_rhs", "wt2_fo_rk4_stage", "wt2_fo_energy",
)

# first-order boundary kinds (src/boundaries1d.jl:59-62)
WT2_BC_DIRICHLET = 1
WT2_BC_SOMMERFELD = 2
WT2_BC_EXCISION = 3
WT2_BC_FULL_DIRICHLET = 4


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32),
        ("M", ctypes.c_int32 * 3),
        ("x0", ctypes.c_double * 3),
        ("x1", ctypes.c_double * 3),
        ("periodic", ctypes.c_int32 * 3),
        ("warp_A", ctypes.c_double),
        ("tau", ctypes.c_double),
        ("eps_ko", ctypes.c_double),
        ("bdry_values", ctypes.c_double * 6),
        ("precision", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("face_kind", ctypes.c_int32 * 6),
        ("sommerfeld_inv_R", ctypes.c_double),
    ]


class BC3d(ctypes.Structure):
    _fields_ = [
        ("kinds", ctypes.c_int32 * 6),
        ("excision_tag", ctypes.c_int32),
        ("sigma", ctypes.c_double),
        ("gPhi", ctypes.c_void_p),
        ("gPi", ctypes.c_void_p),
        ("gD", ctypes.c_void_p),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("npoints", ctypes.c_int64),
        ("Ne", ctypes.c_int32),
        ("curved", ctypes.c_int32),
        ("mu", ctypes.c_double),
        ("h_min", ctypes.c_double),
        ("tau", ctypes.c_double),
        ("max_sym_eig", ctypes.c_double),
        ("device_bytes", ctypes.c_int64),
        ("launches_per_step", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class WT2Error(RuntimeError):
    """Raised for a non-zero return code of libwt2; .code holds the WT2_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__(f"libwt2 error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libwt2.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"libwt2.so not found at {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.c_void_p
        vp = ctypes.c_void_p
        L.wt2_gll.argtypes = [ctypes.c_int32, dp, dp, dp]
        L.wt2_ctx_create.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(vp)]
        L.wt2_ctx_create_mesh.argtypes = [ctypes.POINTER(MeshDesc), ctypes.c_int64, dp, dp, dp, dp, dp,
                                          ctypes.POINTER(vp)]
        L.wt2_ctx_footprint.argtypes = [ctypes.POINTER(MeshDesc), ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]
        L.wt2_ctx_destroy.argtypes = [vp]
        L.wt2_ctx_destroy.restype = None
        L.wt2_ctx_info.argtypes = [vp, ctypes.POINTER(Info)]
        L.wt2_ctx_set_background.argtypes = [vp, dp]
        L.wt2_coords.argtypes = [vp, dp]
        L.wt2_rhs3d.argtypes = [vp, dp, dp, dp, dp, ctypes.c_int32]
        L.wt2_rhs_full.argtypes = [vp, dp, dp, dp, dp, ctypes.c_int32]
        L.wt2_rk4_steps.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                    dp, ctypes.c_int32]
        L.wt2_rk4_steps_noalloc.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, dp]
        L.wt2_energy.argtypes = [vp, dp, dp, dp, ctypes.c_int32]
        L.wt2_set_stream.argtypes = [vp, vp]
        L.wt2_launch_pass.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, dp, dp, ctypes.c_double]
        L.wt2_batch_create.argtypes = [ctypes.POINTER(vp)]
        L.wt2_batch_destroy.argtypes = [vp]
        L.wt2_batch_destroy.restype = None
        L.wt2_batch_add.argtypes = [vp, vp, dp, dp, ctypes.c_double, ctypes.c_int64, ctypes.c_int32,
                                    ctypes.POINTER(ctypes.c_int32)]
        L.wt2_batch_remove.argtypes = [vp, ctypes.c_int32]
        L.wt2_batch_step.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.wt2_batch_step_async.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.wt2_batch_sync.argtypes = [vp]
        L.wt2_batch_slot.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), dp, ctypes.c_int32,
                                     ctypes.POINTER(ctypes.c_int32)]
        L.wt2_batch_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        L.wt2_ctx_autotune.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_int32]
        L.wt2_ctx_tiles.argtypes = [vp, dp, dp]
        L.wt2_ctx_set_tiles.argtypes = [vp, dp, dp]
        L.wt2_fo_coef.argtypes = [vp, dp, dp, ctypes.c_int32]
        L.wt2_fo_rhs.argtypes = [vp, dp, dp, dp, dp, dp, ctypes.c_double, ctypes.POINTER(BC3d), ctypes.c_int32]
        L.wt2_fo_rk4_stage.argtypes = [vp, ctypes.c_int32, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(BC3d)]
        L.wt2_fo_energy.argtypes = [vp, dp, dp, dp, dp, ctypes.c_int32]
        L.wt2_last_error.restype = ctypes.c_char_p
        L.wt2_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != WT2_OK:
        raise WT2Error(rc, lib().wt2_last_error().decode())
    return rc
