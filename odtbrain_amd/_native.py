"""ctypes binding of libodtbrain_hip.so (C ABI in include/odtbrain_hip.h).

The library must have been built in-tree (``python -m odtbrain_amd._build``);
there is no fallback: a missing library raises immediately.
"""
import ctypes
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libodtbrain_hip.so"

ODT_OK = 0
ODT_ERR_INVALID, ODT_ERR_HIP, ODT_ERR_UNSUPPORTED, ODT_ERR_WORKSPACE = 1, 2, 3, 4
ODT_KIND_3D, ODT_KIND_3D_TILTED, ODT_KIND_2D = 0, 1, 2
ODT_FP32, ODT_FP64 = 0, 1
ODT_PAD_NONE, ODT_PAD_EDGE, ODT_PAD_RAMP = 0, 1, 2
ODT_FLAG_IIR_PREFILTER = 1
ODT_SCRATCH_BYTES = 128 << 20  # static workspace of odt_plan_execute_noalloc

EXPORTED = (
    "odt_plan_create", "odt_plan_create_tuned", "odt_plan_workspace_size",
    "odt_plan_execute",
    "odt_plan_execute_noalloc", "odt_plan_destroy", "odt_plan_enable_timing", "odt_plan_last_timing",
    "odt_fft_c2c", "odt_spline_prefilter", "odt_rotate_accumulate",
    "odt_plan_adjoint",
    "odt_batch_create", "odt_batch_create_lanes", "odt_batch_launch",
    "odt_batch_destroy",
    "odt_odt_to_ri", "odt_opt_to_ri",
    "odt_prepare_workspace_size", "odt_sinogram_prepare",
    "odt_apple_workspace_size", "odt_apple_core", "odt_apple_correct",
    "odt_apple_envelope",
    "odt_last_error", "odt_version", "odt_debug_live_allocations",
)


class OdtDesc(ctypes.Structure):
    """Mirror of ``odt_desc`` (include/odtbrain_hip.h)."""
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("A", ctypes.c_int32),
        ("Ny", ctypes.c_int32),
        ("Nx", ctypes.c_int32),
        ("lNy", ctypes.c_int32),
        ("lNx", ctypes.c_int32),
        ("padyl", ctypes.c_int32),
        ("padxl", ctypes.c_int32),
        ("pad_mode", ctypes.c_int32),
        ("intp_order", ctypes.c_int32),
        ("onlyreal", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("angle_param_stride", ctypes.c_int32),
        ("padval_re", ctypes.c_double),
        ("padval_im", ctypes.c_double),
        ("km", ctypes.c_double),
        ("lD", ctypes.c_double),
        ("z0", ctypes.c_double),
        ("ramp_u", ctypes.c_double),
        ("ramp_v", ctypes.c_double),
        ("weights", ctypes.POINTER(ctypes.c_double)),
        ("angle_params", ctypes.POINTER(ctypes.c_double)),
        ("slab_budget_bytes", ctypes.c_int64),
    ]


class OdtTuning(ctypes.Structure):
    """odt_tuning (include/odtbrain_hip.h): launch-shape choices, 0 = the
    planner's default."""
    _fields_ = [("rotate_group", ctypes.c_int32),
                ("rotate_groups_2d", ctypes.c_int32),
                ("t_budget_bytes", ctypes.c_int64),
                ("col_per_angle", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class OdtError(RuntimeError):
    """Error reported by libodtbrain_hip."""


_lib = None


def lib():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise OdtError(
            f"{LIB_PATH} is missing: build it with `python -m odtbrain_amd._build`")
    L = ctypes.CDLL(str(LIB_PATH))
    vp, i32, i64, dp = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                        ctypes.POINTER(ctypes.c_double))
    L.odt_plan_create.argtypes = [ctypes.POINTER(OdtDesc), ctypes.POINTER(vp)]
    L.odt_plan_create_tuned.argtypes = [ctypes.POINTER(OdtDesc),
                                        ctypes.POINTER(OdtTuning),
                                        ctypes.POINTER(vp)]
    L.odt_plan_workspace_size.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    L.odt_plan_execute.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, vp]
    L.odt_plan_execute_noalloc.argtypes = [vp, vp, vp, vp]
    L.odt_plan_destroy.argtypes = [vp]
    L.odt_plan_enable_timing.argtypes = [vp, ctypes.c_int]
    L.odt_plan_last_timing.argtypes = [vp, dp, dp, dp,
                                       ctypes.POINTER(ctypes.c_int)]
    L.odt_fft_c2c.argtypes = [ctypes.c_int, i64, i32, i32, vp, vp, vp]
    L.odt_spline_prefilter.argtypes = [ctypes.c_int, ctypes.c_int, i64, i32,
                                       i64, i64, vp, vp]
    L.odt_rotate_accumulate.argtypes = [ctypes.c_int, ctypes.c_int, i32, i32,
                                        i32, i32, vp, dp, vp, ctypes.c_int,
                                        vp]
    L.odt_plan_adjoint.argtypes = [vp, vp, vp, vp, ctypes.c_size_t, vp]
    vpp = ctypes.POINTER(vp)
    L.odt_batch_create.argtypes = [ctypes.c_int, vpp, vpp, vpp, vp,
                                   ctypes.c_size_t, vp, vpp]
    L.odt_batch_create_lanes.argtypes = [ctypes.c_int, vpp, vpp, vpp, vp,
                                         ctypes.c_size_t, ctypes.c_int, vp, vpp]
    L.odt_batch_launch.argtypes = [vp, vp]
    L.odt_batch_destroy.argtypes = [vp]
    for name in ("odt_odt_to_ri", "odt_opt_to_ri"):
        getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int, i64,
                                     ctypes.c_double, ctypes.c_double, vp, vp,
                                     vp]
    L.odt_prepare_workspace_size.argtypes = [ctypes.c_int, ctypes.c_int, i32,
                                             i32, i32, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_size_t)]
    L.odt_sinogram_prepare.argtypes = [ctypes.c_int, ctypes.c_int, i32, i32,
                                       i32, vp, vp, ctypes.c_int, ctypes.c_int,
                                       vp, vp, ctypes.c_size_t, vp]
    L.odt_apple_workspace_size.argtypes = [i32, i32, i32,
                                           ctypes.POINTER(ctypes.c_size_t)]
    L.odt_apple_core.argtypes = [i32, i32, i32, ctypes.c_double,
                                 ctypes.c_double, vp, vp]
    L.odt_apple_correct.argtypes = [
        ctypes.c_int, i32, i32, i32, vp, ctypes.c_double, ctypes.c_double,
        ctypes.c_int, vp, ctypes.c_double, ctypes.c_double, i32,
        ctypes.c_double, ctypes.c_int, vp, vp, ctypes.c_size_t, vp,
        ctypes.POINTER(i32)]
    L.odt_apple_envelope.argtypes = [i32, i32, i32, vp, vp, vp, vp,
                                     ctypes.c_size_t, vp]
    L.odt_last_error.restype = ctypes.c_char_p
    L.odt_last_error.argtypes = []
    L.odt_version.argtypes = []
    L.odt_debug_live_allocations.argtypes = [ctypes.POINTER(ctypes.c_int64)]
    for name in EXPORTED:
        f = getattr(L, name)
        if name != "odt_last_error":
            f.restype = ctypes.c_int
    _lib = L
    return L


def live_allocations():
    """Device allocations the library holds (see odt_debug_live_allocations
    in include/odtbrain_hip.h: the count a failed call must leave as it
    was)."""
    n = ctypes.c_int64(0)
    check(lib().odt_debug_live_allocations(ctypes.byref(n)),
          "odt_debug_live_allocations")
    return int(n.value)


def check(rc, what):
    if rc != ODT_OK:
        msg = lib().odt_last_error().decode(errors="replace")
        raise OdtError(f"{what} failed (code {rc}): {msg}")


def dptr(arr):
    """ctypes double* of a contiguous float64 numpy array."""
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
