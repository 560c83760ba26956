"""ctypes binding of the C ABI in include/muhrec_hip.h (libmuhrec_hip.so).

The library is built in-tree by ``imagingsuite_amd.build.build()``.  There is
no fallback: if the shared object is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmuhrec_hip.so")

MR_OK = 0
MR_ERR_INVALID = 1
MR_ERR_HIP = 2
MR_ERR_UNSUPPORTED = 3
MR_ERR_INTERNAL = 4
MR_ERR_ABORTED = 5
# int (*mr_slab_fn)(void *user, size_t first_slice, size_t n_slices, const float *slices)
SLAB_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                           ctypes.POINTER(ctypes.c_float))

FILTER_TYPES = {"ram-lak": 0, "shepp-logan": 1, "hanning": 2, "hamming": 3, "butterworth": 4}
ITER_METHODS = {"sirt": 0, "landweber": 1, "cgls": 2}
# mr_batch_plan_create flags
BATCH_TUNE = 1
BATCH_FUSED = 2

# Every symbol include/muhrec_hip.h declares.
EXPORTED_SYMBOLS = (
    "mr_last_error", "mr_version", "mr_device_count",
    "mr_profile_enable", "mr_profile_reset", "mr_profile_query",
    "mr_filter_plan_create", "mr_filter_plan_destroy", "mr_filter_plan_info", "mr_filter_plan_kernel",
    "mr_filter_rows",
    "mr_filter_rows_adjoint", "mr_bp_parallel_adjoint",
    "mr_bp_plan_create", "mr_bp_plan_destroy", "mr_bp_plan_mask", "mr_bp_plan_kernel", "mr_bp_parallel",
    "mr_bp_parallel_noalloc", "mr_fbp_parallel",
    "mr_fdk_window", "mr_fdk_plan_create", "mr_fdk_plan_destroy", "mr_fdk_plan_window", "mr_fdk_bp",
    "mr_joseph_forward", "mr_joseph_adjoint", "mr_linear_forward",
    "mr_iter_plan_create", "mr_iter_run", "mr_iter_plan_destroy",
    "mr_fbp_stream_create", "mr_fbp_stream_run", "mr_fbp_stream_run_slabs", "mr_fbp_stream_destroy", "mr_host_alloc", "mr_host_free",
    "mr_copy_h2d", "mr_copy_d2h", "mr_volume_zxy_to_xyz",
    "mr_lognorm_reference", "mr_lognorm", "mr_lognorm_logdose", "mr_filter_rows_lognorm",
    "mr_lognorm_lut", "mr_lognorm_lut_term",
    "mr_batch_plan_create", "mr_batch_plan_info", "mr_batch_run", "mr_batch_run_graph",
    "mr_batch_run_sequential", "mr_batch_plan_destroy",
)


class JosephGeom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_size_t), ("nz", ctypes.c_size_t), ("n_det", ctypes.c_size_t),
                ("center", ctypes.c_float), ("direction", ctypes.c_int)]


class BatchStack(ctypes.Structure):
    _fields_ = [("width", ctypes.c_size_t), ("slices", ctypes.c_size_t), ("n_proj", ctypes.c_size_t),
                ("center", ctypes.c_float), ("direction", ctypes.c_int),
                ("filter_type", ctypes.c_int), ("cutoff", ctypes.c_float), ("order", ctypes.c_float),
                ("use_bias", ctypes.c_int), ("bias_weight", ctypes.c_float),
                ("padding_doubler", ctypes.c_size_t), ("buffer_size", ctypes.c_int),
                ("d_proj", ctypes.c_void_p), ("angles", ctypes.POINTER(ctypes.c_float)),
                ("weights", ctypes.POINTER(ctypes.c_float)), ("d_volume", ctypes.c_void_p)]


class MrError(RuntimeError):
    """A non-zero mr_status from the library (the reference throws ReconException)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[mr_status {code}] {msg}")
        self.code = code


class BpParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_size_t), ("slices", ctypes.c_size_t),
                ("center", ctypes.c_float), ("direction", ctypes.c_int),
                ("interp", ctypes.c_int), ("buffer_size", ctypes.c_int),
                ("multiprojbp", ctypes.c_int), ("use_roi", ctypes.c_int),
                ("roi", ctypes.c_size_t * 4), ("fused", ctypes.c_int),
                ("correct_tilt", ctypes.c_int), ("tilt_angle", ctypes.c_float),
                ("tilt_pivot", ctypes.c_float), ("tilt_row0", ctypes.c_size_t),
                ("tilt_rows", ctypes.c_size_t)]


class FdkParams(ctypes.Structure):
    _fields_ = [("sod", ctypes.c_float), ("sdd", ctypes.c_float),
                ("pixel", ctypes.c_float * 2), ("voxel", ctypes.c_float * 3),
                ("ppoint", ctypes.c_float * 2), ("center", ctypes.c_float),
                ("direction", ctypes.c_int), ("roi", ctypes.c_size_t * 4),
                ("projection_roi3", ctypes.c_size_t), ("use_roi", ctypes.c_int),
                ("matrix_roi", ctypes.c_size_t * 4), ("module", ctypes.c_int), ("fused", ctypes.c_int),
                ("filter", ctypes.c_int), ("cosine_weight", ctypes.c_int),
                ("correct_tilt", ctypes.c_int), ("tilt_angle", ctypes.c_float),
                ("tilt_pivot", ctypes.c_float), ("tilt_rows", ctypes.c_size_t)]

FDK_MODULES = {"FDKbp_single": 0, "FDKbp": 1}


_lib = None


def load():
    """Load libmuhrec_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with imagingsuite_amd.build.build() "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, f32, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_int
    fp = ctypes.POINTER(ctypes.c_float)
    szp = ctypes.POINTER(ctypes.c_size_t)
    L.mr_last_error.restype = ctypes.c_char_p
    L.mr_version.restype = i32
    L.mr_device_count.restype = i32
    L.mr_profile_enable.argtypes = [i32]
    L.mr_profile_reset.argtypes = []
    L.mr_profile_query.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_long)]
    L.mr_filter_plan_create.argtypes =[ctypes.POINTER(vp), sz, i32, f32, f32, i32, f32, sz]
    L.mr_filter_plan_destroy.argtypes = [vp]
    L.mr_filter_plan_info.argtypes = [vp, szp, szp, fp]
    L.mr_filter_plan_kernel.argtypes = [vp, ctypes.c_char_p, sz]
    L.mr_filter_rows.argtypes = [vp, vp, sz, vp]
    L.mr_filter_rows_adjoint.argtypes = [vp, vp, sz, vp]
    L.mr_bp_parallel_adjoint.argtypes = [vp, vp, sz, fp, fp, vp, vp]
    L.mr_bp_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(BpParams)]
    L.mr_bp_plan_destroy.argtypes = [vp]
    L.mr_bp_plan_mask.argtypes = [vp, szp, szp]
    L.mr_bp_plan_kernel.argtypes = [vp, ctypes.c_char_p, sz]
    L.mr_bp_parallel.argtypes = [vp, vp, sz, fp, fp, vp, i32, vp]
    L.mr_copy_h2d.argtypes = [vp, vp, sz, vp]
    L.mr_copy_d2h.argtypes = [vp, vp, sz, vp]
    L.mr_volume_zxy_to_xyz.argtypes = [vp, vp, sz, sz, vp]
    L.mr_bp_parallel_noalloc.argtypes = [ctypes.POINTER(BpParams), vp, sz, fp, fp, vp, i32, vp]
    L.mr_fbp_parallel.argtypes = [vp, vp, vp, sz, fp, fp, vp, vp]
    L.mr_fdk_window.argtypes = [ctypes.POINTER(FdkParams), szp]
    L.mr_fdk_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(FdkParams)]
    L.mr_fdk_plan_destroy.argtypes = [vp]
    L.mr_fdk_plan_window.argtypes = [vp, szp]
    L.mr_fdk_bp.argtypes = [vp, vp, sz, sz, fp, fp, vp, vp]
    for f in (L.mr_joseph_forward, L.mr_joseph_adjoint):
        f.argtypes = [ctypes.POINTER(JosephGeom), vp, fp, sz, vp, vp]
    L.mr_linear_forward.argtypes = [vp, sz, sz, fp, sz, vp, vp]
    L.mr_iter_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(JosephGeom), fp, sz, i32]
    L.mr_iter_run.argtypes = [vp, vp, vp, i32, f32, i32, ctypes.POINTER(ctypes.c_double), vp]
    L.mr_iter_plan_destroy.argtypes = [vp]
    L.mr_fbp_stream_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(BpParams), sz, sz, sz, i32, f32, f32,
                                       i32, f32, sz]
    L.mr_fbp_stream_run.argtypes = [vp, vp, fp, fp, vp, i32]
    L.mr_fbp_stream_run_slabs.argtypes = [vp, vp, fp, fp, vp, SLAB_FN, vp]
    L.mr_fbp_stream_destroy.argtypes = [vp]
    L.mr_host_alloc.argtypes = [ctypes.POINTER(vp), sz]
    L.mr_host_free.argtypes = [vp]
    L.mr_lognorm_reference.argtypes = [vp, vp, sz, f32, vp]
    L.mr_lognorm.argtypes = [vp, sz, sz, vp, vp, fp, f32, vp]
    L.mr_filter_rows_lognorm.argtypes = [vp, vp, sz, sz, vp, vp, fp, f32, vp]
    L.mr_lognorm_logdose.argtypes = [vp, sz, sz, vp, vp, fp, vp]
    L.mr_lognorm_lut.argtypes = [vp, sz, sz, vp, vp, fp, f32, vp]
    L.mr_lognorm_lut_term.argtypes = [vp, sz, sz, vp, vp, fp, vp]
    L.mr_batch_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(BatchStack), sz, i32]
    L.mr_batch_plan_info.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), szp,
                                     ctypes.POINTER(ctypes.c_double)]
    for n in ("mr_batch_run", "mr_batch_run_graph", "mr_batch_run_sequential"):
        getattr(L, n).argtypes = [vp, vp]
    L.mr_batch_plan_destroy.argtypes = [vp]
    for name in EXPORTED_SYMBOLS:
        if name not in ("mr_last_error", "mr_version", "mr_device_count"):
            getattr(L, name).restype = i32
    _lib = L
    return L


def profile(enable: bool) -> None:
    """Bracket every kernel launch with HIP events (see mr_profile_enable)."""
    check(load().mr_profile_enable(int(bool(enable))))


def profile_reset() -> None:
    check(load().mr_profile_reset())


def profile_query(kernel: str = ""):
    """(total milliseconds, launches) of kernels whose name contains ``kernel``."""
    ms, n = ctypes.c_double(), ctypes.c_long()
    check(load().mr_profile_query(kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def check(rc: int) -> None:
    if rc != MR_OK:
        msg = load().mr_last_error().decode(errors="replace")
        raise MrError(rc, msg)
