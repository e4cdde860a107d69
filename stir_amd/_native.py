"""ctypes bindings of libstir_amd_pmrt.so (the C ABI in include/stir_amd_pmrt.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C stir_amd/csrc``). There is no fallback: if the library is missing
or a call fails, a ``NativeError`` is raised.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from .geometry import GeometryStruct

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstir_amd_pmrt.so")

# every symbol declared in include/stir_amd_pmrt.h
EXPORTED_SYMBOLS = (
    "mi_ctx_create", "mi_ctx_destroy", "mi_last_error", "mi_set_up", "mi_get_info", "mi_subsets_balanced",
    "mi_subset_view_segments", "mi_get_row", "mi_forward_project", "mi_back_project", "mi_forward_project_dev",
    "mi_back_project_dev", "mi_obj_set_data", "mi_obj_compute_sensitivities", "mi_obj_get_sensitivity",
    "mi_obj_set_estimate", "mi_obj_get_estimate", "mi_osem_subset_step", "mi_osem_run", "mi_obj_subset_gradient",
    "mi_obj_log_likelihood", "mi_synchronize", "mi_set_kernel_timing", "mi_get_kernel_timing",
    "mi_symmetry_operations", "mi_set_tuning", "mi_obj_value_and_gradient_dev", "mi_quadratic_prior_dev",
    "mi_osem_subset_step_dev", "mi_osem_subset_step_vjp_dev", "mi_obj_set_use_tof_sensitivities",
    "mi_batch_create", "mi_batch_destroy", "mi_batch_context", "mi_batch_set_data", "mi_batch_set_estimate",
    "mi_batch_get_estimate", "mi_batch_set_launch", "mi_batch_tune", "mi_batch_osem_run", "mi_batch_synchronize",
    "mi_obj_set_multiplicative", "mi_obj_set_use_subset_sensitivities", "mi_osem_set_options", "mi_obj_set_prior",
    "mi_relative_difference_prior_dev", "mi_batch_set_row_groups", "mi_lm_set_events",
    "mi_prior_hessian_times_input_dev", "mi_basic_view_segments_in_subset", "mi_set_image",
    "mi_forward_project_view_segments", "mi_back_project_view_segments", "mi_quadratic_prior_noalloc_dev",
    "mi_obj_hessian_times_input_dev",
)

MI_PRIOR_NONE, MI_PRIOR_QUADRATIC, MI_PRIOR_RELATIVE_DIFFERENCE = 0, 1, 2


class PriorStruct(ctypes.Structure):
    """mi_prior"""
    _fields_ = [
        ("type", ctypes.c_int32),
        ("penalisation_factor", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("only_2D", ctypes.c_int32),
        ("weights", ctypes.c_void_p),
        ("kappa", ctypes.c_void_p),
    ]

MI_TUNE_BP_TILE_X, MI_TUNE_BP_TILE_Y, MI_TUNE_BP_XCD_REMAP, MI_TUNE_BP_SLOT_SKIP, MI_TUNE_FP_XCD_CHUNK = 1, 2, 3, 4, 5
MI_TUNE_BP_SLAB, MI_TUNE_SLAB_TILE_X, MI_TUNE_SLAB_TILE_Y, MI_TUNE_SLAB_COLUMNS_PER_WAVE, MI_TUNE_SLAB_LDS_KB = 6, 7, 8, 9, 10
MI_TUNE_TOF_ROWS = 11
MI_TUNE_BP_QUAD = 12
MI_TUNE_BP_QUAD_UNROLL = 13
MI_TUNE_BP_SPLIT_SUBSETS = 14


class NativeError(RuntimeError):
    pass


class InfoStruct(ctypes.Structure):
    _fields_ = [
        ("num_bins", ctypes.c_int64),
        ("num_voxels", ctypes.c_int64),
        ("num_basic_rows", ctypes.c_int64),
        ("num_basic_elements", ctypes.c_int64),
        ("num_families", ctypes.c_int64),
        ("num_matrix_elements", ctypes.c_int64),
        ("num_table_entries", ctypes.c_int64),
        ("num_nonzeros", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("num_subsets", ctypes.c_int32),
        ("symmetry_flags", ctypes.c_int32 * 5),
        ("nz_padded", ctypes.c_int32),
        ("planes_per_axial_position", ctypes.c_int32),
        ("fp_design_bytes", ctypes.c_int64),
        ("bp_design_bytes", ctypes.c_int64),
        ("slab_back_projection", ctypes.c_int32),
        ("slab_tile_x", ctypes.c_int32),
        ("slab_tile_y", ctypes.c_int32),
        ("slab_segment_groups", ctypes.c_int32),
        ("slab_lds_bytes", ctypes.c_int32),
        ("tof_fp_window", ctypes.c_int32),
        ("tof_bp_window", ctypes.c_int32),
        ("tof_live_span", ctypes.c_int32),
        ("bp_quad", ctypes.c_int32),
        ("num_quad_entries", ctypes.c_int64),
    ]


_lib = None
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def load():
    """Load the native library (raises NativeError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C stir_amd/csrc")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    c_int = ctypes.c_int
    L.mi_last_error.restype = ctypes.c_char_p
    L.mi_last_error.argtypes = []
    L.mi_ctx_create.argtypes = [c_int, ctypes.POINTER(vp)]
    L.mi_ctx_destroy.argtypes = [vp]
    L.mi_ctx_destroy.restype = None
    L.mi_set_up.argtypes = [vp, ctypes.POINTER(GeometryStruct), c_int]
    L.mi_get_info.argtypes = [vp, ctypes.POINTER(InfoStruct)]
    L.mi_subsets_balanced.argtypes = [vp, ctypes.POINTER(c_int)]
    L.mi_subset_view_segments.argtypes = [vp, c_int, _i32p, _i32p, c_int, ctypes.POINTER(c_int)]
    L.mi_get_row.argtypes = [vp, c_int, c_int, c_int, c_int, _i16p, _i16p, _i16p, _f32p, c_int, ctypes.POINTER(c_int)]
    L.mi_forward_project.argtypes = [vp, _f32p, _f32p, c_int]
    L.mi_back_project.argtypes = [vp, _f32p, _f32p, c_int]
    L.mi_forward_project_dev.argtypes = [vp, vp, vp, c_int, vp]
    L.mi_back_project_dev.argtypes = [vp, vp, vp, c_int, vp]
    L.mi_obj_set_data.argtypes = [vp, _f32p, vp, ctypes.c_float, c_int]
    L.mi_obj_compute_sensitivities.argtypes = [vp]
    L.mi_obj_set_use_tof_sensitivities.argtypes = [vp, c_int]
    L.mi_obj_get_sensitivity.argtypes = [vp, c_int, _f32p]
    L.mi_obj_set_estimate.argtypes = [vp, _f32p, c_int]
    L.mi_obj_get_estimate.argtypes = [vp, _f32p]
    L.mi_osem_subset_step.argtypes = [vp, c_int, c_int]
    L.mi_osem_run.argtypes = [vp, c_int, c_int, c_int]
    L.mi_obj_subset_gradient.argtypes = [vp, c_int, c_int, _f32p]
    L.mi_obj_log_likelihood.argtypes = [vp, c_int, ctypes.POINTER(ctypes.c_double)]
    L.mi_synchronize.argtypes = [vp]
    L.mi_symmetry_operations.argtypes = [ctypes.POINTER(GeometryStruct), c_int, _i32p, _i32p, _i32p]
    L.mi_set_tuning.argtypes = [vp, c_int, c_int]
    L.mi_obj_value_and_gradient_dev.argtypes = [vp, vp, c_int, c_int, vp, vp, vp]
    L.mi_quadratic_prior_dev.argtypes = [vp, vp, vp, vp, ctypes.c_float, c_int, vp, vp, vp]
    L.mi_quadratic_prior_noalloc_dev.argtypes = [c_int, c_int, c_int, c_int, ctypes.c_float, ctypes.c_float,
                                                ctypes.c_float, vp, vp, vp, ctypes.c_float, c_int, vp, vp, vp]
    L.mi_osem_subset_step_dev.argtypes = [vp, vp, vp, vp, c_int, c_int, vp]
    L.mi_osem_subset_step_vjp_dev.argtypes = [vp, vp, vp, vp, vp, c_int, c_int, vp]
    L.mi_set_kernel_timing.argtypes = [vp, c_int]
    L.mi_get_kernel_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
    L.mi_batch_create.argtypes = [c_int, c_int, ctypes.POINTER(GeometryStruct), c_int, ctypes.POINTER(vp)]
    L.mi_batch_destroy.argtypes = [vp]
    L.mi_batch_destroy.restype = None
    L.mi_batch_context.argtypes = [vp, c_int, ctypes.POINTER(vp)]
    L.mi_batch_set_data.argtypes = [vp, c_int, _f32p, vp, ctypes.c_float, c_int]
    L.mi_batch_set_estimate.argtypes = [vp, c_int, _f32p, c_int]
    L.mi_batch_get_estimate.argtypes = [vp, c_int, _f32p]
    L.mi_batch_set_launch.argtypes = [vp, c_int, c_int, c_int, c_int, c_int]
    L.mi_batch_tune.argtypes = [vp, ctypes.POINTER(c_int)]
    L.mi_batch_osem_run.argtypes = [vp, c_int, c_int, c_int, c_int]
    L.mi_batch_synchronize.argtypes = [vp]
    L.mi_batch_set_row_groups.argtypes = [vp, c_int]
    L.mi_lm_set_events.argtypes = [vp, _i32p, ctypes.c_int64, vp, ctypes.c_float]
    L.mi_prior_hessian_times_input_dev.argtypes = [vp, c_int, c_int, vp, vp, vp, vp, ctypes.c_float, ctypes.c_float,
                                                   ctypes.c_float, c_int, vp, vp]
    L.mi_obj_hessian_times_input_dev.argtypes = [vp, c_int, vp, vp, c_int, vp, vp]
    L.mi_basic_view_segments_in_subset.argtypes = [vp, c_int, c_int, _i32p, _i32p, c_int, ctypes.POINTER(c_int)]
    L.mi_set_image.argtypes = [vp, _f32p]
    L.mi_forward_project_view_segments.argtypes = [vp, c_int, _i32p, _i32p, _f32p]
    L.mi_back_project_view_segments.argtypes = [vp, c_int, _i32p, _i32p, _f32p, _f32p]
    L.mi_obj_set_multiplicative.argtypes = [vp, vp]
    L.mi_obj_set_use_subset_sensitivities.argtypes = [vp, c_int]
    L.mi_osem_set_options.argtypes = [vp, c_int, ctypes.c_float, ctypes.c_float]
    L.mi_obj_set_prior.argtypes = [vp, ctypes.POINTER(PriorStruct)]
    L.mi_relative_difference_prior_dev.argtypes = [vp, vp, vp, vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                   c_int, vp, vp, vp]
    for name in EXPORTED_SYMBOLS:
        fn = getattr(L, name)
        if name not in ("mi_last_error", "mi_ctx_destroy", "mi_batch_destroy"):
            fn.restype = c_int
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().mi_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed ({rc}): {msg}")


class Context:
    """One device context (mi_ctx)."""

    def __init__(self, device: int = 0):
        L = load()
        h = ctypes.c_void_p()
        check(L.mi_ctx_create(device, ctypes.byref(h)), "mi_ctx_create")
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().mi_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> InfoStruct:
        out = InfoStruct()
        check(load().mi_get_info(self._h, ctypes.byref(out)), "mi_get_info")
        return out

    def set_tuning(self, param: int, value: int) -> None:
        check(load().mi_set_tuning(self._h, int(param), int(value)), "mi_set_tuning")

    def set_kernel_timing(self, enable: bool) -> None:
        check(load().mi_set_kernel_timing(self._h, int(enable)), "mi_set_kernel_timing")

    def kernel_timing(self):
        """(fp_ms, fp_launches, bp_ms, bp_launches) since the last call."""
        f, b = ctypes.c_double(), ctypes.c_double()
        nf, nb = ctypes.c_int64(), ctypes.c_int64()
        check(load().mi_get_kernel_timing(self._h, ctypes.byref(f), ctypes.byref(nf), ctypes.byref(b), ctypes.byref(nb)),
              "mi_get_kernel_timing")
        return f.value, nf.value, b.value, nb.value


def float_ptr_or_none(a: Optional[np.ndarray]):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    return a.ctypes.data_as(ctypes.c_void_p), a


def symmetry_operations(geometry: GeometryStruct, bins: np.ndarray):
    """Host-side (no device) basic bins and symmetry operations of bins (n x 4)."""
    bins = np.ascontiguousarray(bins, dtype=np.int32).reshape(-1, 4)
    basic = np.zeros_like(bins)
    ops = np.zeros((bins.shape[0], 3), np.int32)
    check(load().mi_symmetry_operations(ctypes.byref(geometry), bins.shape[0], bins.reshape(-1), basic.reshape(-1),
                                        ops.reshape(-1)), "mi_symmetry_operations")
    return basic, ops
