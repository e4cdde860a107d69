"""ctypes front end of the CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
never by the stir_amd package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Tuple

import numpy as np

from stir_amd.geometry import (GeometryStruct, ProjDataInfoCylindrical, RayTracingMatrixParameters,
                               VoxelsOnCartesianGrid, make_geometry_struct)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborpmrt.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = load_library(_LIB_PATH)
    return _lib


def load_library(path: str):
    """Load an oracle library and declare its signatures (tests load variants
    with planted mistakes this way, to show that the pins catch them)."""
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.or_plan_create.restype = vp
    L.or_plan_create.argtypes = [ctypes.POINTER(GeometryStruct), ctypes.c_char_p, ctypes.c_int]
    L.or_plan_destroy.argtypes = [vp]
    L.or_symmetry_flags.argtypes = [vp, _i32p]
    L.or_num_bins.restype = ctypes.c_int64
    L.or_num_bins.argtypes = [vp]
    L.or_bin_index.restype = ctypes.c_int64
    L.or_bin_index.argtypes = [vp] + [ctypes.c_int] * 4
    for fn in (L.or_get_row, L.or_trace_row):
        fn.restype = ctypes.c_int
        fn.argtypes = [vp] + [ctypes.c_int] * 4 + [_i16p, _i16p, _i16p, _f32p, ctypes.c_int]
    L.or_get_row_tof.restype = ctypes.c_int
    L.or_get_row_tof.argtypes = [vp] + [ctypes.c_int] * 5 + [_i16p, _i16p, _i16p, _f32p, ctypes.c_int]
    L.or_bin_geometry.argtypes = [vp] + [ctypes.c_int] * 4 + [_f32p]
    L.or_ray_trace.restype = ctypes.c_int
    L.or_ray_trace.argtypes = [_f32p, _f32p, _f32p, ctypes.c_float, _i16p, _i16p, _i16p, _f32p, ctypes.c_int]
    L.or_symmetry_op.argtypes = [vp] + [ctypes.c_int] * 4 + [_i32p, _i32p]
    L.or_subset_view_segments.restype = ctypes.c_int
    L.or_subset_view_segments.argtypes = [vp, ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.c_int]
    L.or_subsets_balanced.restype = ctypes.c_int
    L.or_subsets_balanced.argtypes = [vp, ctypes.c_int]
    for fn in (L.or_forward_project, L.or_forward_project_omp, L.or_back_project, L.or_back_project_omp):
        fn.argtypes = [vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int]
    for fn in (L.or_forward_project_f64, L.or_back_project_f64):
        fn.argtypes = [vp, _f64p, _f64p, ctypes.c_int, ctypes.c_int]
    L.or_nnz.restype = ctypes.c_int64
    L.or_nnz.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.or_set_use_tof_sensitivities.argtypes = [vp, ctypes.c_int]
    L.or_subset_sensitivity.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p]
    L.or_subset_gradient.argtypes = [vp, _f32p, vp, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _f32p]
    L.or_subset_ratio.argtypes = [vp, _f32p, vp, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  _f32p]
    L.or_hessian_times_input.restype = ctypes.c_int
    L.or_hessian_times_input.argtypes = [vp, _f32p, vp, vp, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p]
    L.or_log_likelihood.restype = ctypes.c_double
    L.or_log_likelihood.argtypes = [vp, _f32p, vp, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.or_osem.argtypes = [vp, _f32p, vp, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int]
    L.or_quadratic_prior_weights.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, _f32p]
    L.or_quadratic_prior.restype = ctypes.c_double
    L.or_quadratic_prior.argtypes = [_f32p, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int,
                                     ctypes.c_float, vp]
    L.or_relative_difference_prior.restype = ctypes.c_double
    L.or_relative_difference_prior.argtypes = [_f32p, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p,
                                               ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, vp]
    L.or_lm_gradient.restype = ctypes.c_double
    L.or_lm_gradient.argtypes = [vp, _i32p, ctypes.c_int64, vp, _f32p, ctypes.c_int, ctypes.c_int, vp]
    L.or_lm_osem.argtypes = [vp, _i32p, ctypes.c_int64, vp, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.or_lm_hessian.argtypes = [vp, _i32p, ctypes.c_int64, vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, _f32p]
    L.or_prior_hessian_times_input.argtypes = [ctypes.c_int, _f32p, _f32p, vp, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                               ctypes.c_float, _f32p]
    L.or_set_multiplicative.argtypes = [vp, vp]
    L.or_set_use_subset_sensitivities.argtypes = [vp, ctypes.c_int]
    L.or_set_osem_options.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_float]
    L.or_set_prior.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, vp, vp,
                               ctypes.c_int]
    return L


def quadratic_prior_weights(vox_zyx) -> np.ndarray:
    """QuadraticPrior weights [3][3][3] = vox_x/|offset| (0 at the centre)."""
    w = np.empty(27, np.float32)
    lib().or_quadratic_prior_weights(float(vox_zyx[0]), float(vox_zyx[1]), float(vox_zyx[2]), w)
    return w.reshape(3, 3, 3)


def quadratic_prior(image, vox_zyx, beta, kappa=None, only_2D=False, gradient=True):
    """QuadraticPrior value (float) and gradient (float32 image, or None)."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    nz, ny, nx = img.shape
    w = quadratic_prior_weights(vox_zyx).reshape(-1).copy()
    kp = _opt_ptr(None if kappa is None else np.ascontiguousarray(kappa, np.float32).reshape(img.shape))
    g = np.zeros(img.shape, np.float32) if gradient else None
    v = lib().or_quadratic_prior(img.reshape(-1), kp[0] if kp else None, nz, ny, nx, w, int(only_2D), float(beta),
                                 g.ctypes.data_as(ctypes.c_void_p) if g is not None else None)
    return float(v), g


def relative_difference_prior(image, vox_zyx, beta, gamma=2.0, epsilon=0.0, kappa=None, only_2D=False,
                              gradient=True, weights=None):
    """RelativeDifferencePrior value (float) and gradient (float32 image, or None)."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    nz, ny, nx = img.shape
    w = (quadratic_prior_weights(vox_zyx) if weights is None else np.asarray(weights, np.float32)).reshape(-1).copy()
    kp = _opt_ptr(None if kappa is None else np.ascontiguousarray(kappa, np.float32).reshape(img.shape))
    g = np.zeros(img.shape, np.float32) if gradient else None
    v = lib().or_relative_difference_prior(img.reshape(-1), kp[0] if kp else None, nz, ny, nx, w, int(only_2D),
                                           float(beta), float(gamma), float(epsilon),
                                           g.ctypes.data_as(ctypes.c_void_p) if g is not None else None)
    return float(v), g


def prior_hessian_times_input(kind, image, inp, vox_zyx, beta, gamma=2.0, epsilon=0.0, kappa=None, only_2D=False,
                              weights=None, out=None):
    """kind "quadratic" / "relative difference": accumulate_Hessian_times_input;
    "approximate quadratic": the quadratic prior's add_multiplication_with_approximate_Hessian.
    Returns out (zeros if not given) + the product."""
    k = {"quadratic": 1, "relative difference": 2, "approximate quadratic": 3}[kind]
    x = np.ascontiguousarray(image, dtype=np.float32)
    v = np.ascontiguousarray(inp, dtype=np.float32).reshape(x.shape)
    nz, ny, nx = x.shape
    w = (quadratic_prior_weights(vox_zyx) if weights is None else np.asarray(weights, np.float32)).reshape(-1).copy()
    kp = _opt_ptr(None if kappa is None else np.ascontiguousarray(kappa, np.float32).reshape(x.shape))
    o = np.zeros(x.shape, np.float32) if out is None else np.array(out, np.float32, copy=True)
    lib().or_prior_hessian_times_input(k, x.reshape(-1), v.reshape(-1), kp[0] if kp else None, nz, ny, nx, w,
                                       int(only_2D), float(beta), float(gamma), float(epsilon), o.reshape(-1))
    return o


def ray_trace(start_zyx, stop_zyx, vox_zyx, nc):
    """RayTraceVoxelsOnCartesianGrid on one segment: (z, y, x, w)."""
    st = np.asarray(start_zyx, np.float32); sp = np.asarray(stop_zyx, np.float32)
    vx = np.asarray(vox_zyx, np.float32)
    cap = 4096
    while True:
        z = np.empty(cap, np.int16); y = np.empty(cap, np.int16); x = np.empty(cap, np.int16)
        w = np.empty(cap, np.float32)
        n = lib().or_ray_trace(st, sp, vx, float(nc), z, y, x, w, cap)
        if n >= 0:
            return z[:n].copy(), y[:n].copy(), x[:n].copy(), w[:n].copy()
        cap = -n


def _opt_ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.ctypes.data_as(ctypes.c_void_p), a


class OraclePlan:
    """The oracle for one geometry. Arrays use STIR layouts: images
    [z][y][x] (shape (nz,ny,nx)), projection data flat in ProjDataInMemory order."""

    def __init__(self, pdi: ProjDataInfoCylindrical, img: VoxelsOnCartesianGrid,
                 pm: RayTracingMatrixParameters = RayTracingMatrixParameters(), library=None):
        self.pdi, self.img, self.pm = pdi, img, pm
        self._L = library if library is not None else lib()
        self._g = make_geometry_struct(pdi, img, pm)
        err = ctypes.create_string_buffer(256)
        self._p = self._L.or_plan_create(ctypes.byref(self._g), err, 256)
        if not self._p:
            raise ValueError("oracle: " + err.value.decode())
        self.num_bins = int(self._L.or_num_bins(self._p))

    def __del__(self):
        if getattr(self, "_p", None) and getattr(self, "_L", None) is not None:
            self._L.or_plan_destroy(self._p)
            self._p = None

    # ------------------------------------------------------------------
    def symmetry_flags(self) -> Tuple[int, ...]:
        out = np.zeros(5, np.int32)
        self._L.or_symmetry_flags(self._p, out)
        return tuple(int(v) for v in out)

    def bin_index(self, seg, view, ax, tang) -> int:
        return int(self._L.or_bin_index(self._p, seg, view, ax, tang))

    def _row(self, fn, seg, view, ax, tang):
        cap = 4096
        while True:
            z = np.empty(cap, np.int16); y = np.empty(cap, np.int16); x = np.empty(cap, np.int16)
            w = np.empty(cap, np.float32)
            n = fn(self._p, seg, view, ax, tang, z, y, x, w, cap)
            if n >= 0:
                return z[:n].copy(), y[:n].copy(), x[:n].copy(), w[:n].copy()
            cap = -n

    def row(self, seg, view, ax, tang):
        """Row of a bin via basic bin + symmetry operation: (z, y, x, w)."""
        return self._row(self._L.or_get_row, seg, view, ax, tang)

    def row_tof(self, seg, view, ax, tang, tof):
        """TOF row of a bin for TOF bin `tof` (0-based): (z, y, x, w)."""
        cap = 4096
        while True:
            z = np.empty(cap, np.int16); y = np.empty(cap, np.int16); x = np.empty(cap, np.int16)
            w = np.empty(cap, np.float32)
            n = self._L.or_get_row_tof(self._p, seg, view, ax, tang, tof, z, y, x, w, cap)
            if n >= 0:
                return z[:n].copy(), y[:n].copy(), x[:n].copy(), w[:n].copy()
            cap = -n

    def trace_row(self, seg, view, ax, tang):
        """Row traced directly for this bin (no symmetry)."""
        return self._row(self._L.or_trace_row, seg, view, ax, tang)

    def symmetry_op(self, seg, view, ax, tang):
        """((basic seg, view, ax, tang), (in-plane code, z sign, z add))."""
        b = np.zeros(4, np.int32); o = np.zeros(3, np.int32)
        self._L.or_symmetry_op(self._p, seg, view, ax, tang, b, o)
        return tuple(int(v) for v in b), tuple(int(v) for v in o)

    def bin_geometry(self, seg, view, ax, tang):
        out = np.zeros(5, np.float32)
        self._L.or_bin_geometry(self._p, seg, view, ax, tang, out)
        return out

    def subset_view_segments(self, subset: int, num_subsets: int):
        cap = 4 * self.pdi.num_views * len(self.pdi.num_axial_poss) + 16
        v = np.zeros(cap, np.int32); s = np.zeros(cap, np.int32)
        n = self._L.or_subset_view_segments(self._p, subset, num_subsets, v, s, cap)
        return list(zip(s[:n].tolist(), v[:n].tolist()))

    def subsets_balanced(self, num_subsets: int) -> bool:
        return bool(self._L.or_subsets_balanced(self._p, num_subsets))

    # ------------------------------------------------------------------
    def _img(self, a, dtype=np.float32):
        a = np.ascontiguousarray(a, dtype=dtype)
        assert a.size == self.img.num_voxels
        return a

    def _sino(self, a, dtype=np.float32):
        a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
        assert a.size == self.num_bins
        return a

    def forward_project(self, image, subset=-1, num_subsets=1, omp=True, out=None):
        image = self._img(image)
        if out is None:
            out = np.zeros(self.num_bins, np.float32)
        fn = self._L.or_forward_project_omp if omp else self._L.or_forward_project
        fn(self._p, image.reshape(-1), out, subset, num_subsets)
        return out

    def forward_project_f64(self, image, subset=-1, num_subsets=1):
        image = self._img(image, np.float64)
        out = np.zeros(self.num_bins, np.float64)
        self._L.or_forward_project_f64(self._p, image.reshape(-1), out, subset, num_subsets)
        return out

    def back_project(self, sino, subset=-1, num_subsets=1, omp=False):
        sino = self._sino(sino)
        out = np.zeros(self.img.shape, np.float32)
        fn = self._L.or_back_project_omp if omp else self._L.or_back_project
        fn(self._p, sino, out.reshape(-1), subset, num_subsets)
        return out

    def back_project_f64(self, sino, subset=-1, num_subsets=1):
        sino = self._sino(sino, np.float64)
        out = np.zeros(self.img.shape, np.float64)
        self._L.or_back_project_f64(self._p, sino, out.reshape(-1), subset, num_subsets)
        return out

    def nnz(self, subset=-1, num_subsets=1) -> int:
        return int(self._L.or_nnz(self._p, subset, num_subsets))

    def set_use_tof_sensitivities(self, use: bool) -> None:
        self._L.or_set_use_tof_sensitivities(self._p, int(use))

    def lm_gradient(self, events, x, subset=0, num_subsets=1, additive=None, gradient=True):
        """(value, gradient) of the list-mode objective's subset: events
        int32 n x 5 (segment, view, axial, tangential, timing)."""
        ev = np.ascontiguousarray(events, dtype=np.int32).reshape(-1, 5)
        x = self._img(x)
        add = None if additive is None else self._sino(additive)
        g = np.zeros(self.img.shape, np.float32) if gradient else None
        v = self._L.or_lm_gradient(self._p, ev, ev.shape[0], None if add is None else add.ctypes.data_as(ctypes.c_void_p),
                                 x.reshape(-1), subset, num_subsets,
                                 None if g is None else g.ctypes.data_as(ctypes.c_void_p))
        return float(v), g

    def lm_hessian_times_input(self, events, x, rhs, subset=0, num_subsets=1, additive=None, out=None):
        """out + the list-mode objective's subset Hessian times rhs at x (LM_Hessian event by event)."""
        ev = np.ascontiguousarray(events, dtype=np.int32).reshape(-1, 5)
        x = self._img(x); rhs = self._img(rhs)
        add = None if additive is None else self._sino(additive)
        if out is None:
            out = np.zeros(self.img.shape, np.float32)
        assert out.dtype == np.float32 and out.flags.c_contiguous
        self._L.or_lm_hessian(self._p, ev, ev.shape[0], None if add is None else add.ctypes.data_as(ctypes.c_void_p),
                              x.reshape(-1), rhs.reshape(-1), subset, num_subsets, out.reshape(-1))
        return out

    def lm_osem(self, events, x0, num_subsets, num_subiterations, additive=None, start_subset=0):
        ev = np.ascontiguousarray(events, dtype=np.int32).reshape(-1, 5)
        x = np.array(self._img(x0), dtype=np.float32, copy=True)
        add = None if additive is None else self._sino(additive)
        self._L.or_lm_osem(self._p, ev, ev.shape[0], None if add is None else add.ctypes.data_as(ctypes.c_void_p),
                         x.reshape(-1), num_subsets, num_subiterations, start_subset)
        return x

    def set_multiplicative(self, mult) -> None:
        """Multiplicative factors (kept alive here) or None."""
        self._mult = None if mult is None else self._sino(mult)
        self._L.or_set_multiplicative(self._p, None if self._mult is None else self._mult.ctypes.data_as(ctypes.c_void_p))

    def set_use_subset_sensitivities(self, use: bool) -> None:
        self._L.or_set_use_subset_sensitivities(self._p, int(use))

    def set_osem_options(self, map_model="additive", minimum_relative_change=0.0,
                         maximum_relative_change=float(np.finfo(np.float32).max)) -> None:
        self._L.or_set_osem_options(self._p, int(map_model == "multiplicative"), float(minimum_relative_change),
                                  float(maximum_relative_change))

    def set_prior(self, kind=None, beta=0.0, gamma=2.0, epsilon=0.0, kappa=None, weights=None, only_2D=False) -> None:
        """kind: None, "quadratic" or "relative difference"."""
        t = {None: 0, "quadratic": 1, "relative difference": 2}[kind]
        self._kappa = None if kappa is None else np.ascontiguousarray(kappa, np.float32).reshape(self.img.shape)
        self._w27 = None if weights is None else np.ascontiguousarray(weights, np.float32).reshape(27)
        self._L.or_set_prior(self._p, t, float(beta), float(gamma), float(epsilon),
                           None if self._kappa is None else self._kappa.ctypes.data_as(ctypes.c_void_p),
                           None if self._w27 is None else self._w27.ctypes.data_as(ctypes.c_void_p), int(only_2D))

    def subset_sensitivity(self, subset, num_subsets, zero_end_planes=False):
        out = np.zeros(self.img.shape, np.float32)
        self._L.or_subset_sensitivity(self._p, subset, num_subsets, int(zero_end_planes), out.reshape(-1))
        return out

    def subset_gradient(self, y, x, subset, num_subsets, bkg=None, zero_end_planes=False, add_sensitivity=False):
        y = self._sino(y); x = self._img(x)
        bp = _opt_ptr(None if bkg is None else self._sino(bkg))
        out = np.zeros(self.img.shape, np.float32)
        self._L.or_subset_gradient(self._p, y, bp[0] if bp else None, x.reshape(-1), subset, num_subsets,
                                 int(zero_end_planes), int(add_sensitivity), out.reshape(-1))
        return out

    def subset_ratio(self, y, x, subset, num_subsets, bkg=None, zero_end_planes=False, add_sensitivity=True):
        """The projection data the gradient back-projects: divide_and_truncate(y, A x + b)
        (minus the multiplicative factors or 1 unless add_sensitivity)."""
        y = self._sino(y); x = self._img(x)
        bp = _opt_ptr(None if bkg is None else self._sino(bkg))
        out = np.zeros(self.num_bins, np.float32)
        self._L.or_subset_ratio(self._p, y, bp[0] if bp else None, x.reshape(-1), subset, num_subsets,
                              int(zero_end_planes), int(add_sensitivity), out)
        return out

    def hessian_times_input(self, y, current, inp, subset, num_subsets, bkg=None, approximate=False, out=None):
        """out -= BP(q) for the subset's data-term Hessian product (exact: q =
        divide_and_truncate(y * A inp, (A current + b)^2); approximate: q =
        divide_and_truncate(A inp, y / m / m)). Raises ValueError where the
        reference aborts (a negative value in A inp)."""
        y = self._sino(y); inp = self._img(inp)
        cur = None if approximate else self._img(current)
        bp = _opt_ptr(None if bkg is None else self._sino(bkg))
        if out is None:
            out = np.zeros(self.img.shape, np.float32)
        assert out.dtype == np.float32 and out.flags.c_contiguous
        rc = self._L.or_hessian_times_input(self._p, y, bp[0] if bp else None,
                                            None if cur is None else cur.ctypes.data_as(ctypes.c_void_p),
                                            inp.reshape(-1), subset, num_subsets, int(bool(approximate)),
                                            out.reshape(-1))
        if rc:
            raise ValueError("forward projection of the input has negative values")
        return out

    def log_likelihood(self, y, x, bkg=None, subset=-1, num_subsets=1, zero_end_planes=False) -> float:
        y = self._sino(y); x = self._img(x)
        bp = _opt_ptr(None if bkg is None else self._sino(bkg))
        return float(self._L.or_log_likelihood(self._p, y, bp[0] if bp else None, x.reshape(-1), subset,
                                             num_subsets, int(zero_end_planes)))

    def osem(self, y, x0, num_subsets, num_subiterations, bkg=None, start_subset=0, zero_end_planes=False,
             omp=True):
        y = self._sino(y)
        x = np.array(self._img(x0), dtype=np.float32, copy=True)
        bp = _opt_ptr(None if bkg is None else self._sino(bkg))
        self._L.or_osem(self._p, y, bp[0] if bp else None, x.reshape(-1), num_subsets, num_subiterations,
                      start_subset, int(zero_end_planes), int(omp))
        return x
