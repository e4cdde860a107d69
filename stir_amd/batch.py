"""Batches of independent OSMAPOSL reconstructions on one GPU.

Each problem keeps its own geometry, matrix, data and estimate (one native
context per problem, set up exactly as ``ProjectorByBinPairMI355X.set_up``);
a batched subiteration runs every problem's forward projection with one
launch per lane-group size and likewise the back projection, optionally
replayed from a captured HIP graph. ``run_sequential`` runs the same
subiterations problem by problem through the per-problem entry point
(``mi_osem_run``), the baseline the batched path is compared with.
"""
from __future__ import annotations

import ctypes
from typing import Sequence, Tuple

import numpy as np

from . import _native
from .geometry import ProjDataInfoCylindrical, RayTracingMatrixParameters, VoxelsOnCartesianGrid, make_geometry_struct


class BatchedOSMAPOSL:
    """``len(problems)`` prior-free OSMAPOSL reconstructions advanced together.

    problems: sequence of (ProjDataInfoCylindrical, VoxelsOnCartesianGrid,
    RayTracingMatrixParameters); all use ``num_subsets`` subsets."""

    def __init__(self, problems: Sequence[Tuple[ProjDataInfoCylindrical, VoxelsOnCartesianGrid,
                                                RayTracingMatrixParameters]],
                 num_subsets: int, device: int = 0):
        L = _native.load()
        self.problems = list(problems)
        self.num_subsets = num_subsets
        n = len(self.problems)
        geoms = (_native.GeometryStruct * n)()
        for i, (pdi, img, pm) in enumerate(self.problems):
            geoms[i] = make_geometry_struct(pdi, img, pm)
        h = ctypes.c_void_p()
        _native.check(L.mi_batch_create(device, n, geoms, num_subsets, ctypes.byref(h)), "mi_batch_create")
        self._h = h

    def __len__(self) -> int:
        return len(self.problems)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _native.load().mi_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_data(self, i: int, measured: np.ndarray, additive=0.0, zero_end_planes: bool = False) -> None:
        y = np.ascontiguousarray(measured, dtype=np.float32).reshape(-1)
        if isinstance(additive, np.ndarray):
            a, keep = _native.float_ptr_or_none(additive)
            _native.check(_native.load().mi_batch_set_data(self._h, i, y, a, ctypes.c_float(0.0), int(zero_end_planes)),
                          "mi_batch_set_data")
            del keep
        else:
            _native.check(_native.load().mi_batch_set_data(self._h, i, y, None, ctypes.c_float(float(additive)),
                                                           int(zero_end_planes)), "mi_batch_set_data")

    def set_estimate(self, i: int, image: np.ndarray, enforce_positivity: bool = True) -> None:
        x = np.ascontiguousarray(image, dtype=np.float32).reshape(-1)
        _native.check(_native.load().mi_batch_set_estimate(self._h, i, x, int(enforce_positivity)),
                      "mi_batch_set_estimate")

    def get_estimate(self, i: int) -> np.ndarray:
        img = self.problems[i][1]
        out = np.empty(img.shape, np.float32)
        _native.check(_native.load().mi_batch_get_estimate(self._h, i, out.reshape(-1)), "mi_batch_get_estimate")
        return out

    def set_launch(self, fp_group_max: int = 32, fp_widen: bool = False, bp_pack: int = 2, tile_x: int = 32,
                   tile_y: int = 8, row_groups: int = 32) -> None:
        """row_groups: families of at most this many axial positions run in row
        groups (0: off); the others in lane groups of fp_group_max bins or
        full waves; bp_pack: columns of images of at most 32 planes in full
        waves (0), lane groups (1) or waves split over table entries (2)."""
        L = _native.load()
        _native.check(L.mi_batch_set_launch(self._h, fp_group_max, int(fp_widen), int(bp_pack), tile_x, tile_y),
                      "mi_batch_set_launch")
        _native.check(L.mi_batch_set_row_groups(self._h, row_groups), "mi_batch_set_row_groups")

    def tune(self) -> int:
        """Autotuner: times the candidate launch shapes and keeps the fastest."""
        c = ctypes.c_int()
        _native.check(_native.load().mi_batch_tune(self._h, ctypes.byref(c)), "mi_batch_tune")
        return c.value

    def run(self, num_subiterations: int, start_subiteration: int = 1, start_subset: int = 0,
            use_graph: bool = True) -> None:
        _native.check(_native.load().mi_batch_osem_run(self._h, start_subiteration, num_subiterations, start_subset,
                                                       int(use_graph)), "mi_batch_osem_run")

    def synchronize(self) -> None:
        _native.check(_native.load().mi_batch_synchronize(self._h), "mi_batch_synchronize")

    def context(self, i: int) -> ctypes.c_void_p:
        c = ctypes.c_void_p()
        _native.check(_native.load().mi_batch_context(self._h, i, ctypes.byref(c)), "mi_batch_context")
        return c

    def run_sequential(self, num_subiterations: int, start_subiteration: int = 1, start_subset: int = 0) -> None:
        """The same subiterations, one problem after the other with the
        per-problem kernels (``mi_osem_run`` on each problem's context)."""
        L = _native.load()
        for i in range(len(self)):
            c = self.context(i)
            _native.check(L.mi_osem_run(c, start_subiteration, num_subiterations, start_subset), "mi_osem_run")
            _native.check(L.mi_synchronize(c), "mi_synchronize")
