"""Projector pair and objective function on the MI355X, with STIR's interface.

Mirrors the reference's projector plugin surface
(``ForwardProjectorByBin.h:53-185``, ``BackProjectorByBin.h:50-210``,
``ProjectorByBinPairUsingProjMatrixByBin.cxx:31-107``) and the objective
function / OSMAPOSL entry points used on this path
(``PoissonLogLikelihoodWithLinearModelForMeanAndProjData``,
``OSMAPOSLReconstruction``). All computation happens in the native library
(``libstir_amd_pmrt.so``) on the GPU; arrays cross as numpy (host) or torch
CUDA tensors (device).

Layouts are STIR's: images ``[z][y][x]`` float32 and projection data flat in
``ProjDataInMemory`` order (segments 0,+1,-1,..., then axial, view, tangential).
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Tuple

import numpy as np

from . import _native
from .geometry import (ProjDataInfoCylindrical, RayTracingMatrixParameters, VoxelsOnCartesianGrid,
                       make_geometry_struct)


class ProjMatrixByBinUsingRayTracing:
    """Parameters of the ray-tracing matrix (registered name "Ray Tracing"
    in ``ProjMatrixByBinUsingRayTracing.cxx:67``). The matrix itself is
    built on the device by :class:`ProjectorByBinPairMI355X.set_up`."""

    registered_name = "Ray Tracing"

    def __init__(self, parameters: RayTracingMatrixParameters = RayTracingMatrixParameters(), **kw):
        if kw:
            from dataclasses import replace
            parameters = replace(parameters, **kw)
        self.parameters = parameters

    def set_num_tangential_LORs(self, n: int) -> None:
        from dataclasses import replace
        self.parameters = replace(self.parameters, num_tangential_LORs=n)

    def get_num_tangential_LORs(self) -> int:
        return self.parameters.num_tangential_LORs


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


class ProjectorByBinPairMI355X:
    """Forward/back projector pair sharing one device-resident matrix.

    ``set_up(proj_data_info, image_info, num_subsets)`` replaces
    ``ProjectorByBinPair::set_up`` (``ProjectorByBinPair.cxx:36-47``): it traces
    the basic rows on the device and builds the back-projection table for the
    given number of OSEM subsets (1 for plain projections)."""

    registered_name = "MI355X Ray Tracing"

    def __init__(self, proj_matrix: Optional[ProjMatrixByBinUsingRayTracing] = None, device: int = 0):
        self.proj_matrix = proj_matrix or ProjMatrixByBinUsingRayTracing()
        self._ctx = _native.Context(device)
        self.pdi: Optional[ProjDataInfoCylindrical] = None
        self.img: Optional[VoxelsOnCartesianGrid] = None
        self.num_subsets = 0

    # ------------------------------------------------------------------
    def set_up(self, proj_data_info: ProjDataInfoCylindrical, image_info: VoxelsOnCartesianGrid,
               num_subsets: int = 1) -> None:
        g = make_geometry_struct(proj_data_info, image_info, self.proj_matrix.parameters)
        self._geometry = g
        _native.check(_native.load().mi_set_up(self._ctx.handle, ctypes.byref(g), int(num_subsets)), "mi_set_up")
        self.pdi, self.img, self.num_subsets = proj_data_info, image_info, num_subsets

    @property
    def context(self) -> _native.Context:
        return self._ctx

    def info(self) -> _native.InfoStruct:
        return self._ctx.info()

    def get_symmetry_flags(self) -> Tuple[int, ...]:
        return tuple(self.info().symmetry_flags)

    def subsets_are_approximately_balanced(self) -> bool:
        b = ctypes.c_int()
        _native.check(_native.load().mi_subsets_balanced(self._ctx.handle, ctypes.byref(b)), "mi_subsets_balanced")
        return bool(b.value)

    def subset_view_segments(self, subset: int) -> List[Tuple[int, int]]:
        cap = 4 * self.pdi.num_views * len(self.pdi.num_axial_poss) + 16
        v = np.zeros(cap, np.int32)
        s = np.zeros(cap, np.int32)
        n = ctypes.c_int()
        _native.check(_native.load().mi_subset_view_segments(self._ctx.handle, subset, v, s, cap, ctypes.byref(n)),
                      "mi_subset_view_segments")
        return list(zip(s[:n.value].tolist(), v[:n.value].tolist()))

    def basic_view_segments_in_subset(self, subset: int, num_subsets: int) -> List[Tuple[int, int]]:
        """(view, segment) pairs of ``detail::find_basic_vs_nums_in_subset``
        (``find_basic_vs_nums_in_subset.cxx:31-69``) with this matrix's symmetries."""
        cap = self.pdi.num_views * len(self.pdi.num_axial_poss) + 16
        v = np.zeros(cap, np.int32)
        s = np.zeros(cap, np.int32)
        n = ctypes.c_int()
        _native.check(_native.load().mi_basic_view_segments_in_subset(self._ctx.handle, subset, num_subsets, v, s, cap,
                                                                      ctypes.byref(n)), "mi_basic_view_segments_in_subset")
        return list(zip(v[:n.value].tolist(), s[:n.value].tolist()))

    def related_viewgram_mask(self, basic_view_segments: List[Tuple[int, int]]) -> np.ndarray:
        """Boolean mask over the projection data (ProjDataInMemory order) of the
        RelatedViewgrams of the given basic view-segments, all TOF bins."""
        pdi = self.pdi
        mask = np.zeros(pdi.num_bins, bool)
        ntof = pdi.num_tof_bins
        block = pdi.num_bins // ntof
        offs = pdi.segment_offsets()
        nt, nv = pdi.num_tangential_poss, pdi.num_views
        for view, seg in basic_view_segments:
            for rv, rs in self._related(view, seg):
                nax = pdi.num_axial_poss_of(rs)
                for k in range(ntof):
                    blk = mask[k * block + offs[rs]: k * block + offs[rs] + nax * nv * nt].reshape(nax, nv, nt)
                    blk[:, rv, :] = True
        return mask

    def _related(self, view: int, seg: int) -> List[Tuple[int, int]]:
        # the (view, segment) pairs a basic view-segment's table subset holds:
        # every pair whose basic view-segment is (view, seg)
        key = (view, seg)
        if not hasattr(self, "_related_cache"):
            self._related_cache = {}
            pairs = [(v, s) for s in range(self.pdi.min_segment_num, self.pdi.max_segment_num + 1)
                     for v in range(self.pdi.num_views)]
            bins = np.array([[s, v, 0, 0] for v, s in pairs], np.int32)
            basic, _ = _native.symmetry_operations(self._geometry, bins)
            for (v, s), b in zip(pairs, basic):
                self._related_cache.setdefault((int(b[1]), int(b[0])), []).append((v, s))
        return self._related_cache.get(key, [])

    def get_row(self, segment: int, view: int, axial_pos: int, tang_pos: int):
        """Row of one bin as the device uses it: (z, y, x, w) arrays."""
        cap = 1 << 14
        while True:
            z = np.empty(cap, np.int16); y = np.empty(cap, np.int16); x = np.empty(cap, np.int16)
            w = np.empty(cap, np.float32)
            n = ctypes.c_int()
            _native.check(_native.load().mi_get_row(self._ctx.handle, segment, view, axial_pos, tang_pos, z, y, x, w, cap,
                                                    ctypes.byref(n)), "mi_get_row")
            if n.value <= cap:
                m = n.value
                return z[:m].copy(), y[:m].copy(), x[:m].copy(), w[:m].copy()
            cap = n.value

    # ------------------------------------------------------------------
    def _check_setup(self):
        if self.pdi is None:
            raise RuntimeError("set_up has not been called")

    def forward_project(self, image, subset: int = -1, out=None, stream=None):
        """A x. numpy in -> numpy out (host path); torch CUDA tensor in ->
        torch CUDA tensor out (device path, no host copies). For subset >= 0
        bins outside the subset are zero."""
        self._check_setup()
        if _is_torch(image):
            import torch
            if not image.is_cuda:
                raise ValueError("torch tensors must be on the GPU")
            image = image.contiguous()
            if image.dtype != torch.float32 or image.numel() != self.img.num_voxels:
                raise ValueError("image must be float32 with the image's number of voxels")
            if out is None:
                out = torch.empty(self.pdi.num_bins, dtype=torch.float32, device=image.device)
            st = stream if stream is not None else torch.cuda.current_stream(image.device).cuda_stream
            _native.check(_native.load().mi_forward_project_dev(self._ctx.handle, ctypes.c_void_p(image.data_ptr()),
                                                                ctypes.c_void_p(out.data_ptr()), subset,
                                                                ctypes.c_void_p(st)), "mi_forward_project_dev")
            return out
        image = np.ascontiguousarray(image, dtype=np.float32).reshape(-1)
        if image.size != self.img.num_voxels:
            raise ValueError("image has the wrong number of voxels")
        if out is None:
            out = np.empty(self.pdi.num_bins, np.float32)
        _native.check(_native.load().mi_forward_project(self._ctx.handle, image, out, subset), "mi_forward_project")
        return out

    def back_project(self, proj_data, subset: int = -1, out=None, stream=None):
        """A^T p (overwrites the output image)."""
        self._check_setup()
        if _is_torch(proj_data):
            import torch
            if not proj_data.is_cuda:
                raise ValueError("torch tensors must be on the GPU")
            proj_data = proj_data.contiguous()
            if proj_data.dtype != torch.float32 or proj_data.numel() != self.pdi.num_bins:
                raise ValueError("projection data must be float32 with the data's number of bins")
            if out is None:
                out = torch.empty(self.img.shape, dtype=torch.float32, device=proj_data.device)
            st = stream if stream is not None else torch.cuda.current_stream(proj_data.device).cuda_stream
            _native.check(_native.load().mi_back_project_dev(self._ctx.handle, ctypes.c_void_p(proj_data.data_ptr()),
                                                             ctypes.c_void_p(out.data_ptr()), subset,
                                                             ctypes.c_void_p(st)), "mi_back_project_dev")
            return out
        proj_data = np.ascontiguousarray(proj_data, dtype=np.float32).reshape(-1)
        if proj_data.size != self.pdi.num_bins:
            raise ValueError("projection data has the wrong number of bins")
        if out is None:
            out = np.empty(self.img.shape, np.float32)
        _native.check(_native.load().mi_back_project(self._ctx.handle, proj_data, out.reshape(-1), subset),
                      "mi_back_project")
        return out

    def synchronize(self):
        _native.check(_native.load().mi_synchronize(self._ctx.handle), "mi_synchronize")

    def close(self):
        self._ctx.close()


class ForwardProjectorByBinMI355X:
    """``ForwardProjectorByBin`` plugin view of a :class:`ProjectorByBinPairMI355X`,
    the flow the C++ plugin of the integration notes follows.

    ``set_input`` (``ForwardProjectorByBin.cxx:306-318``) keeps the image on the
    device; ``forward_project(proj_data, subset_num, num_subsets, zero)``
    (``:168-230``) projects the basic view-segments of the subset
    (``find_basic_vs_nums_in_subset``) and nothing else;
    ``forward_project_related_viewgrams`` is one call of
    ``actual_forward_project``'s unit, the RelatedViewgrams of one basic
    view-segment."""

    registered_name = "MI355X Ray Tracing"

    def __init__(self, pair: ProjectorByBinPairMI355X):
        self.pair = pair
        self._has_input = False

    def set_input(self, image) -> None:
        x = np.ascontiguousarray(image, dtype=np.float32).reshape(-1)
        if x.size != self.pair.img.num_voxels:
            raise ValueError("image has the wrong number of voxels")
        _native.check(_native.load().mi_set_image(self.pair.context.handle, x), "mi_set_image")
        self._has_input = True

    def _project(self, vs: List[Tuple[int, int]], proj_data: np.ndarray) -> None:
        if not self._has_input:
            raise RuntimeError("You need to call set_input() before forward_project()")
        views = np.array([v for v, _ in vs], np.int32)
        segs = np.array([s for _, s in vs], np.int32)
        _native.check(_native.load().mi_forward_project_view_segments(self.pair.context.handle, len(vs), views, segs,
                                                                      proj_data), "mi_forward_project_view_segments")

    def forward_project(self, proj_data: Optional[np.ndarray] = None, subset_num: int = 0, num_subsets: int = 1,
                        zero: bool = True) -> np.ndarray:
        if subset_num < 0 or subset_num >= num_subsets:
            raise ValueError(f"forward_project: wrong subset number {subset_num}")
        if proj_data is None:
            proj_data = np.zeros(self.pair.pdi.num_bins, np.float32)
        elif zero and num_subsets > 1:
            proj_data.fill(0.0)
        if proj_data.dtype != np.float32 or not proj_data.flags.c_contiguous or proj_data.size != self.pair.pdi.num_bins:
            raise ValueError("projection data must be a contiguous float32 array of the data's number of bins")
        self._project(self.pair.basic_view_segments_in_subset(subset_num, num_subsets), proj_data.reshape(-1))
        return proj_data

    def forward_project_related_viewgrams(self, view: int, segment: int, proj_data: np.ndarray) -> None:
        self._project([(view, segment)], proj_data.reshape(-1))


class BackProjectorByBinMI355X:
    """``BackProjectorByBin`` plugin view of a :class:`ProjectorByBinPairMI355X`:
    ``start_accumulating_in_new_target`` / ``back_project`` / ``get_output``
    (``BackProjectorByBin.cxx:183-229,300-346``). ``back_project`` stages the
    RelatedViewgrams of the subset's basic view-segments (viewgrams staged
    twice are summed); ``get_output`` back-projects the staged viewgrams only."""

    registered_name = "MI355X Ray Tracing"

    def __init__(self, pair: ProjectorByBinPairMI355X):
        self.pair = pair
        self._staged: Optional[np.ndarray] = None
        self._vs: List[Tuple[int, int]] = []

    def start_accumulating_in_new_target(self) -> None:
        self._staged = np.zeros(self.pair.pdi.num_bins, np.float32)
        self._vs = []
        self._mask = np.zeros(self.pair.pdi.num_bins, bool)

    def back_project(self, proj_data: np.ndarray, subset_num: int = 0, num_subsets: int = 1) -> None:
        if self._staged is None:
            raise RuntimeError("You need to call start_accumulating_in_new_target() before back_project()")
        vs = self.pair.basic_view_segments_in_subset(subset_num, num_subsets)
        m = self.pair.related_viewgram_mask(vs)
        y = np.asarray(proj_data, np.float32).reshape(-1)
        self._staged[m] += y[m]
        self._vs += [p for p in vs if p not in self._vs]

    def get_output(self) -> np.ndarray:
        out = np.empty(self.pair.img.shape, np.float32)
        views = np.array([v for v, _ in self._vs], np.int32)
        segs = np.array([s for _, s in self._vs], np.int32)
        staged = self._staged if self._staged is not None else np.zeros(self.pair.pdi.num_bins, np.float32)
        _native.check(_native.load().mi_back_project_view_segments(self.pair.context.handle, len(self._vs), views, segs,
                                                                   staged, out.reshape(-1)),
                      "mi_back_project_view_segments")
        return out


class PoissonLogLikelihoodWithLinearModelForMeanAndProjData:
    """Prior-free Poisson log-likelihood with data resident on the device.

    ``set_up`` takes the measured data y, an additive term b (array or
    constant) and "zero end planes of segment 0"; sensitivities are
    per-subset (``use_subset_sensitivities`` defaults to true in the reference,
    ``PoissonLogLikelihoodWithLinearModelForMean.cxx:48``)."""

    def __init__(self, projector: ProjectorByBinPairMI355X):
        self.projector = projector
        self._ctx = projector.context

    def set_up(self, measured: np.ndarray, additive=0.0, zero_end_planes: bool = False,
               compute_sensitivities: bool = True, use_tof_sensitivities: bool = False,
               multiplicative: Optional[np.ndarray] = None, use_subset_sensitivities: bool = True) -> None:
        """use_tof_sensitivities: "use time-of-flight sensitivities" (default off:
        TOF data get non-TOF sensitivities, ``...AndProjData.cxx:533-536``).
        multiplicative: the factors m of ybar = m (A x + b) -- the
        normalisation's ``undo`` of ones (``distributable.cxx:197-215``), e.g.
        1 / the "From ProjData" normalisation factors -- or None.
        use_subset_sensitivities: the reference's key (default true); false
        divides every subset's update by total sensitivity / num_subsets."""
        L = _native.load()
        _native.check(L.mi_obj_set_use_tof_sensitivities(self._ctx.handle, int(use_tof_sensitivities)),
                      "mi_obj_set_use_tof_sensitivities")
        _native.check(L.mi_obj_set_use_subset_sensitivities(self._ctx.handle, int(use_subset_sensitivities)),
                      "mi_obj_set_use_subset_sensitivities")
        self.set_multiplicative(multiplicative)
        y = np.ascontiguousarray(measured, dtype=np.float32).reshape(-1)
        if isinstance(additive, np.ndarray):
            ap, keep = _native.float_ptr_or_none(additive)
            self._keep = keep
            _native.check(L.mi_obj_set_data(self._ctx.handle, y, ap, ctypes.c_float(0.0), int(zero_end_planes)),
                          "mi_obj_set_data")
        else:
            _native.check(L.mi_obj_set_data(self._ctx.handle, y, None, ctypes.c_float(float(additive)),
                                            int(zero_end_planes)), "mi_obj_set_data")
        if compute_sensitivities:
            _native.check(L.mi_obj_compute_sensitivities(self._ctx.handle), "mi_obj_compute_sensitivities")

    def set_multiplicative(self, multiplicative: Optional[np.ndarray]) -> None:
        """Multiplicative factors (all TOF bins, measured order) or None; the
        sensitivities must be recomputed (``set_up`` does)."""
        if multiplicative is None:
            _native.check(_native.load().mi_obj_set_multiplicative(self._ctx.handle, None), "mi_obj_set_multiplicative")
            return
        m = np.ascontiguousarray(multiplicative, dtype=np.float32).reshape(-1)
        _native.check(_native.load().mi_obj_set_multiplicative(self._ctx.handle, m.ctypes.data_as(ctypes.c_void_p)),
                      "mi_obj_set_multiplicative")

    def compute_sensitivities(self) -> None:
        _native.check(_native.load().mi_obj_compute_sensitivities(self._ctx.handle), "mi_obj_compute_sensitivities")

    def set_prior(self, prior: Optional["QuadraticPrior"]) -> None:
        """Prior of OSMAPOSL's one-step-late update (``GeneralisedObjectiveFunction::set_prior_sptr``)."""
        self.prior = prior
        if prior is None:
            _native.check(_native.load().mi_obj_set_prior(self._ctx.handle, None), "mi_obj_set_prior")
            return
        st, keep = prior._struct()
        _native.check(_native.load().mi_obj_set_prior(self._ctx.handle, ctypes.byref(st)), "mi_obj_set_prior")
        del keep

    def get_sensitivity(self) -> np.ndarray:
        """Total sensitivity: the subset sensitivities accumulated in subset
        order (``get_sensitivity``, ``...ForMean.cxx:147-153,419-455``)."""
        return self.get_subset_sensitivity(-1)

    def get_subset_sensitivity(self, subset: int) -> np.ndarray:
        out = np.empty(self.projector.img.shape, np.float32)
        _native.check(_native.load().mi_obj_get_sensitivity(self._ctx.handle, subset, out.reshape(-1)),
                      "mi_obj_get_sensitivity")
        return out

    def set_estimate(self, image: np.ndarray, enforce_positivity: bool = False) -> None:
        x = np.ascontiguousarray(image, dtype=np.float32).reshape(-1)
        _native.check(_native.load().mi_obj_set_estimate(self._ctx.handle, x, int(enforce_positivity)),
                      "mi_obj_set_estimate")

    def get_estimate(self) -> np.ndarray:
        out = np.empty(self.projector.img.shape, np.float32)
        _native.check(_native.load().mi_obj_get_estimate(self._ctx.handle, out.reshape(-1)), "mi_obj_get_estimate")
        return out

    def compute_sub_gradient_without_penalty(self, current_estimate: np.ndarray, subset: int = -1,
                                             plus_sensitivity: bool = False) -> np.ndarray:
        self.set_estimate(current_estimate)
        out = np.empty(self.projector.img.shape, np.float32)
        _native.check(_native.load().mi_obj_subset_gradient(self._ctx.handle, subset, int(plus_sensitivity),
                                                            out.reshape(-1)), "mi_obj_subset_gradient")
        return out

    def compute_objective_function_without_penalty(self, current_estimate: np.ndarray, subset: int = -1) -> float:
        """The log-likelihood of one subset, or of all with subset -1."""
        self.set_estimate(current_estimate)
        v = ctypes.c_double()
        _native.check(_native.load().mi_obj_log_likelihood(self._ctx.handle, subset, ctypes.byref(v)),
                      "mi_obj_log_likelihood")
        return float(v.value)

    # With a prior (set_prior) the penalised forms subtract it, as
    # GeneralisedObjectiveFunction.cxx:113-252 does: the prior's value and
    # gradient over num_subsets for one subset, whole for the full objective.
    def _prior_active(self) -> bool:
        prior = getattr(self, "prior", None)
        return prior is not None and prior.penalisation_factor != 0.0

    def compute_penalty(self, current_estimate: np.ndarray, subset: int = -1) -> float:
        """``compute_penalty``: the prior's value (/ num_subsets for one subset); 0 without a prior."""
        if not self._prior_active():
            return 0.0
        p = self.prior.compute_value(current_estimate, self.projector)
        return p / self._num_subsets() if subset >= 0 else p

    def compute_objective_function(self, current_estimate: np.ndarray, subset: int = -1) -> float:
        """``compute_objective_function``: log-likelihood minus the penalty."""
        return (self.compute_objective_function_without_penalty(current_estimate, subset)
                - self.compute_penalty(current_estimate, subset))

    def compute_sub_gradient(self, current_estimate: np.ndarray, subset: int) -> np.ndarray:
        """``compute_sub_gradient``: the subset's gradient minus the prior's gradient / num_subsets."""
        if subset < 0 or subset >= self._num_subsets():
            raise ValueError("compute_sub_gradient subset_num out-of-range error")
        g = self.compute_sub_gradient_without_penalty(current_estimate, subset)
        if self._prior_active():
            pg = self.prior.compute_gradient(current_estimate, self.projector).astype(np.float32)
            g -= pg / np.float32(self._num_subsets())
        return g

    def compute_gradient_without_penalty(self, current_estimate: np.ndarray) -> np.ndarray:
        """Every subset's gradient (all subsets in one pass)."""
        return self.compute_sub_gradient_without_penalty(current_estimate, -1)

    def compute_gradient(self, current_estimate: np.ndarray) -> np.ndarray:
        """``compute_gradient``: every subset's gradient minus the prior's gradient."""
        g = self.compute_gradient_without_penalty(current_estimate)
        if self._prior_active():
            g -= self.prior.compute_gradient(current_estimate, self.projector).astype(np.float32)
        return g

    # ------------------------------------------------------------------
    # Hessian products (GeneralisedObjectiveFunction.cxx:253-470). Each
    # updates ``output`` in place (output -= ...) and returns it; images are
    # numpy arrays or float32 CUDA tensors (then nothing crosses to the host).
    def accumulate_sub_Hessian_times_input_without_penalty(self, output, current_image_estimate, input,
                                                            subset_num: int):
        """``...AndProjData.cxx:1066-1218``: output -= A^T divide_and_truncate(y A input, (A current + b)^2)."""
        return self._hessian(output, current_image_estimate, input, subset_num, False, 0.0)

    def add_multiplication_with_approximate_sub_Hessian_without_penalty(self, output, input, subset_num: int):
        """``...AndProjData.cxx:957-1064``: output -= A^T divide_and_truncate(A input, y / m / m)."""
        return self._hessian(output, None, input, subset_num, True, 0.0)

    def accumulate_sub_Hessian_times_input(self, output, current_image_estimate, input, subset_num: int):
        """The subset term and the prior's Hessian product / num_subsets (``:398-425``)."""
        return self._hessian(output, current_image_estimate, input, subset_num, False, 1.0 / self._num_subsets())

    def add_multiplication_with_approximate_sub_Hessian(self, output, input, subset_num: int):
        """The subset term and the prior's approximate product / num_subsets (``:280-311``)."""
        return self._hessian(output, None, input, subset_num, True, 1.0 / self._num_subsets())

    def accumulate_Hessian_times_input_without_penalty(self, output, current_image_estimate, input):
        """Every subset (``:383-396``), in one pass over the data."""
        return self._hessian(output, current_image_estimate, input, -1, False, 0.0)

    def add_multiplication_with_approximate_Hessian_without_penalty(self, output, input):
        """Every subset (``:313-326``), in one pass over the data."""
        return self._hessian(output, None, input, -1, True, 0.0)

    def accumulate_Hessian_times_input(self, output, current_image_estimate, input):
        """Every subset and the prior's Hessian product once (``:353-381``)."""
        return self._hessian(output, current_image_estimate, input, -1, False, 1.0)

    def add_multiplication_with_approximate_Hessian(self, output, input):
        """Every subset's ``add_multiplication_with_approximate_sub_Hessian`` (``:328-339``):
        the prior's approximate product / num_subsets subtracted num_subsets times."""
        return self._hessian(output, None, input, -1, True, 1.0 / self._num_subsets(), self._num_subsets())

    def _num_subsets(self) -> int:
        return max(1, int(self.projector.num_subsets))

    def _hessian(self, output, current, inp, subset, approximate, prior_scale, prior_repeats=1):
        import torch

        nvox = self.projector.img.num_voxels

        def dev(a, what):
            if _is_torch(a):
                if not a.is_cuda or a.dtype != torch.float32 or a.numel() != nvox:
                    raise ValueError(f"{what} must be a float32 CUDA tensor with the image's number of voxels")
                return a.contiguous()
            a = np.ascontiguousarray(a, dtype=np.float32)
            if a.size != nvox:
                raise ValueError(f"{what} has the wrong number of voxels")
            return torch.from_numpy(a.reshape(-1)).cuda()

        if not _is_torch(output) and not (isinstance(output, np.ndarray) and output.dtype == np.float32
                                          and output.flags.c_contiguous):
            raise ValueError("output must be a contiguous float32 array (it is updated in place)")
        v = dev(inp, "input")
        x = None if approximate else dev(current, "current image estimate")
        out = dev(output, "output")
        if _is_torch(output) and out.data_ptr() != output.data_ptr():
            raise ValueError("output must be contiguous (it is updated in place)")
        st = torch.cuda.current_stream(v.device).cuda_stream
        _native.check(_native.load().mi_obj_hessian_times_input_dev(
            self._ctx.handle, int(approximate), None if x is None else ctypes.c_void_p(x.data_ptr()),
            ctypes.c_void_p(v.data_ptr()), int(subset), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(st)),
            "mi_obj_hessian_times_input_dev")
        prior = getattr(self, "prior", None)
        if prior_scale and prior is not None and prior.penalisation_factor != 0.0:
            tmp = torch.zeros_like(v)
            prior._loss(self.projector).hessian_times_input(x if x is not None else v, v, tmp, approximate)
            tmp = tmp.reshape(-1)
            for _ in range(prior_repeats):
                if prior_scale == 1.0:
                    out.sub_(tmp)
                else:
                    out.sub_(tmp / self._num_subsets())
        if _is_torch(output):
            return output
        output.reshape(-1)[:] = out.cpu().numpy()
        return output


class PoissonLogLikelihoodWithLinearModelForMeanAndListModeDataWithProjMatrixByBin(
        PoissonLogLikelihoodWithLinearModelForMeanAndProjData):
    """List-mode Poisson log-likelihood with the same matrix
    (``PoissonLogLikelihoodWithLinearModelForMeanAndListModeDataWithProjMatrixByBin.cxx``):
    prompt events (int32 n x 5: segment, view, axial, tangential, timing),
    an optional additive projection-data term and multiplicative factors for
    the sensitivities. Gradients and OSMAPOSL updates sum row^T / (row . x + b)
    over the subset's events (every event: the reference caps only the value);
    ``compute_objective_function`` returns the reference's value, -sum log(row . x + b).
    ``accumulate_sub_Hessian_times_input_without_penalty`` (and the prior and
    full-objective versions) give LM_Hessian's sum over the subset's events;
    the approximate products raise, as the reference has none for list mode."""

    def set_up(self, events: np.ndarray, additive=0.0, multiplicative: Optional[np.ndarray] = None,
               use_subset_sensitivities: bool = True, compute_sensitivities: bool = True,
               use_tof_sensitivities: bool = False) -> None:
        L = _native.load()
        _native.check(L.mi_obj_set_use_tof_sensitivities(self._ctx.handle, int(use_tof_sensitivities)),
                      "mi_obj_set_use_tof_sensitivities")
        _native.check(L.mi_obj_set_use_subset_sensitivities(self._ctx.handle, int(use_subset_sensitivities)),
                      "mi_obj_set_use_subset_sensitivities")
        self.set_multiplicative(multiplicative)
        ev = np.ascontiguousarray(events, dtype=np.int32).reshape(-1, 5)
        if isinstance(additive, np.ndarray):
            a = np.ascontiguousarray(additive, dtype=np.float32).reshape(-1)
            _native.check(L.mi_lm_set_events(self._ctx.handle, ev, ev.shape[0], a.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.c_float(0.0)), "mi_lm_set_events")
        else:
            _native.check(L.mi_lm_set_events(self._ctx.handle, ev, ev.shape[0], None, ctypes.c_float(float(additive))),
                          "mi_lm_set_events")
        if compute_sensitivities:
            self.compute_sensitivities()


class QuadraticPrior:
    """``QuadraticPrior`` (``QuadraticPrior.cxx``) as a prior of the objective
    function: penalisation factor, optional 3x3x3 weights (default vox_x /
    distance), kappa image and only_2D. compute_value / compute_gradient run on
    the device of the projector the prior is used with."""

    kind = _native.MI_PRIOR_QUADRATIC

    def __init__(self, penalisation_factor: float = 0.0, only_2D: bool = False, weights=None, kappa=None):
        self.penalisation_factor = float(penalisation_factor)
        self.only_2D = bool(only_2D)
        self.weights = None if weights is None else np.ascontiguousarray(np.asarray(weights, np.float32).reshape(27))
        self.kappa = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float32)
        self.gamma, self.epsilon = 2.0, 0.0

    def _struct(self):
        st = _native.PriorStruct(self.kind, self.penalisation_factor, self.gamma, self.epsilon, int(self.only_2D),
                                 None if self.weights is None else self.weights.ctypes.data,
                                 None if self.kappa is None else self.kappa.ctypes.data)
        return st, (self.weights, self.kappa)

    def _loss(self, projector: "ProjectorByBinPairMI355X"):
        import torch
        from .autograd import QuadraticPriorLoss

        kappa = None if self.kappa is None else torch.from_numpy(self.kappa.reshape(-1)).cuda()
        return QuadraticPriorLoss(projector, self.penalisation_factor, kappa, self.only_2D, self.weights)

    def compute_value(self, image: np.ndarray, projector: "ProjectorByBinPairMI355X") -> float:
        import torch

        x = torch.from_numpy(np.ascontiguousarray(image, dtype=np.float32)).cuda()
        return float(self._loss(projector)(x).item())

    def compute_gradient(self, image: np.ndarray, projector: "ProjectorByBinPairMI355X") -> np.ndarray:
        import torch

        x = torch.from_numpy(np.ascontiguousarray(image, dtype=np.float32)).cuda().requires_grad_(True)
        self._loss(projector)(x).backward()
        return x.grad.cpu().numpy()


class RelativeDifferencePrior(QuadraticPrior):
    """``RelativeDifferencePrior`` (``RelativeDifferencePrior.cxx``): gamma
    (default 2) and epsilon (default 0) besides the quadratic prior's keys."""

    kind = _native.MI_PRIOR_RELATIVE_DIFFERENCE

    def __init__(self, penalisation_factor: float = 0.0, gamma: float = 2.0, epsilon: float = 0.0,
                 only_2D: bool = False, weights=None, kappa=None):
        super().__init__(penalisation_factor, only_2D, weights, kappa)
        self.gamma, self.epsilon = float(gamma), float(epsilon)

    def _loss(self, projector):
        import torch
        from .autograd import RelativeDifferencePriorLoss

        kappa = None if self.kappa is None else torch.from_numpy(self.kappa.reshape(-1)).cuda()
        return RelativeDifferencePriorLoss(projector, self.penalisation_factor, self.gamma, self.epsilon, kappa,
                                           self.only_2D, self.weights)


FLT_MAX = float(np.finfo(np.float32).max)


class OSMAPOSLReconstruction:
    """OSMAPOSL (``OSMAPOSLReconstruction.cxx:266-516``) on the device: prior-free
    EM updates, or one-step-late MAP updates with the objective's prior.
    MAP_model "additive" (default) or "multiplicative"; minimum / maximum
    relative change bound the multiplicative update after subiteration 1.
    ``uniformly_randomise_subset_order``: every cycle of num_subsets
    subiterations visits the subsets in a new random order
    (``IterativeReconstruction.cxx:575-638``; the reference seeds rand() with
    the time, here ``seed`` seeds numpy's generator, None for the time)."""

    def __init__(self, objective: PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                 num_subiterations: int, start_subset: int = 0, enforce_initial_positivity: bool = True,
                 MAP_model: str = "additive", minimum_relative_change: float = 0.0,
                 maximum_relative_change: float = FLT_MAX, start_subiteration_num: int = 1,
                 uniformly_randomise_subset_order: bool = False, seed: Optional[int] = None):
        if MAP_model not in ("additive", "multiplicative"):
            raise ValueError(f"MAP model should have as value 'additive' or 'multiplicative', while it is {MAP_model!r}")
        self.uniformly_randomise_subset_order = bool(uniformly_randomise_subset_order)
        self._rng = np.random.default_rng(seed)
        self._subset_array: List[int] = []
        self.objective = objective
        self.num_subiterations = num_subiterations
        self.start_subset = start_subset
        self.start_subiteration_num = start_subiteration_num
        self.enforce_initial_positivity = enforce_initial_positivity
        self.MAP_model = MAP_model
        self.minimum_relative_change = float(minimum_relative_change)
        self.maximum_relative_change = float(maximum_relative_change)

    def _set_options(self):
        _native.check(_native.load().mi_osem_set_options(self.objective._ctx.handle,
                                                         int(self.MAP_model == "multiplicative"),
                                                         ctypes.c_float(self.minimum_relative_change),
                                                         ctypes.c_float(self.maximum_relative_change)),
                      "mi_osem_set_options")

    def reconstruct(self, initial_image: np.ndarray) -> np.ndarray:
        proj = self.objective.projector
        if not proj.subsets_are_approximately_balanced():
            raise RuntimeError("OSMAPOSL cannot handle unbalanced subsets")
        self._set_options()
        self.objective.set_estimate(initial_image, enforce_positivity=self.enforce_initial_positivity)
        if not self.uniformly_randomise_subset_order:
            _native.check(_native.load().mi_osem_run(self.objective._ctx.handle, self.start_subiteration_num,
                                                     self.num_subiterations, self.start_subset), "mi_osem_run")
            return self.objective.get_estimate()
        for k in range(self.start_subiteration_num, self.start_subiteration_num + self.num_subiterations):
            _native.check(_native.load().mi_osem_subset_step(self.objective._ctx.handle, self.get_subset_num(k), k),
                          "mi_osem_subset_step")
        return self.objective.get_estimate()

    def get_subset_num(self, subiteration_num: int) -> int:
        """``get_subset_num`` (``IterativeReconstruction.cxx:629-638``): (k + start - 1) mod
        num_subsets, or with random order a new permutation at the start of every cycle."""
        n = max(1, int(self.objective.projector.num_subsets))
        if not self.uniformly_randomise_subset_order:
            return (subiteration_num + self.start_subset - 1) % n
        if (subiteration_num - 1) % n == 0 or len(self._subset_array) != n:
            self._subset_array = [int(v) for v in self._rng.permutation(n)]
        return self._subset_array[(subiteration_num - 1) % n]

    def update_estimate(self, subset: int, subiteration_num: int) -> None:
        self._set_options()
        _native.check(_native.load().mi_osem_subset_step(self.objective._ctx.handle, subset, subiteration_num),
                      "mi_osem_subset_step")
