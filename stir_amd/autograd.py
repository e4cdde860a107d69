"""Differentiable projector pair, objective and OSEM map for PyTorch-ROCm.

* ``forward_project`` / ``back_project``: ``torch.autograd.Function`` s whose
  forward is the HIP forward (back) projection and whose backward is the
  hand-written HIP back (forward) projection, its exact transpose.
* ``PoissonLogLikelihoodLoss``: the negative Poisson log-likelihood of
  ``A x + b`` (``...AndProjData.cxx:739-781``). Its forward runs one fused
  forward projection that accumulates the value in FP64 and writes the
  sub-gradient seed, then one back projection; backward only scales the
  stored gradient image.
* ``QuadraticPriorLoss``: ``QuadraticPrior`` value and gradient
  (``QuadraticPrior.cxx:202-367``) in one HIP kernel.
* ``UnrolledOSEM``: OSEM subiterations as a differentiable map of the
  initial image; each subiteration keeps only its input image and update
  image (image checkpoints) and its backward recomputes the forward
  projections (two forward and one back projection).

Images are ``[z][y][x]`` (or a batch ``[B][z][y][x]`` where stated) and
projection data are flat in STIR order, as float32 CUDA tensors; work is
issued on the current stream and nothing is copied to the host. Calls that
share one projector context must run on one stream.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import torch

from . import _native
from .projectors import PoissonLogLikelihoodWithLinearModelForMeanAndProjData, ProjectorByBinPairMI355X


def _stream(t: torch.Tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _as_batch(t: torch.Tensor, n: int, what: str) -> torch.Tensor:
    if not t.is_cuda or t.dtype != torch.float32:
        raise ValueError(f"{what} must be a float32 CUDA tensor")
    if t.numel() == 0 or t.numel() % n != 0:
        raise ValueError(f"{what} has {t.numel()} elements, not a positive multiple of {n}")
    return t.contiguous().reshape(-1, n)


# ---------------------------------------------------------------------------
# projections
class _ForwardProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, pair: ProjectorByBinPairMI355X, subset: int):
        nvox, nbins = pair.img.num_voxels, pair.pdi.num_bins
        xs = _as_batch(x, nvox, "image")
        ctx.pair, ctx.subset, ctx.in_shape = pair, subset, x.shape
        out = torch.empty((xs.shape[0], nbins), dtype=torch.float32, device=x.device)
        for b in range(xs.shape[0]):
            pair.forward_project(xs[b], subset=subset, out=out[b])
        return out if x.numel() != nvox else out[0]

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        pair = ctx.pair
        gs = _as_batch(g, pair.pdi.num_bins, "gradient")
        out = torch.empty((gs.shape[0], pair.img.num_voxels), dtype=torch.float32, device=g.device)
        for b in range(gs.shape[0]):
            pair.back_project(gs[b], subset=ctx.subset, out=out[b])
        return out.reshape(ctx.in_shape), None, None


class _BackProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p: torch.Tensor, pair: ProjectorByBinPairMI355X, subset: int):
        nvox, nbins = pair.img.num_voxels, pair.pdi.num_bins
        ps = _as_batch(p, nbins, "projection data")
        ctx.pair, ctx.subset, ctx.in_shape = pair, subset, p.shape
        out = torch.empty((ps.shape[0],) + tuple(pair.img.shape), dtype=torch.float32, device=p.device)
        for b in range(ps.shape[0]):
            pair.back_project(ps[b], subset=subset, out=out[b])
        return out if p.numel() != nbins else out[0]

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        pair = ctx.pair
        gs = _as_batch(g, pair.img.num_voxels, "gradient")
        out = torch.empty((gs.shape[0], pair.pdi.num_bins), dtype=torch.float32, device=g.device)
        for b in range(gs.shape[0]):
            pair.forward_project(gs[b], subset=ctx.subset, out=out[b])
        return out.reshape(ctx.in_shape), None, None


def forward_project(x: torch.Tensor, pair: ProjectorByBinPairMI355X, subset: int = -1) -> torch.Tensor:
    """A x for an image or a batch of images (differentiable: d/dx is A^T).
    For subset >= 0 only that subset's bins are projected, the others are 0."""
    return _ForwardProject.apply(x, pair, subset)


def back_project(p: torch.Tensor, pair: ProjectorByBinPairMI355X, subset: int = -1) -> torch.Tensor:
    """A^T p for flat projection data or a batch [B][num_bins] (differentiable: d/dp is A)."""
    return _BackProject.apply(p, pair, subset)


# ---------------------------------------------------------------------------
# objective
class _NegLogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, obj: PoissonLogLikelihoodWithLinearModelForMeanAndProjData, subset: int,
                want_grad: bool):
        pair = obj.projector
        nvox = pair.img.num_voxels
        xs = _as_batch(x, nvox, "image")
        values = torch.empty(xs.shape[0], dtype=torch.float64, device=x.device)
        grads = torch.empty_like(xs) if want_grad else None
        L, st = _native.load(), _stream(x)
        for b in range(xs.shape[0]):
            _native.check(L.mi_obj_value_and_gradient_dev(pair.context.handle, _ptr(xs[b]), subset, 0,
                                                          _ptr(grads[b]) if want_grad else None, _ptr(values[b:b + 1]),
                                                          st), "mi_obj_value_and_gradient_dev")
        ctx.in_shape = x.shape
        if want_grad:
            ctx.save_for_backward(grads)
        out = -values
        return out if x.numel() != nvox else out[0]

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        (grads,) = ctx.saved_tensors
        # d(-loglik)/dx = -(sub-gradient)
        scale = (-g.reshape(-1, 1)).to(torch.float32)
        return (grads * scale).reshape(ctx.in_shape), None, None, None


class PoissonLogLikelihoodLoss:
    """Negative Poisson log-likelihood -L(x), L = sum y log(ybar) - ybar with
    ybar = A x + b, thresholded as the reference's
    accumulate_loglikelihood (``recon_array_functions.cxx:336-377``); its
    gradient is minus ``compute_sub_gradient_without_penalty``
    (``...AndProjData.cxx:694-735``). Data and additive term are those given
    to ``objective.set_up``. Returns FP64 (one value per image of a batch)."""

    def __init__(self, objective: PoissonLogLikelihoodWithLinearModelForMeanAndProjData, subset: int = -1):
        self.objective = objective
        self.subset = subset

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        # grad mode is off inside Function.forward: decide here whether the
        # gradient image is computed along with the value
        want_grad = torch.is_grad_enabled() and x.requires_grad
        return _NegLogLikelihood.apply(x, self.objective, self.subset, want_grad)


class _NeighbourhoodPrior(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, prior, want_grad: bool):
        pair = prior.pair
        nvox = pair.img.num_voxels
        xs = _as_batch(x, nvox, "image")
        values = torch.empty(xs.shape[0], dtype=torch.float64, device=x.device)
        grads = torch.empty_like(xs) if want_grad else None
        st = _stream(x)
        for b in range(xs.shape[0]):
            prior._native(xs[b], grads[b] if want_grad else None, values[b:b + 1], st)
        ctx.in_shape = x.shape
        if want_grad:
            ctx.save_for_backward(grads)
        return values if x.numel() != nvox else values[0]

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        (grads,) = ctx.saved_tensors
        return (grads * g.reshape(-1, 1).to(torch.float32)).reshape(ctx.in_shape), None, None


class QuadraticPriorLoss:
    """``QuadraticPrior``: beta * sum_v sum_n w_vn k_v k_n (x_v - x_n)^2 / 4 over
    the 3x3x3 neighbourhood inside the image (``QuadraticPrior.cxx:229-298``),
    gradient beta * sum_n w_vn k_v k_n (x_v - x_n) (``:301-367``). Default
    weights vox_x / distance (``:202-226``); ``weights`` may give a 3x3x3
    array instead. ``kappa``: optional device image. FP64 value per image."""

    def __init__(self, pair: ProjectorByBinPairMI355X, beta: float, kappa: Optional[torch.Tensor] = None,
                 only_2D: bool = False, weights: Optional[Sequence[float]] = None):
        import numpy as np

        self.pair = pair
        self.beta = float(beta)
        self.only_2D = bool(only_2D)
        if kappa is not None:
            kappa = _as_batch(kappa, pair.img.num_voxels, "kappa")[0]
        self.kappa = kappa
        self._w = None if weights is None else np.ascontiguousarray(np.asarray(weights, np.float32).reshape(27))
        self._type = _native.MI_PRIOR_QUADRATIC
        self._gamma_eps = (0.0, 0.0)

    def _weights_arg(self):
        return None if self._w is None else self._w.ctypes.data_as(ctypes.c_void_p)

    def _native(self, x, grad, value, st):
        _native.check(_native.load().mi_quadratic_prior_dev(self.pair.context.handle, _ptr(x), _ptr(self.kappa),
                                                            self._weights_arg(), ctypes.c_float(self.beta),
                                                            int(self.only_2D), _ptr(grad), _ptr(value), st),
                      "mi_quadratic_prior_dev")

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        want_grad = torch.is_grad_enabled() and x.requires_grad
        return _NeighbourhoodPrior.apply(x, self, want_grad)

    def hessian_times_input(self, x: torch.Tensor, v: torch.Tensor, out: Optional[torch.Tensor] = None,
                            approximate: bool = False) -> torch.Tensor:
        """out + H(x) v (``accumulate_Hessian_times_input``), or with approximate
        the quadratic prior's ``add_multiplication_with_approximate_Hessian``."""
        nvox = self.pair.img.num_voxels
        x = _as_batch(x, nvox, "image")[0]
        v = _as_batch(v, nvox, "input")[0]
        out = torch.zeros_like(v) if out is None else _as_batch(out, nvox, "output")[0]
        g, e = self._gamma_eps
        _native.check(_native.load().mi_prior_hessian_times_input_dev(
            self.pair.context.handle, self._type, int(approximate), _ptr(x), _ptr(v), _ptr(self.kappa),
            self._weights_arg(), ctypes.c_float(self.beta), ctypes.c_float(g), ctypes.c_float(e), int(self.only_2D),
            _ptr(out), _stream(v)), "mi_prior_hessian_times_input_dev")
        return out.reshape(self.pair.img.shape)


class RelativeDifferencePriorLoss(QuadraticPriorLoss):
    """``RelativeDifferencePrior``: beta * sum_v sum_n w_vn k_v k_n 0.5 (x_v - x_n)^2 /
    (x_v + x_n + gamma |x_v - x_n| + epsilon) (``RelativeDifferencePrior.cxx:280-373``),
    gradient beta * sum_n w_vn k_v k_n derivative_10(x_v, x_n) (``:286-300,375-433``);
    reference defaults gamma 2, epsilon 0 (``:140-147``)."""

    def __init__(self, pair: ProjectorByBinPairMI355X, beta: float, gamma: float = 2.0, epsilon: float = 0.0,
                 kappa: Optional[torch.Tensor] = None, only_2D: bool = False,
                 weights: Optional[Sequence[float]] = None):
        super().__init__(pair, beta, kappa, only_2D, weights)
        self.gamma = float(gamma)
        self.epsilon = float(epsilon)
        self._type = _native.MI_PRIOR_RELATIVE_DIFFERENCE
        self._gamma_eps = (self.gamma, self.epsilon)

    def _native(self, x, grad, value, st):
        _native.check(_native.load().mi_relative_difference_prior_dev(
            self.pair.context.handle, _ptr(x), _ptr(self.kappa), self._weights_arg(), ctypes.c_float(self.beta),
            ctypes.c_float(self.gamma), ctypes.c_float(self.epsilon), int(self.only_2D), _ptr(grad), _ptr(value), st),
            "mi_relative_difference_prior_dev")


# ---------------------------------------------------------------------------
# unrolled OSEM
FLT_MAX = 3.4028234663852886e38


def _set_bounds(obj: PoissonLogLikelihoodWithLinearModelForMeanAndProjData, lo: float, hi: float) -> None:
    # OSMAPOSL "minimum / maximum relative change" on the objective's context
    _native.check(_native.load().mi_osem_set_options(obj.projector.context.handle, 0, ctypes.c_float(lo),
                                                     ctypes.c_float(hi)), "mi_osem_set_options")


class _OSEMSubiteration(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, obj: PoissonLogLikelihoodWithLinearModelForMeanAndProjData, subset: int,
                subiteration_num: int, bounds):
        pair = obj.projector
        x = _as_batch(x, pair.img.num_voxels, "image")[0].reshape(pair.img.shape)
        out = torch.empty_like(x)
        u = torch.empty_like(x)
        _set_bounds(obj, *bounds)
        _native.check(_native.load().mi_osem_subset_step_dev(pair.context.handle, _ptr(x), _ptr(out), _ptr(u), subset,
                                                             subiteration_num, _stream(x)), "mi_osem_subset_step_dev")
        ctx.obj, ctx.subset, ctx.subiteration_num, ctx.bounds = obj, subset, subiteration_num, bounds
        ctx.save_for_backward(x, u)
        return out

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        x, u = ctx.saved_tensors
        g = g.contiguous()
        gin = torch.empty_like(x)
        _set_bounds(ctx.obj, *ctx.bounds)
        _native.check(_native.load().mi_osem_subset_step_vjp_dev(ctx.obj.projector.context.handle, _ptr(x), _ptr(u),
                                                                 _ptr(g), _ptr(gin), ctx.subset, ctx.subiteration_num,
                                                                 _stream(x)),
                      "mi_osem_subset_step_vjp_dev")
        return gin, None, None, None, None


def osem_subiteration(x: torch.Tensor, objective: PoissonLogLikelihoodWithLinearModelForMeanAndProjData, subset: int,
                      subiteration_num: int, minimum_relative_change: float = 0.0,
                      maximum_relative_change: float = FLT_MAX) -> torch.Tensor:
    """One OSMAPOSL subiteration x * clamp(A_s^T(y/(A_s x + b)) / A_s^T 1)
    (``OSMAPOSLReconstruction.cxx:373-516``), differentiable in x. The clamp
    to [minimum, maximum relative change] applies after subiteration 1
    (``:491-507``); where it acts, the update does not depend on x."""
    bounds = (float(minimum_relative_change), float(maximum_relative_change))
    return _OSEMSubiteration.apply(x, objective, subset, subiteration_num, bounds)


class UnrolledOSEM(torch.nn.Module):
    """``num_subiterations`` OSEM subiterations (subset order
    ``(k + start_subset - 1) % num_subsets``, ``IterativeReconstruction.cxx:630-639``)
    as a differentiable map of the initial image. Memory: two images per
    subiteration (its input and its update); the backward pass recomputes the
    forward projections. Needs ``objective.set_up`` with sensitivities."""

    def __init__(self, objective: PoissonLogLikelihoodWithLinearModelForMeanAndProjData, num_subiterations: int,
                 start_subset: int = 0, minimum_relative_change: float = 0.0,
                 maximum_relative_change: float = FLT_MAX):
        super().__init__()
        self.objective = objective
        self.num_subiterations = int(num_subiterations)
        self.start_subset = int(start_subset)
        self.bounds = (float(minimum_relative_change), float(maximum_relative_change))

    def forward(self, x0: torch.Tensor) -> torch.Tensor:
        nsub = self.objective.projector.num_subsets
        x = x0
        for k in range(1, self.num_subiterations + 1):
            x = osem_subiteration(x, self.objective, (k + self.start_subset - 1) % nsub, k, *self.bounds)
        return x
