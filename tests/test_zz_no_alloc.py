"""Allocation-free prior entry point (mi_quadratic_prior_noalloc_dev), GPU.

Called through the C ABI on two images in a row, the second smaller (the
fixed partial-sum array keeps the first call's contents), and compared with
mi_quadratic_prior_dev on a context of the same image geometry: gradient
bit-identical, FP64 value identical (the same kernels and reduction order).
An image whose prior needs more workgroups than the fixed array holds is
refused on the host before any launch.
"""
import ctypes

import numpy as np
import pytest
import torch

from stir_amd import _native
from stir_amd.geometry import ProjDataInfoCylindrical, Scanner, VoxelsOnCartesianGrid
from stir_amd.projectors import ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _bind():
    L = _native.load()
    f = L.mi_quadratic_prior_noalloc_dev
    vp = ctypes.c_void_p
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                  ctypes.c_float, vp, vp, vp, ctypes.c_float, ctypes.c_int, vp, vp, vp]
    f.restype = ctypes.c_int
    return f


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _existing(img, x, kappa, beta):
    """mi_quadratic_prior_dev on a context set up for this image geometry."""
    sc = Scanner.cylindrical(4, 96, 150.0, 4.0, name="prior")
    pdi = ProjDataInfoCylindrical.construct(sc, span=1, max_delta=2, num_views=48, num_tangential_poss=64)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing())
    P.set_up(pdi, img, num_subsets=1)
    g = torch.empty_like(x)
    v = torch.zeros(1, dtype=torch.float64, device=DEV)
    stream = ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
    _native.check(_native.load().mi_quadratic_prior_dev(P.context.handle, _ptr(x), _ptr(kappa), None, beta, 0, _ptr(g),
                                                        _ptr(v), stream), "mi_quadratic_prior_dev")
    torch.cuda.synchronize()
    P.close()
    return g.cpu().numpy(), float(v.item())


def _noalloc(f, img, x, kappa, beta):
    g = torch.empty_like(x)
    v = torch.zeros(1, dtype=torch.float64, device=DEV)
    stream = ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
    _native.check(f(0, img.nz, img.ny, img.nx, img.vox_z, img.vox_y, img.vox_x, _ptr(x), _ptr(kappa), None, beta, 0,
                    _ptr(g), _ptr(v), stream), "mi_quadratic_prior_noalloc_dev")
    torch.cuda.synchronize()
    return g.cpu().numpy(), float(v.item())


def test_noalloc_prior_equals_context_prior_twice():
    f = _bind()
    rng = np.random.default_rng(71)
    for img, beta, with_kappa in ((VoxelsOnCartesianGrid(31, 128, 128, 2.0, 2.5, 2.5), 0.01, True),
                                  (VoxelsOnCartesianGrid(7, 64, 64, 2.0, 2.0, 2.0), 0.5, False)):
        x = torch.from_numpy(rng.lognormal(0.0, 0.5, img.shape).astype(np.float32)).to(DEV)
        kappa = torch.from_numpy(rng.uniform(0.5, 1.5, img.shape).astype(np.float32)).to(DEV) if with_kappa else None
        g0, v0 = _existing(img, x, kappa, beta)
        g1, v1 = _noalloc(f, img, x, kappa, beta)
        print(f"image {img.shape}: value {v1:.12e} (context {v0:.12e}), gradient max {np.abs(g1).max():.3e}",
              flush=True)
        assert np.array_equal(g0, g1)
        assert v0 == v1


def test_noalloc_prior_refuses_images_beyond_its_partials():
    f = _bind()
    x = torch.zeros(1, dtype=torch.float32, device=DEV)
    g = torch.zeros(1, dtype=torch.float32, device=DEV)
    # 1100 planes of 256 x 256 voxels: 1100 x 4 x 64 workgroups > 65536; refused before any launch
    rc = f(0, 1100, 256, 256, 2.0, 2.0, 2.0, _ptr(x), None, None, 0.01, 0, _ptr(g), None, None)
    assert rc != 0
    assert b"too large" in _native.load().mi_last_error()
