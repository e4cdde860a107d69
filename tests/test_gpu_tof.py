"""Time-of-flight projector pair and OSEM on the GPU against the oracle.

Workload: the 8-ring x 192-crystal scanner (span 1, MRD 3, 96 views x 128
tangential) with 9 TOF bins of 89.04 mm and a 400 ps FWHM kernel
(9 x 540 672 bins), 128x128x15 image at 2 mm, 4 subsets. Bars as without
TOF: forward projection bit-identical to the oracle (same element order and
float order; bars 1e-5 relative); back projection within 1e-5 of the
maximum; OSEM images within 1e-4 relative; objective value within 1e-6 and
sub-gradient within 1e-5 of the maximum. Property checks: the sum over TOF
bins of the TOF forward projection equals the non-TOF projection (1e-5),
and the adjoint identity holds with FP64 sums. Every test runs with the TOF
weights stored at set-up (narrow and wide windows) and computed on the fly.
"""
import numpy as np
import pytest
import torch

from stir_amd import _native, phantoms
from stir_amd.autograd import PoissonLogLikelihoodLoss
from stir_amd.geometry import RayTracingMatrixParameters, config0, config0_tof
from stir_amd.projectors import (OSMAPOSLReconstruction, PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                                 ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing)
from oracle.oracle import OraclePlan

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tof_oracle():
    pdi, img = config0_tof()
    pm = RayTracingMatrixParameters()
    return pdi, img, pm, OraclePlan(pdi, img, pm)


# TOF weights: stored records with the narrowest windows (the default), with
# wide windows, and computed on the fly in every projection
ROW_MODES = {"stored": 1, "stored_wide": 2, "on_the_fly": 0}


@pytest.fixture(scope="module", params=sorted(ROW_MODES))
def tof(request, tof_oracle):
    pdi, img, pm, O = tof_oracle
    print(f"[tof c0 {request.param}] set-up: {pdi.num_bins} bins ({pdi.num_tof_bins} TOF bins), image {img.shape}",
          flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.context.set_tuning(_native.MI_TUNE_TOF_ROWS, ROW_MODES[request.param])
    P.set_up(pdi, img, num_subsets=4)
    P.row_mode = request.param
    yield pdi, img, P, O
    P.close()


def test_tof_setup(tof):
    pdi, img, P, O = tof
    info = P.info()
    assert info.num_bins == pdi.num_bins == O.num_bins
    assert tuple(info.symmetry_flags) == O.symmetry_flags() == (0, 0, 0, 0, 1)
    # 400 ps FWHM (sigma 25.5 mm) against 89 mm bins: the 4-sigma-sqrt(2)
    # support (288 mm) meets at most 5 bins (measured when the rows are stored)
    if P.row_mode != "on_the_fly":
        assert 3 <= info.tof_live_span <= 5
    windows = {"stored": (6, 5), "stored_wide": (14, 13), "on_the_fly": (0, 0)}[P.row_mode]
    assert (info.tof_fp_window, info.tof_bp_window) == windows


def test_tof_forward_projection_bit_exact(tof):
    pdi, img, P, O = tof
    x = np.random.default_rng(1).uniform(0, 1, img.shape).astype(np.float32)
    g = P.forward_project(x)
    r = O.forward_project(x)
    diff = np.abs(g - r)
    print(f"TOF FP: {np.count_nonzero(diff)} of {g.size} bins differ, max rel {diff.max() / np.abs(r).max():.2e}")
    assert diff.max() <= 1e-5 * np.abs(r).max()
    assert np.array_equal(g, r)


def test_tof_back_projection(tof):
    pdi, img, P, O = tof
    y = np.random.default_rng(2).uniform(0, 1, pdi.num_bins).astype(np.float32)
    g = P.back_project(y)
    r = O.back_project_f64(y.astype(np.float64))
    err = np.abs(g - r).max() / np.abs(r).max()
    print(f"TOF BP max err vs FP64 oracle {err:.2e}")
    assert err <= 1e-5


def test_tof_subsets(tof):
    pdi, img, P, O = tof
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, img.shape).astype(np.float32)
    y = rng.uniform(0, 1, pdi.num_bins).astype(np.float32)
    for s in (0, 3):
        g = P.forward_project(x, subset=s)
        r = np.zeros(pdi.num_bins, np.float32)
        O.forward_project(x, subset=s, num_subsets=4, out=r)
        assert np.array_equal(g, r)
        gb = P.back_project(y, subset=s)
        rb = O.back_project_f64(y.astype(np.float64), subset=s, num_subsets=4)
        assert np.abs(gb - rb).max() <= 1e-5 * np.abs(rb).max()


def test_tof_adjoint_fp64(tof):
    pdi, img, P, O = tof
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(4)
    x = torch.rand(img.shape, generator=g, device=dev)
    y = torch.rand(pdi.num_bins, generator=g, device=dev)
    lhs = torch.dot(P.forward_project(x).double(), y.double()).item()
    rhs = torch.dot(x.reshape(-1).double(), P.back_project(y).reshape(-1).double()).item()
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_tof_sum_matches_nontof(tof):
    """Sum over TOF bins = the non-TOF projection with the same (directly
    traced) rows, i.e. with the symmetries TOF disables switched off too (rows
    obtained through symmetries differ from direct ones at the FOV edge)."""
    pdi, img, P, O = tof
    pdi0, img0 = config0()
    P0 = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(RayTracingMatrixParameters(
        do_symmetry_90degrees_min_phi=False, do_symmetry_180degrees_min_phi=False, do_symmetry_swap_segment=False,
        do_symmetry_swap_s=False)))
    P0.set_up(pdi0, img0, num_subsets=1)
    x = np.random.default_rng(5).uniform(0, 1, img.shape).astype(np.float32)
    t = P.forward_project(x).astype(np.float64).reshape(pdi.num_tof_bins, -1).sum(axis=0)
    n = P0.forward_project(x).astype(np.float64)
    P0.close()
    assert np.abs(t - n).max() <= 1e-5 * n.max()


def test_tof_default_sensitivities(tof):
    """Default TOF sensitivities: non-TOF back projections over the non-TOF
    projector's subsets (...AndProjData.cxx:533-536,648-670,884-903)."""
    pdi, img, P, O = tof
    y = np.ones(pdi.num_bins, np.float32)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, 0.1)
    for s in range(4):
        g = obj.get_subset_sensitivity(s)
        r = O.subset_sensitivity(s, 4)
        assert np.abs(g - r).max() <= 1e-5 * r.max()


def test_tof_osem_and_objective(tof):
    """OSEM with TOF sensitivities (the default mixes the non-TOF projector's
    subsets into the TOF update, which leaves a few FOV-edge voxels with a zero
    subset sensitivity and a non-zero update, in the reference as here)."""
    pdi, img, P, O = tof
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project)
    bfull = np.full(pdi.num_bins, bkg, np.float32)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg, use_tof_sensitivities=True)
    got = OSMAPOSLReconstruction(obj, num_subiterations=8).reconstruct(np.ones(img.shape, np.float32))
    O.set_use_tof_sensitivities(True)
    try:
        ref = O.osem(y, np.ones(img.shape, np.float32), 4, 8, bkg=bfull)
    finally:
        O.set_use_tof_sensitivities(False)
    mask = ref > 1e-3 * ref.max()
    rel = (np.abs(got - ref)[mask] / ref[mask]).max()
    print(f"TOF OSEM max rel err {rel:.2e}")
    assert rel <= 1e-4
    xe = np.random.default_rng(6).uniform(0.5, 1.5, img.shape).astype(np.float32)
    L = obj.compute_objective_function(xe)
    Lr = O.log_likelihood(y, xe, bkg=bfull)
    assert abs(L - Lr) <= 1e-6 * abs(Lr)
    gr = obj.compute_sub_gradient_without_penalty(xe, subset=1)
    gro = O.subset_gradient(y, xe, 1, 4, bkg=bfull)
    assert np.abs(gr - gro).max() <= 1e-5 * np.abs(gro).max()
    # autograd loss through the same fused path
    xt = torch.from_numpy(xe).cuda().requires_grad_(True)
    loss = PoissonLogLikelihoodLoss(obj)(xt)
    loss.backward()
    assert abs(-loss.item() - L) <= 1e-12 * abs(L)
    assert np.array_equal((-xt.grad).cpu().numpy(), obj.compute_sub_gradient_without_penalty(xe))
