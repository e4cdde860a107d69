"""The objective's Hessian products on the GPU against the oracle:
accumulate_sub_Hessian_times_input_without_penalty (exact) and
add_multiplication_with_approximate_sub_Hessian_without_penalty
(PoissonLogLikelihoodWithLinearModelForMeanAndProjData.cxx:957-1218), the
full-objective versions and the prior terms (GeneralisedObjectiveFunction.cxx:253-425).

Workload: the 8-ring x 192-crystal scanner (span 1, MRD 3, 96 views x 128
tangential, 540 672 bins), 128x128x15 image at 2 mm, 4 subsets, Poisson data
of mean 5 counts/bin (many zero counts: the truncation at the small value
fires), background 0.1; TOF: the same with 9 TOF bins. The forward
projections are bit-identical to the oracle's and the per-viewgram quotient
uses the same float operations, so the products differ from the oracle only
by the back projection's summation order: bar 1e-5 of the product's maximum.
"Zero end planes of segment 0" must not change the products (the reference
reads the measured data and additive term as given).
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OraclePlan, prior_hessian_times_input
from stir_amd import _native, phantoms
from stir_amd.geometry import RayTracingMatrixParameters, config0, config0_tof
from stir_amd.projectors import (PoissonLogLikelihoodWithLinearModelForMeanAndProjData, ProjectorByBinPairMI355X,
                                 ProjMatrixByBinUsingRayTracing, QuadraticPrior)

pytestmark = pytest.mark.gpu
NSUB = 4


@pytest.fixture(scope="module")
def c0():
    pdi, img = config0()
    pm = RayTracingMatrixParameters()
    print(f"[hessian c0] set-up: {pdi.num_bins} bins, image {img.shape}", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=NSUB)
    O = OraclePlan(pdi, img, pm)
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project)
    rng = np.random.default_rng(31)
    b = rng.uniform(0.05, 0.15, pdi.num_bins).astype(np.float32)
    x = (x_true + 0.5).astype(np.float32)
    v = rng.uniform(0.0, 1.0, img.shape).astype(np.float32)
    yield pdi, img, P, O, y, b, x, v
    P.close()


def _close(got, ref, what):
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"{what}: max err / max {err:.2e}")
    assert err <= 1e-5, what


@pytest.mark.parametrize("zero_end_planes", [False, True])
def test_exact_sub_hessian(c0, zero_end_planes):
    pdi, img, P, O, y, b, x, v = c0
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, b, zero_end_planes=zero_end_planes, compute_sensitivities=False)
    for s in (0, 3):
        print(f"[hessian c0] exact subset {s} zero_end_planes={zero_end_planes}", flush=True)
        start = np.full(img.shape, 3.0, np.float32)
        got = obj.accumulate_sub_Hessian_times_input_without_penalty(start.copy(), x, v, s)
        ref = O.hessian_times_input(y, x, v, s, NSUB, bkg=b, out=start.copy())
        _close(got - start, ref - start, f"exact subset {s}")
        assert np.dot((got - start).reshape(-1).astype(np.float64), v.reshape(-1)) < 0


def test_exact_with_an_additive_constant(c0):
    pdi, img, P, O, y, b, x, v = c0
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, 0.1, compute_sensitivities=False)
    got = obj.accumulate_sub_Hessian_times_input_without_penalty(np.zeros(img.shape, np.float32), x, v, 1)
    ref = O.hessian_times_input(y, x, v, 1, NSUB, bkg=np.full(pdi.num_bins, 0.1, np.float32))
    _close(got, ref, "exact, additive constant")


def test_approximate_with_multiplicative_factors(c0):
    pdi, img, P, O, y, b, x, v = c0
    mult = np.random.default_rng(32).uniform(0.5, 2.0, pdi.num_bins).astype(np.float32)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, b, multiplicative=mult, zero_end_planes=True, compute_sensitivities=False)
    O.set_multiplicative(mult)
    try:
        for s in (0, 2):
            print(f"[hessian c0] approximate subset {s} with factors", flush=True)
            got = obj.add_multiplication_with_approximate_sub_Hessian_without_penalty(
                np.zeros(img.shape, np.float32), v, s)
            ref = O.hessian_times_input(y, x, v, s, NSUB, approximate=True)
            _close(got, ref, f"approximate subset {s}, factors")
    finally:
        O.set_multiplicative(None)
    obj.set_up(y, b, compute_sensitivities=False)
    got = obj.add_multiplication_with_approximate_sub_Hessian_without_penalty(np.zeros(img.shape, np.float32), v, 3)
    ref = O.hessian_times_input(y, x, v, 3, NSUB, approximate=True)
    _close(got, ref, "approximate subset 3, no factors")


def test_full_objective_products(c0):
    """Every subset in one pass equals the reference's loop over subsets."""
    pdi, img, P, O, y, b, x, v = c0
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, b, compute_sensitivities=False)
    for approximate in (False, True):
        print(f"[hessian c0] full objective approximate={approximate}", flush=True)
        ref = np.zeros(img.shape, np.float32)
        for s in range(NSUB):
            O.hessian_times_input(y, x, v, s, NSUB, bkg=b, approximate=approximate, out=ref)
        if approximate:
            got = obj.add_multiplication_with_approximate_Hessian_without_penalty(np.zeros(img.shape, np.float32), v)
        else:
            got = obj.accumulate_Hessian_times_input_without_penalty(np.zeros(img.shape, np.float32), x, v)
        _close(got, ref, f"full, approximate={approximate}")


def test_prior_terms(c0):
    """The prior's Hessian product / num_subsets per subset, once for the
    full exact product, num_subsets times / num_subsets for the full
    approximate one."""
    pdi, img, P, O, y, b, x, v = c0
    beta = 0.5
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, b, compute_sensitivities=False)
    obj.set_prior(QuadraticPrior(beta))
    vox = (img.vox_z, img.vox_y, img.vox_x)
    try:
        ph = prior_hessian_times_input("quadratic", x, v, vox, beta)
        pa = prior_hessian_times_input("approximate quadratic", x, v, vox, beta)
        data = O.hessian_times_input(y, x, v, 2, NSUB, bkg=b)
        got = obj.accumulate_sub_Hessian_times_input(np.zeros(img.shape, np.float32), x, v, 2)
        _close(got, data - ph / np.float32(NSUB), "exact subset 2 with prior")
        full = np.zeros(img.shape, np.float32)
        for s in range(NSUB):
            O.hessian_times_input(y, x, v, s, NSUB, bkg=b, out=full)
        got = obj.accumulate_Hessian_times_input(np.zeros(img.shape, np.float32), x, v)
        _close(got, full - ph, "exact full with prior")
        fa = np.zeros(img.shape, np.float32)
        for s in range(NSUB):
            O.hessian_times_input(y, x, v, s, NSUB, approximate=True, out=fa)
            fa -= pa / np.float32(NSUB)
        got = obj.add_multiplication_with_approximate_Hessian(np.zeros(img.shape, np.float32), v)
        _close(got, fa, "approximate full with prior")
    finally:
        obj.set_prior(None)


def test_device_tensors_match_host_arrays(c0):
    pdi, img, P, O, y, b, x, v = c0
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, b, compute_sensitivities=False)
    host = obj.accumulate_sub_Hessian_times_input_without_penalty(np.ones(img.shape, np.float32), x, v, 1)
    out = torch.ones(img.shape, dtype=torch.float32, device="cuda")
    r = obj.accumulate_sub_Hessian_times_input_without_penalty(out, torch.from_numpy(x).cuda(),
                                                               torch.from_numpy(v).cuda(), 1)
    assert r is out
    assert np.array_equal(out.cpu().numpy(), host)


def test_negative_forward_projection_is_an_error(c0):
    pdi, img, P, O, y, b, x, v = c0
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, b, compute_sensitivities=False)
    w = v.copy()
    w[7, 64, 64] = -1000.0  # every subset holds bins through the centre: about -600 at least
    assert O.forward_project(w, subset=0, num_subsets=NSUB, out=np.zeros(pdi.num_bins, np.float32)).min() < 0
    start = np.full(img.shape, 2.0, np.float32)
    for approximate in (False, True):
        out = start.copy()
        with pytest.raises(_native.NativeError, match="negative"):
            if approximate:
                obj.add_multiplication_with_approximate_sub_Hessian_without_penalty(out, w, 0)
            else:
                obj.accumulate_sub_Hessian_times_input_without_penalty(out, x, w, 0)
        assert np.array_equal(out, start)
    # the context stays usable
    got = obj.accumulate_sub_Hessian_times_input_without_penalty(np.zeros(img.shape, np.float32), x, v, 0)
    _close(got, O.hessian_times_input(y, x, v, 0, NSUB, bkg=b), "exact subset 0 after an error")


def test_tof_products():
    pdi, img = config0_tof()
    pm = RayTracingMatrixParameters()
    print(f"[hessian tof c0] set-up: {pdi.num_bins} bins ({pdi.num_tof_bins} TOF bins)", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=NSUB)
    try:
        O = OraclePlan(pdi, img, pm)
        rng = np.random.default_rng(33)
        x = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
        v = rng.uniform(0.0, 1.0, img.shape).astype(np.float32)
        b = np.full(pdi.num_bins, 0.02, np.float32)
        y = rng.poisson(O.forward_project(x) + b).astype(np.float32)
        obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
        obj.set_up(y, b, zero_end_planes=True, compute_sensitivities=False)
        for approximate in (False, True):
            print(f"[hessian tof c0] subset 1 approximate={approximate}", flush=True)
            if approximate:
                got = obj.add_multiplication_with_approximate_sub_Hessian_without_penalty(
                    np.zeros(img.shape, np.float32), v, 1)
            else:
                got = obj.accumulate_sub_Hessian_times_input_without_penalty(np.zeros(img.shape, np.float32), x, v, 1)
            ref = O.hessian_times_input(y, x, v, 1, NSUB, bkg=b, approximate=approximate)
            _close(got, ref, f"TOF subset 1, approximate={approximate}")
    finally:
        P.close()
