"""Multiplicative factors, total sensitivities, the relative difference prior
and one-step-late MAP OSMAPOSL on the GPU against the oracle.

Workload: the 8-ring x 192-crystal scanner (span 1, MRD 3, 96 views x 128
tangential, 540 672 bins), 128x128x15 image at 2 mm, 4 subsets; phantom and
Poisson data of mean 5 counts/bin, background 0.1; factors uniform in
[0.5, 2]. Bars: OSEM / OSL images within 1e-4 relative of the oracle (voxels
above 1e-3 of the maximum); sensitivities and gradients within 1e-5 of their
maximum; log-likelihood within 1e-6 relative; prior gradients bit-identical
(same float terms, same double accumulation order) and prior values within
1e-12 relative (different order of the double sums).
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OraclePlan, quadratic_prior, relative_difference_prior
from stir_amd import phantoms
from stir_amd.autograd import RelativeDifferencePriorLoss
from stir_amd.geometry import RayTracingMatrixParameters, config0
from stir_amd.projectors import (OSMAPOSLReconstruction, PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                                 ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing, QuadraticPrior,
                                 RelativeDifferencePrior)

pytestmark = pytest.mark.gpu
NSUB = 4


@pytest.fixture(scope="module")
def c0():
    pdi, img = config0()
    pm = RayTracingMatrixParameters()
    print(f"[map c0] set-up: {pdi.num_bins} bins, image {img.shape}", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=NSUB)
    O = OraclePlan(pdi, img, pm)
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project)
    mult = np.random.default_rng(21).uniform(0.5, 2.0, pdi.num_bins).astype(np.float32)
    yield pdi, img, P, O, x_true, y, bkg, mult
    P.close()


def _rel(got, ref):
    mask = ref > 1e-3 * ref.max()
    return float((np.abs(got - ref)[mask] / ref[mask]).max())


def test_multiplicative_objective(c0):
    pdi, img, P, O, x_true, y, bkg, mult = c0
    b = np.full(pdi.num_bins, bkg, np.float32)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg, multiplicative=mult, zero_end_planes=True)
    O.set_multiplicative(mult)
    try:
        for s in (0, 3):
            g, r = obj.get_subset_sensitivity(s), O.subset_sensitivity(s, NSUB, zero_end_planes=True)
            assert np.abs(g - r).max() <= 1e-5 * r.max()
        x = (x_true + 0.5).astype(np.float32)
        L = obj.compute_objective_function(x)
        Lr = O.log_likelihood(y, x, bkg=b, zero_end_planes=True)
        print(f"log-likelihood with factors: {L:.10e} vs oracle {Lr:.10e}")
        assert abs(L - Lr) <= 1e-6 * abs(Lr)
        gr = obj.compute_sub_gradient_without_penalty(x, subset=2)
        gro = O.subset_gradient(y, x, 2, NSUB, bkg=b, zero_end_planes=True)
        assert np.abs(gr - gro).max() <= 1e-5 * np.abs(gro).max()
        got = OSMAPOSLReconstruction(obj, num_subiterations=8).reconstruct(np.ones(img.shape, np.float32))
        ref = O.osem(y, np.ones(img.shape, np.float32), NSUB, 8, bkg=b, zero_end_planes=True)
    finally:
        O.set_multiplicative(None)
        obj.set_up(y, bkg)
    rel = _rel(got, ref)
    print(f"OSEM with factors: max rel {rel:.2e}")
    assert rel <= 1e-4


def test_total_sensitivity_and_relative_change_bounds(c0):
    pdi, img, P, O, x_true, y, bkg, mult = c0
    b = np.full(pdi.num_bins, bkg, np.float32)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg, use_subset_sensitivities=False)
    O.set_use_subset_sensitivities(False)
    O.set_osem_options("additive", 0.5, 1.5)
    try:
        total = sum(O.subset_sensitivity(s, NSUB).astype(np.float64) for s in range(NSUB))
        g0 = obj.get_subset_sensitivity(0)
        assert np.abs(g0 - total / NSUB).max() <= 1e-5 * total.max() / NSUB
        assert np.array_equal(g0, obj.get_subset_sensitivity(3))
        got = OSMAPOSLReconstruction(obj, num_subiterations=8, minimum_relative_change=0.5,
                                     maximum_relative_change=1.5).reconstruct(np.ones(img.shape, np.float32))
        ref = O.osem(y, np.ones(img.shape, np.float32), NSUB, 8, bkg=b)
    finally:
        O.set_use_subset_sensitivities(True)
        O.set_osem_options()
        obj.set_up(y, bkg)
    rel = _rel(got, ref)
    print(f"OSEM, total sensitivity, relative change in [0.5, 1.5]: max rel {rel:.2e}")
    assert rel <= 1e-4


@pytest.mark.parametrize("kind,model", [("quadratic", "additive"), ("relative difference", "additive"),
                                        ("quadratic", "multiplicative"), ("relative difference", "multiplicative")])
def test_one_step_late_osem(c0, kind, model):
    pdi, img, P, O, x_true, y, bkg, mult = c0
    b = np.full(pdi.num_bins, bkg, np.float32)
    kappa = np.random.default_rng(22).uniform(0.5, 1.5, img.shape).astype(np.float32)
    beta = {("quadratic", "additive"): 0.5, ("relative difference", "additive"): 1.0,
            ("quadratic", "multiplicative"): 0.02, ("relative difference", "multiplicative"): 0.05}[(kind, model)]
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg)
    prior = (QuadraticPrior(beta, kappa=kappa) if kind == "quadratic"
             else RelativeDifferencePrior(beta, gamma=2.0, epsilon=1e-3, kappa=kappa))
    obj.set_prior(prior)
    O.set_prior(kind, beta, 2.0, 1e-3, kappa=kappa)
    O.set_osem_options(model)
    try:
        got = OSMAPOSLReconstruction(obj, num_subiterations=8, MAP_model=model).reconstruct(
            np.ones(img.shape, np.float32))
        ref = O.osem(y, np.ones(img.shape, np.float32), NSUB, 8, bkg=b)
    finally:
        obj.set_prior(None)
        O.set_prior(None)
        O.set_osem_options()
    rel = _rel(got, ref)
    free = O.osem(y, np.ones(img.shape, np.float32), NSUB, 8, bkg=b)
    print(f"OSL {kind} {model}: max rel {rel:.2e} (prior changes the image by {_rel(ref, free):.2e})")
    assert rel <= 1e-4
    assert _rel(ref, free) > 1e-3  # the prior matters at this beta


@pytest.mark.parametrize("variant", ["plain", "kappa", "only_2D", "eps"])
def test_relative_difference_prior_kernel(c0, variant):
    pdi, img, P, O, x_true, y, bkg, mult = c0
    xn = np.exp(np.random.default_rng(23).normal(0, 0.5, img.shape)).astype(np.float32)
    xn[0, :4, :4] = 0  # zero pairs (the 0/0 case when eps = 0)
    kappa = np.random.default_rng(24).uniform(0.5, 1.5, img.shape).astype(np.float32) if variant == "kappa" else None
    eps = 0.1 if variant == "eps" else 0.0
    only_2D = variant == "only_2D"
    dev = torch.device("cuda:0")
    loss = RelativeDifferencePriorLoss(P, 0.02, gamma=2.0, epsilon=eps,
                                       kappa=None if kappa is None else torch.from_numpy(kappa.reshape(-1)).to(dev),
                                       only_2D=only_2D)
    x = torch.from_numpy(xn).to(dev).requires_grad_(True)
    v = loss(x)
    v.backward()
    vo, go = relative_difference_prior(xn, (img.vox_z, img.vox_y, img.vox_x), 0.02, 2.0, eps, kappa=kappa,
                                       only_2D=only_2D)
    assert abs(v.item() - vo) <= 1e-12 * abs(vo)
    assert np.array_equal(x.grad.cpu().numpy(), go)


@pytest.mark.parametrize("kind,approximate,use_kappa", [("quadratic", False, False), ("quadratic", True, True),
                                                        ("relative difference", False, False),
                                                        ("relative difference", False, True)])
def test_prior_hessian_times_input(c0, kind, approximate, use_kappa):
    """accumulate_Hessian_times_input / add_multiplication_with_approximate_Hessian
    on the GPU: the same float terms in the same order as the oracle,
    accumulated into a non-zero output -- bit-identical."""
    from oracle.oracle import prior_hessian_times_input
    from stir_amd.autograd import QuadraticPriorLoss
    pdi, img, P, O, x_true, y, bkg, mult = c0
    rng = np.random.default_rng(25)
    xn = np.exp(rng.normal(0, 0.5, img.shape)).astype(np.float32)
    vn = rng.uniform(-1, 1, img.shape).astype(np.float32)
    out0 = rng.uniform(-1, 1, img.shape).astype(np.float32)
    kappa = rng.uniform(0.5, 1.5, img.shape).astype(np.float32) if use_kappa else None
    dev = torch.device("cuda:0")
    kt = None if kappa is None else torch.from_numpy(kappa.reshape(-1)).to(dev)
    loss = (QuadraticPriorLoss(P, 0.03, kappa=kt) if kind == "quadratic"
            else RelativeDifferencePriorLoss(P, 0.03, gamma=2.0, epsilon=0.01, kappa=kt))
    got = loss.hessian_times_input(torch.from_numpy(xn).to(dev), torch.from_numpy(vn).to(dev),
                                   out=torch.from_numpy(out0.copy()).to(dev), approximate=approximate)
    ref = prior_hessian_times_input("approximate quadratic" if approximate else kind, xn, vn,
                                    (img.vox_z, img.vox_y, img.vox_x), 0.03, 2.0, 0.01, kappa=kappa, out=out0)
    assert np.array_equal(got.cpu().numpy(), ref)


@pytest.mark.parametrize("kind", ["quadratic", "relative difference"])
def test_penalised_value_and_gradients(c0, kind):
    """compute_objective_function / compute_penalty / compute_sub_gradient /
    compute_gradient with a prior (GeneralisedObjectiveFunction.cxx:113-252):
    the likelihood terms minus the prior's value and gradient, over
    num_subsets for one subset. Against the oracle's likelihood and prior:
    values within 1e-6 relative, gradients within 1e-5 of their maximum."""
    pdi, img, P, O, x_true, y, bkg, mult = c0
    b = np.full(pdi.num_bins, bkg, np.float32)
    vox = (img.vox_z, img.vox_y, img.vox_x)
    x = (x_true + 0.5).astype(np.float32)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg)
    if kind == "quadratic":
        beta = 0.5
        obj.set_prior(QuadraticPrior(beta))
        pv, pg = quadratic_prior(x, vox, beta)
    else:
        beta = 0.05
        obj.set_prior(RelativeDifferencePrior(beta, gamma=2.0, epsilon=1e-3))
        pv, pg = relative_difference_prior(x, vox, beta, 2.0, 1e-3)
    try:
        s = 1
        L = obj.compute_objective_function(x, subset=s)
        ref = O.log_likelihood(y, x, bkg=b, subset=s, num_subsets=NSUB) - pv / NSUB
        assert abs(L - ref) <= 1e-6 * abs(ref)
        assert abs(obj.compute_penalty(x, subset=s) - pv / NSUB) <= 1e-6 * abs(pv)
        Lf = obj.compute_objective_function(x)
        reff = O.log_likelihood(y, x, bkg=b) - pv
        assert abs(Lf - reff) <= 1e-6 * abs(reff)
        g = obj.compute_sub_gradient(x, s)
        gref = O.subset_gradient(y, x, s, NSUB, bkg=b) - pg / np.float32(NSUB)
        err = np.abs(g - gref).max() / np.abs(gref).max()
        print(f"{kind}: penalised subset gradient max err {err:.2e} of max")
        assert err <= 1e-5
        gf = obj.compute_gradient(x)
        gfref = O.subset_gradient(y, x, -1, 1, bkg=b) - pg
        err = np.abs(gf - gfref).max() / np.abs(gfref).max()
        print(f"{kind}: penalised gradient max err {err:.2e} of max")
        assert err <= 1e-5
        with pytest.raises(ValueError):
            obj.compute_sub_gradient(x, NSUB)
    finally:
        obj.set_prior(None)
