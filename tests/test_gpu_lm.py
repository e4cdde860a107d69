"""List-mode objective and OSMAPOSL on the GPU against the oracle's
event-by-event restatement.

Workload: the 8-ring x 192-crystal scanner (span 1, MRD 3, 96 x 128 bins),
128x128x15 image, Poisson data of mean 5 counts/bin as one event per count in
a seeded random order (about 2.7 M events), additive term 0.1, 4 subsets;
and the same geometry with 3 TOF bins (events of a mean 0.02 x the TOF
forward projection). Bars: gradients and Hessian products (LM_Hessian: the
device sums the events of one bin as one term, the oracle event by event)
within 1e-5 of their maximum, the value within 1e-9 relative (FP64 sums in a
different order), OSEM images within 1e-4 relative.
"""
import numpy as np
import pytest

from oracle.oracle import OraclePlan
from stir_amd import _native, phantoms
from stir_amd.geometry import RayTracingMatrixParameters, config0, config0_tof
from stir_amd.projectors import (OSMAPOSLReconstruction,
                                 PoissonLogLikelihoodWithLinearModelForMeanAndListModeDataWithProjMatrixByBin,
                                 ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing)

pytestmark = pytest.mark.gpu
NSUB = 4


@pytest.fixture(scope="module")
def lm():
    pdi, img = config0()
    pm = RayTracingMatrixParameters()
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=NSUB)
    O = OraclePlan(pdi, img, pm)
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project)
    ev = pdi.events_from_counts(y, seed=3)
    print(f"[lm] {ev.shape[0]} events", flush=True)
    add = np.full(pdi.num_bins, bkg, np.float32)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndListModeDataWithProjMatrixByBin(P)
    obj.set_up(ev, add)
    yield pdi, img, P, O, x_true, ev, add, obj
    P.close()


def test_lm_gradient_and_value(lm):
    pdi, img, P, O, x_true, ev, add, obj = lm
    x = (x_true + 0.2).astype(np.float32)
    for s, with_sens in ((1, True), (2, False)):
        g = obj.compute_sub_gradient_without_penalty(x, subset=s, plus_sensitivity=with_sens)
        _, r = O.lm_gradient(ev, x, s, NSUB, additive=add)
        if not with_sens:
            r = r - O.subset_sensitivity(s, NSUB)
        err = np.abs(g - r).max() / np.abs(r).max()
        print(f"LM gradient subset {s} (plus sensitivity {with_sens}): max err {err:.2e} of max")
        assert err <= 1e-5
    v = obj.compute_objective_function(x, subset=1)
    vr, _ = O.lm_gradient(ev, x, 1, NSUB, additive=add, gradient=False)
    assert abs(v - vr) <= 1e-9 * abs(vr)


def test_lm_osem(lm):
    pdi, img, P, O, x_true, ev, add, obj = lm
    got = OSMAPOSLReconstruction(obj, num_subiterations=8).reconstruct(np.ones(img.shape, np.float32))
    ref = O.lm_osem(ev, np.ones(img.shape, np.float32), NSUB, 8, additive=add)
    m = ref > 1e-3 * ref.max()
    rel = float((np.abs(got - ref)[m] / ref[m]).max())
    print(f"LM OSEM max rel {rel:.2e}")
    assert rel <= 1e-4


def test_lm_hessian(lm):
    pdi, img, P, O, x_true, ev, add, obj = lm
    x = (x_true + 0.2).astype(np.float32)
    v = np.random.default_rng(7).uniform(0.0, 1.0, img.shape).astype(np.float32)
    for s in (1, -1):
        print(f"[lm] Hessian product, subset {s}", flush=True)
        start = np.full(img.shape, 2.0, np.float32)
        if s >= 0:
            got = obj.accumulate_sub_Hessian_times_input_without_penalty(start.copy(), x, v, s)
            ref = O.lm_hessian_times_input(ev, x, v, s, NSUB, additive=add, out=start.copy())
        else:
            got = obj.accumulate_Hessian_times_input_without_penalty(start.copy(), x, v)
            ref = O.lm_hessian_times_input(ev, x, v, 0, 1, additive=add, out=start.copy())
        err = np.abs(got - ref).max() / np.abs(ref - start).max()
        print(f"LM Hessian product subset {s}: max err {err:.2e} of max")
        assert err <= 1e-5
    with pytest.raises(_native.NativeError, match="approximate"):
        obj.add_multiplication_with_approximate_sub_Hessian_without_penalty(np.zeros(img.shape, np.float32), v, 0)


def test_lm_capped_events(lm):
    """Every event capped (x = 1e-9, no additive term: 1 > 1e4 fwd): the
    gradient still back-projects y/fwd (the reference's gradient
    instantiation caps nothing), the value takes log(1e-4) per event. Runs
    last on the module's context: it replaces the events."""
    pdi, img, P, O, x_true, ev, add, obj = lm
    sub = np.ascontiguousarray(ev[:50000])
    obj2 = PoissonLogLikelihoodWithLinearModelForMeanAndListModeDataWithProjMatrixByBin(P)
    obj2.set_up(sub, 0.0)
    x = np.full(img.shape, 1e-9, np.float32)
    g = obj2.compute_sub_gradient_without_penalty(x, subset=2, plus_sensitivity=True)
    vr, r = O.lm_gradient(sub, x, 2, NSUB)
    err = np.abs(g - r).max() / np.abs(r).max()
    print(f"capped LM gradient: max err {err:.2e} of max {np.abs(r).max():.3e}")
    assert np.abs(r).max() > 0 and err <= 1e-5
    v = obj2.compute_objective_function(x, subset=2)
    assert abs(v - vr) <= 1e-9 * abs(vr)


def test_lm_tof_gradient_and_hessian():
    pdi, img = config0_tof(3)
    pm = RayTracingMatrixParameters()
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=NSUB)
    O = OraclePlan(pdi, img, pm)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    y = rng.poisson(np.maximum(O.forward_project(x), 0) * 0.02).astype(np.float32)
    ev = pdi.events_from_counts(y, seed=2)
    add = rng.uniform(0.005, 0.02, pdi.num_bins).astype(np.float32)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndListModeDataWithProjMatrixByBin(P)
    obj.set_up(ev, add, use_tof_sensitivities=True)
    g = obj.compute_sub_gradient_without_penalty(x, subset=1, plus_sensitivity=True)
    _, r = O.lm_gradient(ev, x, 1, NSUB, additive=add)
    v = rng.uniform(0.0, 1.0, img.shape).astype(np.float32)
    h = obj.accumulate_sub_Hessian_times_input_without_penalty(np.zeros(img.shape, np.float32), x, v, 1)
    hr = O.lm_hessian_times_input(ev, x, v, 1, NSUB, additive=add)
    P.close()
    err = np.abs(g - r).max() / np.abs(r).max()
    print(f"TOF LM gradient: max err {err:.2e} of max")
    assert err <= 1e-5
    err = np.abs(h - hr).max() / np.abs(hr).max()
    print(f"TOF LM Hessian product: max err {err:.2e} of max")
    assert err <= 1e-5
