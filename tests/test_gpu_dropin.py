"""The drop-in projector plugins' subset path on the GPU, driven the way
ForwardProjectorByBin::forward_project(ProjData&, subset_num, num_subsets)
and BackProjectorByBin::back_project(ProjData&, subset_num, num_subsets)
drive a plugin (ForwardProjectorByBin.cxx:168-230, BackProjectorByBin.cxx:
183-229): set_input, then the RelatedViewgrams of the subset's basic
view-segments only.

Workloads: the 8-ring x 192-crystal scanner (span 1, MRD 3, 96 x 128 bins,
128x128x15 image) with a table of 4 subsets, asked for subsets of 4 (the
table's own) and of 8 (any other division); the same geometry with 3 TOF
bins. Bars: forward projections bit-exact against the oracle's subset
projection and not one other bin written (a sentinel value survives outside
the subset); back projections within 1e-5 of the maximum of the oracle's.
"""
import numpy as np
import pytest

from oracle.oracle import OraclePlan
from stir_amd.geometry import RayTracingMatrixParameters, config0, config0_tof
from stir_amd.projectors import (BackProjectorByBinMI355X, ForwardProjectorByBinMI355X, ProjectorByBinPairMI355X,
                                 ProjMatrixByBinUsingRayTracing)

pytestmark = pytest.mark.gpu
SENTINEL = np.float32(-7.0)


def _setup(pdi, img, nsub):
    pm = RayTracingMatrixParameters()
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=nsub)
    return P, OraclePlan(pdi, img, pm)


@pytest.fixture(scope="module")
def c0():
    pdi, img = config0()
    P, O = _setup(pdi, img, 4)
    yield pdi, img, P, O
    P.close()


@pytest.mark.parametrize("num_subsets", [4, 8])
def test_forward_subset_through_the_plugin(c0, num_subsets):
    pdi, img, P, O = c0
    rng = np.random.default_rng(31)
    x = rng.uniform(0, 1, img.shape).astype(np.float32)
    F = ForwardProjectorByBinMI355X(P)
    F.set_input(x)
    for s in range(num_subsets):
        got = np.full(pdi.num_bins, SENTINEL, np.float32)
        F.forward_project(got, s, num_subsets, zero=False)
        ref = O.forward_project(x, subset=s, num_subsets=num_subsets, out=np.full(pdi.num_bins, SENTINEL, np.float32))
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), s
        written = int((ref != SENTINEL).sum())
        assert 0 < written < pdi.num_bins
    # zero=True: everything outside the subset is 0
    got = F.forward_project(np.full(pdi.num_bins, SENTINEL, np.float32), 1, num_subsets, zero=True)
    ref = O.forward_project(x, subset=1, num_subsets=num_subsets, out=np.zeros(pdi.num_bins, np.float32))
    assert np.array_equal(got, ref)


def test_forward_one_basic_view_segment_at_a_time(c0):
    """actual_forward_project's unit: one basic view-segment's RelatedViewgrams
    per call (the plugin projects lazily, one call per unit); the calls of a
    subset together equal the subset call."""
    pdi, img, P, O = c0
    x = np.random.default_rng(32).uniform(0, 1, img.shape).astype(np.float32)
    F = ForwardProjectorByBinMI355X(P)
    F.set_input(x)
    whole = F.forward_project(np.full(pdi.num_bins, SENTINEL, np.float32), 2, 4, zero=False)
    parts = np.full(pdi.num_bins, SENTINEL, np.float32)
    vs = P.basic_view_segments_in_subset(2, 4)
    assert len(vs) > 1
    for view, seg in vs:
        F.forward_project_related_viewgrams(view, seg, parts)
    assert np.array_equal(parts.view(np.uint32), whole.view(np.uint32))


@pytest.mark.parametrize("num_subsets", [4, 8])
def test_back_subset_through_the_plugin(c0, num_subsets):
    pdi, img, P, O = c0
    y = np.random.default_rng(33).uniform(0, 1, pdi.num_bins).astype(np.float32)
    B = BackProjectorByBinMI355X(P)
    for s in (0, num_subsets - 1):
        B.start_accumulating_in_new_target()
        B.back_project(y, s, num_subsets)
        got = B.get_output()
        ref = O.back_project(y, subset=s, num_subsets=num_subsets)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"plugin back projection, subset {s} of {num_subsets}: max err {err:.2e} of max")
        assert err <= 1e-5
    # two subsets staged before get_output: their sum
    B.start_accumulating_in_new_target()
    B.back_project(y, 0, num_subsets)
    B.back_project(y, 1, num_subsets)
    got = B.get_output()
    ref = (O.back_project(y, subset=0, num_subsets=num_subsets).astype(np.float64)
           + O.back_project(y, subset=1, num_subsets=num_subsets))
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_tof_subset_through_the_plugin():
    pdi, img = config0_tof(3)
    P, O = _setup(pdi, img, 4)
    rng = np.random.default_rng(34)
    x = rng.uniform(0, 1, img.shape).astype(np.float32)
    y = rng.uniform(0, 1, pdi.num_bins).astype(np.float32)
    F = ForwardProjectorByBinMI355X(P)
    F.set_input(x)
    got = F.forward_project(np.full(pdi.num_bins, SENTINEL, np.float32), 3, 4, zero=False)
    ref = O.forward_project(x, subset=3, num_subsets=4, out=np.full(pdi.num_bins, SENTINEL, np.float32))
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    B = BackProjectorByBinMI355X(P)
    B.start_accumulating_in_new_target()
    B.back_project(y, 3, 4)
    got = B.get_output()
    ref = O.back_project(y, subset=3, num_subsets=4)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    P.close()
