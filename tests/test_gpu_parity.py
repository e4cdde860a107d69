"""Parity of the HIP projector pair and OSEM against the CPU oracle.

Bars: matrix rows (voxel index sets and values) bit-exact; forward
projection within 1e-5 relative (bit-exact is expected: same row order, no
FMA contraction); back projection within 1e-5 of the maximum; OSEM images
within 1e-4 relative. Workload: the 8-ring x 192-crystal scanner, span 1,
MRD 3, 96 views x 128 tangential, 128x128x15 image at 2 mm, plus the
reference's RPTsens_seg3_PM golden sensitivity image.
"""
import os

import numpy as np
import pytest

from stir_amd.geometry import RayTracingMatrixParameters, config0, rpt_sensitivity_case
from stir_amd.projectors import (OSMAPOSLReconstruction, PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                                 ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing)
from stir_amd import phantoms

from oracle.oracle import OraclePlan

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _pair(pdi, img, pm, nsub):
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=nsub)
    return P


@pytest.fixture(scope="module")
def c0():
    pdi, img = config0()
    pm = RayTracingMatrixParameters()
    print(f"[c0] set-up: {pdi.num_bins} bins, image {img.shape}", flush=True)
    P = _pair(pdi, img, pm, 4)
    O = OraclePlan(pdi, img, pm)
    yield pdi, img, pm, P, O
    P.close()


def _sample_bins(pdi, rng, n):
    out = []
    for _ in range(n):
        seg = int(rng.integers(pdi.min_segment_num, pdi.max_segment_num + 1))
        out.append((seg, int(rng.integers(0, pdi.num_views)), int(rng.integers(0, pdi.num_axial_poss_of(seg))),
                    int(rng.integers(pdi.min_tangential_pos_num, pdi.max_tangential_pos_num + 1))))
    return out


def test_setup_info(c0):
    pdi, img, pm, P, O = c0
    info = P.info()
    assert info.num_bins == pdi.num_bins == O.num_bins
    assert tuple(info.symmetry_flags) == O.symmetry_flags()
    assert info.num_table_entries > 0
    # nnz (matrix elements whose voxel lies inside the image), the numerator
    # of the gather-model bandwidth, counted by the oracle bin by bin
    assert info.num_nonzeros == O.nnz()
    assert info.bp_design_bytes > 4 * (info.num_bins + info.num_voxels)
    assert info.fp_design_bytes > 4 * (info.num_bins + info.num_voxels)


def test_rows_bit_exact(c0):
    pdi, img, pm, P, O = c0
    rng = np.random.default_rng(1)
    bins = _sample_bins(pdi, rng, 400)
    bins += [(s, v, a, t // 3) for (s, v, a, t) in _sample_bins(pdi, rng, 400)]
    # include all tangential positions of a few views/segments (basic and symmetric)
    for seg in (0, 1, -1, pdi.max_segment_num, pdi.min_segment_num):
        for view in (0, 1, 23, 24, 25, 47, 48, 49, 71, 72, 73, 95):
            for tang in range(pdi.min_tangential_pos_num, pdi.max_tangential_pos_num + 1, 7):
                bins.append((seg, view, 0, tang))
    mism = 0
    nonempty = 0
    for b in bins:
        zo, yo, xo, wo = O.row(*b)
        zg, yg, xg, wg = P.get_row(*b)
        nonempty += len(zo) > 0
        same = (np.array_equal(zo, zg) and np.array_equal(yo, yg) and np.array_equal(xo, xg)
                and np.array_equal(wo.view(np.uint32), wg.view(np.uint32)))
        if not same:
            mism += 1
            if mism < 5:
                print("row mismatch", b, len(zo), len(zg))
    assert mism == 0, f"{mism} of {len(bins)} rows differ"
    # bins with |s| beyond the FOV radius have empty rows; most sampled ones do not
    assert nonempty > len(bins) // 4


def test_forward_projection(c0):
    pdi, img, pm, P, O = c0
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1, img.shape).astype(np.float32)
    ref = O.forward_project(x)
    got = P.forward_project(x)
    d = np.abs(got - ref)
    scale = np.maximum(np.abs(ref), 1e-30)
    assert np.max(d / scale) <= 1e-5
    exact = np.mean(got.view(np.uint32) == ref.view(np.uint32))
    print(f"forward projection bit-exact fraction {exact:.6f}")
    assert exact == 1.0


def test_back_projection(c0):
    pdi, img, pm, P, O = c0
    rng = np.random.default_rng(3)
    y = rng.uniform(0, 1, pdi.num_bins).astype(np.float32)
    ref = O.back_project(y)
    got = P.back_project(y)
    ref64 = O.back_project_f64(y.astype(np.float64))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    e32 = np.abs(got - ref64).max() / np.abs(ref64).max()
    print(f"back projection max err vs oracle {err:.3e}, vs FP64 {e32:.3e}")
    assert err <= 1e-5
    mask = ref64 > 1e-3 * ref64.max()
    assert np.max(np.abs(got[mask] - ref[mask]) / ref[mask]) <= 1e-5


def test_adjoint_identity(c0):
    pdi, img, pm, P, O = c0
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, img.shape).astype(np.float32)
    y = rng.uniform(0, 1, pdi.num_bins).astype(np.float32)
    lhs = np.dot(P.forward_project(x).astype(np.float64), y.astype(np.float64))
    rhs = np.dot(x.reshape(-1).astype(np.float64), P.back_project(y).reshape(-1).astype(np.float64))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_subset_projections(c0):
    pdi, img, pm, P, O = c0
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, img.shape).astype(np.float32)
    y = rng.uniform(0, 1, pdi.num_bins).astype(np.float32)
    total = np.zeros(img.shape, np.float64)
    for s in range(4):
        ref = O.forward_project(x, subset=s, num_subsets=4, out=np.zeros(pdi.num_bins, np.float32))
        got = P.forward_project(x, subset=s)
        assert np.array_equal(got, ref)
        bref = O.back_project(y, subset=s, num_subsets=4)
        bgot = P.back_project(y, subset=s)
        assert np.abs(bgot - bref).max() <= 1e-5 * np.abs(bref).max()
        total += bgot
    full = P.back_project(y)
    assert np.abs(total - full).max() <= 1e-5 * np.abs(full).max()


def test_rpt_sensitivity_golden():
    """The reference's own golden image: OSMAPOSL_test_PM_for_sensitivity.par
    (RPT, 2 tangential LORs, zero end planes, segments -3..3, 12 subsets,
    total sensitivity), compared with compare_image's tolerance 5e-4 of max."""
    pdi, img, pm = rpt_sensitivity_case()
    P = _pair(pdi, img, pm, 12)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(np.ones(pdi.num_bins, np.float32), 0.0, zero_end_planes=True)
    sens = sum(obj.get_subset_sensitivity(s).astype(np.float64) for s in range(12))
    ref = np.fromfile(os.path.join(GOLDEN, "RPTsens_seg3_PM.v"), dtype="<f4").reshape(img.shape)
    err = np.abs(sens - ref).max() / np.abs(ref).max()
    print(f"RPT sensitivity error {err:.3e} of max")
    assert err <= 5e-4
    assert int((sens != 0).sum()) == int((ref != 0).sum())
    P.close()


def test_osem_matches_oracle(c0):
    pdi, img, pm, P, O = c0
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg)
    recon = OSMAPOSLReconstruction(obj, num_subiterations=8)
    got = recon.reconstruct(np.ones(img.shape, np.float32))
    ref = O.osem(y, np.ones(img.shape, np.float32), 4, 8, bkg=np.full(pdi.num_bins, bkg, np.float32))
    mask = ref > 1e-3 * ref.max()
    rel = np.abs(got - ref)[mask] / ref[mask]
    print(f"OSEM max rel err {rel.max():.3e}")
    assert rel.max() <= 1e-4
    assert np.abs(got - ref).max() <= 1e-4 * ref.max()


def test_osem_in_random_subset_order(c0):
    """"uniformly randomise subset order": a new permutation every cycle of 4
    subiterations; the oracle applies the same subsets one subiteration at a
    time (default relative-change bounds: no thresholding), within 1e-4."""
    pdi, img, pm, P, O = c0
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg)
    order = [OSMAPOSLReconstruction(obj, 8, uniformly_randomise_subset_order=True, seed=5)]
    order = [order[0].get_subset_num(k) for k in range(1, 9)]
    print("subset order", order)
    assert sorted(order[:4]) == [0, 1, 2, 3] and sorted(order[4:]) == [0, 1, 2, 3]
    assert order != [0, 1, 2, 3, 0, 1, 2, 3]
    recon = OSMAPOSLReconstruction(obj, num_subiterations=8, uniformly_randomise_subset_order=True, seed=5)
    got = recon.reconstruct(np.ones(img.shape, np.float32))
    ref = np.ones(img.shape, np.float32)
    b = np.full(pdi.num_bins, bkg, np.float32)
    for s in order:
        ref = O.osem(y, ref, 4, 1, bkg=b, start_subset=s)
    mask = ref > 1e-3 * ref.max()
    rel = np.abs(got - ref)[mask] / ref[mask]
    print(f"random-order OSEM max rel err {rel.max():.3e}")
    assert rel.max() <= 1e-4


def test_gradient_and_loglik(c0):
    pdi, img, pm, P, O = c0
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg)
    rng = np.random.default_rng(6)
    x = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    bfull = np.full(pdi.num_bins, bkg, np.float32)
    for s in (0, 3):
        g = obj.compute_sub_gradient_without_penalty(x, subset=s)
        gr = O.subset_gradient(y, x, s, 4, bkg=bfull)
        assert np.abs(g - gr).max() <= 1e-5 * np.abs(gr).max()
    L = obj.compute_objective_function(x)
    Lr = O.log_likelihood(y, x, bkg=bfull)
    assert abs(L - Lr) <= 1e-6 * abs(Lr)


def test_truncation_branches_fire(c0):
    """Data that make the HIP epilogues take every branch the oracle pins in
    test_objective_oracle_pins.py: non-integer and negative counts; an
    estimate that is zero outside r = 60 mm with an additive term of 1e-7
    where A x = 0, so y > 1e4 ybar (quotient capped at 1e4 on non-empty rows)
    and the log-likelihood floor max(ybar, y / 1e4) acts; one count of 1e8 in
    some viewgrams (their small value becomes 100, so most of their bins are
    thresholded); OSEM with relative-change bounds 0.8 / 1.25. HIP against
    the oracle: gradients within 1e-5 of the maximum, the value within 1e-6
    relative, OSEM images within 1e-4 relative."""
    pdi, img, pm, P, O = c0
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project, seed=11)
    rng = np.random.default_rng(12)
    y = (y + rng.uniform(0, 1, y.shape)).astype(np.float32)
    zz, yy, xx = np.meshgrid(np.arange(img.nz), (np.arange(img.ny) + img.min_y) * img.vox_y,
                             (np.arange(img.nx) + img.min_x) * img.vox_x, indexing="ij")
    x = np.where(yy ** 2 + xx ** 2 < 60.0 ** 2, x_true + 0.5, 0.0).astype(np.float32)
    ax = O.forward_project(x)
    b = np.full(pdi.num_bins, bkg, np.float32)
    b[ax == 0] = 1e-7
    nvg = 0
    for view in range(0, pdi.num_views, 11):
        y[pdi.bin_index(0, view, 3, 5)] = 1e8
        nvg += 1
    y[rng.choice(pdi.num_bins, 200, replace=False)] = -2.0
    capped = (y > 1e4 * (ax + b)) & (ax == 0)
    print(f"bins with A x = 0: {(ax == 0).sum()}, capped: {capped.sum()}, viewgrams with a 1e8 count: {nvg}")
    assert capped.sum() > 1000
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, b)
    for s, with_sens in ((1, True), (2, False)):
        g = obj.compute_sub_gradient_without_penalty(x, subset=s, plus_sensitivity=with_sens)
        gr = O.subset_gradient(y, x, s, 4, bkg=b, add_sensitivity=with_sens)
        err = np.abs(g - gr).max() / np.abs(gr).max()
        print(f"gradient subset {s} (plus sensitivity {with_sens}): max err {err:.2e} of max")
        assert err <= 1e-5
    L = obj.compute_objective_function(x)
    Lr = O.log_likelihood(y, x, bkg=b)
    print(f"log-likelihood {L:.10e} oracle {Lr:.10e}")
    assert abs(L - Lr) <= 1e-6 * abs(Lr)
    recon = OSMAPOSLReconstruction(obj, num_subiterations=8, minimum_relative_change=0.8,
                                   maximum_relative_change=1.25)
    got = recon.reconstruct(np.ones(img.shape, np.float32))
    O.set_osem_options(minimum_relative_change=0.8, maximum_relative_change=1.25)
    try:
        ref = O.osem(y, np.ones(img.shape, np.float32), 4, 8, bkg=b)
    finally:
        O.set_osem_options()
    mask = ref > 1e-3 * ref.max()
    rel = (np.abs(got - ref)[mask] / ref[mask]).max()
    print(f"bounded OSEM max rel err {rel:.3e}")
    assert rel <= 1e-4
