"""Tables whose segments all have one image plane per axial position (span
3, no single-ring-difference segment), and a mixed one, GPU against the
oracle: the kernels' range-checked run and column reads (a plane an entry
does not reach reads 0).

Scanners: 8 rings x 192 crystals (15-plane image, one z slot per lane; a
128 x 128 and a 127 x 127 image) and
40 rings x 64 crystals (79-plane image, two z slots per lane), span 3, MRD 4;
the 8-ring scanner with MRD 5 (outer segments of one ring difference, 2
planes per axial position: a mixed table); 2 mm voxels. Bars: forward projection within 1e-5 relative (bit-exact
expected), back projection within 1e-5 of the maximum, OSEM within 1e-4.
"""
import numpy as np
import pytest

from oracle.oracle import OraclePlan
from stir_amd import _native
from stir_amd.geometry import ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, VoxelsOnCartesianGrid
from stir_amd.projectors import (OSMAPOSLReconstruction, PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                                 ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing)

pytestmark = pytest.mark.gpu

# name: scanner, views, tangential positions, image, maximum ring difference,
# planes per axial position shared by all segments (0: mixed)
CASES = {
    "8rings": (Scanner.cylindrical(8, 192, 300.0, 4.0, name="cyl8x192"), 96, 128,
               VoxelsOnCartesianGrid(15, 128, 128, 2.0, 2.0, 2.0), 4, 1),
    "40rings": (Scanner.cylindrical(40, 64, 100.0, 4.0, name="cyl40x64"), 32, 48,
                VoxelsOnCartesianGrid(79, 48, 48, 2.0, 2.0, 2.0), 4, 1),
    # MRD 5: the outer segments hold ring difference 5 only (2 planes per axial position)
    "8rings_mixed": (Scanner.cylindrical(8, 192, 300.0, 4.0, name="cyl8x192"), 96, 128,
                     VoxelsOnCartesianGrid(15, 128, 128, 2.0, 2.0, 2.0), 5, 0),
    # an odd number of columns in x and y: the last 2 x 2 blocks hold one column of each pair
    "8rings_odd": (Scanner.cylindrical(8, 192, 300.0, 4.0, name="cyl8x192"), 96, 128,
                   VoxelsOnCartesianGrid(15, 127, 127, 2.0, 2.0, 2.0), 4, 1),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    sc, nv, nt, img, mrd, ppa = CASES[request.param]
    pdi = ProjDataInfoCylindrical.construct(sc, span=3, max_delta=mrd, num_views=nv, num_tangential_poss=nt)
    pm = RayTracingMatrixParameters()
    print(f"[span3] {request.param}: {pdi.num_bins} bins, image {img.shape}", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=4)
    O = OraclePlan(pdi, img, pm)
    yield request.param, pdi, img, P, O
    P.close()


def test_one_plane_per_axial_position(case):
    name, pdi, img, P, O = case
    info = P.info()
    assert info.planes_per_axial_position == CASES[name][5]
    assert info.nz_padded > 64 if name == "40rings" else info.nz_padded <= 64


def test_forward_and_back_projection(case):
    name, pdi, img, P, O = case
    rng = np.random.default_rng(11)
    x = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    got, ref = P.forward_project(x), O.forward_project(x)
    rel = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"{name} forward: max err {rel:.2e} of max, bit-exact bins {np.mean(got == ref):.4f}")
    assert rel <= 1e-5
    y = rng.uniform(0.0, 2.0, pdi.num_bins).astype(np.float32)
    for s in (-1, 2):
        got, ref = P.back_project(y, subset=s), O.back_project(y, subset=s, num_subsets=4 if s >= 0 else 1)
        err = float(np.abs(got - ref).max() / np.abs(ref).max())
        print(f"{name} back (subset {s}): max err {err:.2e} of max")
        assert err <= 1e-5


def test_osem(case):
    name, pdi, img, P, O = case
    rng = np.random.default_rng(12)
    x_true = np.zeros(img.shape, np.float32)
    zz, yy, xx = np.meshgrid(*[np.arange(n) - (n - 1) / 2 for n in img.shape], indexing="ij")
    x_true[(xx ** 2 + yy ** 2) < (0.35 * img.shape[2]) ** 2] = 1.0
    fwd = O.forward_project(x_true)
    scale = 5.0 / max(float(fwd[fwd > 0].mean()), 1e-12)
    y = rng.poisson(fwd * scale).astype(np.float32)
    bkg = np.full(pdi.num_bins, 0.1, np.float32)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, 0.1)
    got = OSMAPOSLReconstruction(obj, num_subiterations=8).reconstruct(np.ones(img.shape, np.float32))
    ref = O.osem(y, np.ones(img.shape, np.float32), 4, 8, bkg=bkg)
    mask = ref > 1e-3 * ref.max()
    rel = float((np.abs(got - ref)[mask] / ref[mask]).max())
    print(f"{name} OSEM: max rel {rel:.2e}")
    assert rel <= 1e-4


@pytest.mark.parametrize("mrd,ppa", [(4, 1), (5, 0)])
def test_tof_span3(mrd, ppa):
    """Time of flight on a one-plane table and on a mixed one (the 8-ring
    scanner with 9 TOF bins of 89.04 mm, 400 ps FWHM, MRD 4 and 5), with the
    stored TOF rows: forward bit-exact, back projection within 1e-5 of the
    FP64 oracle's maximum, OSEM with TOF sensitivities within 1e-4."""
    from dataclasses import replace
    from stir_amd import phantoms
    sc = replace(Scanner.cylindrical(8, 192, 300.0, 4.0, name="cyl8x192_tof"), timing_resolution_ps=400.0,
                 size_of_timing_pos_ps=594.0, max_num_timing_poss=9)
    pdi = ProjDataInfoCylindrical.construct(sc, span=3, max_delta=mrd, num_views=96, num_tangential_poss=128).with_tof(1)
    img = VoxelsOnCartesianGrid(15, 128, 128, 2.0, 2.0, 2.0)
    pm = RayTracingMatrixParameters()
    print(f"[span3 tof] MRD {mrd}: {pdi.num_bins} bins", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=4)
    try:
        info = P.info()
        assert info.planes_per_axial_position == ppa
        assert (info.tof_fp_window, info.tof_bp_window) == (6, 5)
        O = OraclePlan(pdi, img, pm)
        rng = np.random.default_rng(13)
        x = rng.uniform(0, 1, img.shape).astype(np.float32)
        assert np.array_equal(P.forward_project(x), O.forward_project(x))
        y = rng.uniform(0, 1, pdi.num_bins).astype(np.float32)
        g, r = P.back_project(y), O.back_project_f64(y.astype(np.float64))
        err = float(np.abs(g - r).max() / np.abs(r).max())
        print(f"span3 TOF back: max err {err:.2e} of max")
        assert err <= 1e-5
        x_true, yd, bkg = phantoms.config0_data(pdi, img, O.forward_project)
        obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
        obj.set_up(yd, bkg, use_tof_sensitivities=True)
        got = OSMAPOSLReconstruction(obj, num_subiterations=8).reconstruct(np.ones(img.shape, np.float32))
        O.set_use_tof_sensitivities(True)
        ref = O.osem(yd, np.ones(img.shape, np.float32), 4, 8, bkg=np.full(pdi.num_bins, bkg, np.float32))
        mask = ref > 1e-3 * ref.max()
        rel = float((np.abs(got - ref)[mask] / ref[mask]).max())
        print(f"span3 TOF OSEM: max rel {rel:.2e}")
        assert rel <= 1e-4
    finally:
        P.close()


def test_block_back_projection_bit_identical(case):
    """With MI_TUNE_BP_QUAD one-plane tables back-project over 2 x 2 column
    blocks; the sums equal the column kernel's bit for bit (plain, in one
    launch and in one launch per subset, and OSEM-update epilogues)."""
    name, pdi, img, P, O = case
    assert P.info().bp_quad == 0  # off by default
    Q = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(RayTracingMatrixParameters()))
    ctx = Q.context
    ctx.set_tuning(_native.MI_TUNE_BP_QUAD, 1)
    Q.set_up(pdi, img, num_subsets=4)
    try:
        info = Q.info()
        assert info.bp_quad == (1 if CASES[name][5] == 1 else 0)
        if not info.bp_quad:
            return
        assert 0 < info.num_quad_entries < info.num_table_entries
        print(f"{name}: {info.num_table_entries} column entries, {info.num_quad_entries} block entries "
              f"({info.num_quad_entries / info.num_table_entries:.3f})")
        y = np.random.default_rng(13).uniform(0.0, 2.0, pdi.num_bins).astype(np.float32)
        for subset in (1, -1):
            for split in (1, 0):
                ctx.set_tuning(_native.MI_TUNE_BP_SPLIT_SUBSETS, split)
                blk = Q.back_project(y, subset=subset)
                ctx.set_tuning(_native.MI_TUNE_BP_QUAD, 0)
                col = Q.back_project(y, subset=subset)
                ctx.set_tuning(_native.MI_TUNE_BP_QUAD, 1)
                assert np.array_equal(blk.view(np.uint32), col.view(np.uint32)), (subset, split)
                assert np.array_equal(col.view(np.uint32), P.back_project(y, subset=subset).view(np.uint32))
        ctx.set_tuning(_native.MI_TUNE_BP_SPLIT_SUBSETS, 0)
        meas = np.random.default_rng(14).poisson(Q.forward_project(np.ones(img.shape, np.float32)) + 0.1)
        obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(Q)
        obj.set_up(meas.astype(np.float32), 0.1)
        outs = []
        for quad in (1, 0):
            ctx.set_tuning(_native.MI_TUNE_BP_QUAD, quad)
            outs.append(OSMAPOSLReconstruction(obj, num_subiterations=2).reconstruct(np.ones(img.shape, np.float32)))
        assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    finally:
        Q.close()
