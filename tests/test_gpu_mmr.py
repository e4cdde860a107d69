"""The projector pair and OSEM at mMR size (64 rings x 504 crystals, span 11,
MRD 60, 837 x 252 x 344 bins; 344 x 344 x 127 image, 21 subsets), GPU.

The production shape itself is checked against the oracle:
  * matrix rows at span 11 bit for bit (voxel index sets and weights):
    segment 0 (one LOR per axial position, add_adjacent_z's overlap weights,
    the offset_in_z rule for direct rays), the oblique segments (z + 1
    copies merged), every symmetry operation, first / middle / last axial
    positions;
  * the forward projection of a subset bit for bit (np.array_equal);
  * the back projection of two subsets within 1e-5 of the maximum and
    1e-5 relative where the value exceeds 1e-3 of the maximum;
  * over all subsets, the adjoint identity <A x, y> = <x, A^T y> in FP64;
  * OSMAPOSL, 21 subsets x 3 iterations from a uniform image on the
    torso phantom's Poisson data (mean 2/bin + 0.05), against the oracle's
    OSEM: within 1e-4 relative where the image exceeds 1e-3 of its maximum;
  * the differentiable objective at this size: a batch of two log-normal
    (0, 0.5) images, negative log-likelihood of one subset plus
    0.01 * QuadraticPrior through torch autograd; the likelihood value within
    1e-6 relative and the sub-gradient within 1e-5 of the maximum of the
    oracle's, the prior's gradient bit-identical to the oracle's and its
    value within 1e-10 (FP64 sums of 15 M terms in another order), and the
    batch's gradient their sum.
"""
import time

import numpy as np
import pytest
import torch

from oracle.oracle import OraclePlan, quadratic_prior
from stir_amd import _native, phantoms
from stir_amd.geometry import RayTracingMatrixParameters, config1
from stir_amd.projectors import (OSMAPOSLReconstruction, PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                                 ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing)

pytestmark = pytest.mark.gpu
NSUB = 21


@pytest.fixture(scope="module")
def mmr():
    pdi, img = config1()
    pm = RayTracingMatrixParameters()
    print(f"[mmr] set-up: {pdi.num_bins} bins, image {img.shape}", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=NSUB)
    t0 = time.time()
    O = OraclePlan(pdi, img, pm)
    print(f"[mmr] oracle plan {time.time() - t0:.1f} s", flush=True)
    yield pdi, img, pm, P, O
    P.close()


@pytest.fixture(scope="module")
def mmr_data(mmr):
    pdi, img, pm, P, O = mmr
    x_true, y, bkg = phantoms.config1_data(pdi, img, P.forward_project)
    return y, bkg


def test_rows_bit_exact_span11(mmr):
    pdi, img, pm, P, O = mmr
    rng = np.random.default_rng(11)
    bins = []
    for seg in range(pdi.min_segment_num, pdi.max_segment_num + 1):
        nax = pdi.num_axial_poss_of(seg)
        for view in (0, 1, 62, 63, 64, 125, 126, 127, 188, 189, 251):
            for ax in (0, nax // 2, nax - 1):
                for tang in (-171, -100, -37, -1, 0, 1, 55, 120, 170, 172):
                    if pdi.min_tangential_pos_num <= tang <= pdi.max_tangential_pos_num:
                        bins.append((seg, view, ax, tang))
    for _ in range(300):
        seg = int(rng.integers(pdi.min_segment_num, pdi.max_segment_num + 1))
        bins.append((seg, int(rng.integers(0, pdi.num_views)), int(rng.integers(0, pdi.num_axial_poss_of(seg))),
                     int(rng.integers(pdi.min_tangential_pos_num, pdi.max_tangential_pos_num + 1))))
    mism, nonempty, nelem = 0, 0, 0
    for b in bins:
        zo, yo, xo, wo = O.row(*b)
        zg, yg, xg, wg = P.get_row(*b)
        nonempty += len(zo) > 0
        nelem += len(zo)
        same = (np.array_equal(zo, zg) and np.array_equal(yo, yg) and np.array_equal(xo, xg)
                and np.array_equal(wo.view(np.uint32), wg.view(np.uint32)))
        if not same:
            mism += 1
            if mism < 5:
                print("row mismatch", b, len(zo), len(zg))
    print(f"{len(bins)} rows ({nonempty} non-empty, {nelem} elements) compared, {mism} differ")
    assert mism == 0
    assert nonempty > len(bins) // 2


def test_nonzero_count(mmr):
    """nnz at mMR size (3.0e10), counted by the device at set-up, equals the
    oracle's bin-by-bin count of matrix elements inside the image."""
    pdi, img, pm, P, O = mmr
    t0 = time.time()
    ref = O.nnz()
    got = P.info().num_nonzeros
    print(f"[mmr] nnz {got} (oracle {ref}, {time.time() - t0:.1f} s)", flush=True)
    assert got == ref


def test_subset_projections_match_oracle(mmr):
    pdi, img, pm, P, O = mmr
    rng = np.random.default_rng(21)
    x = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    y = rng.uniform(0.0, 2.0, pdi.num_bins).astype(np.float32)
    for s in (7, 20):
        print(f"[mmr] oracle forward projection of subset {s}", flush=True)
        ref = O.forward_project(x, subset=s, num_subsets=NSUB, out=np.zeros(pdi.num_bins, np.float32))
        got = P.forward_project(x, subset=s)
        print(f"forward subset {s}: bit-exact bins {np.mean(got.view(np.uint32) == ref.view(np.uint32)):.6f}",
              flush=True)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        print(f"[mmr] oracle back projection of subset {s}", flush=True)
        bref = O.back_project(y, subset=s, num_subsets=NSUB, omp=True)
        bgot = P.back_project(y, subset=s)
        err = float(np.abs(bgot - bref).max() / np.abs(bref).max())
        m = bref > 1e-3 * bref.max()
        rel = float((np.abs(bgot - bref)[m] / bref[m]).max())
        print(f"back subset {s}: max err {err:.2e} of max, {rel:.2e} relative", flush=True)
        assert err <= 1e-5 and rel <= 1e-5


def test_block_back_projection_bit_identical(mmr):
    """The back projection over 2 x 2 column blocks (MI_TUNE_BP_QUAD) and the
    one in one launch per subset (MI_TUNE_BP_SPLIT_SUBSETS) equal the default
    (column kernel, one launch for all subsets) bit for bit."""
    pdi, img, pm, P, O = mmr
    assert P.info().bp_quad == 0
    y = np.random.default_rng(23).uniform(0.0, 2.0, pdi.num_bins).astype(np.float32)
    ref = {s: P.back_project(y, subset=s) for s in (7, -1)}
    ctx = P.context
    try:
        ctx.set_tuning(_native.MI_TUNE_BP_SPLIT_SUBSETS, 1)
        assert np.array_equal(P.back_project(y).view(np.uint32), ref[-1].view(np.uint32))
    finally:
        ctx.set_tuning(_native.MI_TUNE_BP_SPLIT_SUBSETS, 0)
    Q = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    Q.context.set_tuning(_native.MI_TUNE_BP_QUAD, 1)
    Q.set_up(pdi, img, num_subsets=NSUB)
    try:
        info = Q.info()
        assert info.bp_quad == 1
        print(f"mMR: {info.num_table_entries} column entries, {info.num_quad_entries} block entries "
              f"({info.num_quad_entries / info.num_table_entries:.3f})")
        for s in (7, -1):
            same = np.mean(Q.back_project(y, subset=s).view(np.uint32) == ref[s].view(np.uint32))
            print(f"subset {s}: bit-identical voxels {same:.6f}", flush=True)
            assert same == 1.0
    finally:
        Q.close()


def test_adjoint_identity(mmr):
    pdi, img, pm, P, O = mmr
    rng = np.random.default_rng(22)
    x = rng.uniform(0.0, 1.0, img.shape).astype(np.float32)
    y = rng.uniform(0.0, 1.0, pdi.num_bins).astype(np.float32)
    ax = P.forward_project(x)
    aty = P.back_project(y)
    lhs = float(np.dot(ax.astype(np.float64), y.astype(np.float64)))
    rhs = float(np.dot(x.ravel().astype(np.float64), aty.ravel().astype(np.float64)))
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"adjoint at mMR size: <Ax,y> {lhs:.9e} <x,ATy> {rhs:.9e} rel {rel:.2e}")
    assert rel <= 1e-5


def test_osem_21_subsets_3_iterations(mmr, mmr_data):
    pdi, img, pm, P, O = mmr
    y, bkg = mmr_data
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg)
    t0 = time.time()
    got = OSMAPOSLReconstruction(obj, num_subiterations=3 * NSUB).reconstruct(np.ones(img.shape, np.float32))
    print(f"[mmr] GPU OSEM 21 x 3: {time.time() - t0:.1f} s; oracle OSEM (OpenMP) ...", flush=True)
    t0 = time.time()
    ref = O.osem(y, np.ones(img.shape, np.float32), NSUB, 3 * NSUB, bkg=np.full(pdi.num_bins, bkg, np.float32))
    print(f"[mmr] oracle OSEM {time.time() - t0:.1f} s", flush=True)
    mask = ref > 1e-3 * ref.max()
    rel = float((np.abs(got - ref)[mask] / ref[mask]).max())
    print(f"mMR OSEM 21 x 3: max rel err {rel:.3e} over {int(mask.sum())} voxels, "
          f"max abs {np.abs(got - ref).max() / ref.max():.2e} of max")
    assert rel <= 1e-4
    assert np.abs(got - ref).max() <= 1e-4 * ref.max()


def test_batch_of_two_gradient_nll_plus_quadratic_prior(mmr, mmr_data):
    from stir_amd.autograd import PoissonLogLikelihoodLoss, QuadraticPriorLoss

    pdi, img, pm, P, O = mmr
    y, bkg = mmr_data
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg)
    subset, beta = 5, 0.01
    rng = np.random.default_rng(202)
    xs = rng.lognormal(0.0, 0.5, (2,) + tuple(img.shape)).astype(np.float32)
    dev = torch.device("cuda:0")
    xt = torch.from_numpy(xs).to(dev).requires_grad_(True)
    nll = PoissonLogLikelihoodLoss(obj, subset)
    prior = QuadraticPriorLoss(P, beta)
    vn = nll(xt)
    vp = prior(xt)
    (vn + vp).sum().backward()
    g = xt.grad.cpu().numpy()
    bfull = np.full(pdi.num_bins, bkg, np.float32)
    vox = (img.vox_z, img.vox_y, img.vox_x)
    for b in range(2):
        print(f"[mmr] oracle gradient of subset {subset}, image {b}", flush=True)
        g_o = O.subset_gradient(y, xs[b], subset, NSUB, bkg=bfull)
        v_o = O.log_likelihood(y, xs[b], bkg=bfull, subset=subset, num_subsets=NSUB)
        vp_o, gp_o = quadratic_prior(xs[b], vox, beta)
        rel_v = abs(-vn[b].item() - v_o) / abs(v_o)
        # the loss gradient is minus the likelihood's plus the prior's
        g_nll = g[b] - gp_o
        err = float(np.abs(-g_nll - g_o).max() / np.abs(g_o).max())
        print(f"image {b}: NLL value rel {rel_v:.2e}, gradient err {err:.2e} of max, "
              f"prior value {vp[b].item():.9e} (oracle {vp_o:.9e})", flush=True)
        assert rel_v <= 1e-6
        assert err <= 1e-5
        assert abs(vp[b].item() - vp_o) <= 1e-10 * abs(vp_o)  # FP64 sums of 15 M terms in another order
        xb = torch.from_numpy(xs[b]).to(dev).requires_grad_(True)
        prior(xb).backward()
        assert np.array_equal(xb.grad.cpu().numpy(), gp_o)
