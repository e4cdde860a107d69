"""Pins of the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible transcription mistake fails it:
the reference's own golden sensitivity image, closed forms of the Siddon
tracer, an independent FP64 interval tracer, the reference's
symmetric-versus-direct row invariant, the adjoint identity in FP64,
finite differences of the log-likelihood, EM monotonicity and the
reference's subset-balance table.
"""
import math
import os

import numpy as np
import pytest

from stir_amd.geometry import (ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, VoxelsOnCartesianGrid,
                               config0, rpt_sensitivity_case)
from oracle.oracle import OraclePlan, ray_trace

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def small_case(views=48, tang=64, rings=4, mrd=2, nxy=64, nz=7):
    sc = Scanner.cylindrical(rings, 2 * views, 150.0, 4.0, name="small")
    pdi = ProjDataInfoCylindrical.construct(sc, span=1, max_delta=mrd, num_views=views, num_tangential_poss=tang)
    img = VoxelsOnCartesianGrid(nz, nxy, nxy, 2.0, 2.0, 2.0)
    return pdi, img


# ---------------------------------------------------------------------------
# reference-held golden vector
def test_rpt_sensitivity_golden_image():
    """recon_test_pack/RPTsens_seg3_PM.v: PMRT total sensitivity (2 tangential LORs,
    zero end planes, segments -3..3); compare_image tolerance 5e-4 of max."""
    pdi, img, pm = rpt_sensitivity_case()
    P = OraclePlan(pdi, img, pm)
    sens = sum(P.subset_sensitivity(s, 12, zero_end_planes=True).astype(np.float64) for s in range(12))
    ref = np.fromfile(os.path.join(GOLDEN, "RPTsens_seg3_PM.v"), dtype="<f4").reshape(img.shape)
    err = np.abs(sens - ref).max() / np.abs(ref).max()
    assert err <= 5e-4
    # observed agreement is at the level of float summation order
    assert err <= 2e-5
    assert int((sens != 0).sum()) == int((ref != 0).sum()) == 85591


def test_golden_image_detects_planted_mistake():
    """The fixture rejects a plausible mistake: one tangential LOR instead of two."""
    pdi, img, pm = rpt_sensitivity_case()
    wrong = OraclePlan(pdi, img, RayTracingMatrixParameters(num_tangential_LORs=1))
    sens = sum(wrong.subset_sensitivity(s, 12, zero_end_planes=True).astype(np.float64) for s in range(12))
    ref = np.fromfile(os.path.join(GOLDEN, "RPTsens_seg3_PM.v"), dtype="<f4").reshape(img.shape)
    assert np.abs(sens - ref).max() / np.abs(ref).max() > 5e-4


# ---------------------------------------------------------------------------
# closed forms of the tracer (RayTraceVoxelsOnCartesianGrid.cxx)
def test_axis_aligned_ray_closed_form():
    vox = (2.0, 2.0, 2.0)
    z, y, x, w = ray_trace((2.0, -10.3, 3.0), (2.0, 7.6, 3.0), vox, 0.5)
    # start and end voxels are counted whole: every voxel gets vox_y * nc
    assert list(y) == list(range(-10, 9))
    assert set(z.tolist()) == {2} and set(x.tolist()) == {3}
    np.testing.assert_allclose(w, 2.0 * 0.5, rtol=1e-5)


def test_ray_between_voxel_planes_is_split():
    z, y, x, w = ray_trace((2.0, -4.2, 3.5), (2.0, 4.2, 3.5), (2.0, 2.0, 2.0), 1.0)
    assert set(x.tolist()) == {3, 4}
    np.testing.assert_allclose(w, 1.0, rtol=1e-5)  # half of vox_y * nc
    # the union is sorted by (z, y, x)
    keys = list(zip(z.tolist(), y.tolist(), x.tolist()))
    assert keys == sorted(keys)


def _exact_intervals(start, stop, vox, nc):
    """FP64 exact line/voxel intersection lengths from start to stop (voxel units,
    voxels centred on integers): list of ((z,y,x), length)."""
    s = np.asarray(start, np.float64); e = np.asarray(stop, np.float64); v = np.asarray(vox, np.float64)
    d = e - s
    ts = [0.0, 1.0]
    for ax in range(3):
        if abs(d[ax]) < 1e-12:
            continue
        lo, hi = sorted((s[ax], e[ax]))
        for k in range(math.floor(lo - 0.5), math.ceil(hi + 0.5) + 1):
            t = (k + 0.5 - s[ax]) / d[ax]
            if 0.0 < t < 1.0:
                ts.append(t)
    ts = np.unique(np.asarray(ts))
    length = float(np.sqrt(np.sum((d * v) ** 2)))
    out = []
    for a, b in zip(ts[:-1], ts[1:]):
        mid = s + 0.5 * (a + b) * d
        out.append((tuple(int(math.floor(m + 0.5)) for m in mid), (b - a) * length * nc))
    return out


def test_siddon_matches_fp64_brute_force():
    rng = np.random.default_rng(11)
    vox = (2.03125, 2.08626, 2.08626)
    checked = 0
    for _ in range(300):
        start = rng.uniform(-20, 20, 3)
        stop = rng.uniform(-20, 20, 3)
        if np.any(np.abs(stop - start) < 0.5):
            continue
        if stop[0] < start[0]:
            start, stop = stop, start
        nc = 1.0 / vox[2]
        z, y, x, w = ray_trace(start, stop, vox, nc)
        ref = _exact_intervals(np.float32(start), np.float32(stop), vox, nc)
        # merge consecutive segments in the same voxel (corner crossings)
        merged = []
        for c, l in ref:
            if merged and merged[-1][0] == c:
                merged[-1] = (c, merged[-1][1] + l)
            else:
                merged.append((c, l))
        sid = [((int(a), int(b), int(c)), float(d)) for a, b, c, d in zip(z, y, x, w) if d > 1e-6]
        # interior voxels (the tracer counts start and end voxels whole)
        assert [c for c, _ in sid[1:-1]] == [c for c, _ in merged[1:-1]]
        for (c1, l1), (c2, l2) in zip(sid[1:-1], merged[1:-1]):
            assert abs(l1 - l2) <= 1e-4 * max(1.0, l2)
        checked += 1
    assert checked > 200


# ---------------------------------------------------------------------------
# the reference's symmetry invariant (test_DataSymmetriesForBins_PET_CartesianGrid.cxx:86-197)
def _rows_equal_tolerant(r1, r2):
    """ProjMatrixElemsForOneBin::operator== on sorted rows (ProjMatrixElemsForOneBin.cxx:414-475)."""
    def srt(r):
        z, y, x, w = r
        o = np.lexsort((x, y, z))
        return list(zip(z[o].tolist(), y[o].tolist(), x[o].tolist())), w[o]
    c1, w1 = srt(r1)
    c2, w2 = srt(r2)
    tol = (w1.max() if len(w1) else 0.0) * 0.002
    i = j = 0
    while i < len(c1) and j < len(c2):
        if c1[i] == c2[j]:
            if abs(w1[i] - w2[j]) > tol:
                return False
        else:
            if w1[i] < tol:
                i += 1
                continue
            if w2[j] < tol:
                j += 1
                continue
            return False
        i += 1
        j += 1
    return all(w1[i:] <= tol) and all(w2[j:] <= tol)


@pytest.mark.parametrize("case", ["config0", "rpt"])
def test_symmetric_rows_equal_direct_rows(case):
    if case == "config0":
        pdi, img = config0()
        pm = RayTracingMatrixParameters()
    else:
        pdi, img, pm = rpt_sensitivity_case()
    P = OraclePlan(pdi, img, pm)
    rng = np.random.default_rng(3)
    bad = 0
    n = 0
    for _ in range(400):
        seg = int(rng.integers(pdi.min_segment_num, pdi.max_segment_num + 1))
        b = (seg, int(rng.integers(0, pdi.num_views)), int(rng.integers(0, pdi.num_axial_poss_of(seg))),
             int(rng.integers(pdi.min_tangential_pos_num // 2, pdi.max_tangential_pos_num // 2 + 1)))
        sym = P.row(*b)
        direct = P.trace_row(*b)
        n += len(direct[0]) > 0
        bad += not _rows_equal_tolerant(sym, direct)
    assert n > 100
    assert bad == 0


def test_symmetric_rows_detect_wrong_operation():
    """A plausible mistake (swapped x/y mirror) is caught by the invariant."""
    pdi, img = config0()
    P = OraclePlan(pdi, img)
    z, y, x, w = P.row(1, 60, 2, 10)
    direct = P.trace_row(1, 60, 2, 10)
    assert _rows_equal_tolerant((z, y, x, w), direct)
    assert not _rows_equal_tolerant((z, x, y, w), direct)


# ---------------------------------------------------------------------------
# algebraic properties
def test_adjoint_identity_fp64():
    pdi, img = small_case()
    P = OraclePlan(pdi, img)
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, img.shape)
    y = rng.uniform(0, 1, pdi.num_bins)
    lhs = float(np.dot(P.forward_project_f64(x), y))
    rhs = float(np.dot(x.reshape(-1), P.back_project_f64(y).reshape(-1)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_forward_projection_float_vs_fp64():
    pdi, img = small_case()
    P = OraclePlan(pdi, img)
    x = np.random.default_rng(6).uniform(0, 1, img.shape).astype(np.float32)
    f32 = P.forward_project(x).astype(np.float64)
    f64 = P.forward_project_f64(x.astype(np.float64))
    assert np.max(np.abs(f32 - f64)) <= 1e-5 * np.abs(f64).max()


def _loglik_fp64(P, y, x, b):
    ybar = P.forward_project_f64(x) + b
    return float(np.sum(y * np.log(ybar) - ybar))


def test_loglikelihood_and_gradient_vs_fp64_finite_differences():
    pdi, img = small_case()
    P = OraclePlan(pdi, img)
    rng = np.random.default_rng(7)
    x = rng.uniform(0.5, 1.5, img.shape)
    b = 0.1
    ybar = P.forward_project_f64(x) + b
    y = rng.poisson(ybar).astype(np.float32)
    y = np.maximum(y, 1.0).astype(np.float32)  # keep every bin away from the zero-count branch
    bfull = np.full(pdi.num_bins, b, np.float32)
    L = P.log_likelihood(y, x.astype(np.float32), bkg=bfull)
    L64 = _loglik_fp64(P, y.astype(np.float64), x, b)
    assert abs(L - L64) <= 1e-6 * abs(L64)
    g = P.subset_gradient(y, x.astype(np.float32), 0, 1, bkg=bfull)
    # FP64 central differences on 64 voxels that the data sees
    sens = P.back_project_f64(np.ones(pdi.num_bins))
    cand = np.flatnonzero(sens.reshape(-1) > 0.05 * sens.max())
    idx = rng.choice(cand, 64, replace=False)
    eps = 1e-4
    gmax = np.abs(g).max()
    for j in idx:
        xp = x.copy().reshape(-1); xp[j] += eps
        xm = x.copy().reshape(-1); xm[j] -= eps
        fd = (_loglik_fp64(P, y, xp.reshape(img.shape), b) - _loglik_fp64(P, y, xm.reshape(img.shape), b)) / (2 * eps)
        assert abs(fd - g.reshape(-1)[j]) <= 1e-3 * gmax


def test_em_increases_likelihood():
    pdi, img = small_case()
    P = OraclePlan(pdi, img)
    rng = np.random.default_rng(8)
    truth = rng.uniform(0, 2, img.shape).astype(np.float32)
    b = 0.1
    y = rng.poisson(P.forward_project(truth).astype(np.float64) + b).astype(np.float32)
    bfull = np.full(pdi.num_bins, b, np.float32)
    x = np.ones(img.shape, np.float32)
    prev = P.log_likelihood(y, x, bkg=bfull)
    for k in range(1, 6):
        x = P.osem(y, x, 1, 1, bkg=bfull, start_subset=0)
        cur = P.log_likelihood(y, x, bkg=bfull)
        assert cur >= prev - 1e-6 * abs(prev)
        prev = cur


def test_osem_subsets_sum_to_full_sensitivity():
    pdi, img = config0()
    P = OraclePlan(pdi, img)
    total = sum(P.subset_sensitivity(s, 4).astype(np.float64) for s in range(4))
    full = P.subset_sensitivity(0, 1)
    assert np.abs(total - full).max() <= 1e-5 * full.max()


# ---------------------------------------------------------------------------
# subsets (find_basic_vs_nums_in_subset.cxx, ...AndProjData.cxx:490-529)
@pytest.mark.parametrize("views,nsub,balanced", [(96, 4, True), (96, 8, True), (192, 8, True), (252, 21, True),
                                                 (144, 8, False)])
def test_subset_balance_table(views, nsub, balanced):
    sc = Scanner.cylindrical(2, 2 * views, 200.0, 4.0)
    pdi = ProjDataInfoCylindrical.construct(sc, span=1, max_delta=1, num_views=views, num_tangential_poss=16)
    img = VoxelsOnCartesianGrid(3, 16, 16, 2.0, 2.0, 2.0)
    assert OraclePlan(pdi, img).subsets_balanced(nsub) == balanced


def test_subsets_partition_view_segments():
    pdi, img = config0()
    P = OraclePlan(pdi, img)
    seen = []
    for s in range(4):
        seen += P.subset_view_segments(s, 4)
    assert sorted(seen) == sorted((seg, v) for seg in pdi.segments for v in range(pdi.num_views))
