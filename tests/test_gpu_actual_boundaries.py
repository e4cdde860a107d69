"""The ray-tracing matrix with "use actual detector boundaries"
(ProjMatrixByBinUsingRayTracing.cxx:608-634,760-762) on the GPU against the
oracle.

Workload: the 8-ring x 192-crystal scanner (span 1, MRD 3, 96 views x 128
tangential, non-arc-corrected: the option applies), 128x128x15 image at 2 mm,
4 subsets, with one and with two tangential LORs per bin (the latter at twice
the nominal spacing). Bars as without the option: rows bit for bit (voxel
indices and weights), forward projection bit for bit, back projection within
1e-5 of the maximum; the rows differ from the nominal matrix's.
"""
import numpy as np
import pytest

from oracle.oracle import OraclePlan
from stir_amd.geometry import RayTracingMatrixParameters, config0
from stir_amd.projectors import ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[1, 2])
def actual(request):
    pdi, img = config0()
    pm = RayTracingMatrixParameters(use_actual_detector_boundaries=True, num_tangential_LORs=request.param)
    print(f"[actual boundaries] {request.param} tangential LOR(s): set-up", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=4)
    O = OraclePlan(pdi, img, pm)
    yield pdi, img, P, O, request.param
    P.close()


def _bins(pdi, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        seg = int(rng.integers(pdi.min_segment_num, pdi.max_segment_num + 1))
        out.append((seg, int(rng.integers(0, pdi.num_views)), int(rng.integers(0, pdi.num_axial_poss_of(seg))),
                    int(rng.integers(pdi.min_tangential_pos_num, pdi.max_tangential_pos_num + 1))))
    return out


def test_rows_bit_exact(actual):
    pdi, img, P, O, nt = actual
    nominal = OraclePlan(pdi, img, RayTracingMatrixParameters(num_tangential_LORs=nt))
    mism, nonempty, changed = 0, 0, 0
    bins = _bins(pdi, 300, 5) + [(s, v, 0, t) for s in (0, 1, -3) for v in (0, 1, 47, 95) for t in range(-60, 60, 7)]
    for b in bins:
        zo, yo, xo, wo = O.row(*b)
        zg, yg, xg, wg = P.get_row(*b)
        nonempty += len(zo) > 0
        if not (np.array_equal(zo, zg) and np.array_equal(yo, yg) and np.array_equal(xo, xg)
                and np.array_equal(wo.view(np.uint32), wg.view(np.uint32))):
            mism += 1
            if mism < 5:
                print("row mismatch", b, len(zo), len(zg))
        zn, yn, xn, wn = nominal.row(*b)
        changed += not (np.array_equal(zo, zn) and np.array_equal(wo.view(np.uint32), wn.view(np.uint32)))
    print(f"{len(bins)} rows, {nonempty} non-empty, {changed} differ from the nominal matrix")
    assert mism == 0
    assert nonempty > len(bins) // 4 and changed > len(bins) // 4


def test_projections(actual):
    pdi, img, P, O, nt = actual
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, img.shape).astype(np.float32)
    got, ref = P.forward_project(x), O.forward_project(x)
    print(f"forward projection: bit-exact fraction {np.mean(got.view(np.uint32) == ref.view(np.uint32)):.6f}")
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    y = rng.uniform(0, 1, pdi.num_bins).astype(np.float32)
    bg, br = P.back_project(y), O.back_project(y)
    err = np.abs(bg - br).max() / np.abs(br).max()
    print(f"back projection: max err {err:.2e} of max")
    assert err <= 1e-5
