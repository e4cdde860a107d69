"""Oracle pins for "use actual detector boundaries" of the ray-tracing matrix
(ProjMatrixByBinUsingRayTracing.cxx:300-387,608-634,760-762), CPU.

With the option each bin's LOR is the line between its two detectors
(ProjDataInfoCylindricalNoArcCorr.cxx:164-200): det1 + det2 = 2 view + N/2
for an even tangential position and one less for an odd one, det1 - det2 =
tang - N/2 (modulo N). So the LOR of an even position is the nominal one
(phi = view * pi / views + tilt, s = R sin(tang pi / N)), and the LOR of an
odd position is the nominal one turned by -pi/N (half a view). Where the
pair wraps round detector 0, phi comes out pi further and s changes sign;
with the segment's own tan(theta) that line is the z-mirror of the nominal
LOR, i.e. the nominal LOR of segment -seg (the reference warns, :623-626, and
keeps it; so does the oracle). The pins
compare rows (identical voxel sets, weights within 1e-4 of the row's
maximum: the two paths round phi differently) with two plans that do
not use the option: the same geometry, and the same geometry with the
intrinsic tilt lowered by pi/N; all symmetries are off, so every row is
traced directly. The option is ignored, as in the reference, for
axially compressed (span > 1) and for arc-corrected data. A planted mistake
in the detector pair (det2 without its rounding of negative halves) fails.
Parity unpinned: the doubled spacing of several tangential LORs (:760-762)
has no independent check here (the GPU rows are compared with the oracle
for it).
"""
import math
import os
import shutil
import subprocess
from dataclasses import replace

import numpy as np
import pytest

from oracle import oracle as oracle_mod
from oracle.oracle import OraclePlan
from stir_amd.geometry import ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, VoxelsOnCartesianGrid

NOSYM = dict(do_symmetry_90degrees_min_phi=False, do_symmetry_180degrees_min_phi=False,
             do_symmetry_swap_segment=False, do_symmetry_swap_s=False, do_symmetry_shift_z=False)
SC = Scanner.cylindrical(4, 64, 80.0, 4.0)
IMG = VoxelsOnCartesianGrid(7, 40, 40, 2.0, 2.0, 2.0)


def _plan(scanner, span, actual, arc_corrected=False, library=None):
    pdi = ProjDataInfoCylindrical.construct(scanner, span=span, max_delta=3, num_views=32, num_tangential_poss=40,
                                            arc_corrected=arc_corrected)
    pm = RayTracingMatrixParameters(use_actual_detector_boundaries=actual, **NOSYM)
    kw = {} if library is None else {"library": library}
    return pdi, OraclePlan(pdi, IMG, pm, **kw)


def _row_dict(O, b):
    z, y, x, w = O.row(*b)
    return {(int(a), int(c), int(d)): float(v) for a, c, d, v in zip(z, y, x, w)}


def _same(ra, rb):
    if set(ra) != set(rb):
        return False
    if not ra:
        return True
    m = max(abs(v) for v in rb.values())
    # phi is rounded to float once from double here and formed in float in the
    # nominal plan: one float step of phi moves crossing lengths by ~1e-5 mm
    return max(abs(ra[k] - rb[k]) for k in ra) <= 1e-4 * m


def _bins(pdi, n=160, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        seg = int(rng.integers(pdi.min_segment_num, pdi.max_segment_num + 1))
        out.append((seg, int(rng.integers(0, pdi.num_views)), int(rng.integers(0, pdi.num_axial_poss_of(seg))),
                    int(rng.integers(-18, 18))))
    return out


def _wraps(view, tang, ndet):
    det1 = (view + (tang >> 1) + ndet) % ndet
    det2 = (view - ((tang + 1) >> 1) + ndet // 2) % ndet
    return det1 + det2 != 2 * view + ndet // 2 - (tang & 1)


def _pin_rows(A):
    pdi, OA = A
    _, OC = _plan(SC, 1, False)
    _, OB = _plan(replace(SC, intrinsic_tilt=float(np.float32(-math.pi / SC.num_detectors_per_ring))), 1, False)
    checked = {0: 0, 1: 0, "wrapped": 0}
    for b in _bins(pdi, 240):
        seg, view, ax, tang = b
        wrapped = _wraps(view, tang, SC.num_detectors_per_ring)
        ra = _row_dict(OA, b)
        ref = _row_dict(OC if tang % 2 == 0 else OB, (-seg if wrapped else seg, view, ax, tang))
        assert _same(ra, ref), b
        checked[tang % 2] += bool(ra)
        checked["wrapped"] += bool(ra) and wrapped and seg != 0
    assert checked[0] > 20 and checked[1] > 20 and checked["wrapped"] > 5, checked


def test_rows_follow_the_detector_pairs():
    _pin_rows(_plan(SC, 1, True))


def test_odd_positions_differ_from_the_nominal_rows():
    """The option changes something: an odd position's row is not the nominal one."""
    pdi, OA = _plan(SC, 1, True)
    _, OC = _plan(SC, 1, False)
    diff = [b for b in _bins(pdi) if b[3] % 2 and not _same(_row_dict(OA, b), _row_dict(OC, b))]
    assert len(diff) > 20


@pytest.mark.parametrize("span, arc", [(3, False), (1, True)])
def test_ignored_for_compressed_or_arc_corrected_data(span, arc):
    pdi, OA = _plan(SC, span, True, arc_corrected=arc)
    _, OC = _plan(SC, span, False, arc_corrected=arc)
    for b in _bins(pdi, 60, seed=3):
        za, ya, xa, wa = OA.row(*b)
        zc, yc, xc, wc = OC.row(*b)
        assert np.array_equal(za, zc) and np.array_equal(ya, yc) and np.array_equal(xa, xc)
        assert np.array_equal(wa.view(np.uint32), wc.view(np.uint32))


def test_planted_detector_pair_mistake_fails(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    src_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    src = open(os.path.join(src_dir, "pmrt_oracle.c")).read()
    old = "  const int det2 = (view - ((tang + 1) >> 1) + ndet / 2) % ndet;"
    assert src.count(old) == 1
    (tmp_path / "pmrt_oracle.c").write_text(src.replace(old, "  const int det2 = (view - tang / 2 + ndet / 2) % ndet;"))
    shutil.copy(os.path.join(src_dir, "pmrt_oracle.h"), tmp_path / "pmrt_oracle.h")
    so = tmp_path / "liborpmrt_mutant.so"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
                    "-o", str(so), str(tmp_path / "pmrt_oracle.c"), "-lm"], check=True)
    with pytest.raises(AssertionError):
        _pin_rows(_plan(SC, 1, True, library=oracle_mod.load_library(str(so))))
