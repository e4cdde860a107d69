"""Closed-form pins of the oracle's objective-function branches, CPU.

The estimate is the zero image, so ybar = A x + b equals the additive term b
exactly and every branch can be driven with hand-picked values:

* divide_and_truncate (recon_array_functions.cxx:172-266): per-viewgram
  threshold small = max(max(y) * 1e-6, 0) with y <= small -> 0 (including
  y < 0 and y == small), the quotient cap y > 1e4 ybar -> exactly 1e4
  (including ybar = 0 and ybar < 0), y / ybar otherwise; "zero end planes
  of segment 0" (distributable.cxx:218-223); the gradient form y/ybar - 1
  (RPC, ...AndProjData.cxx:1450-1465);
* the log-likelihood floor max(ybar, y / 1e4) and the y <= small branch
  (recon_array_functions.cxx:336-377);
* OSMAPOSL: divide(u, sens, 0) with 0/0 = 0 (numerics/divide.inl:26-47) and
  the relative-change bounds, applied after subiteration 1 only
  (OSMAPOSLReconstruction.cxx:491-507).

Each pin is then run against oracle builds with one plausible mistake
planted in the C source (cap 1e3, a sinogram-wide threshold, `<` for
`<=`, no floor, 0/0 left to the division, clamping in subiteration 1,
swapped bounds): every planted mistake must fail its pin.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as oracle_mod
from oracle.oracle import OraclePlan
from stir_amd import phantoms
from stir_amd.geometry import (ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, VoxelsOnCartesianGrid)

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
F32 = np.float32


def _geometry():
    sc = Scanner.cylindrical(4, 64, 80.0, 4.0)
    pdi = ProjDataInfoCylindrical.construct(sc, span=1, max_delta=3, num_views=32, num_tangential_poss=32)
    img = VoxelsOnCartesianGrid(7, 32, 32, 2.0, 2.0, 2.0)
    return pdi, img


@pytest.fixture(scope="module")
def geom():
    return _geometry()


# ---------------------------------------------------------------------------
# the pins (each raises AssertionError when the oracle breaks its branch)
def _ratio_case(pdi):
    """y, b and the hand-computed quotient at every bin that is not 0."""
    nb = pdi.num_bins
    y = np.zeros(nb, F32)
    b = np.ones(nb, F32)
    want = {}

    def put(seg, view, ax, tang, yv, bv, q):
        i = pdi.bin_index(seg, view, ax, tang)
        y[i], b[i] = yv, bv
        if q != 0:
            want[i] = F32(q)

    # viewgram (0, 3): max y = 5, small = 5e-6
    put(0, 3, 2, 0, 5.0, 1e-5, 1e4)        # 5 > 1e4 * 1e-5: capped at 1e4 exactly
    put(0, 3, 2, 1, 5.0, 2.0, 2.5)         # 5 / 2
    put(0, 3, 4, 1, 3.0, 0.0, 1e4)         # non-zero over zero: capped
    put(0, 3, 4, 2, 3.0, -1.0, 1e4)        # negative denominator: capped
    put(0, 3, 4, 3, -3.0, 1.0, 0.0)        # negative numerator: 0
    put(0, 3, 3, 4, 4e-6, 1.0, 0.0)        # below the viewgram's small value: 0
    # viewgram (0, 7): max y = 1e7, small = float(1e7 * 1e-6f) = 10
    put(0, 7, 1, 0, 1e7, 1.0, 1e4)         # capped
    put(0, 7, 1, 1, 9.5, 1.0, 0.0)         # below small
    put(0, 7, 1, 2, 10.0, 1.0, 0.0)        # equal to small: 0 (y <= small)
    put(0, 7, 1, 3, 10.5, 1.0, 10.5)       # above small
    # viewgram (0, 8): max y = 9.5, small = 9.5e-6: the threshold of
    # viewgram (0, 7) does not reach it
    put(0, 8, 1, 1, 9.5, 1.0, 9.5)
    # viewgram (1, 8): an end plane of segment 1 is not zeroed
    put(1, 8, 0, 2, 6.0, 3.0, 2.0)
    # segment 0, first axial position: zeroed with "zero end planes"
    put(0, 5, 0, 2, 6.0, 3.0, 2.0)
    return y, b, want, pdi.bin_index(0, 5, 0, 2)


def check_divide_and_truncate(O, pdi, img):
    y, b, want, end_bin = _ratio_case(pdi)
    x0 = np.zeros(img.shape, F32)
    q = O.subset_ratio(y, x0, -1, 1, bkg=b, add_sensitivity=True)
    expect = np.zeros(pdi.num_bins, F32)
    for i, v in want.items():
        expect[i] = v
    bad = np.flatnonzero(q.view(np.uint32) != expect.view(np.uint32))
    assert bad.size == 0, f"{bad.size} bins differ, first {bad[:4]}: {q[bad[:4]]} vs {expect[bad[:4]]}"
    # the gradient form subtracts 1 (no multiplicative factors) everywhere
    g = O.subset_ratio(y, x0, -1, 1, bkg=b, add_sensitivity=False)
    assert np.array_equal(g, expect - F32(1))
    # zero end planes of segment 0: that bin's y and b are zeroed, its quotient is 0
    qz = O.subset_ratio(y, x0, -1, 1, bkg=b, add_sensitivity=True, zero_end_planes=True)
    expect_z = expect.copy()
    expect_z[end_bin] = 0
    assert np.array_equal(qz, expect_z)


def check_log_likelihood_floor(O, pdi, img):
    """b = 1e-5 everywhere, one bin y = 5: that bin's estimate is floored to
    y / 1e4 = 5e-4 (> b); every other bin (y = 0 <= small) adds -b. One bin
    y = 2e-6 in the same viewgram (below small = 5e-6) adds -max(b, 2e-10) = -b."""
    nb = pdi.num_bins
    y = np.zeros(nb, F32)
    b = np.full(nb, 1e-5, F32)
    y[pdi.bin_index(0, 3, 2, 5)] = 5.0
    y[pdi.bin_index(0, 3, 1, 5)] = 2e-6
    v = O.log_likelihood(y, np.zeros(img.shape, F32), bkg=b)
    ne = float(F32(5.0) / F32(1e4))
    bb = float(F32(1e-5))
    expect = 5.0 * math.log(ne) - ne - (nb - 1) * bb
    assert abs(v - expect) <= 1e-12 * abs(expect), (v, expect)


def check_osem_bounds_and_zero_over_zero(O, pdi, img):
    """One subset. Subiteration 1 is not clamped, subiteration 2 is clamped to
    [0.95, 1.05] exactly; voxels with zero sensitivity and zero update become
    0 (0/0 = 0), not NaN."""
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project, seed=4)
    b = np.full(pdi.num_bins, bkg, F32)
    x0 = np.random.default_rng(8).uniform(0.5, 1.5, img.shape).astype(F32)
    sens = O.subset_sensitivity(0, 1)
    O.set_osem_options()
    free1 = O.osem(y, x0, 1, 1, bkg=b, omp=False)
    O.set_osem_options(minimum_relative_change=0.95, maximum_relative_change=1.05)
    try:
        x1 = O.osem(y, x0, 1, 1, bkg=b, omp=False)
        x2 = O.osem(y, x0, 1, 2, bkg=b, omp=False)
    finally:
        O.set_osem_options()
    zero = sens == 0
    assert zero.any() and np.all(free1[zero] == 0)
    assert np.array_equal(x1, free1)
    assert np.mean(np.abs(free1[~zero] / x0[~zero] - 1) > 0.06) > 0.05  # the bounds would have acted
    u = O.subset_gradient(y, x1, 0, 1, bkg=b, add_sensitivity=True)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.where((sens == 0) & (u == 0), F32(0), u / sens).astype(F32)
    expect = (x1 * np.clip(f, F32(0.95), F32(1.05))).astype(F32)
    assert np.array_equal(x2, expect)
    ratio = x2[~zero] / x1[~zero]
    assert np.mean(ratio <= 0.9500001) > 0.01 and np.mean(ratio >= 1.0499999) > 0.01


PINS = {"ratio": check_divide_and_truncate, "loglik": check_log_likelihood_floor,
        "osem": check_osem_bounds_and_zero_over_zero}


@pytest.mark.parametrize("pin", sorted(PINS))
def test_pin_holds(geom, pin):
    pdi, img = geom
    PINS[pin](OraclePlan(pdi, img, RayTracingMatrixParameters()), pdi, img)


# ---------------------------------------------------------------------------
# planted mistakes: (pin that must catch it, source text, replacement)
MUTANTS = {
    "cap_1e3": ("ratio", "        const float max_quotient = 10000.F;\n        const float denom = ybar[i];",
                "        const float max_quotient = 1000.F;\n        const float denom = ybar[i];"),
    "threshold_strict": ("ratio", "      if (num <= small_value) num = 0;", "      if (num < small_value) num = 0;"),
    "threshold_whole_data": (
        "ratio",
        "  const float small_value = maxf_(mx * 0.000001F, 0.F);\n  for (int a = 0; a < nax; ++a)\n"
        "    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {\n      const int64_t i",
        "  for (int64_t ii = 0; ii < p->num_bins * p->ntof; ++ii) if (y[ii] > mx) mx = y[ii];\n"
        "  const float small_value = maxf_(mx * 0.000001F, 0.F);\n  for (int a = 0; a < nax; ++a)\n"
        "    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {\n      const int64_t i"),
    "no_floor": ("loglik", "          const float ne = maxf_(est, yv / 10000.F);", "          const float ne = est;"),
    "floor_1e3": ("loglik", "          const float ne = maxf_(est, yv / 10000.F);",
                  "          const float ne = maxf_(est, yv / 1000.F);"),
    "zero_over_zero": ("osem",
                       "        if (fabsf(sens[subset][i]) <= small_value && fabsf(u[i]) <= small_value) u[i] = 0;\n"
                       "        else u[i] /= sens[subset][i];",
                       "        u[i] /= sens[subset][i];"),
    "clamp_subiteration_1": ("osem", "     * (OSMAPOSLReconstruction.cxx:491-507, thresholding.h:45-54) */\n    if (k != 1)",
                             "     * (OSMAPOSLReconstruction.cxx:491-507, thresholding.h:45-54) */\n    if (k != 0)"),
    "bounds_swapped": ("osem",
                       "thresholding.h:45-54) */\n    if (k != 1)\n      for (size_t i = 0; i < nv; ++i) {\n"
                       "        if (u[i] > p->max_rel) u[i] = p->max_rel;\n"
                       "        else if (p->min_rel > u[i]) u[i] = p->min_rel;",
                       "thresholding.h:45-54) */\n    if (k != 1)\n      for (size_t i = 0; i < nv; ++i) {\n"
                       "        if (u[i] > p->min_rel) u[i] = p->min_rel;\n"
                       "        else if (p->max_rel > u[i]) u[i] = p->max_rel;"),
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    return tmp_path_factory.mktemp("oracle_mutants")


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_fails_its_pin(geom, mutant_dir, name):
    pin, old, new = MUTANTS[name]
    src = open(os.path.join(ORACLE_SRC, "pmrt_oracle.c")).read()
    assert src.count(old) == 1, f"mutation site of {name} not found exactly once"
    d = mutant_dir / name
    d.mkdir()
    (d / "pmrt_oracle.c").write_text(src.replace(old, new))
    shutil.copy(os.path.join(ORACLE_SRC, "pmrt_oracle.h"), d / "pmrt_oracle.h")
    so = d / "liborpmrt_mutant.so"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
                    "-o", str(so), str(d / "pmrt_oracle.c"), "-lm"], check=True)
    pdi, img = geom
    O = OraclePlan(pdi, img, RayTracingMatrixParameters(), library=oracle_mod.load_library(str(so)))
    with pytest.raises(AssertionError):
        PINS[pin](O, pdi, img)
