"""Batched reconstructions (stir_amd.batch) on the GPU.

Workload: four problems of the uneven-batch family covering every lane-group
size: 8 rings x 192 crystals with a 96x96x15 image (families of <= 8 axial
positions, 15 planes), 16 rings x 192 with 128x128x31, 24 rings x 192 span 3
with 96x96x47, and 8 rings x 288 (180-degree symmetry only) with 96x96x15;
radius 350 mm, 8 subsets, 2 iterations, cylinder-and-spheres phantom with
Poisson data (mean 5/bin + 0.1).

Bars: batched images identical to the per-problem path (both sum in the same
order) for every launch shape that keeps the table order (row groups, lane
groups, full waves, column tiles), within 1e-5 relative for back projections
split over table entries (parts summed in a fixed part order), HIP-graph
replays identical to eager launches, and one problem within 1e-4 relative of
the oracle's OSEM.
"""
import numpy as np
import pytest

from stir_amd import phantoms
from stir_amd.batch import BatchedOSMAPOSL
from stir_amd.geometry import (ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, VoxelsOnCartesianGrid)
from stir_amd.projectors import ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing
from oracle.oracle import OraclePlan

pytestmark = pytest.mark.gpu
NSUB = 8


def _problem(rings, crystals, xy):
    sc = Scanner.cylindrical(rings, crystals, 350.0, 4.0, name=f"cyl{rings}x{crystals}")
    pdi = ProjDataInfoCylindrical.construct(sc, span=1 if rings <= 16 else 3, max_delta=rings - 1,
                                            num_views=crystals // 2, num_tangential_poss=crystals // 2)
    img = VoxelsOnCartesianGrid(2 * rings - 1, xy, xy, 2.0, 2.0, 2.0)
    return pdi, img, RayTracingMatrixParameters(do_symmetry_90degrees_min_phi=crystals != 288)


@pytest.fixture(scope="module")
def batch():
    problems = [_problem(8, 192, 96), _problem(16, 192, 128), _problem(24, 192, 96), _problem(8, 288, 96)]
    print("[batch] set-up of 4 problems", flush=True)
    B = BatchedOSMAPOSL(problems, NSUB)
    data = []
    for i, (pdi, img, pm) in enumerate(problems):
        P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
        P.set_up(pdi, img, num_subsets=1)
        x_true, y, bkg = phantoms.config0_data(pdi, img, P.forward_project, seed=i)
        P.close()
        B.set_data(i, y, bkg)
        data.append((y, bkg))
    yield B, problems, data
    B.close()


def _reset(B, problems):
    for i, (pdi, img, pm) in enumerate(problems):
        B.set_estimate(i, np.ones(img.shape, np.float32))


def _images(B):
    return [B.get_estimate(i) for i in range(len(B))]


def _max_rel(a, b):
    m = a > 1e-3 * a.max()
    return float((np.abs(a - b)[m] / a[m]).max())


def test_batched_equals_per_problem(batch):
    """Full waves and lane groups sum every voxel in table order, as the
    per-problem kernels: identical images."""
    B, problems, data = batch
    _reset(B, problems)
    B.run_sequential(2 * NSUB)
    seq = _images(B)
    for bp_pack in (0, 1):
        B.set_launch(32, False, bp_pack, 32, 8)
        _reset(B, problems)
        B.run(2 * NSUB, use_graph=False)
        bat = _images(B)
        for i, (a, b) in enumerate(zip(seq, bat)):
            print(f"bp_pack {bp_pack} problem {i}: max |batched - per-problem| {np.abs(a - b).max():.3e}")
            assert np.array_equal(a, b)
    # waves split over table entries: parts summed in part order
    B.set_launch()
    _reset(B, problems)
    B.run(2 * NSUB, use_graph=False)
    for i, (a, b) in enumerate(zip(seq, _images(B))):
        rel = _max_rel(a, b)
        print(f"split waves problem {i}: max rel {rel:.2e}")
        assert rel <= 1e-5


def test_graph_replays_equal_eager(batch):
    B, problems, data = batch
    B.set_launch()
    _reset(B, problems)
    B.run(2 * NSUB, use_graph=False)
    eager = _images(B)
    for rep in range(3):
        _reset(B, problems)  # eager work between replays
        B.run(2 * NSUB, use_graph=True)
        got = _images(B)
        for a, b in zip(eager, got):
            assert np.array_equal(a, b), f"replay {rep}"


def test_set_estimate_right_after_run(batch):
    """run() returns without waiting; set_estimate() (and set_data) must wait
    for the queued subiterations before writing the estimate: a run, a
    reset with no get_estimate in between, a second run equals a clean
    reset and run."""
    B, problems, data = batch
    B.set_launch()
    _reset(B, problems)
    B.run(NSUB, use_graph=True)
    ref = _images(B)
    _reset(B, problems)
    B.run(2 * NSUB, use_graph=True)
    _reset(B, problems)  # no synchronisation with the run above
    B.run(NSUB, use_graph=True)
    for a, b in zip(ref, _images(B)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("launch", [(32, True, 1, 32, 8, 0), (0, False, 0, 1, 1, 0), (16, False, 1, 8, 8, 16),
                                    (64, False, 1, 32, 8, 64), (32, False, 0, 32, 8, 32), (32, False, 2, 32, 8, 32),
                                    (0, False, 2, 8, 8, 0)])
def test_launch_shapes_do_not_change_results(batch, launch):
    """Exact for shapes that keep the table order, 1e-5 for split waves."""
    B, problems, data = batch
    B.set_launch(32, False, 0, 32, 8)
    _reset(B, problems)
    B.run(NSUB, use_graph=False)
    ref = _images(B)
    B.set_launch(*launch)
    _reset(B, problems)
    B.run(NSUB, use_graph=False)
    got = _images(B)
    B.set_launch()
    for a, b in zip(ref, got):
        if launch[2] == 2:
            assert _max_rel(a, b) <= 1e-5
        else:
            assert np.array_equal(a, b)


def test_tuner_and_oracle(batch):
    B, problems, data = batch
    choice = B.tune()
    assert choice >= 0
    _reset(B, problems)
    B.run(2 * NSUB, use_graph=True)
    got = B.get_estimate(0)
    pdi, img, pm = problems[0]
    y, bkg = data[0]
    ref = OraclePlan(pdi, img, pm).osem(y, np.ones(img.shape, np.float32), NSUB, 2 * NSUB,
                                        bkg=np.full(pdi.num_bins, bkg, np.float32))
    mask = ref > 1e-3 * ref.max()
    rel = (np.abs(got - ref)[mask] / ref[mask]).max()
    print(f"batched problem 0 vs oracle OSEM: max rel {rel:.2e} (tuner choice {choice})")
    assert rel <= 1e-4
