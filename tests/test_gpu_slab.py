"""Back projection through projection-data slabs in LDS (one plane per
axial position), GPU.

Workloads: the 40-ring x 64-crystal scanner (span 3, MRD 4: 79-plane image,
two 64-plane slots, entries reaching one or both slots) and the 8-ring x
192-crystal one (span 3, MRD 4: 15 planes, one slot), 2 mm voxels, 4
subsets; at mMR size in tests/test_gpu_mmr.py.

Bars: the slab path (MI_TUNE_BP_SLAB 1) and the column-table path agree
within 1e-5 of the maximum for full and subset back projections and for the
OSEM update; launch shapes (tiles, columns per wave, LDS budgets that split
the segments into groups) give results within 1e-5 and each shape is
deterministic (two runs bit-identical); the slab path meets the oracle
within 1e-5 of the maximum.
"""
import numpy as np
import pytest

from oracle.oracle import OraclePlan
from stir_amd import _native
from stir_amd.geometry import ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, VoxelsOnCartesianGrid
from stir_amd.projectors import (OSMAPOSLReconstruction, PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                                 ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing)

pytestmark = pytest.mark.gpu

CASES = {
    "40rings": (Scanner.cylindrical(40, 64, 100.0, 4.0, name="cyl40x64"), 32, 48,
                VoxelsOnCartesianGrid(79, 48, 48, 2.0, 2.0, 2.0)),
    "8rings": (Scanner.cylindrical(8, 192, 300.0, 4.0, name="cyl8x192"), 96, 128,
               VoxelsOnCartesianGrid(15, 128, 128, 2.0, 2.0, 2.0)),
}


def _pair(pdi, img, tuning=()):
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(RayTracingMatrixParameters()))
    P.context.set_tuning(_native.MI_TUNE_BP_SLAB, 1)
    for k, v in tuning:
        P.context.set_tuning(k, v)
    P.set_up(pdi, img, num_subsets=4)
    return P


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    sc, nv, nt, img = CASES[request.param]
    pdi = ProjDataInfoCylindrical.construct(sc, span=3, max_delta=4, num_views=nv, num_tangential_poss=nt)
    print(f"[slab] {request.param}: {pdi.num_bins} bins, image {img.shape}", flush=True)
    S = _pair(pdi, img)
    C = _pair(pdi, img, [(_native.MI_TUNE_BP_SLAB, 0)])
    si = S.info()
    print(f"[slab] tile {si.slab_tile_x} x {si.slab_tile_y}, {si.slab_segment_groups} segment groups, "
          f"{si.slab_lds_bytes} B of LDS, {si.num_table_entries} entries", flush=True)
    assert si.slab_back_projection == 1 and C.info().slab_back_projection == 0
    yield request.param, pdi, img, S, C
    S.close()
    C.close()


def _err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_slab_against_column_table(case):
    name, pdi, img, S, C = case
    rng = np.random.default_rng(41)
    y = rng.uniform(0.0, 2.0, pdi.num_bins).astype(np.float32)
    for subset in (-1, 0, 3):
        a, b = S.back_project(y, subset=subset), C.back_project(y, subset=subset)
        print(f"{name} subset {subset}: slab vs column table {_err(a, b):.2e}", flush=True)
        assert _err(a, b) <= 1e-5
        assert np.array_equal(a, S.back_project(y, subset=subset))


def test_slab_against_oracle(case):
    name, pdi, img, S, C = case
    y = np.random.default_rng(42).uniform(0.0, 1.0, pdi.num_bins).astype(np.float32)
    O = OraclePlan(pdi, img, RayTracingMatrixParameters())
    ref = O.back_project_f64(y.astype(np.float64))
    got = S.back_project(y)
    print(f"{name}: slab vs FP64 oracle {_err(got, ref):.2e}", flush=True)
    assert _err(got, ref) <= 1e-5


@pytest.mark.parametrize("shape", [(16, 8, 8, 160), (8, 4, 2, 80), (8, 4, 2, 24), (4, 8, 8, 16)])
def test_launch_shapes(case, shape):
    name, pdi, img, S, C = case
    tx, ty, cpw, kb = shape
    P = _pair(pdi, img, [(_native.MI_TUNE_SLAB_TILE_X, tx), (_native.MI_TUNE_SLAB_TILE_Y, ty),
                         (_native.MI_TUNE_SLAB_COLUMNS_PER_WAVE, cpw), (_native.MI_TUNE_SLAB_LDS_KB, kb)])
    y = np.random.default_rng(43).uniform(0.0, 2.0, pdi.num_bins).astype(np.float32)
    info = P.info()
    assert info.slab_back_projection == 1 and (info.slab_tile_x, info.slab_tile_y) == (tx, ty)
    assert info.slab_lds_bytes <= kb * 1024 or info.slab_segment_groups == 1
    a = P.back_project(y, subset=1)
    assert np.array_equal(a, P.back_project(y, subset=1))
    b = S.back_project(y, subset=1)
    P.close()
    print(f"{name} shape {shape}: {_err(a, b):.2e} from the default shape", flush=True)
    assert _err(a, b) <= 1e-5


def test_osem_update_slab_against_column_table(case):
    name, pdi, img, S, C = case
    rng = np.random.default_rng(44)
    x = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    y = rng.poisson(3.0, pdi.num_bins).astype(np.float32)
    out = []
    for P in (S, C):
        obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
        obj.set_up(y, 0.1)
        obj.set_estimate(x)
        OSMAPOSLReconstruction(obj, 1).update_estimate(2, 1)
        out.append(obj.get_estimate())
    print(f"{name}: OSEM update slab vs column table {_err(out[0], out[1]):.2e}", flush=True)
    assert _err(out[0], out[1]) <= 1e-5
