"""The projector pair at mMR size (64 rings x 504 crystals, span 11, MRD 60,
837 x 252 x 344 bins; 344 x 344 x 127 image, 21 subsets), GPU.

The oracle projects one subset at this size in well under a minute, so the
production shape itself is checked against it: the forward projection of a
subset bit for bit, its back projection within 1e-5 of the maximum. Over
all subsets, the adjoint identity <A x, y> = <x, A^T y> in FP64 sums.
"""
import numpy as np
import pytest

from oracle.oracle import OraclePlan
from stir_amd.geometry import RayTracingMatrixParameters, config1
from stir_amd.projectors import ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing

pytestmark = pytest.mark.gpu
NSUB = 21


@pytest.fixture(scope="module")
def mmr():
    pdi, img = config1()
    pm = RayTracingMatrixParameters()
    print(f"[mmr] set-up: {pdi.num_bins} bins, image {img.shape}", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=NSUB)
    yield pdi, img, pm, P
    P.close()


def test_subset_projections_match_oracle(mmr):
    pdi, img, pm, P = mmr
    O = OraclePlan(pdi, img, pm)
    rng = np.random.default_rng(21)
    x = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    s = 7
    print("[mmr] oracle forward projection of one subset", flush=True)
    ref = O.forward_project(x, subset=s, num_subsets=NSUB, out=np.zeros(pdi.num_bins, np.float32))
    got = P.forward_project(x, subset=s)
    print(f"forward subset {s}: bit-exact bins {np.mean(got == ref):.6f}, "
          f"max err {np.abs(got - ref).max() / np.abs(ref).max():.2e} of max", flush=True)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    y = rng.uniform(0.0, 2.0, pdi.num_bins).astype(np.float32)
    print("[mmr] oracle back projection of one subset", flush=True)
    bref = O.back_project(y, subset=s, num_subsets=NSUB, omp=True)
    bgot = P.back_project(y, subset=s)
    err = float(np.abs(bgot - bref).max() / np.abs(bref).max())
    print(f"back subset {s}: max err {err:.2e} of max", flush=True)
    assert err <= 1e-5


def test_adjoint_identity(mmr):
    pdi, img, pm, P = mmr
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
