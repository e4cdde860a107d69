"""Time-of-flight projector pair at mMR size, GPU against the oracle.

Workload: 64 rings x 504 crystals, span 11, MRD 60, 837 x 252 x 344 bins
with 11 TOF bins of 89.04 mm (594 ps) and a 400 ps FWHM kernel
(11 x 72 557 856 = 798 136 416 bins); 344 x 344 x 127 image at
2.08626 x 2.08626 x 2.03125 mm; 21 subsets. With TOF only the z-shift
symmetry stays on, so every (segment, view, tangential position) has its
own directly traced row.

Bars: the forward projection of a subset bit for bit (np.array_equal; rows
and TOF weights in the reference's element and float order), the back
projection of the same subset within 1e-5 of the maximum and 1e-5 relative
where the value exceeds 1e-3 of the maximum, and over all subsets the
adjoint identity <A x, y> = <x, A^T y> in FP64 (1e-5).
"""
import time

import numpy as np
import pytest
import torch

from oracle.oracle import OraclePlan
from stir_amd.geometry import RayTracingMatrixParameters, config3
from stir_amd.projectors import ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing

pytestmark = pytest.mark.gpu
NSUB = 21


@pytest.fixture(scope="module")
def tof_mmr():
    pdi, img = config3()
    pm = RayTracingMatrixParameters()
    print(f"[tof mmr] set-up: {pdi.num_bins} bins ({pdi.num_tof_bins} TOF bins), image {img.shape}", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=NSUB)
    t0 = time.time()
    O = OraclePlan(pdi, img, pm)
    print(f"[tof mmr] oracle plan {time.time() - t0:.1f} s", flush=True)
    yield pdi, img, P, O
    P.close()


def test_tof_mmr_setup(tof_mmr):
    pdi, img, P, O = tof_mmr
    assert pdi.num_tof_bins == 11
    assert P.info().num_bins == pdi.num_bins == O.num_bins == 11 * 72557856
    assert tuple(P.info().symmetry_flags) == O.symmetry_flags() == (0, 0, 0, 0, 1)
    # the projections run from the stored TOF rows (narrow windows)
    info = P.info()
    print(f"[tof mmr] live span {info.tof_live_span}, windows {info.tof_fp_window}/{info.tof_bp_window}", flush=True)
    assert (info.tof_fp_window, info.tof_bp_window) == (6, 5)


def test_tof_mmr_subset_projections_match_oracle(tof_mmr):
    pdi, img, P, O = tof_mmr
    rng = np.random.default_rng(31)
    x = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    s = 13
    print(f"[tof mmr] oracle forward projection of subset {s}", flush=True)
    ref = O.forward_project(x, subset=s, num_subsets=NSUB, out=np.zeros(pdi.num_bins, np.float32))
    got = P.forward_project(x, subset=s)
    same = float(np.mean(got.view(np.uint32) == ref.view(np.uint32)))
    print(f"TOF forward subset {s}: {np.count_nonzero(ref)} non-zero bins, bit-exact fraction {same:.8f}",
          flush=True)
    assert np.count_nonzero(ref) > 0
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    del ref, got
    y = rng.uniform(0.0, 2.0, pdi.num_bins).astype(np.float32)
    print(f"[tof mmr] oracle back projection of subset {s}", flush=True)
    bref = O.back_project(y, subset=s, num_subsets=NSUB, omp=True)
    bgot = P.back_project(y, subset=s)
    err = float(np.abs(bgot - bref).max() / np.abs(bref).max())
    m = bref > 1e-3 * bref.max()
    rel = float((np.abs(bgot - bref)[m] / bref[m]).max())
    print(f"TOF back subset {s}: max err {err:.2e} of max, {rel:.2e} relative", flush=True)
    assert err <= 1e-5 and rel <= 1e-5


def test_tof_mmr_adjoint_identity_fp64(tof_mmr):
    pdi, img, P, O = tof_mmr
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(32)
    x = torch.rand(tuple(img.shape), generator=g, device=dev)
    y = torch.rand(pdi.num_bins, generator=g, device=dev)
    lhs = torch.dot(P.forward_project(x).double(), y.double()).item()
    rhs = torch.dot(x.reshape(-1).double(), P.back_project(y).reshape(-1).double()).item()
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"TOF adjoint at mMR size: <Ax,y> {lhs:.9e} <x,ATy> {rhs:.9e} rel {rel:.2e}")
    assert rel <= 1e-5
