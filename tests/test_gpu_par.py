"""OSMAPOSL parameter files run end to end on the GPU (stir_amd.osmaposl).

1. The reference's own OSMAPOSL_test_PM_for_sensitivity.par on its
   Utahscat600k_ca_seg4.hs header, unchanged: the sensitivity image it writes
   matches the reference's RPTsens_seg3_PM.v within compare_image's 5e-4 of the
   maximum, with the same support.
2. A .par written here for the 8-ring x 192-crystal workload (span 1, MRD 3,
   96 x 128 bins, 128x128x15 image at 2 mm through "zoom"): Interfile data, additive sinogram, "From
   ProjData" normalisation, quadratic prior with a kappa image, the
   multiplicative MAP model, relative-change bounds, 4 subsets and 6
   subiterations saved every 2: the saved images exist and the last one is
   within 1e-4 relative of the oracle's one-step-late OSEM with the same
   settings (factors = 1 / normalisation).
"""
import os
import shutil

import numpy as np
import pytest

from oracle.oracle import OraclePlan
from stir_amd import interfile as I
from stir_amd import phantoms
from stir_amd.geometry import RayTracingMatrixParameters, config0
from stir_amd.osmaposl import run_par

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_reference_sensitivity_par(tmp_path):
    for f in ("OSMAPOSL_test_PM_for_sensitivity.par", "Utahscat600k_ca_seg4.hs"):
        shutil.copy(os.path.join(GOLDEN, f), tmp_path / f)
    print("[par] OSMAPOSL_test_PM_for_sensitivity.par", flush=True)
    _, s = run_par(str(tmp_path / "OSMAPOSL_test_PM_for_sensitivity.par"), output_dir=str(tmp_path))
    s["projector"].close()
    img, got = I.read_image(str(tmp_path / "my_RPTsens_seg3_PM.hv"))
    _, ref = I.read_image(os.path.join(GOLDEN, "RPTsens_seg3_PM.hv"))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"sensitivity from the reference .par: max err {err:.2e} of max")
    assert img.shape == ref.shape
    assert err <= 5e-4
    assert int((got != 0).sum()) == int((ref != 0).sum())


PAR = """OSMAPOSLParameters :=
objective function type:= PoissonLogLikelihoodWithLinearModelForMeanAndProjData
PoissonLogLikelihoodWithLinearModelForMeanAndProjData Parameters:=
input file := data.hs
zero end planes of segment 0 := 1
zoom := 2.454369306564331
xy output image size (in pixels) := 128
additive sinogram := additive.hs
Bin Normalisation type := From ProjData
  Bin Normalisation From ProjData :=
  normalisation projdata filename := norm.hs
  End Bin Normalisation From ProjData :=
projector pair type := Separate Projectors
  Projector Pair Using Separate Projectors Parameters :=
    Forward projector type := MI355X Ray Tracing
      Forward Projector Using MI355X Ray Tracing Parameters :=
        Ray Tracing Matrix Parameters :=
        End Ray Tracing Matrix Parameters :=
      End Forward Projector Using MI355X Ray Tracing Parameters :=
    Back projector type := MI355X Ray Tracing
      Back Projector Using MI355X Ray Tracing Parameters :=
        Ray Tracing Matrix Parameters :=
        End Ray Tracing Matrix Parameters :=
      End Back Projector Using MI355X Ray Tracing Parameters :=
  End Projector Pair Using Separate Projectors Parameters :=
prior type := Quadratic
  Quadratic Prior Parameters :=
  penalisation factor := 0.02
  kappa filename := kappa.hv
  END Quadratic Prior Parameters :=
end PoissonLogLikelihoodWithLinearModelForMeanAndProjData Parameters:=
MAP_model := multiplicative
minimum relative change := 0.2
maximum relative change := 5
output filename prefix := recon
number of subsets:= 4
number of subiterations:= 6
save estimates at subiteration intervals:= 2
END :=
"""


def test_written_par_with_norm_prior_and_bounds(tmp_path):
    pdi, img = config0()
    pm = RayTracingMatrixParameters()
    O = OraclePlan(pdi, img, pm)
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project)
    rng = np.random.default_rng(31)
    norm = rng.uniform(0.5, 2.0, pdi.num_bins).astype(np.float32)
    add = np.full(pdi.num_bins, bkg, np.float32) * rng.uniform(0.5, 1.5, pdi.num_bins).astype(np.float32)
    kappa = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    I.write_proj_data(str(tmp_path / "data.hs"), pdi, y)
    I.write_proj_data(str(tmp_path / "additive.hs"), pdi, add)
    I.write_proj_data(str(tmp_path / "norm.hs"), pdi, norm)
    I.write_image(str(tmp_path / "kappa.hv"), img, kappa)
    (tmp_path / "recon.par").write_text(PAR)
    print("[par] written .par with normalisation, prior, bounds", flush=True)
    x, s = run_par(str(tmp_path / "recon.par"), output_dir=str(tmp_path))
    s["projector"].close()
    assert s["image"] == img
    for k in (2, 4, 6):
        assert (tmp_path / f"recon_{k}.hv").exists()
    _, saved = I.read_image(str(tmp_path / "recon_6.hv"))
    assert np.array_equal(saved, x)
    mult = (np.float32(1) / norm).astype(np.float32)
    O.set_multiplicative(mult)
    O.set_prior("quadratic", 0.02, kappa=kappa)
    O.set_osem_options("multiplicative", 0.2, 5.0)
    try:
        ref = O.osem(y, np.ones(img.shape, np.float32), 4, 6, bkg=add, zero_end_planes=True)
    finally:
        O.set_multiplicative(None)
        O.set_prior(None)
        O.set_osem_options()
    mask = ref > 1e-3 * ref.max()
    rel = float((np.abs(x - ref)[mask] / ref[mask]).max())
    print(f"written .par OSL vs oracle: max rel {rel:.2e}")
    assert rel <= 1e-4
