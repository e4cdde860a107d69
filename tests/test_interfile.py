"""Interfile projection data / images and OSMAPOSL .par files (CPU).

Fixtures are the reference's own recon_test_pack headers: the RPT sinogram
header Utahscat600k_ca_seg4.hs (segments listed -4..4, sinogram order, scanner
named only by "originating system"), Siemens_mMR_seg2.hs and
scatter_cylinder.hs (viewgram order, a "Scanner parameters" block),
RPTsens_seg3_PM.hv/.v and OSMAPOSL_test_PM_for_sensitivity.par. Their data
files are not in the reference (apart from the image), so data placement is
checked with synthetic files whose every value encodes its (TOF bin, segment,
view, axial, tangential) position.
"""
import os
from dataclasses import replace

import numpy as np
import pytest

from stir_amd import interfile as I
from stir_amd.geometry import (ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, config0, config0_tof,
                               config1, config4_problems, rpt_sensitivity_case)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_keyword_standardisation():
    n = I._norm_key
    assert n("!matrix size [1]") == n("matrix size[1]") == "matrix size [1]"
    assert n("do_symmetry_90degrees_min_phi") == "do symmetry 90degrees min phi"
    assert n("  Number of   rings ") == "number of rings"


def test_par_file_gives_the_rpt_sensitivity_problem():
    pdi, img, pm, keys = I.reconstruction_setup(os.path.join(GOLDEN, "OSMAPOSL_test_PM_for_sensitivity.par"))
    rpdi, rimg, rpm = rpt_sensitivity_case()
    assert pdi == replace(rpdi, scanner=replace(rpdi.scanner, name="PRT-1"))
    assert img == rimg
    assert pm == rpm
    assert keys == {"num_subsets": 12, "num_subiterations": 1, "start_subset": 0, "start_subiteration": 2,
                    "zero_end_planes": True, "use_subset_sensitivities": False}
    par = I.parse_par(os.path.join(GOLDEN, "OSMAPOSL_test_PM_for_sensitivity.par"))
    assert I.find_key(par, "projector pair type") == "Matrix"
    assert I.find_key(par, "matrix type") == "Ray Tracing"
    assert I.find_key(par, "objective function type") == "PoissonLogLikelihoodWithLinearModelForMeanAndProjData"


def test_par_projector_keys():
    text = """OSMAPOSLParameters :=
    projector pair type := Matrix
      Projector Pair Using Matrix Parameters :=
      Matrix type := Ray Tracing
      Ray tracing matrix parameters :=
        restrict to cylindrical FOV := 0
        number of rays in tangential direction to trace for each bin := 3
        use actual detector boundaries := 0
        do_symmetry_90degrees_min_phi := 0
        do_symmetry_swap_s := 0
      End Ray tracing matrix parameters :=
      End Projector Pair Using Matrix Parameters :=
    END :=
    """
    pm = I.ray_tracing_matrix_parameters(I.parse_par(text))
    assert pm == RayTracingMatrixParameters(restrict_to_cylindrical_FOV=False, num_tangential_LORs=3,
                                            do_symmetry_90degrees_min_phi=False,
                                            do_symmetry_swap_s=False)


def test_known_scanner_headers():
    """Scanner from "originating system" with header overrides (RPT), and from a
    "Scanner parameters" block (mMR, an unknown scanner)."""
    pdi, h = I.read_proj_data_info(os.path.join(GOLDEN, "Utahscat600k_ca_seg4.hs"))
    assert h["_segment_sequence"] == [-4, -3, -2, -1, 0, 1, 2, 3, 4]
    assert pdi.num_axial_poss == (12, 13, 14, 15, 16, 15, 14, 13, 12) and pdi.min_segment_num == -4
    assert pdi.arc_corrected and pdi.bin_size == pdi.scanner.default_bin_size
    assert pdi.scanner.ring_spacing == np.float32(6.75)
    assert pdi.scanner.intrinsic_tilt == float(np.float32(-0.46875 * np.pi / 180))
    mmr, _ = I.read_proj_data_info(os.path.join(GOLDEN, "Siemens_mMR_seg2.hs"))
    ref = ProjDataInfoCylindrical.construct(Scanner.Siemens_mMR(328.0, 7.0), span=1, max_delta=2, num_views=252,
                                            num_tangential_poss=344)
    assert mmr == replace(ref, scanner=replace(ref.scanner, name="Siemens mMR"))
    cyl, h = I.read_proj_data_info(os.path.join(GOLDEN, "scatter_cylinder.hs"))
    assert (cyl.scanner.num_rings, cyl.scanner.num_detectors_per_ring) == (8, 64)
    assert cyl.scanner.inner_ring_radius == np.float32(443.1)
    assert cyl.scanner.average_depth_of_interaction == np.float32(8.4)
    assert cyl.scanner.ring_spacing == np.float32(19.62)
    assert (cyl.num_views, cyl.num_tangential_poss, cyl.num_axial_poss) == (32, 35, (8,))
    assert I._file_axes(h["_labels"]) == "viewgram"


def test_unknown_scanner_without_parameters_is_an_error(tmp_path):
    text = open(os.path.join(GOLDEN, "Utahscat600k_ca_seg4.hs")).read().replace("PRT-1", "no such scanner")
    text = text.replace("distance between rings (cm) := 0.675", "")
    p = tmp_path / "x.hs"
    p.write_text(text)
    with pytest.raises(ValueError):
        I.read_proj_data_info(str(p))


def _code(k, s, v, a, t):
    """Distinct integer below 2**20 (exact in float32) for a position with
    TOF bin k < 16, segment |s| < 8, view < 16, axial < 32, tangential < 8."""
    return (((k * 16 + s + 8) * 16 + v) * 32 + a) * 8 + t


def _encoded_file(path, pdi, listed_segments, order):
    """Raw file with segments in `listed_segments` order, each in sinogram or
    viewgram order, TOF bins outermost; value = code of its position."""
    nv, nt = pdi.num_views, pdi.num_tangential_poss
    chunks = []
    for k in range(pdi.num_tof_bins):
        for s in listed_segments:
            na = pdi.num_axial_poss_of(s)
            a, v, t = np.meshgrid(np.arange(na), np.arange(nv), np.arange(nt), indexing="ij")
            code = _code(k, s, v, a, t).astype(np.float64)
            if order == "viewgram":
                code = code.transpose(1, 0, 2)
            chunks.append(code.reshape(-1))
    np.concatenate(chunks).astype("<f8").tofile(path)


@pytest.mark.parametrize("order", ["sinogram", "viewgram"])
def test_segment_and_axis_placement(tmp_path, order):
    """A header listing segments -4..4 (the Utah header) and one in
    viewgram order: every value lands at ProjDataInMemory's index of its
    (segment, view, axial, tangential) position."""
    text = open(os.path.join(GOLDEN, "Utahscat600k_ca_seg4.hs")).read()
    text = text.replace("!number of bytes per pixel := 4", "!number of bytes per pixel := 8")
    # shrink to 8 views x 6 tangential positions
    text = text.replace("!matrix size [2] := 96", "!matrix size [2] := 8").replace("!matrix size [1] := 128",
                                                                                  "!matrix size [1] := 6")
    if order == "viewgram":
        text = (text.replace("matrix axis label [2] := view", "matrix axis label [2] := axial coordinate@")
                .replace("matrix axis label [3] := axial coordinate", "matrix axis label [3] := view")
                .replace("axial coordinate@", "axial coordinate")
                .replace("!matrix size [2] := 8", "!matrix size [2] := AX")
                .replace("!matrix size [3] := { 12,13,14,15,16,15,14,13,12}", "!matrix size [3] := 8")
                .replace("!matrix size [2] := AX", "!matrix size [2] := { 12,13,14,15,16,15,14,13,12}"))
    hs = tmp_path / "enc.hs"
    hs.write_text(text.replace("Utahscat600k_ca_seg4.s", "enc.s"))
    pdi, _ = I.read_proj_data_info(str(hs))
    assert (pdi.num_views, pdi.num_tangential_poss) == (8, 6)
    _encoded_file(tmp_path / "enc.s", pdi, [-4, -3, -2, -1, 0, 1, 2, 3, 4], order)
    _, y = I.read_proj_data(str(hs))
    rng = np.random.default_rng(0)
    for _ in range(300):
        s = int(rng.integers(-4, 5))
        v = int(rng.integers(0, 8))
        a = int(rng.integers(0, pdi.num_axial_poss_of(s)))
        t = int(rng.integers(0, 6))
        want = _code(0, s, v, a, t)
        assert y[pdi.bin_index(s, v, a, t + pdi.min_tangential_pos_num)] == np.float32(want)


def test_tof_placement(tmp_path):
    pdi, _ = config0_tof()
    small = replace(pdi, num_views=4, num_tangential_poss=6)
    I.write_proj_data(str(tmp_path / "t.hs"), small, np.zeros(small.num_bins, np.float32))
    text = (tmp_path / "t.hs").read_text().replace("!number of bytes per pixel := 4", "!number of bytes per pixel := 8")
    (tmp_path / "t.hs").write_text(text)
    _encoded_file(tmp_path / "t.s", small, small.segment_sequence(), "sinogram")
    q, y = I.read_proj_data(str(tmp_path / "t.hs"))
    assert q == small and q.num_tof_bins == 9
    block = small.num_bins // small.num_tof_bins
    for k in (0, 4, 8):
        for s in small.segments:
            want = _code(k, s, 3, 1, 5)
            assert y[k * block + small.bin_index(s, 3, 1, 5 + small.min_tangential_pos_num)] == np.float32(want)


@pytest.mark.parametrize("which", ["config0", "config1", "tof", "rpt", "batch0", "batch1", "batch2"])
def test_proj_data_round_trip(tmp_path, which):
    if which == "config0":
        pdi = config0()[0]
    elif which == "config1":
        pdi = config1()[0]
    elif which == "tof":
        pdi = config0_tof()[0]
    elif which == "rpt":
        pdi = rpt_sensitivity_case()[0]
    else:
        pdi = config4_problems()[int(which[-1])][0]
    x = np.random.default_rng(1).standard_normal(pdi.num_bins).astype(np.float32)
    I.write_proj_data(str(tmp_path / "p.hs"), pdi, x)
    q, y = I.read_proj_data(str(tmp_path / "p.hs"))
    assert q == pdi
    assert np.array_equal(x, y)


def test_golden_image_and_image_round_trip(tmp_path):
    img, data = I.read_image(os.path.join(GOLDEN, "RPTsens_seg3_PM.hv"))
    ref_img = rpt_sensitivity_case()[1]
    assert img.shape == ref_img.shape == (31, 60, 60)
    assert abs(img.vox_x - ref_img.vox_x) <= 1e-5 * ref_img.vox_x and img.vox_z == ref_img.vox_z
    raw = np.fromfile(os.path.join(GOLDEN, "RPTsens_seg3_PM.v"), dtype="<f4").reshape(31, 60, 60)
    assert np.array_equal(data, raw)
    I.write_image(str(tmp_path / "i.hv"), ref_img, data)
    img2, data2 = I.read_image(str(tmp_path / "i.hv"))
    assert img2 == ref_img and np.array_equal(data2, data)


def test_image_number_formats(tmp_path):
    """Big-endian signed shorts with a rescale slope."""
    vals = np.arange(-5, 19, dtype=np.int16).reshape(2, 3, 4)
    vals.astype(">i2").tofile(tmp_path / "s.v")
    (tmp_path / "s.hv").write_text("""!INTERFILE :=
name of data file := s.v
imagedata byte order := BIGENDIAN
!number format := signed integer
!number of bytes per pixel := 2
number of dimensions := 3
matrix axis label [1] := x
!matrix size [1] := 4
scaling factor (mm/pixel) [1] := 2
matrix axis label [2] := y
!matrix size [2] := 3
scaling factor (mm/pixel) [2] := 2
matrix axis label [3] := z
!matrix size [3] := 2
scaling factor (mm/pixel) [3] := 3
data rescale slope := 0.5
!END OF INTERFILE :=
""")
    img, data = I.read_image(str(tmp_path / "s.hv"))
    assert img.shape == (2, 3, 4) and (img.vox_x, img.vox_z) == (2.0, 3.0)
    assert np.array_equal(data, vals.astype(np.float32) * np.float32(0.5))


def test_par_nesting_of_blocks_closed_by_end_keys():
    """Blocks whose start key does not end in "parameters" ("Bin Normalisation
    From ProjData") are closed by their own End key, so later keys stay in the
    enclosing block."""
    from tests.test_gpu_par import PAR
    t = I.parse_par(PAR)
    obj = t["osmaposlparameters"]["poissonloglikelihoodwithlinearmodelformeanandprojdata parameters"]
    assert obj["bin normalisation from projdata"] == {"normalisation projdata filename": "norm.hs"}
    assert obj["projector pair type"] == "Separate Projectors"
    assert obj["quadratic prior parameters"] == {"penalisation factor": 0.02, "kappa filename": "kappa.hv"}
    assert t["osmaposlparameters"]["map model"] == "multiplicative"
    assert I.find_key(t, "forward projector type") == "MI355X Ray Tracing"


def test_nested_list_values():
    assert I._parse_value("{{{0,1,0},{1,0,1},{0,1,0}}}") == [[[0, 1, 0], [1, 0, 1], [0, 1, 0]]]
    assert I._parse_value("{ 12,13,14}") == [12, 13, 14]


def test_default_image_size_covers_the_fov():
    """xy size -1: 2 ceil(R_fov / voxel) + 1 (VoxelsOnCartesianGrid.cxx:252-267):
    mMR 344 tangential positions at R = 328 mm reach 288.0 mm -> 2*139+1."""
    from stir_amd.geometry import VoxelsOnCartesianGrid
    pdi = config1()[0]
    assert VoxelsOnCartesianGrid.from_proj_data_info(pdi).shape == (127, 279, 279)


def test_image_grid_zooms_and_z_offset():
    """VoxelsOnCartesianGrid(proj_data_info, zooms, origin, sizes)
    (VoxelsOnCartesianGrid.cxx:230-288): voxel size = sampling / zoom per axis,
    the default z size independent of the z zoom, origin z = "Z offset (in mm)"."""
    from stir_amd.geometry import ProjDataInfoCylindrical, Scanner, VoxelsOnCartesianGrid

    sc = Scanner.cylindrical(4, 64, 80.0, 4.0)
    pdi = ProjDataInfoCylindrical.construct(sc, span=1, max_delta=3, num_views=32, num_tangential_poss=32)
    base = VoxelsOnCartesianGrid.from_proj_data_info(pdi)
    img = VoxelsOnCartesianGrid.from_proj_data_info(pdi, zoom=1.0, zoom_z=2.0, z_offset=3.0)
    assert base.vox_z == 2.0 and img.vox_z == 1.0
    assert img.nz == base.nz == 2 * 4 - 1
    assert img.origin_z == 3.0 and base.origin_z == 0.0
    assert (img.vox_y, img.nx) == (base.vox_y, base.nx)
    with pytest.raises(ValueError):
        VoxelsOnCartesianGrid.from_proj_data_info(pdi, zoom_z=0.0)
