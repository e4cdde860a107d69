"""Host-side checks that need no GPU: geometry construction, the C ABI
library's exports and error path, and the host plan's symmetry operations
against the oracle's (an independent transcription of the same tables)."""
import ctypes
import os
import re

import numpy as np
import pytest

from stir_amd import _native
from stir_amd.geometry import (ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, VoxelsOnCartesianGrid,
                               config0, config1, make_geometry_struct, rpt_sensitivity_case)
from oracle.oracle import OraclePlan

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mmr_span11_layout():
    """Appendix numbers of the mMR layout: span 11, MRD 60 -> 11 segments,
    127,115,115,93,93,71,71,49,49,27,27 axial positions, 837 sinograms."""
    pdi, img = config1()
    assert pdi.min_segment_num == -5 and pdi.max_segment_num == 5
    seg_nax = {s: pdi.num_axial_poss_of(s) for s in pdi.segments}
    assert seg_nax[0] == 127
    assert [seg_nax[s] for s in (1, 2, 3, 4, 5)] == [115, 93, 71, 49, 27]
    assert all(seg_nax[s] == seg_nax[-s] for s in range(1, 6))
    assert pdi.num_sinograms == 837
    assert pdi.num_bins == 72557856
    assert img.num_voxels == 15028672
    # ring differences of segment 1 are 6..16 (ProjDataInfo.cxx:509-574)
    i = 1 - pdi.min_segment_num
    assert (pdi.min_ring_diff[i], pdi.max_ring_diff[i]) == (6, 16)
    assert pdi.max_ring_diff[-1] == 60


def test_config0_and_rpt_layouts():
    pdi, img = config0()
    assert pdi.num_sinograms == 44 and pdi.num_bins == 540672 and img.shape == (15, 128, 128)
    pdi, img, pm = rpt_sensitivity_case()
    assert [pdi.num_axial_poss_of(s) for s in pdi.segments] == [13, 14, 15, 16, 15, 14, 13]
    assert img.shape == (31, 60, 60)
    assert abs(img.vox_x - 4.44114) < 1e-5 and img.vox_z == 3.375
    assert pm.num_tangential_LORs == 2


def test_segment_sequence_and_bin_index_match_oracle():
    pdi, img = config0()
    P = OraclePlan(pdi, img)
    assert pdi.segment_sequence() == [0, 1, -1, 2, -2, 3, -3]
    rng = np.random.default_rng(0)
    for _ in range(200):
        seg = int(rng.integers(-3, 4))
        b = (seg, int(rng.integers(0, 96)), int(rng.integers(0, pdi.num_axial_poss_of(seg))), int(rng.integers(-64, 64)))
        assert pdi.bin_index(*b) == P.bin_index(*b)


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "stir_amd_pmrt.h")).read()
    declared = set(re.findall(r"\b(mi_[a-z_0-9]+)\s*\(", header))
    assert declared == set(_native.EXPORTED_SYMBOLS)
    lib = _native.load()
    for name in declared:
        assert hasattr(lib, name), name


def test_library_fails_loudly_without_device():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = _native.load().mi_ctx_create(0, ctypes.byref(h))
    assert rc != 0
    assert _native.load().mi_last_error().decode()


def _all_bins(pdi, ax_max=None):
    out = []
    for seg in pdi.segments:
        for view in range(pdi.num_views):
            for ax in range(pdi.num_axial_poss_of(seg) if ax_max is None else min(ax_max, pdi.num_axial_poss_of(seg))):
                for tang in range(pdi.min_tangential_pos_num, pdi.max_tangential_pos_num + 1):
                    out.append((seg, view, ax, tang))
    return np.array(out, np.int32)


@pytest.mark.parametrize("case", ["config0", "rpt", "config0_no_z_shift", "config0_180_only", "mmr_basic_bins"])
def test_host_symmetry_operations_match_oracle(case):
    pm = RayTracingMatrixParameters()
    ax_max = None
    if case == "rpt":
        pdi, img, pm = rpt_sensitivity_case()
    elif case == "mmr_basic_bins":
        pdi, img = config1()
        img = VoxelsOnCartesianGrid(127, 32, 32, img.vox_z, img.vox_y, img.vox_x)  # rows are not traced here
        ax_max = 2
    else:
        pdi, img = config0()
        if case == "config0_no_z_shift":
            pm = RayTracingMatrixParameters(do_symmetry_shift_z=False)
        elif case == "config0_180_only":
            pm = RayTracingMatrixParameters(do_symmetry_90degrees_min_phi=False)
    bins = _all_bins(pdi, ax_max)
    g = make_geometry_struct(pdi, img, pm)
    basic, ops = _native.symmetry_operations(g, bins)
    P = OraclePlan(pdi, img, pm)
    rng = np.random.default_rng(1)
    sel = rng.choice(len(bins), min(len(bins), 60000), replace=False)
    for i in sel:
        b, o = P.symmetry_op(*bins[i].tolist())
        assert tuple(basic[i].tolist()) == b, (bins[i], basic[i], b)
        assert tuple(ops[i].tolist()) == o, (bins[i], ops[i], o)


def test_unsupported_geometry_is_rejected_on_host():
    pdi, img = config0()
    img = VoxelsOnCartesianGrid(15, 128, 128, 2.0, 2.0, 2.0, origin_x=1.0)
    g = make_geometry_struct(pdi, img, RayTracingMatrixParameters())
    with pytest.raises(_native.NativeError, match="origin"):
        _native.symmetry_operations(g, np.zeros((1, 4), np.int32))


def test_random_subset_order_is_a_permutation_per_cycle():
    """get_subset_num (IterativeReconstruction.cxx:629-638) without a device:
    (k + start - 1) mod n in order, a fresh permutation every n subiterations
    with "uniformly randomise subset order", reproducible with a seed."""
    from stir_amd.projectors import OSMAPOSLReconstruction

    class _Pair:
        num_subsets = 5

    class _Obj:
        projector = _Pair()

    plain = OSMAPOSLReconstruction(_Obj(), 10, start_subset=2)
    assert [plain.get_subset_num(k) for k in range(1, 11)] == [(k + 1) % 5 for k in range(1, 11)]
    r1 = OSMAPOSLReconstruction(_Obj(), 15, uniformly_randomise_subset_order=True, seed=7)
    seq = [r1.get_subset_num(k) for k in range(1, 16)]
    for c in range(3):
        assert sorted(seq[5 * c:5 * c + 5]) == list(range(5))
    assert len({tuple(seq[0:5]), tuple(seq[5:10]), tuple(seq[10:15])}) > 1
    r2 = OSMAPOSLReconstruction(_Obj(), 15, uniformly_randomise_subset_order=True, seed=7)
    assert [r2.get_subset_num(k) for k in range(1, 16)] == seq


def test_hessian_entry_point_rejects_null_arguments():
    """mi_obj_hessian_times_input_dev checks its arguments before any device
    call: a NULL context or image is MI_ERR_ARGUMENT."""
    L = _native.load()
    buf = (ctypes.c_float * 4)()
    rc = L.mi_obj_hessian_times_input_dev(None, 0, ctypes.cast(buf, ctypes.c_void_p), ctypes.cast(buf, ctypes.c_void_p),
                                          0, ctypes.cast(buf, ctypes.c_void_p), None)
    assert rc != 0 and "NULL" in L.mi_last_error().decode()
