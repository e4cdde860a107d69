"""Interfile images and projection data, and STIR ``.par`` files.

* ``read_header`` / ``parse_par``: STIR's ``key := value`` syntax (keys are
  case-insensitive, a leading ``!`` or ``%`` is ignored, ``;`` starts a
  comment, ``{a, b}`` is a list; ``.par`` blocks nest until their
  ``End ... :=`` line).
* ``read_image`` / ``write_image``: ``.hv`` + raw data as VoxelsOnCartesianGrid
  (``InterfileHeader.cxx``; x fastest, then y, then z).
* ``read_proj_data`` / ``write_proj_data``: ``.hs`` + raw data as
  ProjDataInfoCylindrical and a flat array in ProjDataInMemory order. The
  file's segments are numbered from their average ring difference and read in
  the listed order (``find_segment_sequence``, ``InterfileHeader.cxx:795-935``);
  sinogram (axial slower than view) and viewgram (view slower than axial)
  storage are both read; TOF bins outermost.
* ``ray_tracing_matrix_parameters`` / ``reconstruction_setup``: the projector
  and reconstruction keys of an OSMAPOSL ``.par`` file
  (``ProjMatrixByBinUsingRayTracing.cxx:76-104``, ``OSMAPOSLReconstruction``,
  ``PoissonLogLikelihoodWithLinearModelForMeanAndProjData`` keys).
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import replace
from typing import Any, Dict, List, Optional, Tuple

import numpy as np

from .geometry import (ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, VoxelsOnCartesianGrid, f32)


# ---------------------------------------------------------------------------
# key := value syntax
def _norm_key(k: str) -> str:
    """Keyword as STIR standardises it (interfile_keyword_functions.cxx:28-66):
    lower case, runs of white space, '_' and '!' collapsed to one space,
    trimmed; an index ``[i]`` is kept as `` [i]``."""
    k = k.strip()
    idx = ""
    m = re.match(r"^(.*?)\[\s*(\d+)\s*\]\s*$", k)
    if m:
        k, idx = m.group(1), f" [{int(m.group(2))}]"
    return re.sub(r"[\s_!%]+", " ", k).strip().lower() + idx


def _parse_value(v: str):
    """Scalar, or (nested) list for {a, b} / {{a, b}, {c, d}}."""
    v = v.strip()
    if not v.startswith("{"):
        return _parse_scalar(v)
    stack: List[list] = [[]]
    tok = ""
    for ch in v:
        if ch == "{":
            stack.append([])
        elif ch in ",}":
            if tok.strip():
                stack[-1].append(_parse_scalar(tok))
            tok = ""
            if ch == "}":
                inner = stack.pop()
                stack[-1].append(inner)
        else:
            tok += ch
    return stack[0][0] if stack[0] else []


def _parse_scalar(v: str):
    v = v.strip()
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v)
    except ValueError:
        return v


def _lines(text: str):
    for raw in text.splitlines():
        line = raw.split(";", 1)[0].strip()
        if not line or ":=" not in line:
            continue
        k, v = line.split(":=", 1)
        yield _norm_key(k), v.strip()


def read_header(path: str) -> Dict[str, Any]:
    """Flat dict of an Interfile header (last occurrence of a key wins)."""
    with open(path) as f:
        text = f.read()
    out: Dict[str, Any] = {}
    for k, v in _lines(text):
        out[k] = _parse_value(v) if v else ""
    return out


def parse_par(path_or_text: str) -> Dict[str, Any]:
    """Nested dict of a STIR parameter file. A key whose value names a type
    (``projector pair type := Matrix``) keeps that value; the following
    ``<Type> Parameters :=`` block becomes a nested dict under the block's key."""
    text = path_or_text
    if os.path.exists(path_or_text):
        with open(path_or_text) as f:
            text = f.read()
    lines = list(_lines(text))
    # a key without value opens a block when a matching "End <key>" follows or
    # its name ends in "parameters" (closed by "End ..." or the final "END")
    closers = {k[4:] for k, v in lines if k.startswith("end ") and v == ""}
    root: Dict[str, Any] = {}
    stack: List[Dict[str, Any]] = [root]
    for k, v in lines:
        if k == "end":
            break
        if k.startswith("end ") and v == "":
            if len(stack) > 1:
                stack.pop()
            continue
        if v == "" and (k in closers or k.endswith("parameters")):
            child: Dict[str, Any] = {}
            stack[-1][k] = child
            stack.append(child)
            continue
        stack[-1][k] = _parse_value(v) if v else ""
    return root


def find_key(tree: Dict[str, Any], key: str, default=None):
    """First value of `key` anywhere in a nested ``parse_par`` dict."""
    key = _norm_key(key)
    if key in tree:
        return tree[key]
    for v in tree.values():
        if isinstance(v, dict):
            r = find_key(v, key, None)
            if r is not None:
                return r
    return default


# ---------------------------------------------------------------------------
# images
def _short(v: float) -> str:
    """Shortest decimal that reads back as the same float32."""
    return np.format_float_positional(np.float32(v), unique=True, trim="-")


_NUMBER_FORMATS = {
    ("float", 4): np.float32,
    ("float", 8): np.float64,
    ("short float", 4): np.float32,
    ("long float", 8): np.float64,
    ("unsigned integer", 1): np.uint8,
    ("unsigned integer", 2): np.uint16,
    ("unsigned integer", 4): np.uint32,
    ("signed integer", 1): np.int8,
    ("signed integer", 2): np.int16,
    ("signed integer", 4): np.int32,
}


def _dtype(h: Dict[str, Any]) -> np.dtype:
    fmt = str(h.get("number format", "float")).lower()
    nb = int(h.get("number of bytes per pixel", 4))
    try:
        dt = np.dtype(_NUMBER_FORMATS[(fmt, nb)])
    except KeyError as e:
        raise ValueError(f"unsupported number format {fmt!r} with {nb} bytes") from e
    order = str(h.get("imagedata byte order", "littleendian")).lower()
    return dt.newbyteorder("<" if order.startswith("little") else ">")


def _data_path(header_path: str, h: Dict[str, Any]) -> str:
    name = str(h["name of data file"])
    return name if os.path.isabs(name) else os.path.join(os.path.dirname(os.path.abspath(header_path)), name)


def _scaled(h: Dict[str, Any], data: np.ndarray) -> np.ndarray:
    s = h.get("data rescale slope", h.get("scale factor", 1))
    o = h.get("data rescale offset", 0)
    if isinstance(s, list):
        s = s[0]
    out = data.astype(np.float32)
    if s != 1 or o != 0:
        out = (out * np.float32(s) + np.float32(o)).astype(np.float32)
    return out


def read_image(path: str) -> Tuple[VoxelsOnCartesianGrid, np.ndarray]:
    """(image geometry, array [z][y][x] float32) of an ``.hv`` file."""
    h = read_header(path)
    labels = {int(m.group(1)): str(v).strip().lower() for k, v in h.items()
              for m in [re.match(r"matrix axis label \[(\d)\]", k)] if m}
    sizes = {int(m.group(1)): int(v) for k, v in h.items() for m in [re.match(r"matrix size \[(\d)\]", k)] if m}
    if [labels.get(i) for i in (1, 2, 3)] != ["x", "y", "z"]:
        raise ValueError("only x, y, z images (x fastest) are supported")
    scale = {i: float(h.get(f"scaling factor (mm/pixel) [{i}]", 1.0)) for i in (1, 2, 3)}
    nx, ny, nz = sizes[1], sizes[2], sizes[3]
    offset = int(h.get("data offset in bytes [1]", 0))
    raw = np.fromfile(_data_path(path, h), dtype=_dtype(h), count=nx * ny * nz, offset=offset)
    if raw.size != nx * ny * nz:
        raise ValueError(f"{path}: data file is shorter than the header says")
    img = VoxelsOnCartesianGrid(nz, ny, nx, f32(scale[3]), f32(scale[2]), f32(scale[1]))
    return img, _scaled(h, raw).reshape(nz, ny, nx)


def write_image(path: str, img: VoxelsOnCartesianGrid, data: np.ndarray) -> None:
    """Write ``path`` (.hv) and its float32 little-endian data file (.v)."""
    a = np.ascontiguousarray(data, dtype="<f4").reshape(img.shape)
    data_name = os.path.splitext(os.path.basename(path))[0] + ".v"
    a.tofile(os.path.join(os.path.dirname(os.path.abspath(path)), data_name))
    lines = ["!INTERFILE  :=", f"name of data file := {data_name}", "!GENERAL DATA :=", "!GENERAL IMAGE DATA :=",
             "!type of data := PET", "imagedata byte order := LITTLEENDIAN", "!PET STUDY (General) :=",
             "!PET data type := Image", "process status := Reconstructed", "!number format := float",
             "!number of bytes per pixel := 4", "number of dimensions := 3"]
    for i, (lab, n, s) in enumerate((("x", img.nx, img.vox_x), ("y", img.ny, img.vox_y), ("z", img.nz, img.vox_z)), 1):
        lines += [f"matrix axis label [{i}] := {lab}", f"!matrix size [{i}] := {n}",
                  f"scaling factor (mm/pixel) [{i}] := {_short(s)}"]
    lines += ["number of time frames := 1", "!END OF INTERFILE :="]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ---------------------------------------------------------------------------
# projection data
# scanners recognised from "originating system" (Scanner.cxx name lists)
_KNOWN_SCANNERS = {
    "prt-1": Scanner.RPT,
    "rpt": Scanner.RPT,
    "siemens mmr": lambda: Scanner.Siemens_mMR(328.0, 7.0),
    "mmr": lambda: Scanner.Siemens_mMR(328.0, 7.0),
    "2008": lambda: Scanner.Siemens_mMR(328.0, 7.0),
}


def _fget(h: Dict[str, Any], key: str, default: float) -> float:
    v = h.get(key, default)
    return float(v) if v != "" else default


def _scanner_from_header(h: Dict[str, Any]) -> Scanner:
    """Scanner of a projection-data header (InterfileHeader.cxx:1040-1385):
    keys missing from the header are taken from the scanner named by
    "originating system" (converted to cm in float, as the reference does),
    then every length is rounded to float from the double value in mm."""
    name = str(h.get("originating system", "")).strip()
    guess = _KNOWN_SCANNERS.get(name.lower())
    g = guess() if guess else None
    rings = int(h.get("number of rings", -1))
    ndet = int(h.get("number of detectors per ring", -1))
    diam = _fget(h, "inner ring diameter (cm)", -1.0)
    doi = _fget(h, "average depth of interaction (cm)", -1.0)
    rs = _fget(h, "distance between rings (cm)", -1.0)
    bs = _fget(h, "default bin size (cm)", -1.0)
    nonarc = int(h.get("maximum number of non-arc-corrected bins", 0))
    arcb = int(h.get("default number of arc-corrected bins", 0))
    ntof = int(h.get("maximum number of (unmashed) tof time bins", h.get("number of tof time bins", -1)))
    tsize = _fget(h, "size of unmashed tof time bins (ps)", _fget(h, "size of timing bin (ps)", -1.0))
    tres = _fget(h, "tof timing resolution (ps)", _fget(h, "timing resolution (ps)", -1.0))
    if g is not None:
        f = np.float32
        rings = g.num_rings if rings < 1 else rings
        ndet = g.num_detectors_per_ring if ndet < 1 else ndet
        diam = float(f(f(g.inner_ring_radius) * f(2))) / 10.0 if diam < 0 else diam
        doi = float(f(f(g.average_depth_of_interaction) / f(10))) if doi < 0 else doi
        rs = float(f(f(g.ring_spacing) / f(10))) if rs < 0 else rs
        bs = float(f(f(g.default_bin_size) / f(10))) if bs < 0 else bs
        nonarc = g.max_num_non_arccorrected_bins if nonarc <= 0 else nonarc
        arcb = g.default_num_arccorrected_bins if arcb <= 0 else arcb
        ntof = g.max_num_timing_poss if ntof < 0 else ntof
        tsize = g.size_of_timing_pos_ps if tsize < 0 else tsize
        tres = g.timing_resolution_ps if tres < 0 else tres
    if rings < 1 or ndet < 1 or diam <= 0 or rs <= 0:
        raise ValueError(f"scanner {name!r} is not known and the header does not describe it")
    view_offset = _fget(h, "view offset (degrees)", 0.0)
    return Scanner(name=name or "interfile", num_rings=rings, num_detectors_per_ring=ndet,
                   inner_ring_radius=f32(diam * 10.0 / 2), average_depth_of_interaction=f32(max(doi, 0.0) * 10),
                   ring_spacing=f32(rs * 10.0), default_bin_size=f32(bs * 10), intrinsic_tilt=f32(view_offset * math.pi / 180),
                   max_num_non_arccorrected_bins=nonarc, default_num_arccorrected_bins=arcb,
                   timing_resolution_ps=max(tres, 0.0), size_of_timing_pos_ps=f32(max(tsize, 0.0)),
                   max_num_timing_poss=max(ntof, 0))


def _segment_sequence(min_rd: List[int], max_rd: List[int]) -> List[int]:
    """Segment number of each listed segment: listed segments sorted by
    min+max ring difference, numbered relative to the one holding 0."""
    sums = [mn + mx for mn, mx in zip(min_rd, max_rd)]
    order = sorted(range(len(sums)), key=lambda i: (sums[i], i))
    zero = next((r for r, i in enumerate(order) if abs(sums[i]) <= 1e-3), None)
    if zero is None:
        raise ValueError("projection data without segment 0")
    seg = [0] * len(sums)
    for r, i in enumerate(order):
        seg[i] = r - zero
    return seg


def read_proj_data_info(path: str) -> Tuple[ProjDataInfoCylindrical, Dict[str, Any]]:
    """ProjDataInfoCylindrical of an ``.hs`` header (and the parsed header)."""
    h = read_header(path)
    labels = {int(m.group(1)): str(v).strip().lower() for k, v in h.items()
              for m in [re.match(r"matrix axis label \[(\d)\]", k)] if m}
    sizes = {int(m.group(1)): v for k, v in h.items() for m in [re.match(r"matrix size \[(\d)\]", k)] if m}
    by_label = {lab: sizes[i] for i, lab in labels.items()}
    min_rd = h["minimum ring difference per segment"]
    max_rd = h["maximum ring difference per segment"]
    min_rd = min_rd if isinstance(min_rd, list) else [min_rd]
    max_rd = max_rd if isinstance(max_rd, list) else [max_rd]
    nax = by_label["axial coordinate"]
    nax = nax if isinstance(nax, list) else [nax]
    seq = _segment_sequence(min_rd, max_rd)
    sc = _scanner_from_header(h)
    order = sorted(range(len(seq)), key=lambda i: seq[i])
    arc = "arc correction" in str(h.get("applied corrections", "")).lower()
    bin_size = sc.default_bin_size
    if arc:
        # InterfileHeader.cxx:1413-1431
        eff = _fget(h, "effective central bin size (cm)", -1.0)
        if eff <= 0:
            eff = float(np.float32(np.float32(sc.default_bin_size) / np.float32(10)))
        bin_size = f32(eff * 10.0)
    tof = int(h.get("number of dimensions", 4)) > 4
    if tof and labels.get(5) != "timing positions":
        raise ValueError("TOF projection data need 'matrix axis label [5] := timing positions'")
    pdi = ProjDataInfoCylindrical(scanner=sc, min_ring_diff=tuple(int(min_rd[i]) for i in order),
                                  max_ring_diff=tuple(int(max_rd[i]) for i in order),
                                  num_axial_poss=tuple(int(nax[i]) for i in order), min_segment_num=seq[order[0]],
                                  num_views=int(by_label["view"]), num_tangential_poss=int(by_label["tangential coordinate"]),
                                  arc_corrected=arc, bin_size=bin_size)
    if tof:
        pdi = pdi.with_tof(int(h.get("tof mashing factor", 1)))
        if pdi.num_tof_bins != int(by_label["timing positions"]):
            raise ValueError("number of timing positions does not match the scanner and TOF mashing factor")
    h["_labels"], h["_segment_sequence"] = labels, seq
    return pdi, h


def _file_axes(labels: Dict[int, str]) -> str:
    mid = (labels.get(2), labels.get(3))
    if mid == ("view", "axial coordinate"):
        return "sinogram"   # [segment][axial][view][tang]
    if mid == ("axial coordinate", "view"):
        return "viewgram"   # [segment][view][axial][tang]
    raise ValueError(f"unsupported axis order {labels}")


def read_proj_data(path: str) -> Tuple[ProjDataInfoCylindrical, np.ndarray]:
    """(geometry, flat float32 array in ProjDataInMemory order) of an ``.hs`` file."""
    pdi, h = read_proj_data_info(path)
    layout = _file_axes(h["_labels"])
    seq = h["_segment_sequence"]
    nv, nt = pdi.num_views, pdi.num_tangential_poss
    block = pdi.num_bins // pdi.num_tof_bins
    offset = int(h.get("data offset in bytes [1]", 0))
    raw = np.fromfile(_data_path(path, h), dtype=_dtype(h), count=pdi.num_bins, offset=offset)
    if raw.size != pdi.num_bins:
        raise ValueError(f"{path}: data file is shorter than the header says")
    raw = _scaled(h, raw)
    out = np.empty(pdi.num_bins, np.float32)
    offs = pdi.segment_offsets()
    pos = 0
    for k in range(pdi.num_tof_bins):
        for s in seq:
            na = pdi.num_axial_poss_of(s)
            chunk = raw[pos:pos + na * nv * nt]
            pos += na * nv * nt
            if layout == "sinogram":
                sino = chunk.reshape(na, nv, nt)
            else:
                sino = chunk.reshape(nv, na, nt).transpose(1, 0, 2)
            out[k * block + offs[s]:k * block + offs[s] + na * nv * nt] = sino.reshape(-1)
    return pdi, out


def restrict_proj_data(pdi: ProjDataInfoCylindrical, data: np.ndarray, max_abs_segment: int):
    """(restricted geometry, its data): the segments |s| <= max_abs_segment of
    flat data in ProjDataInMemory order ("maximum absolute segment number to
    process")."""
    sub = pdi.restrict_segments(max_abs_segment)
    if sub == pdi:
        return pdi, data
    full = np.asarray(data, np.float32).reshape(pdi.num_tof_bins, -1)
    offs, subo = pdi.segment_offsets(), sub.segment_offsets()
    out = np.empty((sub.num_tof_bins, sub.num_bins // sub.num_tof_bins), np.float32)
    vt = pdi.num_views * pdi.num_tangential_poss
    for seg in sub.segments:
        n = pdi.num_axial_poss_of(seg) * vt
        out[:, subo[seg]:subo[seg] + n] = full[:, offs[seg]:offs[seg] + n]
    return sub, out.reshape(-1)


def write_proj_data(path: str, pdi: ProjDataInfoCylindrical, data: np.ndarray) -> None:
    """Write ``path`` (.hs) and its float32 little-endian data (.s) in
    ProjDataInMemory order (sinograms, segments 0, +1, -1, ...)."""
    a = np.ascontiguousarray(data, dtype="<f4").reshape(-1)
    if a.size != pdi.num_bins:
        raise ValueError("data size does not match the projection data")
    data_name = os.path.splitext(os.path.basename(path))[0] + ".s"
    a.tofile(os.path.join(os.path.dirname(os.path.abspath(path)), data_name))
    seq = pdi.segment_sequence()
    sc = pdi.scanner
    fmt = lambda v: "{" + ",".join(str(x) for x in v) + "}"
    i = lambda s: s - pdi.min_segment_num
    dims = 5 if pdi.is_tof_data else 4
    lines = ["!INTERFILE  :=", f"name of data file := {data_name}", f"originating system := {sc.name}",
             "!GENERAL DATA :=", "!GENERAL IMAGE DATA :=", "!type of data := PET",
             "imagedata byte order := LITTLEENDIAN", "!PET STUDY (General) :=", "!PET data type := Emission",
             "applied corrections := {arc correction}" if pdi.arc_corrected else "applied corrections := {None}",
             "!number format := float", "!number of bytes per pixel := 4", f"number of dimensions := {dims}",
             "matrix axis label [4] := segment", f"!matrix size [4] := {len(seq)}",
             "matrix axis label [2] := view", f"!matrix size [2] := {pdi.num_views}",
             "matrix axis label [3] := axial coordinate",
             f"!matrix size [3] := {fmt(pdi.num_axial_poss_of(s) for s in seq)}",
             "matrix axis label [1] := tangential coordinate", f"!matrix size [1] := {pdi.num_tangential_poss}"]
    if pdi.is_tof_data:
        lines += ["matrix axis label [5] := timing positions", f"!matrix size [5] := {pdi.num_tof_bins}",
                  f"TOF mashing factor := {pdi.tof_mash_factor}"]
    lines += [f"minimum ring difference per segment := {fmt(pdi.min_ring_diff[i(s)] for s in seq)}",
              f"maximum ring difference per segment := {fmt(pdi.max_ring_diff[i(s)] for s in seq)}",
              "Scanner parameters :=", f"Scanner type := {sc.name}",
              f"Number of rings := {sc.num_rings}", f"Number of detectors per ring := {sc.num_detectors_per_ring}",
              f"Inner ring diameter (cm) := {float(np.float32(sc.inner_ring_radius)) * 2 / 10!r}",
              f"Average depth of interaction (cm) := {float(np.float32(sc.average_depth_of_interaction)) / 10!r}",
              f"Distance between rings (cm) := {float(np.float32(sc.ring_spacing)) / 10!r}",
              f"Default bin size (cm) := {float(np.float32(sc.default_bin_size)) / 10!r}",
              f"View offset (degrees) := {float(np.float32(sc.intrinsic_tilt)) * 180 / math.pi!r}",
              f"Maximum number of non-arc-corrected bins := {sc.max_num_non_arccorrected_bins}",
              f"Default number of arc-corrected bins := {sc.default_num_arccorrected_bins}"]
    if sc.is_tof_ready:
        lines += [f"Maximum number of (unmashed) TOF time bins := {sc.max_num_timing_poss}",
                  f"Size of unmashed TOF time bins (ps) := {sc.size_of_timing_pos_ps!r}",
                  f"TOF timing resolution (ps) := {sc.timing_resolution_ps!r}"]
    lines += ["end scanner parameters :="]
    if pdi.arc_corrected:
        lines.append(f"effective central bin size (cm) := {float(np.float32(pdi.bin_size)) / 10!r}")
    lines += ["number of time frames := 1", "!END OF INTERFILE :="]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ---------------------------------------------------------------------------
# .par files
def ray_tracing_matrix_parameters(par: Dict[str, Any]) -> RayTracingMatrixParameters:
    """The ``Ray Tracing Matrix Parameters`` keys (ProjMatrixByBinUsingRayTracing.cxx:76-104)."""
    def flag(key, default):
        v = find_key(par, key, None)
        return default if v is None or v == "" else bool(int(v))

    n = find_key(par, "number of rays in tangential direction to trace for each bin", 1)
    return RayTracingMatrixParameters(
        restrict_to_cylindrical_FOV=flag("restrict to cylindrical FOV", True),
        num_tangential_LORs=int(n) if n != "" else 1,
        use_actual_detector_boundaries=flag("use actual detector boundaries", False),
        do_symmetry_90degrees_min_phi=flag("do_symmetry_90degrees_min_phi", True),
        do_symmetry_180degrees_min_phi=flag("do_symmetry_180degrees_min_phi", True),
        do_symmetry_swap_segment=flag("do_symmetry_swap_segment", True),
        do_symmetry_swap_s=flag("do_symmetry_swap_s", True),
        do_symmetry_shift_z=flag("do_symmetry_shift_z", True))


def reconstruction_setup(par_path: str, data_dir: Optional[str] = None):
    """Problem of an OSMAPOSL ``.par`` file: (ProjDataInfoCylindrical of its
    input file, restricted to "maximum absolute segment number to process";
    VoxelsOnCartesianGrid from the zooms, sizes and z offset; RayTracingMatrixParameters;
    dict of the reconstruction keys)."""
    par = parse_par(par_path)
    base = data_dir or os.path.dirname(os.path.abspath(par_path))
    input_file = str(find_key(par, "input file"))
    pdi, _ = read_proj_data_info(os.path.join(base, input_file))
    max_seg = find_key(par, "maximum absolute segment number to process", -1)
    if isinstance(max_seg, int) and max_seg >= 0:
        pdi = pdi.restrict_segments(max_seg)
    zoom = float(find_key(par, "zoom", 1.0))
    xy = int(find_key(par, "xy output image size (in pixels)", -1))
    img = VoxelsOnCartesianGrid.from_proj_data_info(pdi, zoom=zoom, xy_size=xy,
                                                    z_size=int(find_key(par, "z output image size (in pixels)", -1)),
                                                    zoom_z=float(find_key(par, "z zoom", 1.0)),
                                                    z_offset=float(find_key(par, "z offset (in mm)", 0.0)))
    keys = {
        "num_subsets": int(find_key(par, "number of subsets", 1)),
        "num_subiterations": int(find_key(par, "number of subiterations", 1)),
        "start_subset": int(find_key(par, "start at subset", 0)),
        "start_subiteration": int(find_key(par, "start at subiteration number", 1)),
        "zero_end_planes": bool(int(find_key(par, "zero end planes of segment 0", 0))),
        "use_subset_sensitivities": bool(int(find_key(par, "use subset sensitivities", 1))),
    }
    return pdi, img, ray_tracing_matrix_parameters(par), keys
