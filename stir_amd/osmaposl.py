"""Run an OSMAPOSL parameter file with the MI355X projector pair.

``python -m stir_amd.osmaposl recon.par`` mirrors the reference's ``OSMAPOSL``
utility (``OSMAPOSL.cxx``, ``OSMAPOSLReconstruction.cxx``,
``IterativeReconstruction.cxx``) for the keys of this path:

* objective ``PoissonLogLikelihoodWithLinearModelForMeanAndProjData``:
  ``input file``, ``maximum absolute segment number to process``, ``zero end
  planes of segment 0``, ``additive sinogram``, ``Bin Normalisation type`` (``None``
  or ``From ProjData``: the factors' reciprocals are the model's multiplicative
  factors), ``use_subset_sensitivities``, ``sensitivity filename`` (written),
  ``use time-of-flight sensitivities``, ``prior type`` (``None``, ``Quadratic``,
  ``Relative Difference Prior`` with ``penalisation factor``, ``only 2D``,
  ``kappa filename``, ``weights``, ``gamma value``, ``epsilon value``);
* projector pair ``Matrix`` with ``Matrix type := Ray Tracing`` or ``Separate
  Projectors`` with the ``MI355X Ray Tracing`` forward and back projectors, and
  the ``Ray Tracing Matrix Parameters`` keys;
* image: ``zoom``, ``Z zoom``, ``XY output image size (in pixels)``, ``Z output
  image size (in pixels)``, ``Z offset (in mm)``;
* reconstruction: ``number of subsets``, ``number of subiterations`` (the last
  subiteration number), ``start at subiteration number``, ``start at subset``,
  ``save estimates at subiteration intervals``, ``initial estimate``, ``enforce
  initial positivity condition``, ``MAP_model``, ``maximum relative change``,
  ``minimum relative change``, ``output filename prefix``, ``disable output``,
  ``uniformly randomise subset order`` (a new random permutation of the
  subsets every cycle, IterativeReconstruction.cxx:575-638).

Keys this path does not implement (filters, other projectors, normalisations
or priors) raise an error instead of being ignored.
Sensitivities are always computed.
"""
from __future__ import annotations

import os
import sys
from typing import Any, Dict, Optional, Tuple

import numpy as np

from . import interfile as I
from .geometry import VoxelsOnCartesianGrid

REGISTERED_PROJECTOR = "MI355X Ray Tracing"


def _get(par: Dict[str, Any], key: str, default=None):
    v = I.find_key(par, key, None)
    return default if v is None or v == "" else v


def _truthy(v) -> bool:
    return bool(int(v)) if not isinstance(v, str) else v.strip() not in ("", "0")


def _check_projector(par: Dict[str, Any]) -> None:
    pair = str(_get(par, "projector pair type", "Matrix")).strip()
    if pair.lower() == "matrix":
        mt = str(_get(par, "matrix type", "Ray Tracing")).strip()
        if mt.lower() != "ray tracing":
            raise ValueError(f"Matrix type {mt!r} is not supported (only Ray Tracing)")
    elif pair.lower() == "separate projectors":
        for key in ("forward projector type", "back projector type"):
            t = str(_get(par, key, "")).strip()
            if t.lower() != REGISTERED_PROJECTOR.lower():
                raise ValueError(f"{key} {t!r} is not supported (only {REGISTERED_PROJECTOR})")
    else:
        raise ValueError(f"Projector pair type {pair!r} is not supported")


def _unsupported(par: Dict[str, Any]) -> None:
    for key in ("inter-update filter type", "inter-iteration filter type", "post-filter type"):
        v = _get(par, key)
        if v is not None and str(v).strip().lower() not in ("", "none"):
            raise ValueError(f"{key} := {v} is not supported")


def _path(base: str, name) -> str:
    name = str(name).strip()
    return name if os.path.isabs(name) else os.path.join(base, name)


def _weights(v) -> Optional[np.ndarray]:
    if v is None:
        return None
    w = np.asarray(v, np.float32)
    if w.size != 27:
        raise ValueError("prior weights must be 3x3x3")
    return w.reshape(27)


def _prior(par: Dict[str, Any], base: str, shape):
    from .projectors import QuadraticPrior, RelativeDifferencePrior

    kind = str(_get(par, "prior type", "None")).strip().lower()
    if kind in ("", "none"):
        return None
    beta = float(_get(par, "penalisation factor", 0.0))
    only_2D = _truthy(_get(par, "only 2D", 0))
    w = _weights(_get(par, "weights"))
    kf = _get(par, "kappa filename")
    kappa = None
    if kf is not None:
        _, kappa = I.read_image(_path(base, kf))
        if kappa.shape != tuple(shape):
            raise ValueError("kappa image does not match the reconstructed image")
    if kind == "quadratic":
        return QuadraticPrior(beta, only_2D=only_2D, weights=w, kappa=kappa)
    if kind == "relative difference prior":
        return RelativeDifferencePrior(beta, gamma=float(_get(par, "gamma value", 2.0)),
                                       epsilon=float(_get(par, "epsilon value", 0.0)), only_2D=only_2D, weights=w,
                                       kappa=kappa)
    raise ValueError(f"prior type {kind!r} is not supported")


def _multiplicative(par: Dict[str, Any], base: str, pdi, max_seg) -> Optional[np.ndarray]:
    kind = str(_get(par, "bin normalisation type", "None")).strip().lower()
    if kind in ("", "none"):
        return None
    if kind != "from projdata":
        raise ValueError(f"Bin Normalisation type {kind!r} is not supported")
    npdi, n = I.read_proj_data(_path(base, _get(par, "normalisation projdata filename")))
    if max_seg is not None:
        npdi, n = I.restrict_proj_data(npdi, n, max_seg)
    if npdi.num_axial_poss != pdi.num_axial_poss or npdi.num_views != pdi.num_views or \
            npdi.num_tangential_poss != pdi.num_tangential_poss:
        raise ValueError("normalisation data do not match the emission data")
    if pdi.num_tof_bins > 1 and npdi.num_tof_bins == 1:
        n = np.tile(n, pdi.num_tof_bins)  # non-TOF factors for every TOF bin (BinNormalisationFromProjData.cxx:89-110)
    # undo() of ones: 1 / factor (BinNormalisationFromProjData.cxx:142-150)
    with np.errstate(divide="ignore"):
        return (np.float32(1) / n.astype(np.float32)).astype(np.float32)


def set_up_from_par(par_path: str, device: int = 0):
    """Parse an OSMAPOSL .par file and set up the projector pair, the objective
    function and the reconstruction. Returns a dict with ``projector``,
    ``objective``, ``reconstruction``, ``image`` (geometry), ``initial``
    (initial image), ``keys`` (reconstruction keys) and ``par`` (parse tree)."""
    from .projectors import (OSMAPOSLReconstruction, PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                             ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing, FLT_MAX)

    par = I.parse_par(par_path)
    base = os.path.dirname(os.path.abspath(par_path))
    if "osmaposlparameters" not in par:
        raise ValueError("not an OSMAPOSL parameter file (no OSMAPOSLParameters block)")
    obj_type = str(_get(par, "objective function type", "")).strip()
    if obj_type != "PoissonLogLikelihoodWithLinearModelForMeanAndProjData":
        raise ValueError(f"objective function type {obj_type!r} is not supported")
    _check_projector(par)
    _unsupported(par)
    pm = I.ray_tracing_matrix_parameters(par)
    input_path = _path(base, _get(par, "input file"))
    pdi_full, _ = I.read_proj_data_info(input_path)
    max_seg = _get(par, "maximum absolute segment number to process")
    max_seg = None if max_seg is None or int(max_seg) < 0 else int(max_seg)
    pdi = pdi_full if max_seg is None else pdi_full.restrict_segments(max_seg)
    img = VoxelsOnCartesianGrid.from_proj_data_info(pdi, zoom=float(_get(par, "zoom", 1.0)),
                                                    xy_size=int(_get(par, "xy output image size (in pixels)", -1)),
                                                    z_size=int(_get(par, "z output image size (in pixels)", -1)),
                                                    zoom_z=float(_get(par, "z zoom", 1.0)),
                                                    z_offset=float(_get(par, "z offset (in mm)", 0.0)))
    nsub = int(_get(par, "number of subsets", 1))
    keys = {
        "num_subsets": nsub,
        "num_subiterations": int(_get(par, "number of subiterations", 1)),
        "start_subiteration": int(_get(par, "start at subiteration number", 1)),
        "start_subset": int(_get(par, "start at subset", 0)),
        "save_interval": int(_get(par, "save estimates at subiteration intervals", 1)),
        "output_prefix": str(_get(par, "output filename prefix", "")).strip(),
        "disable_output": _truthy(_get(par, "disable output", 0)),
        "sensitivity_filename": _get(par, "sensitivity filename"),
        "zero_end_planes": _truthy(_get(par, "zero end planes of segment 0", 0)),
    }
    # measured data (only read when iterations are run: a sensitivity-only file
    # needs the geometry alone)
    n_updates = keys["num_subiterations"] - keys["start_subiteration"] + 1
    if n_updates > 0:
        _, y = I.read_proj_data(input_path)
        if max_seg is not None:
            _, y = I.restrict_proj_data(pdi_full, y, max_seg)
    else:
        y = np.zeros(pdi.num_bins, np.float32)
    additive = 0.0
    add_file = _get(par, "additive sinogram")
    if add_file is not None and str(add_file).strip() not in ("", "0"):
        apdi, additive = I.read_proj_data(_path(base, add_file))
        if max_seg is not None:
            _, additive = I.restrict_proj_data(apdi, additive, max_seg)
    mult = _multiplicative(par, base, pdi, max_seg)

    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm), device=device)
    P.set_up(pdi, img, num_subsets=nsub)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, additive, zero_end_planes=keys["zero_end_planes"], multiplicative=mult,
               use_subset_sensitivities=_truthy(_get(par, "use_subset_sensitivities", 1)),
               use_tof_sensitivities=_truthy(_get(par, "use time-of-flight sensitivities", 0)))
    obj.set_prior(_prior(par, base, img.shape))
    recon = OSMAPOSLReconstruction(obj, keys["num_subiterations"], start_subset=keys["start_subset"],
                                   enforce_initial_positivity=_truthy(_get(par, "enforce initial positivity condition", 1)),
                                   MAP_model=str(_get(par, "map_model", "additive")).strip(),
                                   minimum_relative_change=float(_get(par, "minimum relative change", 0.0)),
                                   maximum_relative_change=float(_get(par, "maximum relative change", FLT_MAX)),
                                   start_subiteration_num=keys["start_subiteration"],
                                   uniformly_randomise_subset_order=_truthy(_get(par, "uniformly randomise subset order",
                                                                                 0)))
    init_file = _get(par, "initial estimate")
    if init_file is not None and str(init_file).strip():
        iimg, initial = I.read_image(_path(base, init_file))
        if iimg.shape != img.shape:
            raise ValueError("initial estimate does not match the image size")
    else:
        initial = np.ones(img.shape, np.float32)
    return {"projector": P, "objective": obj, "reconstruction": recon, "image": img, "initial": initial,
            "keys": keys, "par": par, "proj_data_info": pdi}


def run_par(par_path: str, output_dir: Optional[str] = None, device: int = 0) -> Tuple[np.ndarray, Dict[str, Any]]:
    """Reconstruct as the reference's OSMAPOSL utility does: sensitivities
    (written to "sensitivity filename" if given, total sensitivity), then
    subiterations start..number of subiterations, saving
    <prefix>_<subiteration>.hv every "save estimates at subiteration
    intervals" and at the last one (IterativeReconstruction.cxx:334-346,
    427-433,568-572). Returns the final estimate and the set-up dict."""
    s = set_up_from_par(par_path, device)
    out = output_dir or os.getcwd()
    k, obj, recon, img = s["keys"], s["objective"], s["reconstruction"], s["image"]
    if k["sensitivity_filename"] is not None and not k["disable_output"]:
        I.write_image(_path(out, _with_ext(k["sensitivity_filename"])), img, obj.get_sensitivity())
    x = s["initial"]
    recon._set_options()
    obj.set_estimate(x, enforce_positivity=recon.enforce_initial_positivity)
    nsub = k["num_subsets"]
    last = k["num_subiterations"]
    for it in range(k["start_subiteration"], last + 1):
        recon.update_estimate(recon.get_subset_num(it), it)
        if (it % k["save_interval"] == 0 or it == last) and not k["disable_output"] and k["output_prefix"]:
            I.write_image(os.path.join(out, f"{k['output_prefix']}_{it}.hv"), img, obj.get_estimate())
    return obj.get_estimate(), s


def _with_ext(name) -> str:
    name = str(name).strip()
    return name if name.endswith(".hv") else os.path.splitext(name)[0] + ".hv"


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m stir_amd.osmaposl file.par", file=sys.stderr)
        return 1
    x, s = run_par(argv[0])
    print(f"OSMAPOSL: image {s['image'].shape}, max {float(x.max()):.6g}")
    s["projector"].close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
