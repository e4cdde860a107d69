"""Runs the host plan and the oracle, built with AddressSanitizer and UBSan,
on the workload geometries; started by tests/test_host_sanitizers.py in a
process of its own with the sanitizer runtime preloaded. Prints one line per
case and exits non-zero on any mismatch (the sanitizers abort on their own
findings)."""
import ctypes
import sys

import numpy as np

sys.path.insert(0, sys.argv[3])
from oracle import oracle as OR  # noqa: E402
from stir_amd.geometry import (RayTracingMatrixParameters, config0, config1, config3, config4_problems,  # noqa: E402
                               make_geometry_struct, rpt_sensitivity_case)


def main():
    plan_lib = ctypes.CDLL(sys.argv[1])
    plan_lib.plan_probe.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    cases = [("C0", *config0(), RayTracingMatrixParameters(), 4), ("C1", *config1(), RayTracingMatrixParameters(), 21),
             ("C3", *config3(), RayTracingMatrixParameters(), 21)]
    pdi, img, pm = rpt_sensitivity_case()
    cases.append(("RPT", pdi, img, pm, 1))
    for i, (pdi, img, pm) in enumerate(config4_problems()[:6]):
        cases.append((f"C4-{i}", pdi, img, pm, 8))
    for name, pdi, img, pm, nsub in cases:
        g = make_geometry_struct(pdi, img, pm)
        out = np.zeros(7, np.int64)
        err = ctypes.create_string_buffer(256)
        rc = plan_lib.plan_probe(ctypes.byref(g), nsub, out.ctypes.data, err, 256)
        if rc != 0:
            print(f"{name}: plan_probe failed {rc}: {err.value.decode()}")
            sys.exit(1)
        nb = pdi.num_bins // pdi.num_tof_bins
        ok = out[0] == nb and out[2] == out[6] and out[1] > 0 and out[4] >= out[2] // max(1, pdi.num_tangential_poss)
        print(f"{name}: bins {out[0]}, basic rays {out[1]}, families {out[2]}, related view-segments {out[4]}, "
              f"{'ok' if ok else 'MISMATCH'}", flush=True)
        if not ok:
            sys.exit(1)
    # the oracle on the 8-ring workload: rows, a subset's projections, the objective pieces
    lib = OR.load_library(sys.argv[2])
    pdi, img = config0()
    O = OR.OraclePlan(pdi, img, RayTracingMatrixParameters(), library=lib)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    y = rng.poisson(3.0, pdi.num_bins).astype(np.float32)
    for b in ((0, 0, 0, 0), (3, 95, 4, -64), (-3, 47, 0, 63)):
        O.row(*b)
    fp = O.forward_project(x, subset=1, num_subsets=4, out=np.zeros(pdi.num_bins, np.float32))
    bp = O.back_project(y, subset=2, num_subsets=4)
    g = O.subset_gradient(y, x, 3, 4, bkg=np.full(pdi.num_bins, 0.1, np.float32))
    ll = O.log_likelihood(y, x, bkg=np.full(pdi.num_bins, 0.1, np.float32))
    print(f"oracle: fp sum {fp.sum():.6e}, bp sum {bp.sum():.6e}, gradient sum {g.sum():.6e}, loglik {ll:.6e}", flush=True)
    if not (np.isfinite(fp).all() and np.isfinite(bp).all() and np.isfinite(g).all() and np.isfinite(ll)):
        sys.exit(1)


if __name__ == "__main__":
    main()
