"""Benchmark of the MI355X ray-tracing-matrix projector pair at mMR size.

A step is one forward projection of the whole image into all 72.6M bins plus
one back projection of all bins (a projector pair), with image and
projection data resident in HBM (STIR layouts, torch device tensors) and the
matrix built beforehand. Prints one JSON line.

Also measured (extra fields): OSEM iterations/s (21 subsets), per-kernel
times from HIP events on the kernels' stream, the HBM roofline of the
dominant kernel, and the CPU oracle (OpenMP) on a bounded sample of the same
workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mMR-size fwd+back proj pairs/s, OSEM iter/s, autograd grad/s; HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0


TRAFFIC_FILE = "profiles/pair_mMR_pmc_traffic.txt"
SOURCE_GLOBS = ("stir_amd/csrc/*.hip", "stir_amd/csrc/*.cpp", "stir_amd/csrc/*.hpp", "include/*.h")


def source_digest():
    """SHA-256 over the HIP / C++ sources of the library (sorted paths and
    contents): the stamp that ties a counter summary to the kernels it measured."""
    import glob
    import hashlib

    h = hashlib.sha256()
    for pat in SOURCE_GLOBS:
        for path in sorted(glob.glob(os.path.join(ROOT, pat))):
            h.update(os.path.relpath(path, ROOT).encode())
            with open(path, "rb") as f:
                h.update(f.read())
    return h.hexdigest()


def measured_traffic(kernel):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary: (calibrated FETCH_SIZE + WRITE_SIZE bytes, raw counter bytes),
    or (None, None) when the summary is missing or was measured on other
    sources than these (its source digest differs)."""
    stamp, cal, raw = None, None, None
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            for line in f:
                parts = line.split()
                if parts[:2] == ["#", "source_digest"]:
                    stamp = parts[2]
                elif parts[:3] == ["#", "traffic_bytes_per_launch", kernel]:
                    cal = int(parts[3])
                elif parts[:3] == ["#", "raw_counter_bytes_per_launch", kernel]:
                    raw = int(parts[3])
    except (OSError, ValueError, IndexError):
        return None, None
    if stamp != source_digest():
        return None, None
    return cal, raw


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--osem-iterations", type=int, default=1)
    ap.add_argument("--skip-osem", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-autograd", action="store_true")
    ap.add_argument("--skip-tof", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-map", action="store_true")
    ap.add_argument("--skip-listmode", action="store_true")
    ap.add_argument("--skip-hessian", action="store_true")
    ap.add_argument("--tof-steps", type=int, default=3)
    ap.add_argument("--autograd-steps", type=int, default=3)
    ap.add_argument("--cpu-subsets", type=int, default=14,
                    help="subsets of 21 in the CPU baseline sample (about 10-15 s on 16 cores)")
    ap.add_argument("--radius", type=float, default=328.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB read + 256 MiB write element-wise kernel first (known traffic for rocprof counters)")
    ap.add_argument("--bp-quad", action="store_true",
                    help="pair: back-project over 2 x 2 column blocks (MI_TUNE_BP_QUAD; same sums, for comparison)")
    ap.add_argument("--bp-split", action="store_true",
                    help="pair: back-project all subsets in one launch per subset (MI_TUNE_BP_SPLIT_SUBSETS)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "--gpus must be 1",
                          "detail": "the projector pair runs on one GPU; sharding across GPUs is designed, not launched"}))
        sys.exit(2)

    import numpy as np
    import torch

    from stir_amd import _native, phantoms
    from stir_amd.geometry import RayTracingMatrixParameters, config1
    from stir_amd.projectors import (OSMAPOSLReconstruction, PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                                     ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing)

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        # one element-wise kernel that reads 256 MiB and writes 256 MiB: the
        # reference for scaling rocprof FETCH_SIZE / WRITE_SIZE to bytes
        src = torch.ones(64 * 2**20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.add(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    pdi, img = config1(args.radius)
    pm = RayTracingMatrixParameters()
    nsub = 21
    log(f"set-up: {pdi.num_bins} bins, image {img.shape}, {nsub} subsets")
    t0 = time.time()
    pair = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    if args.bp_quad:
        pair.context.set_tuning(_native.MI_TUNE_BP_QUAD, 1)
    if args.bp_split:
        pair.context.set_tuning(_native.MI_TUNE_BP_SPLIT_SUBSETS, 1)
    pair.set_up(pdi, img, num_subsets=nsub)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    info = pair.info()
    log(f"set-up {setup_s:.2f}s basic rows {info.num_basic_rows} elements {info.num_basic_elements} "
        f"nnz {info.num_nonzeros} table {info.num_table_entries} device {info.device_bytes / 2**30:.2f} GiB")

    # synthetic, seeded inputs resident in HBM
    x_true = phantoms.config1_phantom(img)
    x = torch.from_numpy(x_true).to(dev)
    sino = torch.empty(pdi.num_bins, dtype=torch.float32, device=dev)
    back = torch.empty(img.shape, dtype=torch.float32, device=dev)

    def step():
        pair.forward_project(x, out=sino)
        pair.back_project(sino, out=back)

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx = pair.context
    ctx.kernel_timing()  # reset
    ctx.set_kernel_timing(True)
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fp_ms, nfp, bp_ms, nbp = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    fp_avg = fp_ms / max(nfp, 1)
    bp_avg = bp_ms / max(nbp, 1)
    log(f"pair {ms_per_step:.3f} ms (fp kernel {fp_avg:.3f} ms, bp kernel {bp_avg:.3f} ms)")

    nnz = int(info.num_nonzeros)
    nbins = int(pdi.num_bins)
    nvox = int(img.num_voxels)
    dominant = "bp" if bp_avg >= fp_avg else "fp"
    dom_ms = bp_avg if dominant == "bp" else fp_avg
    kernel = ("bp_quad_kernel" if args.bp_quad else "bp_kernel") if dominant == "bp" else "fp_kernel"
    # HBM bytes per launch under this design: the matrix data the kernel
    # streams (back: table + offsets; forward: rows per task + records) plus
    # one read of the input and one write of the output (mi_info)
    design = {"fp": int(info.fp_design_bytes), "bp": int(info.bp_design_bytes)}
    achieved = design[dominant] / (dom_ms * 1e-3) / 1e9
    traffic, traffic_raw = measured_traffic(kernel)
    # the gather model: one 4-byte operand read per non-zero, each output
    # written once, each input streamed once; re-reads served by L2 / MALL
    # count as HBM bytes here, so it can exceed the HBM peak
    b_gather = 4 * nnz + 4 * nbins + 4 * nvox
    fp_gbs = b_gather / (fp_avg * 1e-3) / 1e9
    bp_gbs = b_gather / (bp_avg * 1e-3) / 1e9

    result = {
        "metric": METRIC,
        "value": round(1000.0 / ms_per_step, 4),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded ellipsoidal torso phantom (10 hot spheres, lung 0.2), Poisson data mean 2/bin + 0.05",
        "config": {
            "workload": "mMR-size projector pair: 64 rings x 504 crystals, R=%.0f mm, ring spacing 4.0625 mm, "
                        "span 11, MRD 60, 837 sinograms x 252 views x 344 tangential; image 344x344x127 at "
                        "2.08626x2.08626x2.03125 mm; ray-tracing matrix, all symmetries" % args.radius,
            "num_bins": nbins,
            "num_voxels": nvox,
            "nnz": nnz,
            "subsets": nsub,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": design[dominant],
            "bytes_model": "design bytes: matrix data streamed once + input read once + output written once",
            "traffic_raw_counter_bytes": traffic_raw,
            "traffic_source": TRAFFIC_FILE if traffic is not None else None,
        },
        "gather_model": {
            "bytes_per_launch": b_gather,
            "fp_GBps": round(fp_gbs, 1),
            "bp_GBps": round(bp_gbs, 1),
            "bp_frac_of_peak": round(bp_gbs / HBM_PEAK_GBS, 4),
            "note": "4 nnz + 4 N_bins + 4 N_vox per launch; counts cache-served re-reads",
        },
        "kernels": {
            "fp_kernel_ms": round(fp_avg, 4),
            "bp_kernel_ms": round(bp_avg, 4),
            "fp_design_GBps": round(design["fp"] / (fp_avg * 1e-3) / 1e9, 1),
            "bp_design_GBps": round(design["bp"] / (bp_avg * 1e-3) / 1e9, 1),
            "compulsory_bytes_per_projection": 4 * (nbins + nvox),
        },
        "source_digest": source_digest(),
        "setup_s": round(setup_s, 3),
    }

    obj = None
    if not (args.skip_osem and args.skip_autograd and args.skip_map and args.skip_hessian):
        fwd = pair.forward_project(x).cpu().numpy()
        y, scale = phantoms.poisson_data(fwd, 2.0, 0.05, seed=101)
        del fwd
        obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(pair)
        t0 = time.time()
        obj.set_up(y, 0.05)
        torch.cuda.synchronize()
        sens_s = time.time() - t0

    if not args.skip_osem:
        recon = OSMAPOSLReconstruction(obj, num_subiterations=nsub)
        obj.set_estimate(np.ones(img.shape, np.float32), enforce_positivity=True)
        # one warm-up subiteration, then timed full iterations
        recon.update_estimate(0, 1)
        torch.cuda.synchronize()
        ctx.kernel_timing()
        t0 = time.perf_counter()
        k = 2
        for _ in range(args.osem_iterations):
            for s in range(nsub):
                recon.update_estimate((k - 1) % nsub, k)
                k += 1
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        it_s = (t1 - t0) / args.osem_iterations
        result["osem"] = {"iter_per_s": round(1.0 / it_s, 4), "ms_per_iteration": round(it_s * 1000, 3),
                          "subsets": nsub, "sensitivity_s": round(sens_s, 3)}
        log(f"OSEM iteration {it_s * 1000:.2f} ms")

    if not args.skip_map:
        result["map"] = map_bench(obj, img, nsub)

    if not args.skip_autograd:
        result["autograd"] = autograd_bench(pair, obj, img, pdi, dev, args.autograd_steps, ms_per_step)

    if not args.skip_hessian:
        result["hessian"] = hessian_bench(obj, img, dev, ms_per_step)

    # the list-mode objective takes over the pair's context (its data become
    # the events' histogram): it runs after everything that uses the
    # projection-data objective
    if not args.skip_listmode:
        result["listmode"] = listmode_bench(pair, pdi, img, x, nsub)

    if not args.skip_tof:
        result["tof"] = tof_bench(dev, x, args.tof_steps, nsub, args.radius)

    if not args.skip_batch:
        result["batch"] = batch_bench(dev)

    if not args.skip_cpu:
        result["cpu_baseline"] = cpu_baseline(pdi, img, pm, x_true, nsub, args.cpu_subsets)
    print(json.dumps(result), flush=True)


def hessian_bench(obj, img, dev, pair_ms, steps=3):
    """The objective's Hessian-vector products over all subsets on device
    images (accumulate_Hessian_times_input_without_penalty: two forward
    projections, the per-viewgram quotient, one back projection; the
    approximate product: one forward projection fewer), after one warm-up."""
    import numpy as np
    import torch

    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.uniform(0.5, 1.5, img.shape).astype(np.float32)).to(dev)
    v = torch.from_numpy(rng.uniform(0.0, 1.0, img.shape).astype(np.float32)).to(dev)
    out = torch.zeros(img.shape, dtype=torch.float32, device=dev)
    res = {}
    for name, call in (("exact", lambda: obj.accumulate_Hessian_times_input_without_penalty(out, x, v)),
                       ("approximate", lambda: obj.add_multiplication_with_approximate_Hessian_without_penalty(out, v))):
        call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            call()
        torch.cuda.synchronize()
        res[f"{name}_ms"] = round((time.perf_counter() - t0) * 1000.0 / steps, 3)
    res["exact_vs_pair"] = round(res["exact_ms"] / pair_ms, 3)
    log(f"Hessian products: exact {res['exact_ms']:.2f} ms, approximate {res['approximate_ms']:.2f} ms")
    return res


def autograd_bench(pair, obj, img, pdi, dev, steps, pair_ms):
    """Gradient of NLL + 0.01 * QuadraticPrior for a batch of 2 log-normal(0, 0.5)
    images through the autograd layer, plus the FP64 adjoint identity."""
    import numpy as np
    import torch

    from stir_amd.autograd import PoissonLogLikelihoodLoss, QuadraticPriorLoss, back_project, forward_project

    rng = np.random.default_rng(202)
    xs = torch.from_numpy(rng.lognormal(0.0, 0.5, (2,) + tuple(img.shape)).astype(np.float32)).to(dev)
    nll = PoissonLogLikelihoodLoss(obj)
    prior = QuadraticPriorLoss(pair, 0.01)

    def grad_step():
        xg = xs.detach().requires_grad_(True)
        (nll(xg) + prior(xg)).sum().backward()
        return xg.grad

    grad_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        grad_step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    per_image_ms = dt * 1000 / xs.shape[0]
    log(f"autograd: {per_image_ms:.3f} ms per image gradient ({per_image_ms / pair_ms:.3f} x a pair)")
    g = torch.Generator(device=dev).manual_seed(7)
    xa = torch.rand(tuple(img.shape), generator=g, device=dev)
    pa = torch.rand(pdi.num_bins, generator=g, device=dev)
    with torch.no_grad():
        lhs = torch.dot(forward_project(xa, pair).double(), pa.double()).item()
        rhs = torch.dot(xa.reshape(-1).double(), back_project(pa, pair).reshape(-1).double()).item()
    return {"grad_per_s": round(1000.0 / per_image_ms, 4), "ms_per_gradient": round(per_image_ms, 3),
            "gradient_vs_pair": round(per_image_ms / pair_ms, 4), "batch": int(xs.shape[0]),
            "loss": "Poisson NLL (mean 2/bin, b=0.05) + 0.01 * QuadraticPrior, log-normal(0, 0.5) images",
            "adjoint_rel_err_fp64": abs(lhs - rhs) / abs(lhs)}


def tof_bench(dev, x, steps, nsub, radius):
    """TOF projector pairs at the mMR-like geometry with 11 TOF bins of 89 mm
    (400 ps FWHM), same phantom; one OSEM iteration on Poisson data of mean
    0.3/bin (+0.05)."""
    import numpy as np
    import torch

    from stir_amd import phantoms
    from stir_amd.geometry import RayTracingMatrixParameters, config3
    from stir_amd.projectors import (OSMAPOSLReconstruction, PoissonLogLikelihoodWithLinearModelForMeanAndProjData,
                                     ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing)

    pdi, img = config3(radius)
    t0 = time.time()
    pair = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(RayTracingMatrixParameters()))
    pair.set_up(pdi, img, num_subsets=nsub)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    info = pair.info()
    log(f"TOF set-up {setup_s:.2f}s: {pdi.num_bins} bins, table {info.num_table_entries}, "
        f"device {info.device_bytes / 2**30:.2f} GiB")
    sino = torch.empty(pdi.num_bins, dtype=torch.float32, device=dev)
    back = torch.empty(img.shape, dtype=torch.float32, device=dev)
    pair.forward_project(x, out=sino)
    pair.back_project(sino, out=back)
    torch.cuda.synchronize()
    ctx = pair.context
    ctx.kernel_timing()
    ctx.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        pair.forward_project(x, out=sino)
        pair.back_project(sino, out=back)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    fp_ms, nfp, bp_ms, nbp = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    out = {"pairs_per_s": round(1.0 / dt, 4), "ms_per_pair": round(dt * 1000, 3),
           "fp_kernel_ms": round(fp_ms / max(nfp, 1), 3), "bp_kernel_ms": round(bp_ms / max(nbp, 1), 3),
           "num_bins": int(pdi.num_bins), "tof_bins": int(pdi.num_tof_bins), "setup_s": round(setup_s, 3),
           "weights": ("stored windows of %d (forward) and %d (back) TOF bins" % (info.tof_fp_window, info.tof_bp_window)
                       if info.tof_fp_window else "computed on the fly"),
           "live_span": int(info.tof_live_span)}
    log(f"TOF pair {dt * 1000:.2f} ms (fp {out['fp_kernel_ms']} ms, bp {out['bp_kernel_ms']} ms)")
    y, _ = phantoms.poisson_data(sino.cpu().numpy(), 0.3, 0.05, seed=303)
    del sino, back
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(pair)
    obj.set_up(y, 0.05)
    del y
    recon = OSMAPOSLReconstruction(obj, num_subiterations=nsub)
    obj.set_estimate(np.ones(img.shape, np.float32), enforce_positivity=True)
    recon.update_estimate(0, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(2, nsub + 2):
        recon.update_estimate((k - 1) % nsub, k)
    torch.cuda.synchronize()
    it_s = time.perf_counter() - t0
    out["osem_ms_per_iteration"] = round(it_s * 1000, 3)
    log(f"TOF OSEM iteration {it_s * 1000:.2f} ms")
    pair.close()
    return out


def map_bench(obj, img, nsub):
    """One-step-late OSMAPOSL at mMR size: one full iteration (21 subsets)
    with the relative difference prior (beta 0.1, gamma 2, epsilon 1e-3),
    additive MAP model."""
    import numpy as np
    import torch

    from stir_amd.projectors import OSMAPOSLReconstruction, RelativeDifferencePrior

    obj.set_prior(RelativeDifferencePrior(0.1, gamma=2.0, epsilon=1e-3))
    recon = OSMAPOSLReconstruction(obj, num_subiterations=nsub)
    obj.set_estimate(np.ones(img.shape, np.float32), enforce_positivity=True)
    recon.update_estimate(0, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(2, nsub + 2):
        recon.update_estimate((k - 1) % nsub, k)
    torch.cuda.synchronize()
    it_s = time.perf_counter() - t0
    obj.set_prior(None)
    log(f"OSL (relative difference prior) iteration {it_s * 1000:.2f} ms")
    return {"prior": "relative difference (beta 0.1, gamma 2, epsilon 1e-3)", "ms_per_iteration": round(it_s * 1000, 3),
            "iter_per_s": round(1.0 / it_s, 4), "subsets": nsub}


def listmode_bench(pair, pdi, img, x, nsub):
    """List-mode OSMAPOSL at mMR size: Poisson events of mean 0.5 per bin of
    the phantom's forward projection (about 36 M prompts), additive 0.05;
    event hand-over (host histogram) and one full iteration (21 subsets)."""
    import numpy as np
    import torch

    from stir_amd import phantoms
    from stir_amd.projectors import (OSMAPOSLReconstruction,
                                     PoissonLogLikelihoodWithLinearModelForMeanAndListModeDataWithProjMatrixByBin)

    fwd = pair.forward_project(x).cpu().numpy()
    counts, _ = phantoms.poisson_data(fwd, 0.5, 0.0, seed=103)
    del fwd
    ev = pdi.events_from_counts(counts, seed=104)
    del counts
    lm = PoissonLogLikelihoodWithLinearModelForMeanAndListModeDataWithProjMatrixByBin(pair)
    t0 = time.perf_counter()
    lm.set_up(ev, 0.05, compute_sensitivities=False)
    handover_s = time.perf_counter() - t0
    lm.compute_sensitivities()
    nev = int(ev.shape[0])
    del ev
    recon = OSMAPOSLReconstruction(lm, num_subiterations=nsub)
    lm.set_estimate(np.ones(img.shape, np.float32), enforce_positivity=True)
    recon.update_estimate(0, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(2, nsub + 2):
        recon.update_estimate((k - 1) % nsub, k)
    torch.cuda.synchronize()
    it_s = time.perf_counter() - t0
    log(f"list mode: {nev} events, hand-over {handover_s:.2f} s, OSEM iteration {it_s * 1000:.2f} ms")
    return {"events": nev, "handover_s": round(handover_s, 3), "ms_per_iteration": round(it_s * 1000, 3),
            "events_per_s_per_iteration": round(nev / it_s, 1), "subsets": nsub}


def batch_bench(dev):
    """24 uneven OSEM problems (8 subsets x 2 iterations each): batched (tuned
    launch shape, HIP graph) against the same problems run one after the
    other through the per-problem path."""
    import numpy as np
    import torch

    from stir_amd import phantoms
    from stir_amd.batch import BatchedOSMAPOSL
    from stir_amd.geometry import config4_problems
    from stir_amd.projectors import ProjectorByBinPairMI355X, ProjMatrixByBinUsingRayTracing

    problems = config4_problems()
    nsub, nsubiter = 8, 16
    t0 = time.time()
    B = BatchedOSMAPOSL(problems, nsub)
    setup_s = time.time() - t0
    nbins = 0
    for i, (pdi, img, pm) in enumerate(problems):
        P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
        P.set_up(pdi, img, num_subsets=1)
        _, y, bkg = phantoms.config0_data(pdi, img, P.forward_project, seed=40 + i)
        P.close()
        B.set_data(i, y, bkg)
        nbins += pdi.num_bins
    log(f"batch: {len(problems)} problems, {nbins} bins, set-up {setup_s:.2f}s")

    def reset():
        for i, (pdi, img, pm) in enumerate(problems):
            B.set_estimate(i, np.ones(img.shape, np.float32))
        torch.cuda.synchronize()

    # sequential per-problem path (warm-up run first)
    reset()
    B.run_sequential(nsubiter)
    reset()
    t0 = time.perf_counter()
    B.run_sequential(nsubiter)
    seq_s = time.perf_counter() - t0
    seq_imgs = [B.get_estimate(i) for i in range(len(B))]
    choice = B.tune()
    reset()
    B.run(nsubiter, use_graph=True)  # capture + warm-up
    B.synchronize()
    reset()
    t0 = time.perf_counter()
    B.run(nsubiter, use_graph=True)
    B.synchronize()
    bat_s = time.perf_counter() - t0
    worst = 0.0
    for i in range(len(B)):
        a, b = seq_imgs[i], B.get_estimate(i)
        m = a > 1e-3 * a.max()
        worst = max(worst, float((np.abs(a - b)[m] / a[m]).max()))
    B.close()
    log(f"batch: sequential {seq_s * 1000:.1f} ms, batched {bat_s * 1000:.1f} ms ({seq_s / bat_s:.2f}x), "
        f"max rel image difference {worst:.2e}, tuner choice {choice}")
    return {"problems": len(problems), "subsets": nsub, "subiterations": nsubiter, "total_bins": int(nbins),
            "sequential_ms": round(seq_s * 1000, 3), "batched_ms": round(bat_s * 1000, 3),
            "speedup": round(seq_s / bat_s, 3), "reconstructions_per_s": round(len(problems) / bat_s, 3),
            "max_rel_image_diff": worst, "tuner_choice": choice}


def cpu_baseline(pdi, img, pm, x_true, nsub, nsample):
    """The oracle's OpenMP forward + back projection (per-thread images
    reduced serially, as the reference) on `nsample` of the 21 subsets,
    scaled to whole pairs/s."""
    import numpy as np

    from oracle.oracle import OraclePlan

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log(f"cpu baseline: oracle plan ({threads} threads)")
    t0 = time.time()
    plan = OraclePlan(pdi, img, pm)
    plan_s = time.time() - t0
    sino = np.zeros(pdi.num_bins, np.float32)
    t0 = time.time()
    for s in range(nsample):
        plan.forward_project(x_true, subset=s, num_subsets=nsub, omp=True, out=sino)
        plan.back_project(sino, subset=s, num_subsets=nsub, omp=True)
    dt = time.time() - t0
    frac = nsample / nsub
    pairs_per_s = frac / dt
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    log(f"cpu baseline {dt:.2f}s for {nsample}/{nsub} of a pair -> {pairs_per_s:.5f} pairs/s")
    return {"value": round(pairs_per_s, 6), "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"oracle OpenMP forward+back projection of {nsample} of {nsub} subsets "
                      f"({frac * 100:.1f}% of the bins), scaled to whole pairs; cache fill {plan_s:.1f}s excluded; "
                      f"CPU {cpu}"}


if __name__ == "__main__":
    main()
