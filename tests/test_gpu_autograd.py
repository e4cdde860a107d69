"""The differentiable layer (stir_amd.autograd) on the GPU.

Workload: the 8-ring x 192-crystal scanner (span 1, MRD 3, 96 views x 128
tangential), 128x128x15 image at 2 mm, 4 subsets; log-normal(0, 0.5)
images; Poisson data of the cylinder phantom (mean 5/bin, background 0.1).

Checks and bars:
  * the autograd Functions run the HIP pair: forward values and gradients
    bit-identical to direct projector calls (and batched == one by one);
  * adjoint identity <Ax, y> = <x, A^T y> with FP64 sums, rel <= 1e-5;
  * negative log-likelihood: value equal to the objective function (rel
    1e-12) and to the oracle (rel 1e-6); gradient bit-identical to
    compute_sub_gradient_without_penalty and within 1e-5 of the maximum of
    the oracle's;
  * FP64 finite differences (eps 1e-4, 64 random voxels) of NLL + quadratic
    prior (beta 0.01) against the autograd gradient, max|g|/1000 (the
    reference's ObjectiveFunctionTests recipe);
  * prior kernel: gradient bit-identical to the oracle, value rel 1e-12;
  * unrolled OSEM: images within 1e-4 relative of the oracle's OSEM; the
    hand-written vector-Jacobian product against torch autograd through
    the projection Functions (rel 1e-4) and against FP64 central
    differences of one subiteration computed with the oracle.
"""
import numpy as np
import pytest
import torch

from stir_amd import phantoms
from stir_amd.autograd import (PoissonLogLikelihoodLoss, QuadraticPriorLoss, UnrolledOSEM, back_project,
                               forward_project, osem_subiteration)
from stir_amd.geometry import RayTracingMatrixParameters, config0
from stir_amd.projectors import (PoissonLogLikelihoodWithLinearModelForMeanAndProjData, ProjectorByBinPairMI355X,
                                 ProjMatrixByBinUsingRayTracing)
from oracle.oracle import OraclePlan, quadratic_prior

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def c0():
    pdi, img = config0()
    pm = RayTracingMatrixParameters()
    print(f"[autograd c0] set-up: {pdi.num_bins} bins, image {img.shape}", flush=True)
    P = ProjectorByBinPairMI355X(ProjMatrixByBinUsingRayTracing(pm))
    P.set_up(pdi, img, num_subsets=4)
    O = OraclePlan(pdi, img, pm)
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project)
    obj = PoissonLogLikelihoodWithLinearModelForMeanAndProjData(P)
    obj.set_up(y, bkg)
    yield pdi, img, P, O, obj, y, bkg
    P.close()


def _lognormal(shape, seed):
    return np.random.default_rng(seed).lognormal(0.0, 0.5, shape).astype(np.float32)


def _t(a, grad=False):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV).requires_grad_(grad)


def test_projection_functions_are_the_hip_pair(c0):
    pdi, img, P, O, obj, y, bkg = c0
    rng = np.random.default_rng(11)
    x = _t(rng.uniform(0, 1, img.shape).astype(np.float32), grad=True)
    p = _t(rng.uniform(0, 1, pdi.num_bins).astype(np.float32))
    for subset in (-1, 2):
        x.grad = None
        q = forward_project(x, P, subset)
        assert torch.equal(q, P.forward_project(x.detach(), subset=subset))
        (q * p).sum().backward()
        assert torch.equal(x.grad, P.back_project(p, subset=subset))
    pr = p.clone().requires_grad_(True)
    xb = back_project(pr, P)
    assert torch.equal(xb, P.back_project(p))
    (xb * x.detach()).sum().backward()
    assert torch.equal(pr.grad, P.forward_project(x.detach()))


def test_batched_matches_one_by_one(c0):
    pdi, img, P, O, obj, y, bkg = c0
    xs = _t(np.stack([_lognormal(img.shape, 1), _lognormal(img.shape, 2)]), grad=True)
    q = forward_project(xs, P)
    assert q.shape == (2, pdi.num_bins)
    for b in range(2):
        assert torch.equal(q[b], P.forward_project(xs[b].detach()))
    loss = PoissonLogLikelihoodLoss(obj)
    prior = QuadraticPriorLoss(P, 0.01)
    total = loss(xs) + prior(xs)
    assert total.shape == (2,) and total.dtype == torch.float64
    total.sum().backward()
    for b in range(2):
        xb = xs[b].detach().clone().requires_grad_(True)
        (loss(xb) + prior(xb)).backward()
        assert torch.equal(xs.grad[b], xb.grad)


def test_adjoint_identity_fp64(c0):
    pdi, img, P, O, obj, y, bkg = c0
    g = torch.Generator(device=DEV).manual_seed(7)
    x = torch.rand(img.shape, generator=g, device=DEV)
    p = torch.rand(pdi.num_bins, generator=g, device=DEV)
    lhs = torch.dot(forward_project(x, P).double(), p.double()).item()
    rhs = torch.dot(x.reshape(-1).double(), back_project(p, P).reshape(-1).double()).item()
    print(f"adjoint rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_neg_log_likelihood_value_and_gradient(c0):
    pdi, img, P, O, obj, y, bkg = c0
    xn = _lognormal(img.shape, 4)
    bfull = np.full(pdi.num_bins, bkg, np.float32)
    for subset in (-1, 1):
        x = _t(xn, grad=True)
        L = PoissonLogLikelihoodLoss(obj, subset)(x)
        L.backward()
        ref_value = obj.compute_objective_function(xn, subset=subset)
        assert abs(-L.item() - ref_value) <= 1e-12 * abs(ref_value)
        o_value = O.log_likelihood(y, xn, bkg=bfull, subset=subset, num_subsets=4)
        assert abs(-L.item() - o_value) <= 1e-6 * abs(o_value)
        g = (-x.grad).cpu().numpy()
        g_api = obj.compute_sub_gradient_without_penalty(xn, subset=subset)
        assert np.array_equal(g, g_api)
        if subset >= 0:
            g_o = O.subset_gradient(y, xn, subset, 4, bkg=bfull)
            assert np.abs(g - g_o).max() <= 1e-5 * np.abs(g_o).max()
    # value only (no graph): same number
    with torch.no_grad():
        v = PoissonLogLikelihoodLoss(obj)(_t(xn))
    assert abs(-v.item() - obj.compute_objective_function(xn)) <= 1e-12 * abs(v.item())


def test_finite_differences_fp64(c0):
    """Autograd gradient of NLL + 0.01 * QuadraticPrior against FP64 central
    differences (eps 1e-4) at 64 random voxels inside the field of view,
    the model evaluated with the oracle's FP64 forward projector."""
    pdi, img, P, O, obj, y, bkg = c0
    beta, eps = 0.01, 1e-4
    xn = _lognormal(img.shape, 9)
    x = _t(xn, grad=True)
    (PoissonLogLikelihoodLoss(obj)(x) + QuadraticPriorLoss(P, beta)(x)).backward()
    g = x.grad.cpu().numpy()
    x64 = xn.astype(np.float64)
    ybar = O.forward_project_f64(x64) + bkg
    yd = y.astype(np.float64)
    sens = O.subset_sensitivity(-1, 1)
    rng = np.random.default_rng(21)
    cand = np.flatnonzero(sens.reshape(-1) > 0.05 * sens.max())
    vox_zyx = (img.vox_z, img.vox_y, img.vox_x)
    worst = 0.0
    tol = np.abs(g).max() / 1000
    for j in rng.choice(cand, 64, replace=False):
        e = np.zeros(img.num_voxels)
        e[j] = 1.0
        a = O.forward_project_f64(e.reshape(img.shape))
        nzb = np.flatnonzero(a)
        yp, ym, yb = ybar[nzb] + eps * a[nzb], ybar[nzb] - eps * a[nzb], yd[nzb]
        d_nll = np.sum((yp - ym) - yb * (np.log(yp) - np.log(ym)))
        # prior difference from its local neighbourhood (FP64, pair-wise)
        z, yy, xx = np.unravel_index(j, img.shape)
        zs, ys, xs = (slice(max(c - 1, 0), c + 2) for c in (z, yy, xx))
        patch = x64[zs, ys, xs].copy()
        c = (z - zs.start, yy - ys.start, xx - xs.start)
        w = _prior_weights64(vox_zyx)
        d_pr = 0.0
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    n = (c[0] + dz, c[1] + dy, c[2] + dx)
                    if (dz, dy, dx) == (0, 0, 0) or not all(0 <= n[i] < patch.shape[i] for i in range(3)):
                        continue
                    wv = w[dz + 1, dy + 1, dx + 1]
                    d0 = patch[c] - patch[n]
                    # the pair appears twice in the sum: 2 * w * d^2 / 4
                    d_pr += wv * ((d0 + eps) ** 2 - (d0 - eps) ** 2) / 2
        fd = (d_nll + beta * d_pr) / (2 * eps)
        gj = g.reshape(-1)[j]
        worst = max(worst, abs(fd - gj) / max(abs(fd), 1e-12))
        assert abs(fd - gj) <= tol, (j, fd, gj)
    print(f"FD worst rel err {worst:.2e} (tolerance {tol:.3e} abs)")


def _prior_weights64(vox_zyx):
    vz, vy, vx = vox_zyx
    w = np.zeros((3, 3, 3))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if (dz, dy, dx) != (0, 0, 0):
                    w[dz + 1, dy + 1, dx + 1] = vx / np.sqrt((dx * vx) ** 2 + (dy * vy) ** 2 + (dz * vz) ** 2)
    return w


@pytest.mark.parametrize("variant", ["plain", "kappa", "only_2D", "weights", "beta0"])
def test_prior_kernel_matches_oracle(c0, variant):
    pdi, img, P, O, obj, y, bkg = c0
    xn = _lognormal(img.shape, 13)
    kappa = np.random.default_rng(14).uniform(0.5, 1.5, img.shape).astype(np.float32) if variant == "kappa" else None
    beta = 0.0 if variant == "beta0" else 0.01
    weights = None
    vox = (img.vox_z, img.vox_y, img.vox_x)
    if variant == "weights":
        weights = np.random.default_rng(15).uniform(0, 2, 27).astype(np.float32)
        weights[13] = 0
    prior = QuadraticPriorLoss(P, beta, kappa=None if kappa is None else _t(kappa), only_2D=variant == "only_2D",
                               weights=weights)
    x = _t(xn, grad=True)
    v = prior(x)
    v.backward()
    if weights is None:
        vo, go = quadratic_prior(xn, vox, beta, kappa=kappa, only_2D=variant == "only_2D")
    else:
        from oracle import oracle as OR
        import ctypes
        go = np.zeros(img.shape, np.float32)
        vo = OR.lib().or_quadratic_prior(xn.reshape(-1), None, *img.shape, weights, 0, beta,
                                         go.ctypes.data_as(ctypes.c_void_p))
    assert abs(v.item() - vo) <= 1e-12 * max(abs(vo), 1e-300)
    assert np.array_equal(x.grad.cpu().numpy(), go)


def test_unrolled_osem_matches_oracle_osem(c0):
    pdi, img, P, O, obj, y, bkg = c0
    obj.set_up(y, bkg)
    x0 = _t(np.ones(img.shape, np.float32))
    with torch.no_grad():
        x8 = UnrolledOSEM(obj, 8)(x0).cpu().numpy()
    ref = O.osem(y, np.ones(img.shape, np.float32), 4, 8, bkg=np.full(pdi.num_bins, bkg, np.float32))
    mask = ref > 1e-3 * ref.max()
    rel = (np.abs(x8 - ref)[mask] / ref[mask]).max()
    print(f"unrolled OSEM max rel err {rel:.3e}")
    assert rel <= 1e-4


def _torch_subiteration(x, P, y_t, bkg, sens, subset):
    """The same OSEM subiteration through the projection Functions and torch ops."""
    ybar = forward_project(x, P, subset) + bkg
    r = torch.where(y_t > 0, y_t / ybar, torch.zeros_like(ybar))
    u = back_project(r, P, subset)
    return x * torch.where(sens > 0, u / torch.where(sens > 0, sens, torch.ones_like(sens)), torch.zeros_like(u))


@pytest.mark.parametrize("num_subiterations", [3, 10])
def test_osem_vjp_matches_torch_autograd(c0, num_subiterations):
    """The unrolled map's hand-written VJP (image checkpoints, recomputed
    forward projections) against torch autograd through the projection
    Functions, over 3 and over 10 subiterations (2.5 passes over the 4
    subsets)."""
    pdi, img, P, O, obj, y, bkg = c0
    obj.set_up(y, bkg)
    y_t = _t(y)
    rng = np.random.default_rng(17)
    x0n = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    wimg = _t(rng.uniform(-1, 1, img.shape).astype(np.float32))
    x = _t(x0n, grad=True)
    out = UnrolledOSEM(obj, num_subiterations)(x)
    (out * wimg).sum().backward()
    g_hand = x.grad.clone()
    xr = _t(x0n, grad=True)
    xc = xr
    ones = torch.ones(pdi.num_bins, device=DEV)
    for k in range(1, num_subiterations + 1):
        s = (k - 1) % 4
        sens = P.back_project(ones, subset=s).clone()
        xc = _torch_subiteration(xc, P, y_t, bkg, sens, s)
    assert torch.allclose(xc.detach(), out.detach(), rtol=1e-5, atol=1e-6 * out.abs().max().item())
    (xc * wimg).sum().backward()
    err = (g_hand - xr.grad).abs().max().item() / xr.grad.abs().max().item()
    print(f"OSEM VJP ({num_subiterations} subiterations) vs torch autograd: max err / max {err:.2e}")
    assert err <= 1e-4


def test_osem_vjp_fp64_central_difference(c0):
    """<VJP(g), v> against <g, (F(x + h v) - F(x - h v)) / 2h> for one
    subiteration F evaluated in FP64 with the oracle's projectors."""
    pdi, img, P, O, obj, y, bkg = c0
    obj.set_up(y, bkg)
    rng = np.random.default_rng(19)
    x0 = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    gimg = rng.uniform(-1, 1, img.shape)
    v = rng.uniform(-1, 1, img.shape)
    subset = 2
    x = _t(x0, grad=True)
    osem_subiteration(x, obj, subset, 2).backward(_t(gimg.astype(np.float32)))
    lhs = float(np.sum(x.grad.cpu().numpy().astype(np.float64) * v))
    yd = y.astype(np.float64)
    sens = O.back_project_f64(np.ones(pdi.num_bins), subset=subset, num_subsets=4)

    def F(xx):
        ybar = O.forward_project_f64(xx, subset=subset, num_subsets=4) + bkg
        r = np.where(yd > 0, yd / ybar, 0.0)
        u = O.back_project_f64(r, subset=subset, num_subsets=4)
        return xx * np.where(sens > 0, u / np.where(sens > 0, sens, 1.0), 0.0)

    h = 1e-3
    xd = x0.astype(np.float64)
    rhs = float(np.sum(gimg * (F(xd + h * v) - F(xd - h * v)) / (2 * h)))
    print(f"OSEM VJP vs FP64 central difference: {lhs:.8e} vs {rhs:.8e}")
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)


def test_osem_step_and_vjp_with_relative_change_bounds(c0):
    """Bounds 0.9 / 1.1 at subiteration 2 (they act after subiteration 1):
    the out-of-place step equals mi_osem_subset_step bit for bit, clamps
    some voxels at each bound, and its VJP (zero through the clamp where it
    acts) matches FP64 central differences of the clamped subiteration."""
    from stir_amd.projectors import OSMAPOSLReconstruction

    pdi, img, P, O, obj, y, bkg = c0
    obj.set_up(y, bkg)
    lo, hi, subset = 0.9, 1.1, 1
    rng = np.random.default_rng(23)
    x0 = rng.uniform(0.5, 1.5, img.shape).astype(np.float32)
    # scaled so that A x0 + b carries the data's counts: the update factors
    # then fall on both sides of 1 and both bounds act
    x0 *= np.float32((float(y.sum()) - bkg * y.size) / float(P.forward_project(x0).astype(np.float64).sum()))
    obj.set_estimate(x0)
    OSMAPOSLReconstruction(obj, 1, minimum_relative_change=lo, maximum_relative_change=hi).update_estimate(subset, 2)
    in_place = obj.get_estimate()
    x = _t(x0, grad=True)
    out = osem_subiteration(x, obj, subset, 2, lo, hi)
    assert np.array_equal(out.detach().cpu().numpy(), in_place)
    u = in_place / x0
    print(f"clamped at lo {np.mean(u <= np.float32(lo) * 1.0000001):.3f}, at hi {np.mean(u >= np.float32(hi) * 0.9999999):.3f}")
    assert np.mean(np.isclose(u, lo, rtol=1e-6)) > 0.01 and np.mean(np.isclose(u, hi, rtol=1e-6)) > 0.01
    # subiteration 1: no clamp
    out1 = osem_subiteration(_t(x0), obj, subset, 1, lo, hi).cpu().numpy()
    assert np.mean(np.abs(out1 / x0 - 1) > 0.11) > 0.01
    gimg = rng.uniform(-1, 1, img.shape)
    v = rng.uniform(-1, 1, img.shape)
    out.backward(_t(gimg.astype(np.float32)))
    lhs = float(np.sum(x.grad.cpu().numpy().astype(np.float64) * v))
    yd = y.astype(np.float64)
    sens = O.back_project_f64(np.ones(pdi.num_bins), subset=subset, num_subsets=4)

    def F(xx):
        ybar = O.forward_project_f64(xx, subset=subset, num_subsets=4) + bkg
        r = np.where(yd > 0, yd / ybar, 0.0)
        f = O.back_project_f64(r, subset=subset, num_subsets=4)
        f = np.where(sens > 0, f / np.where(sens > 0, sens, 1.0), 0.0)
        return xx * np.clip(f, lo, hi)

    h = 1e-5
    xd = x0.astype(np.float64)
    rhs = float(np.sum(gimg * (F(xd + h * v) - F(xd - h * v)) / (2 * h)))
    print(f"clamped OSEM VJP vs FP64 central difference: {lhs:.8e} vs {rhs:.8e}")
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)
