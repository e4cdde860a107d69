"""Oracle pins for the objective function's multiplicative factors, the total
sensitivity mode, the relative difference prior and the one-step-late MAP
update of OSMAPOSL (CPU).

Pins: hand-computed values of the relative difference prior on two voxels;
its gradient against central differences of its value; its large-epsilon
limit against the quadratic prior (pinned elsewhere); exact power-of-two
scale equivariance of the model ybar = m (A x + b) in OSEM, the log-likelihood
and the gradient (a factor applied in the wrong place breaks it); the
log-likelihood gradient with random factors against a directional central
difference; the total-sensitivity mode against the pinned subset
sensitivities; and the one-step-late fixed point: after many iterations the
penalised gradient BP(y/ybar) - sens - beta grad R vanishes on the object,
while the same residual with the prior's sign flipped does not.
"""
import numpy as np
import pytest

from oracle.oracle import OraclePlan, quadratic_prior, relative_difference_prior
from stir_amd import phantoms
from stir_amd.geometry import (ProjDataInfoCylindrical, RayTracingMatrixParameters, Scanner, VoxelsOnCartesianGrid)

VOX = (2.0, 2.0, 2.0)


@pytest.fixture(scope="module")
def tiny():
    sc = Scanner.cylindrical(4, 64, 80.0, 4.0)
    pdi = ProjDataInfoCylindrical.construct(sc, span=1, max_delta=3, num_views=32, num_tangential_poss=32)
    img = VoxelsOnCartesianGrid(7, 32, 32, 2.0, 2.0, 2.0)
    return pdi, img, RayTracingMatrixParameters()


def test_rdp_two_voxels_closed_form():
    """x = (1, 3), gamma 2, eps 0, one neighbour of weight 1 each way:
    value = 2 * 0.5 * 4 / 8 = 0.5; d/dx1 = (1-3)(4+1+9)/64 = -0.4375,
    d/dx2 = (3-1)(4+3+3)/64 = 0.3125."""
    img = np.array([[[1.0, 3.0]]], np.float32)
    v, g = relative_difference_prior(img, VOX, beta=1.0)
    assert v == 0.5
    assert g.reshape(-1).tolist() == [-0.4375, 0.3125]
    v, g = relative_difference_prior(img, VOX, beta=3.0, kappa=np.array([[[2.0, 1.0]]], np.float32))
    assert v == 3.0 and g.reshape(-1).tolist() == [-2.625, 1.875]
    # 0/0 at x = y = 0 with eps = 0 contributes nothing; eps > 0 is finite
    z = np.zeros((1, 1, 2), np.float32)
    v, g = relative_difference_prior(z, VOX, beta=1.0)
    assert v == 0.0 and not g.any()
    v, g = relative_difference_prior(np.array([[[0.0, 2.0]]], np.float32), VOX, beta=1.0, epsilon=1.0)
    # (0-2)^2 * 0.5 / (2 + 4 + 1) per ordered pair
    assert abs(v - 4.0 / 7.0) < 1e-12


@pytest.mark.parametrize("gamma,eps,use_kappa,only_2D", [(2.0, 0.0, False, False), (0.0, 0.1, True, False),
                                                          (5.0, 0.01, False, True), (2.0, 0.0, True, False)])
def test_rdp_gradient_is_the_derivative_of_the_value(gamma, eps, use_kappa, only_2D):
    rng = np.random.default_rng(7)
    x = rng.uniform(0.2, 2.0, (4, 5, 6)).astype(np.float32)
    kappa = rng.uniform(0.5, 1.5, x.shape).astype(np.float32) if use_kappa else None
    beta = 0.7
    _, g = relative_difference_prior(x, VOX, beta, gamma, eps, kappa=kappa, only_2D=only_2D)
    for v in [(0, 0, 0), (1, 2, 3), (3, 4, 5), (2, 0, 5)]:
        h = np.float32(1e-2) * x[v]
        xp, xm = x.copy(), x.copy()
        xp[v] += h
        xm[v] -= h
        fp, _ = relative_difference_prior(xp, VOX, beta, gamma, eps, kappa=kappa, only_2D=only_2D, gradient=False)
        fm, _ = relative_difference_prior(xm, VOX, beta, gamma, eps, kappa=kappa, only_2D=only_2D, gradient=False)
        fd = (fp - fm) / (float(xp[v]) - float(xm[v]))
        assert abs(fd - g[v]) <= 2e-3 * max(abs(fd), 1e-3), (v, fd, g[v])


def test_rdp_large_epsilon_is_the_quadratic_prior():
    """gamma 0, eps E >> x: w 0.5 (a-b)^2 / E per ordered pair, i.e. the
    quadratic prior's w (a-b)^2 / 4 times 2/E; gradients likewise."""
    rng = np.random.default_rng(8)
    x = rng.uniform(0.5, 1.5, (3, 6, 5)).astype(np.float32)
    E = 1e6
    vr, gr = relative_difference_prior(x, VOX, 1.0, gamma=0.0, epsilon=E)
    vq, gq = quadratic_prior(x, VOX, 1.0)
    assert abs(vr * E / 2 - vq) <= 1e-5 * vq
    assert np.abs(gr * E / 2 - gq).max() <= 1e-5 * np.abs(gq).max()


def _data(tiny, seed=0):
    pdi, img, pm = tiny
    O = OraclePlan(pdi, img, pm)
    x_true, y, bkg = phantoms.config0_data(pdi, img, O.forward_project, seed=seed)
    return O, x_true, y, bkg


def test_multiplicative_power_of_two_equivariance(tiny):
    """m = 2 everywhere with the background halved: m (A x/2 + b/2) = A x + b,
    so OSEM images halve, the log-likelihood at x/2 equals the one at x without
    factors and the gradient doubles -- exactly (the quotient y/(A x + b)
    stays below the 1e4 cap of divide_and_truncate)."""
    O, x_true, y, bkg = _data(tiny)
    pdi, img, _ = tiny
    x0 = np.ones(img.shape, np.float32)
    b = np.full(pdi.num_bins, bkg, np.float32)
    ref = O.osem(y, x0, 4, 6, bkg=b, omp=False)  # fixed summation order
    L1 = O.log_likelihood(y, x_true, bkg=b)
    g1 = O.subset_gradient(y, x_true, 1, 4, bkg=b)
    O.set_multiplicative(np.full(pdi.num_bins, 2.0, np.float32))
    try:
        got = O.osem(y, x0 / 2, 4, 6, bkg=b / 2, omp=False)
        L2 = O.log_likelihood(y, x_true / 2, bkg=b / 2)
        g2 = O.subset_gradient(y, x_true / 2, 1, 4, bkg=b / 2)
    finally:
        O.set_multiplicative(None)
    assert np.array_equal(got, ref / 2)
    assert L2 == L1
    assert np.array_equal(g2, 2 * g1)


def test_multiplicative_gradient_directional_derivative(tiny):
    O, x_true, y, bkg = _data(tiny, 1)
    pdi, img, _ = tiny
    rng = np.random.default_rng(2)
    m = rng.uniform(0.5, 2.0, pdi.num_bins).astype(np.float32)
    b = np.full(pdi.num_bins, bkg, np.float32)
    x = (x_true + 0.5).astype(np.float32)
    d = (rng.uniform(-1, 1, img.shape) * x).astype(np.float32)
    O.set_multiplicative(m)
    try:
        g = O.subset_gradient(y, x, -1, 1, bkg=b).astype(np.float64)
        t = 1e-3
        Lp = O.log_likelihood(y, (x + t * d).astype(np.float32), bkg=b)
        Lm = O.log_likelihood(y, (x - t * d).astype(np.float32), bkg=b)
    finally:
        O.set_multiplicative(None)
    fd = (Lp - Lm) / (2 * t)
    an = float((g * d).sum())
    assert abs(fd - an) <= 1e-3 * abs(an), (fd, an)
    # the factors change the gradient: not the one without them
    g0 = O.subset_gradient(y, x, -1, 1, bkg=b).astype(np.float64)
    assert abs(float((g0 * d).sum()) - an) > 1e-2 * abs(an)


def test_multiplicative_sensitivity_is_back_projection_of_factors(tiny):
    O, _, _, _ = _data(tiny)
    pdi, img, _ = tiny
    m = np.random.default_rng(3).uniform(0.5, 2.0, pdi.num_bins).astype(np.float32)
    O.set_multiplicative(m)
    try:
        s = O.subset_sensitivity(2, 4)
    finally:
        O.set_multiplicative(None)
    assert np.array_equal(s, O.back_project(m, subset=2, num_subsets=4, omp=False))


def test_total_sensitivity_mode(tiny):
    """use subset sensitivities := 0: subset k's update is divided by
    total/num_subsets instead of sens_k, so one subiteration differs from the
    default by the voxelwise factor sens_0 / (total / 4)."""
    O, _, y, bkg = _data(tiny)
    pdi, img, _ = tiny
    x0 = np.ones(img.shape, np.float32)
    b = np.full(pdi.num_bins, bkg, np.float32)
    a = O.osem(y, x0, 4, 1, bkg=b)
    O.set_use_subset_sensitivities(False)
    try:
        t = O.osem(y, x0, 4, 1, bkg=b)
    finally:
        O.set_use_subset_sensitivities(True)
    sens = [O.subset_sensitivity(s, 4) for s in range(4)]
    total = sens[0].copy()
    for s in sens[1:]:
        total += s
    mask = (sens[0] > 1e-3 * sens[0].max()) & (a > 1e-3 * a.max())
    ratio = t[mask] / a[mask]
    want = sens[0][mask] / (total[mask] / 4)
    assert np.abs(ratio - want).max() <= 1e-5 * want.max()
    assert np.abs(ratio - 1).max() > 1e-3  # the two modes differ


@pytest.mark.parametrize("kind", ["quadratic", "relative difference"])
def test_one_step_late_fixed_point(tiny, kind):
    O, x_true, y, bkg = _data(tiny, 4)
    pdi, img, _ = tiny
    b = np.full(pdi.num_bins, bkg, np.float32)
    sens = O.subset_sensitivity(0, 1)
    prior = quadratic_prior if kind == "quadratic" else relative_difference_prior
    _, g_true = prior(x_true + 0.1, VOX, 1.0)
    beta = 0.3 * float(np.median(sens[sens > 0])) / float(np.abs(g_true).max())
    O.set_prior(kind, beta)
    try:
        x = O.osem(y, np.ones(img.shape, np.float32), 1, 300, bkg=b)
    finally:
        O.set_prior(None)
    u = O.subset_gradient(y, x, -1, 1, bkg=b, add_sensitivity=True)
    _, gR = prior(x, VOX, beta)
    on = x > 0.05 * x.max()
    res = np.abs(u - sens - gR)[on]
    flipped = np.abs(u - sens + gR)[on]
    scale = float(np.median(sens[on]))
    print(f"{kind}: median |residual| {np.median(res) / scale:.2e}, flipped sign {np.median(flipped) / scale:.2e}")
    assert np.median(res) <= 0.2 * np.median(flipped)
    assert np.median(res) <= 0.02 * scale


def test_rdp_hessian_two_voxels_closed_form():
    """x = (1, 3), gamma 2, eps 0, input (1, 0): H in = (w f20(1,3), w f11(3,1))
    = (2 (2*3)^2 / 8^3, -2 (2*3)(2*1) / 8^3) = (0.140625, -0.046875)."""
    from oracle.oracle import prior_hessian_times_input
    x = np.array([[[1.0, 3.0]]], np.float32)
    out = prior_hessian_times_input("relative difference", x, np.array([[[1.0, 0.0]]], np.float32), VOX, 1.0)
    assert out.reshape(-1).tolist() == [0.140625, -0.046875]


@pytest.mark.parametrize("kind,use_kappa,only_2D", [("quadratic", False, False), ("quadratic", True, True),
                                                    ("relative difference", False, False),
                                                    ("relative difference", True, False)])
def test_prior_hessian_is_the_derivative_of_the_gradient(kind, use_kappa, only_2D):
    from oracle.oracle import prior_hessian_times_input
    rng = np.random.default_rng(9)
    x = rng.uniform(0.5, 2.0, (4, 5, 6)).astype(np.float32)
    d = rng.uniform(-1, 1, x.shape).astype(np.float32)
    kappa = rng.uniform(0.5, 1.5, x.shape).astype(np.float32) if use_kappa else None
    beta, eps = 0.7, 0.01
    hv = prior_hessian_times_input(kind, x, d, VOX, beta, 2.0, eps, kappa=kappa, only_2D=only_2D)
    prior = quadratic_prior if kind == "quadratic" else relative_difference_prior
    kw = {} if kind == "quadratic" else {"gamma": 2.0, "epsilon": eps}
    t = 1e-3  # the relative difference prior's curvature changes fast (|x_j - x_k|)
    _, gp = prior((x + t * d).astype(np.float32), VOX, beta, kappa=kappa, only_2D=only_2D, **kw)
    _, gm = prior((x - t * d).astype(np.float32), VOX, beta, kappa=kappa, only_2D=only_2D, **kw)
    fd = (gp.astype(np.float64) - gm) / (2 * t)
    assert np.abs(hv - fd).max() <= 2e-3 * np.abs(fd).max()


def test_quadratic_hessian_is_the_gradient_operator_and_approximate_hessian():
    """Quadratic prior: H x = grad R(x) (R is quadratic); the approximate
    Hessian of constant input 1 is beta * (sum of the neighbour weights) inside."""
    from oracle.oracle import prior_hessian_times_input, quadratic_prior_weights
    rng = np.random.default_rng(10)
    x = rng.uniform(0.5, 2.0, (5, 6, 7)).astype(np.float32)
    hx = prior_hessian_times_input("quadratic", np.zeros_like(x), x, VOX, 0.3)
    _, g = quadratic_prior(x, VOX, 0.3)
    assert np.abs(hx - g).max() <= 1e-5 * np.abs(g).max()
    ones = np.ones_like(x)
    a = prior_hessian_times_input("approximate quadratic", x, ones, VOX, 0.3)
    w = quadratic_prior_weights(VOX)
    assert abs(float(a[2, 3, 3]) - 0.3 * float(w.sum())) <= 1e-6 * float(w.sum())
