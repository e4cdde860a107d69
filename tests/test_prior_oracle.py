"""Pins of the oracle's QuadraticPrior (QuadraticPrior.cxx:202-367): closed
forms, an independent pair-wise brute force, the reference's numerical
gradient recipe (ObjectiveFunctionTests.h:86-148) and a planted mistake."""
import itertools

import numpy as np
import pytest

from oracle.oracle import quadratic_prior, quadratic_prior_weights

VOX = (2.03125, 2.08626, 2.08626)


def _offsets(only_2D=False):
    dzs = (0,) if only_2D else (-1, 0, 1)
    return [(dz, dy, dx) for dz in dzs for dy in (-1, 0, 1) for dx in (-1, 0, 1) if (dz, dy, dx) != (0, 0, 0)]


def test_weights_closed_form():
    w = quadratic_prior_weights(VOX)
    assert w[1, 1, 1] == 0
    for dz, dy, dx in _offsets():
        d = np.sqrt((dx * VOX[2]) ** 2 + (dy * VOX[1]) ** 2 + (dz * VOX[0]) ** 2)
        assert abs(w[dz + 1, dy + 1, dx + 1] - VOX[2] / d) <= 1e-7 * VOX[2] / d
    # in-plane x neighbours have weight exactly 1 and the table is symmetric
    assert w[1, 1, 0] == 1 and w[1, 1, 2] == 1
    assert np.array_equal(w, w[::-1, ::-1, ::-1])


def test_constant_image_is_flat():
    x = np.full((5, 6, 7), 3.25, np.float32)
    v, g = quadratic_prior(x, VOX, 0.7)
    assert v == 0.0 and not g.any()


@pytest.mark.parametrize("where", ["interior", "corner", "face"])
def test_single_voxel_closed_form(where):
    beta = 0.01
    x = np.zeros((5, 6, 7), np.float32)
    c = {"interior": (2, 3, 3), "corner": (0, 0, 0), "face": (4, 2, 3)}[where]
    x[c] = 2.0
    w = quadratic_prior_weights(VOX).astype(np.float64)
    inside = [o for o in _offsets() if all(0 <= c[i] + o[i] < x.shape[i] for i in range(3))]
    wsum = sum(w[o[0] + 1, o[1] + 1, o[2] + 1] for o in inside)
    v, g = quadratic_prior(x, VOX, beta)
    # each of the pairs {c, n} appears twice with (2-0)^2/4 = 1
    assert v == pytest.approx(beta * 2 * wsum * 1.0, rel=1e-6)
    assert g[c] == pytest.approx(beta * wsum * 2.0, rel=1e-6)
    for o in inside:
        n = tuple(c[i] + o[i] for i in range(3))
        assert g[n] == pytest.approx(-beta * w[o[0] + 1, o[1] + 1, o[2] + 1] * 2.0, rel=1e-6)
    assert np.count_nonzero(g) == len(inside) + 1


def _brute_force(x, kappa, beta, only_2D=False):
    """Unordered neighbour pairs {v, n}: value += w k_v k_n (x_v - x_n)^2 / 2,
    grad_v += w k_v k_n (x_v - x_n), grad_n -= the same (FP64)."""
    nz, ny, nx = x.shape
    w = quadratic_prior_weights(VOX).astype(np.float64)
    xd = x.astype(np.float64)
    kd = np.ones_like(xd) if kappa is None else kappa.astype(np.float64)
    val, g = 0.0, np.zeros_like(xd)
    for z, y, xx in itertools.product(range(nz), range(ny), range(nx)):
        for o in _offsets(only_2D):
            n = (z + o[0], y + o[1], xx + o[2])
            if not all(0 <= n[i] < x.shape[i] for i in range(3)) or n <= (z, y, xx):
                continue
            ww = w[o[0] + 1, o[1] + 1, o[2] + 1] * kd[z, y, xx] * kd[n]
            d = xd[z, y, xx] - xd[n]
            val += ww * d * d / 2
            g[z, y, xx] += ww * d
            g[n] -= ww * d
    return beta * val, beta * g


@pytest.mark.parametrize("kappa,only_2D", [(False, False), (True, False), (False, True)])
def test_matches_pairwise_brute_force(kappa, only_2D):
    rng = np.random.default_rng(3)
    x = rng.lognormal(0, 0.5, (4, 5, 6)).astype(np.float32)
    k = rng.uniform(0.5, 1.5, x.shape).astype(np.float32) if kappa else None
    v, g = quadratic_prior(x, VOX, 0.01, kappa=k, only_2D=only_2D)
    bv, bg = _brute_force(x, k, 0.01, only_2D)
    assert v == pytest.approx(bv, rel=1e-6)
    np.testing.assert_allclose(g, bg, rtol=1e-5, atol=1e-6 * np.abs(bg).max())


def test_numerical_gradient_recipe():
    """ObjectiveFunctionTests::test_gradient: one-sided differences with
    eps = 1e-4 * max(x), tolerance max|g| / 1000."""
    rng = np.random.default_rng(5)
    x = rng.lognormal(0, 0.5, (3, 6, 6)).astype(np.float32)
    beta = 1.0
    v0, g = quadratic_prior(x, VOX, beta)
    eps = np.float32(1e-4 * x.max())
    tol = max(abs(g.min()), g.max()) / 1000
    for idx in itertools.product(*(range(s) for s in x.shape)):
        xp = x.copy()
        xp[idx] += eps
        v1, _ = quadratic_prior(xp, VOX, beta, gradient=False)
        num = (v1 - v0) / float(xp[idx] - x[idx])
        assert abs(num - g[idx]) <= tol, (idx, num, g[idx])


def test_planted_mistake_fails_closed_form():
    """Weights 1/distance^2 instead of vox_x/distance (a plausible slip)
    must fail the single-voxel closed form used above."""
    from oracle import oracle as O
    import ctypes

    x = np.zeros((5, 6, 7), np.float32)
    x[2, 3, 3] = 2.0
    w_good = quadratic_prior_weights(VOX)
    w_bad = np.zeros(27, np.float32)
    for dz, dy, dx in _offsets():
        d2 = (dx * VOX[2]) ** 2 + (dy * VOX[1]) ** 2 + (dz * VOX[0]) ** 2
        w_bad[((dz + 1) * 3 + dy + 1) * 3 + dx + 1] = VOX[2] ** 2 / d2
    g = np.zeros(x.shape, np.float32)
    v_bad = O.lib().or_quadratic_prior(x.reshape(-1), None, 5, 6, 7, w_bad, 0, 0.01, g.ctypes.data_as(ctypes.c_void_p))
    assert v_bad != pytest.approx(0.01 * 2 * w_good.astype(np.float64).sum(), rel=1e-3)
