"""Pins the oracle's sigma-coordinate metrics (A13): the reference's closed
form (tests/unit/test_sigma_coordinate.cpp:71-89), the transform identity
checked by finite differences, and the dry-column guards."""
import numpy as np


def test_reference_closed_form(orc):
    H, Hi, dx, dy, dt = orc.sigma_metrics([-0.5], [1.0], [10.0], [0.1], [0.0], [0.2], [0.0])
    Ht = 11.0
    assert H[0] == Ht and abs(Hi[0] - 1 / Ht) < 1e-16
    expected = -(1.0 / Ht) * ((1.0 - 0.5) * 0.1 + (-0.5) * 0.2)
    assert abs(dx[0] - expected) < 1e-12 and dy[0] == 0.0 and dt[0] == 0.0


def test_metric_matches_finite_difference_of_transform(orc):
    """sigma(x, z) = (z - eta(x)) / (eta(x) + h(x)) differentiated at fixed z"""
    rng = np.random.default_rng(1)
    n = 50
    x0 = rng.uniform(0, 1, n)
    eta = lambda x: 0.3 * np.sin(3 * x)
    h = lambda x: 10 + 4 * np.cos(2 * x)
    deta = lambda x: 0.9 * np.cos(3 * x)
    dh = lambda x: -8 * np.sin(2 * x)
    s0 = rng.uniform(-1, 0, n)
    z0 = eta(x0) + s0 * (eta(x0) + h(x0))
    sig = lambda x: (z0 - eta(x)) / (eta(x) + h(x))
    eps = 1e-6
    fd = (sig(x0 + eps) - sig(x0 - eps)) / (2 * eps)
    _, _, dx, _, _ = orc.sigma_metrics(s0, eta(x0), h(x0), deta(x0), np.zeros(n), dh(x0),
                                       np.zeros(n))
    np.testing.assert_allclose(dx, fd, rtol=1e-6, atol=1e-9)
    # time derivative: d sigma/dt at fixed z with eta(t) = eta0 + c t, h fixed
    c = 0.01
    sig_t = lambda t: (z0 - (eta(x0) + c * t)) / (eta(x0) + c * t + h(x0))
    fdt = (sig_t(eps) - sig_t(-eps)) / (2 * eps)
    _, _, _, _, dt = orc.sigma_metrics(s0, eta(x0), h(x0), deta(x0), np.zeros(n), dh(x0),
                                       np.zeros(n), np.full(n, c))
    np.testing.assert_allclose(dt, fdt, rtol=1e-6, atol=1e-9)


def test_dry_guards(orc):
    H, Hi, dx, dy, dt = orc.sigma_metrics([-0.5, -0.2], [0.0, -1.0], [1e-11, 1.0], [1, 1],
                                          [1, 1], [1, 1], [1, 1], [1, 1])
    assert Hi[0] == 0 and dx[0] == 0 and dy[0] == 0 and dt[0] == 0
    assert H[1] == 0.0 and Hi[1] == 0 and dx[1] == 0
