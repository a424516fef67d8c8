"""Pins the oracle's hydrostatic pressure and sigma-coordinate pressure
gradient (oracle/pressure.c; the algorithm of
include/physics/pressure_gradient.hpp:46-367, which does not compile in the
reference) with closed forms computed independently from numpy's Legendre
routines: node coordinates, weights and the exact derivatives of
polynomials.  The vertical direction is the element's zeta index."""
import numpy as np
import pytest

G, RHO0 = 9.81, 1025.0


def _lgl(order):
    c = np.zeros(order + 1)
    c[-1] = 1.0
    inner = np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c))
    x = np.r_[-1.0, np.sort(inner), 1.0]
    w = 2.0 / (order * (order + 1) * np.polynomial.legendre.legval(x, c) ** 2)
    return x, w


def _grid(order, E=3):
    """xi, eta, zeta at every node, element-major (E, nd), d = i + n (j + n k)."""
    n = order + 1
    x, _ = _lgl(order)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    xi, et, ze = x[i].reshape(-1), x[j].reshape(-1), x[k].reshape(-1)
    return [np.tile(a, E) for a in (xi, et, ze)]


def _partial_sums(order, from_top):
    """c_k = sum_{k' >= k} w_k' / 2 (from the top) or sum_{k' <= k} (from the bottom)."""
    _, w = _lgl(order)
    sw = 0.5 * w
    return np.cumsum(sw[::-1])[::-1] if from_top else np.cumsum(sw)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_hydrostatic_pressure_closed_forms(orc, order):
    n = order + 1
    xi, et, ze = _grid(order)
    H = np.full(xi.size, 37.0)
    p = orc.hydrostatic_pressure(order, np.full(xi.size, 1020.0), H, G)
    c = _partial_sums(order, True)
    want = G * 1020.0 * 37.0 * c[(np.arange(xi.size) % n ** 3) // n ** 2]
    np.testing.assert_allclose(p, want, rtol=1e-14)
    # at the bottom node the weights sum to one: p = g rho H
    assert np.allclose(p[:n * n], G * 1020.0 * 37.0, rtol=1e-14)
    # density linear in zeta: the weighted partial sums of rho
    x, w = _lgl(order)
    rho = 1020.0 + 3.0 * ze
    p = orc.hydrostatic_pressure(order, rho, H, G)
    k = (np.arange(xi.size) % n ** 3) // n ** 2
    want = np.array([G * 37.0 * np.sum(0.5 * w[kk:] * (1020.0 + 3.0 * x[kk:])) for kk in k])
    np.testing.assert_allclose(p, want, rtol=1e-14)


@pytest.mark.parametrize("method", [0, 1])
def test_uniform_density_flat_column_has_no_baroclinic_force(orc, method):
    order = 3
    xi, et, ze = _grid(order)
    sigma = 0.5 * (ze - 1.0)
    px, py = orc.pressure_gradient(order, method, 2, np.zeros(xi.size), np.full(xi.size, 1025.0),
                                   np.full(xi.size, 50.0), sigma)
    assert np.abs(px).max() < 1e-9 and np.abs(py).max() < 1e-9


@pytest.mark.parametrize("order", [2, 3, 4])
def test_xi_linear_density(orc, order):
    """rho = rho_a + a xi over a flat column of depth H, sigma = (zeta - 1) / 2:
    standard: p = g H (rho_a + a xi) c_k, so pg_x = -g H a c_k / rho0;
    density Jacobian: J_x = a z_zeta = a H / 2, pg_x = -(g / rho0) a H / 2 sum_{k'<=k} sw_k'."""
    n = order + 1
    a, Hc = 0.8, 40.0
    xi, et, ze = _grid(order)
    k = (np.arange(xi.size) % n ** 3) // n ** 2
    rho = 1024.0 + a * xi
    sigma = 0.5 * (ze - 1.0)
    eta = np.zeros(xi.size)
    H = np.full(xi.size, Hc)
    px, py = orc.pressure_gradient(order, 0, 2, eta, rho, H, sigma)
    ctop = _partial_sums(order, True)
    np.testing.assert_allclose(px, -G * Hc * a * ctop[k] / RHO0, rtol=1e-11, atol=1e-14)
    assert np.abs(py).max() < 1e-12
    px, py = orc.pressure_gradient(order, 1, 2, eta, rho, H, sigma)
    cbot = _partial_sums(order, False)
    want = -(G / RHO0) * a * Hc / 2.0 * cbot[k]
    np.testing.assert_allclose(px, want, rtol=1e-11, atol=1e-14)
    assert np.abs(py).max() < 1e-12
    # planted mistake: integrating the Jacobian from the top instead fails
    assert not np.allclose(px, -(G / RHO0) * a * Hc / 2.0 * ctop[k], rtol=1e-6)


def test_barotropic_gradient_of_a_linear_surface(orc):
    order = 4
    xi, et, ze = _grid(order)
    eta = 0.02 * xi - 0.03 * et
    H = 30.0 + 5.0 * ze
    px, py = orc.pressure_gradient(order, 0, 1, eta, np.full(xi.size, 1025.0), H,
                                   0.5 * (ze - 1.0))
    np.testing.assert_allclose(px, -G * H * 0.02, rtol=1e-12)
    np.testing.assert_allclose(py, -G * H * -0.03, rtol=1e-12)


@pytest.mark.parametrize("order", [3, 4])
def test_sigma_correction_over_a_sloping_bottom(orc, order):
    """Uniform density, H constant in the element, dh/dx = s, flat surface:
    p = g rho H c_k, D_xi p = 0, and dp/dx|_z = (D_zeta p) sigma s / H, so
    pg_x = -(g rho H (dc/dzeta)_k) sigma s / (H rho0), with dc/dzeta the
    exact derivative of the interpolant of c_k."""
    n = order + 1
    x, _ = _lgl(order)
    ctop = _partial_sums(order, True)
    dc = np.polynomial.polynomial.polyval(
        x, np.polynomial.polynomial.polyder(np.polynomial.polynomial.polyfit(x, ctop, order)))
    xi, et, ze = _grid(order)
    k = (np.arange(xi.size) % n ** 3) // n ** 2
    sigma = 0.5 * (ze - 1.0)
    slope, Hc, rho = 0.01, 60.0, 1026.0
    dh = np.full(xi.size, slope)
    px, py = orc.pressure_gradient(order, 0, 2, np.zeros(xi.size), np.full(xi.size, rho),
                                   np.full(xi.size, Hc), sigma, dh_dx=dh)
    want = -(G * rho * Hc * dc[k]) * (sigma * slope / Hc) / RHO0
    np.testing.assert_allclose(px, want, rtol=1e-9, atol=1e-12)
    assert np.abs(py).max() < 1e-12
    # planted mistake: the opposite sign of the correction is not the oracle's
    assert not np.allclose(px, -want, rtol=1e-6)
