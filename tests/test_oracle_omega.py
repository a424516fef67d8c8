"""Pins the oracle's vertical velocity diagnosis (oracle/omega.c, restating
VerticalVelocityDiagnosis and ColumnVerticalVelocity,
src/physics/vertical_velocity.cpp:55-275).  The reference has no tests of
these classes, so the checks are closed forms: the trapezoidal column
integral is exact for a divergence linear along the column, the column is the
run of n consecutive DOFs (node_index_3d, the xi index of the basis layout),
the dry guard, the kinematic surface and bottom values of w, and the node
tables of ColumnVerticalVelocity."""
import numpy as np
import pytest


def _x(orc, order):
    return orc.lgl(order + 1)[0]


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6])
def test_linear_divergence_is_integrated_exactly(orc, order):
    n = order + 1
    nd = n ** 3
    x = _x(orc, order)
    rng = np.random.default_rng(order)
    E = 3
    a, b = rng.standard_normal((E, n * n)), rng.standard_normal((E, n * n))
    Hc = rng.uniform(5, 50, size=(E, n * n))
    # column i = DOFs i*n .. i*n + n - 1, level k = position in the column
    div = (a[:, :, None] + b[:, :, None] * x[None, None, :]).reshape(E, nd)
    H = np.repeat(Hc[:, :, None], n, axis=2).reshape(E, nd)
    om = orc.diagnose_omega(order, div, H).reshape(E, n * n, n)
    # -(1/H) * 0.5 * int_{-1}^{x_k} (a + b s) ds  (dsigma = dzeta / 2)
    want = -(0.5 * (a[:, :, None] * (x + 1) + b[:, :, None] * (x ** 2 - 1) / 2)) / Hc[:, :, None]
    np.testing.assert_allclose(om, want, rtol=0, atol=1e-13 * np.abs(want).max())
    assert np.all(om[:, :, 0] == 0.0)  # omega(-1) = 0


def test_column_is_the_consecutive_dofs(orc):
    """A divergence that varies only with d // n (the column index) gives a
    column-wise constant integrand; one that varies with d % n does not."""
    order, n = 3, 4
    nd = n ** 3
    d = np.arange(nd)
    H = np.full((1, nd), 10.0)
    om_col = orc.diagnose_omega(order, (1.0 + (d // n)).reshape(1, nd).astype(float), H)
    x = _x(orc, order)
    want = -(1.0 + (d // n)) * 0.5 * (x[d % n] + 1) / 10.0
    np.testing.assert_allclose(om_col[0], want, rtol=0, atol=1e-14)
    # varying along the column: the trapezoid of a non-constant integrand
    om_lev = orc.diagnose_omega(order, (1.0 + (d % n)).reshape(1, nd).astype(float), H)
    assert not np.allclose(om_lev[0], -(1.0 + (d % n)) * 0.5 * (x[d % n] + 1) / 10.0)


def test_dry_columns_give_zero(orc):
    order, n = 2, 3
    nd = n ** 3
    div = np.ones((1, nd))
    H = np.full((1, nd), 5.0)
    H[0, :n] = 0.0  # column 0 dry
    om = orc.diagnose_omega(order, div, H)[0]
    assert np.all(om[:n] == 0.0)
    assert np.all(om[n:].reshape(-1, n)[:, 1:] < 0)


def test_from_velocity_sums_the_divergence_first(orc):
    order = 3
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal((2, 4, 64))
    H = rng.uniform(1, 3, size=(4, 64))
    np.testing.assert_array_equal(orc.diagnose_omega_from_velocity(order, a, b, H),
                                  orc.diagnose_omega(order, a + b, H))


def test_w_kinematic_values_and_the_two_routes_agree(orc):
    rng = np.random.default_rng(2)
    m = 200
    sig = rng.uniform(-1, 0, m)
    sig[:3] = [0.0, -1.0, -0.5]
    detadt = rng.standard_normal(m)
    z = np.zeros(m)
    # no omega, no velocity: w = deta/dt (1 + sigma): deta/dt at the
    # surface, 0 at the bottom
    w = orc.omega_to_w_simple(z, z, z, rng.uniform(-1, 1, m), rng.uniform(5, 9, m), sig, z, z,
                              detadt)
    assert w[0] == detadt[0] and w[1] == 0.0
    np.testing.assert_allclose(w, detadt * (1 + sig), rtol=1e-15, atol=0)
    # omega_to_physical_w with dz/dx = deta/dx (1 + sigma) etc. is the same w
    om, u, v, eta, h, ex, ey = rng.standard_normal((7, m))
    ws = orc.omega_to_w_simple(om, u, v, eta, h, sig, ex, ey, detadt)
    wp = orc.omega_to_w(om, u, v, eta + h, ex * (1 + sig), ey * (1 + sig), detadt * (1 + sig))
    np.testing.assert_array_equal(ws, wp)


def test_column_nodes(orc):
    # 3..5 levels: LGL nodes mapped to [-1, 0]
    for order in (2, 3, 4):
        n = order + 1
        P = np.polynomial.legendre.Legendre.basis(n - 1).deriv()
        lgl = np.concatenate([[-1.0], np.sort(P.roots().real), [1.0]])
        np.testing.assert_allclose(orc.column_sigma_nodes(order), 0.5 * (lgl - 1), atol=1e-15)
    np.testing.assert_array_equal(orc.column_sigma_nodes(1), [-1.0, 0.0])
    # 6 or more levels: uniform spacing (the reference's fallback, not LGL)
    np.testing.assert_allclose(orc.column_sigma_nodes(6), np.linspace(-1, 0, 7), atol=1e-15)


@pytest.mark.parametrize("order", [1, 3, 4, 6])
def test_column_diagnosis_closed_form(orc, order):
    n = order + 1
    s = orc.column_sigma_nodes(order)
    rng = np.random.default_rng(order + 5)
    ncol = 7
    a, b = rng.standard_normal(ncol), rng.standard_normal(ncol)
    Hc = rng.uniform(2, 20, ncol)
    ob = rng.standard_normal(ncol)
    div = a[:, None] + b[:, None] * s[None, :]
    H = np.repeat(Hc[:, None], n, axis=1)
    H[0] = 0.0  # dry column keeps omega_bottom
    om = orc.diagnose_column(order, div, H, ob)
    integ = a[:, None] * (s - s[0]) + b[:, None] * (s ** 2 - s[0] ** 2) / 2
    want = ob[:, None] - integ / Hc[:, None]
    want[0] = ob[0]
    np.testing.assert_allclose(om, want, rtol=0, atol=1e-13 * np.abs(want).max())
    assert np.all(orc.diagnose_column(order, div, H)[:, 0] == 0.0)


def test_continuity_residual(orc):
    """Zero for an exact balance div + d(H omega)/dsigma = 0; the RMS of a
    constant imbalance c is |c|."""
    rng = np.random.default_rng(4)
    div = rng.standard_normal((5, 27))
    np.testing.assert_array_equal(orc.continuity_residual(2, div, -div), np.zeros(5))
    np.testing.assert_allclose(orc.continuity_residual(2, div, 0.75 - div), np.full(5, 0.75),
                               rtol=1e-15)
