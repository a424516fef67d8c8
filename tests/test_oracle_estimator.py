"""Pins the oracle's error estimator / marking (A19) and modal coefficients
(A20) against the reference's tests and closed forms computed in numpy.

Reference tests restated: tests/unit/test_amr.cpp:321-371 (constant field ->
zero indicator; i-ramp -> non-zero)."""
import numpy as np
import pytest


def uniform_bounds(nx, ny, nz, L=(1.0, 1.0, 1.0)):
    b = []
    # element order of build_uniform: x fastest (the values only need the sizes)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                dx, dy, dz = L[0] / nx, L[1] / ny, L[2] / nz
                b.append([i * dx, (i + 1) * dx, j * dy, (j + 1) * dy, k * dz, (k + 1) * dz])
    return np.array(b)


def test_constant_field_zero_indicator(orc):
    n = 3
    b = uniform_bounds(2, 2, 2)
    v, m, r, c = orc.estimate(2, np.ones((8, n ** 3)), b)
    assert np.all(v == 0.0) and np.all(m == 0) and not r.any() and not c.any()


def test_ramp_closed_form(orc):
    """solution(i) = i: variation ndof-1, h = cbrt(0.5^3); var along x = n-1,
    y = (n-1) n, z = (n-1) n^2 -> only Z exceeds half the maximum."""
    order, n = 2, 3
    b = uniform_bounds(2, 2, 2)
    u = np.tile(np.arange(n ** 3, dtype=np.float64), (8, 1))
    v, m, r, c = orc.estimate(order, u, b)
    assert np.all(v == 1.0)  # all equal -> normalised to 1
    assert np.all(m == 4) and r.all() and not c.any()


def _numpy_estimate(order, u, b, rt=0.1, ct=0.01, sf=0.5):
    n = order + 1
    U = u.reshape(-1, n, n, n)  # [e][k][j][i]
    val = (U.max(axis=(1, 2, 3)) - U.min(axis=(1, 2, 3))) * np.cbrt(
        (b[:, 1] - b[:, 0]) * (b[:, 3] - b[:, 2]) * (b[:, 5] - b[:, 4]))
    vx = np.ptp(U, axis=3).max(axis=(1, 2))
    vy = np.ptp(U, axis=2).max(axis=(1, 2))
    vz = np.ptp(U, axis=1).max(axis=(1, 2))
    mv = np.maximum(np.maximum(vx, vy), vz)
    mask = ((vx > sf * mv) * 1 + (vy > sf * mv) * 2 + (vz > sf * mv) * 4) * (mv >= 1e-14)
    mx = val.max()
    if mx > 1e-14:
        val = val / mx
        return val, mask, val > rt, val < ct
    return val, mask, np.zeros(len(val), bool), np.zeros(len(val), bool)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_estimate_matches_numpy_closed_form(orc, order):
    rng = np.random.default_rng(order)
    n = order + 1
    E = 200
    b = uniform_bounds(10, 5, 4, L=(3.0, 2.0, 0.5))
    # anisotropic separable fields with a spread of amplitudes per element
    x = np.linspace(-1, 1, n)
    amp = 10.0 ** rng.uniform(-4, 0, size=(E, 3))
    U = (amp[:, 0, None, None, None] * np.sin(2 * x)[None, None, None, :]
         + amp[:, 1, None, None, None] * x[None, None, :, None] ** 2
         + amp[:, 2, None, None, None] * np.cos(x)[None, :, None, None])
    u = U.reshape(E, -1) + rng.normal(size=(E, 1))
    v, m, r, c = orc.estimate(order, u, b)
    ev, em, er, ec = _numpy_estimate(order, u, b)
    np.testing.assert_allclose(v, ev, rtol=2e-16 * 4, atol=0)
    np.testing.assert_array_equal(m, em)
    np.testing.assert_array_equal(r.astype(bool), er)
    np.testing.assert_array_equal(c.astype(bool), ec)
    assert 0 < r.sum() < E and 0 < c.sum() < E  # thresholds exercised


def test_tiny_max_skips_normalisation(orc):
    b = uniform_bounds(2, 1, 1)
    u = np.zeros((2, 8))
    u[0, 0] = 1e-15  # value 1e-15 * 0.5^(1/3)... < 1e-14
    v, m, r, c = orc.estimate(1, u, b)
    assert v[0] > 0 and v[0] < 1e-14 and not r.any() and not c.any()


def test_mark_elements_rules(orc):
    refine = np.array([1, 1, 0, 0, 1, 0], np.uint8)
    coarsen = np.array([0, 0, 1, 1, 0, 0], np.uint8)
    mask = np.array([0, 3, 0, 0, 5, 0], np.uint8)
    levels = np.array([[0, 0, 0], [1, 2, 0], [0, 0, 0], [0, 1, 0], [8, 0, 0], [2, 2, 2]])
    a, mo = orc.mark_elements(refine, coarsen, mask, levels, max_level=8, min_level=0)
    np.testing.assert_array_equal(a, [1, 1, 0, 2, 0, 0])
    np.testing.assert_array_equal(mo, [7, 3, 0, 0, 0, 0])


def _orthonormal_legendre(p, x):
    c = np.zeros(p + 1)
    c[p] = 1.0
    return np.sqrt((2 * p + 1) / 2.0) * np.polynomial.legendre.legval(x, c)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6, 8])
def test_modal_vandermonde_and_unit_modes(orc, order):
    n = order + 1
    x, _ = orc.lgl(n)
    V, Vi = orc.modal_matrices(order)
    Vn = np.polynomial.legendre.legvander(x, order) * np.sqrt((2 * np.arange(n) + 1) / 2.0)
    np.testing.assert_allclose(V, Vn, atol=1e-14)
    np.testing.assert_allclose(Vi @ V, np.eye(n), atol=1e-12)
    rng = np.random.default_rng(order)
    for _ in range(4):
        p, q, r = rng.integers(0, n, 3)
        u = np.einsum("k,j,i->kji", _orthonormal_legendre(r, x), _orthonormal_legendre(q, x),
                      _orthonormal_legendre(p, x)).reshape(1, -1)
        coef = orc.modal_coeffs(order, u)[0]
        e = np.zeros(n ** 3)
        e[p + n * (q + n * r)] = 1.0
        np.testing.assert_allclose(coef, e, atol=1e-12)


@pytest.mark.parametrize("order", [2, 3, 5])
def test_modal_parseval(orc, order):
    """sum of squared orthonormal coefficients = exact integral of u^2"""
    n = order + 1
    x, _ = orc.lgl(n)
    rng = np.random.default_rng(10 + order)
    C = rng.normal(size=(n, n, n))  # [r][q][p]
    P = np.stack([_orthonormal_legendre(p, x) for p in range(n)])  # [p][i]
    u = np.einsum("rqp,pi,qj,rk->kji", C, P, P, P).reshape(1, -1)
    coef = orc.modal_coeffs(order, u)[0]
    np.testing.assert_allclose(coef, C.reshape(-1), atol=1e-11)
    gx, gw = np.polynomial.legendre.leggauss(n + 1)
    Pg = np.stack([_orthonormal_legendre(p, gx) for p in range(n)])
    ug = np.einsum("rqp,pi,qj,rk->kji", C, Pg, Pg, Pg)
    integral = np.einsum("kji,i,j,k->", ug ** 2, gw, gw, gw)
    assert abs(integral - (coef ** 2).sum()) < 1e-10 * integral
