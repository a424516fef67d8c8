"""Pins the oracle's 1D bases (A1) against closed forms and independent
libraries, and checks that the product's host basis is bit-identical.

Reference tests restated: tests/unit/test_basis_hexahedron.cpp:16-102 and the
closed-form nodes in tests/test_utils.hpp:43-72.
"""
import math

import numpy as np
import pytest


def test_lgl_closed_forms(orc):
    x, w = orc.lgl(4)
    s = math.sqrt(1.0 / 5.0)
    np.testing.assert_allclose(x, [-1, -s, s, 1], rtol=0, atol=1e-15)
    np.testing.assert_allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6], rtol=0, atol=1e-15)
    x, w = orc.lgl(5)
    s = math.sqrt(3.0 / 7.0)
    np.testing.assert_allclose(x, [-1, -s, 0, s, 1], rtol=0, atol=1e-15)
    np.testing.assert_allclose(w, [0.1, 49 / 90, 32 / 45, 49 / 90, 0.1], rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", range(2, 10))
def test_lgl_weights_and_endpoints(orc, n):
    x, w = orc.lgl(n)
    assert x[0] == -1.0 and x[-1] == 1.0
    assert abs(w.sum() - 2.0) < 1e-14
    assert np.all(np.diff(x) > 0)
    # interior LGL nodes are the roots of P'_{n-1}
    N = n - 1
    if n > 2:
        dP = np.polynomial.legendre.Legendre.basis(N).deriv()
        assert np.max(np.abs(dP(x[1:-1]))) < 1e-12 * max(1.0, np.max(np.abs(dP.coef)))
    # LGL quadrature is exact to degree 2n-3
    for deg in range(0, 2 * n - 2):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-13


@pytest.mark.parametrize("n", range(1, 10))
def test_gl_matches_numpy(orc, n):
    x, w = orc.gl(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, xr, rtol=0, atol=1e-14)
    np.testing.assert_allclose(w, wr, rtol=0, atol=1e-14)
    assert abs(w.sum() - 2.0) < 1e-14


@pytest.mark.parametrize("order", range(1, 9))
def test_derivative_exact_on_polynomials(orc, order):
    n = order + 1
    x, _ = orc.lgl(n)
    D = orc.dmat(x)
    rng = np.random.default_rng(order)
    for deg in range(0, order + 1):
        c = rng.standard_normal(deg + 1)
        p = np.polynomial.Polynomial(c)
        np.testing.assert_allclose(D @ p(x), p.deriv()(x), rtol=0, atol=2e-12 * (1 + np.abs(c).sum()))
    # rows of D annihilate constants
    assert np.max(np.abs(D.sum(axis=1))) < 1e-12


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_kronecker_structure(orc, order):
    n = order + 1
    x, _ = orc.lgl(n)
    D = orc.dmat(x)
    Dxi, Deta, Dzeta = orc.kron(order, D)
    I = np.eye(n)
    # d = i + n (j + n k): xi is the fastest index
    assert np.array_equal(Dxi, np.kron(I, np.kron(I, D)))
    assert np.array_equal(Deta, np.kron(I, np.kron(D, I)))
    assert np.array_equal(Dzeta, np.kron(D, np.kron(I, I)))


def test_lgl_rejects_single_node(orc):
    # std::invalid_argument("LGL requires at least 2 nodes"), basis_hexahedron.cpp:74-76
    x = np.zeros(1)
    w = np.zeros(1)
    assert orc.load().orc_lgl_nodes(1, orc._p(x), orc._p(w)) != 0


@pytest.mark.parametrize("order", range(1, 9))
def test_product_basis_bit_identical(orc, order):
    from drifter_amd import synth

    x, w, D = synth.lgl_nodes(order)
    xo, wo = orc.lgl(order + 1)
    assert np.array_equal(x, xo)
    assert np.array_equal(w, wo)
    assert np.array_equal(D, orc.dmat(xo))
