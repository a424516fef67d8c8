"""Pins the oracle's SigmaCoordinateOperators (oracle/sigmaops.c, restating
src/dg/operators_3d.cpp:513-662) by closed forms: reference-coordinate
derivatives of polynomials of degree <= N are exact, sigma at the nodes is
(zeta - 1) / 2 of the node's third coordinate, the metric update's own sign
convention ((sigma dH/dx - deta/dx) / H, the opposite sign on the sigma term
to SigmaCoordinate::dsigma_dx, :608), and every operator against its formula
applied to the exact polynomial derivatives."""
import numpy as np
import pytest


def _poly_field(orc, order, rng, E):
    """E elements of a random polynomial of degree <= N in (xi, eta, zeta)
    with its exact reference derivatives."""
    n = order + 1
    x = orc.lgl(n)[0]
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    X, Y, Z = X.reshape(-1), Y.reshape(-1), Z.reshape(-1)
    U, dX, dY, dZ = [], [], [], []
    for _ in range(E):
        a = rng.standard_normal((3, n))
        px, py, pz = (np.polynomial.Polynomial(c) for c in a)
        U.append(px(X) * py(Y) * pz(Z))
        dX.append(px.deriv()(X) * py(Y) * pz(Z))
        dY.append(px(X) * py.deriv()(Y) * pz(Z))
        dZ.append(px(X) * py(Y) * pz.deriv()(Z))
    return [np.array(t) for t in (U, dX, dY, dZ)], Z


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6])
def test_gradients_and_divergence_exact_on_polynomials(orc, order):
    rng = np.random.default_rng(order)
    E = 4
    (U, dX, dY, dZ), _ = _poly_field(orc, order, rng, E)
    nd = U.shape[1]
    dsdx, dsdy, Hinv, w = rng.standard_normal((4, E, nd))
    tol = 1e-11 * max(1.0, np.abs(dZ).max())
    gx, gy = orc.sigma_ops_apply(order, orc.SIGMA_OP_HGRAD, U, dsdx=dsdx, dsdy=dsdy)
    np.testing.assert_allclose(gx, dX - dsdx * dZ, rtol=0, atol=tol)
    np.testing.assert_allclose(gy, dY - dsdy * dZ, rtol=0, atol=tol)
    gz = orc.sigma_ops_apply(order, orc.SIGMA_OP_VGRAD, U, Hinv=Hinv)
    np.testing.assert_allclose(gz, Hinv * dZ, rtol=0, atol=tol)
    ale = orc.sigma_ops_apply(order, orc.SIGMA_OP_ALE, U, a=w, Hinv=Hinv)
    np.testing.assert_allclose(ale, -w * Hinv * dZ, rtol=0, atol=tol)
    # divergence of three fields: d(Hu)/dxi + d(Hv)/deta + d(omega)/dzeta
    (V, vX, vY, vZ), _ = _poly_field(orc, order, rng, E)
    (W, wX, wY, wZ), _ = _poly_field(orc, order, rng, E)
    div = orc.sigma_ops_apply(order, orc.SIGMA_OP_DIV, U, a=V, b=W)
    np.testing.assert_allclose(div, dX + vY + wZ, rtol=0,
                               atol=1e-11 * max(1.0, np.abs(dX + vY + wZ).max()))
    # material derivative
    dUdt, u, v, om, wm = rng.standard_normal((5, E, nd))
    md = orc.sigma_ops_apply(order, orc.SIGMA_OP_MATDER, U, a=dUdt, b=u, c=v, f=om, g=wm,
                             dsdx=dsdx, dsdy=dsdy, Hinv=Hinv)
    want = dUdt + u * (dX - dsdx * dZ) + v * (dY - dsdy * dZ) + (om - wm * Hinv) * dZ
    np.testing.assert_allclose(md, want, rtol=0, atol=10 * tol)


def test_metrics_sigma_and_sign_convention(orc):
    order, n = 3, 4
    nd = n ** 3
    E = 2
    x = orc.lgl(n)[0]
    rng = np.random.default_rng(3)
    eta = rng.uniform(-0.5, 0.5, (E, nd))
    h = rng.uniform(10, 20, (E, nd))
    z = np.zeros((E, nd))
    m = orc.sigma_ops_metrics(order, eta, h, z, z, z, z)
    kz = np.arange(nd) // (n * n)
    np.testing.assert_array_equal(m["sigma"][0], 0.5 * (x[kz] - 1.0))
    np.testing.assert_array_equal(m["H"], eta + h)
    # flat surface and bottom: no slope of the sigma surfaces
    assert np.all(m["dsdx"] == 0) and np.all(m["dsdy"] == 0) and np.all(m["dsdt"] == 0)
    # sloping bottom only: this variant gives +sigma dh/dx / H (the tested
    # SigmaCoordinate formula gives -sigma dh/dx / H)
    dhx = np.full((E, nd), 0.01)
    m2 = orc.sigma_ops_metrics(order, eta, h, z, z, dhx, z)
    np.testing.assert_allclose(m2["dsdx"], m["sigma"] * 0.01 / (eta + h), rtol=1e-15)
    # surface slope only: -deta/dx (1 + ... ) -> (sigma - 1) deta/dx / H
    ex = np.full((E, nd), 0.002)
    m3 = orc.sigma_ops_metrics(order, eta, h, ex, z, z, z)
    np.testing.assert_allclose(m3["dsdx"], (m["sigma"] - 1.0) * 0.002 / (eta + h), rtol=1e-14)
    # time derivative: -(1 + sigma) deta/dt / H, zero at the bottom node
    et = np.full((E, nd), 0.3)
    m4 = orc.sigma_ops_metrics(order, eta, h, z, z, z, z, deta_dt=et)
    np.testing.assert_allclose(m4["dsdt"], -(1 + m["sigma"]) * 0.3 / (eta + h), rtol=1e-15)
    assert np.all(m4["dsdt"][:, : n * n] == 0)
    # dry node: Hinv = 0 and every metric 0
    eta2 = eta.copy()
    eta2[0, 5] = -h[0, 5]
    m5 = orc.sigma_ops_metrics(order, eta2, h, ex, ex, dhx, dhx, deta_dt=et)
    assert m5["Hinv"][0, 5] == 0 and m5["dsdx"][0, 5] == 0 and m5["dsdt"][0, 5] == 0
