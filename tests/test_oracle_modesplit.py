"""Pins the oracle's mode-split step (A11/A12) by closed forms.

* Quiescent column (u = v = 0, eta = 0): the step leaves eta = 0 and u = v = 0,
  sets H(ih*n + k) = h(ih) -- the bathymetry at DOF index ih, the reference's
  .head(n_h) layout (src/physics/mode_splitting.cpp:515) -- and rescales
  T = HT / H_new, HT = H_new T, with rho from the linear EOS (:591-598).
* Uniform transport, flat eta: the barotropic RHS reduces to the Coriolis
  rotation, so each Heun substep (:293-317) multiplies w = HU + i HV by
  1 + z + z^2/2 (z = -i f dt/n_sub).  After n_sub substeps u, v carry the
  rotated depth average (recombine :411-417), while Hu, Hv keep the
  un-rotated H (U_bar + u') of update_baroclinic (:531-536): recombine does
  not refresh them.
Reference invariant tests restated: tests/unit/test_mode_splitting.cpp:52-213.
"""
import numpy as np
import pytest

NF = 12
HU, HV, ETA, HT, HS, H, U, V, T, S, OM, RHO = range(12)


def _full(order, E, fill):
    nd = (order + 1) ** 3
    F = np.zeros((E, NF, nd))
    for k, val in fill.items():
        F[:, k] = val
    return F


@pytest.mark.parametrize("order", [1, 3, 4])
def test_quiescent_column(orc, order):
    nd = (order + 1) ** 3
    nh = (order + 1) ** 2
    E = 3
    rng = np.random.default_rng(order)
    h = 40.0 + 10.0 * rng.random((E, nd))
    Tn = 15.0 + rng.random((E, nd))
    Sn = 33.0 + rng.random((E, nd))
    F = _full(order, E, {H: h, HT: h * Tn, HS: h * Sn, T: Tn, S: Sn})
    out = orc.mode_split_step(order, 600.0, 7, F.reshape(-1).copy(), h.reshape(-1).copy(),
                              np.full(E * nd, 1e-4)).reshape(E, NF, nd)
    n = order + 1
    Hnew = np.repeat(h[:, :nh], n, axis=1)  # H(ih*n + k) = 0 + h(ih)
    assert np.max(np.abs(out[:, ETA])) == 0.0
    assert np.max(np.abs(out[:, [HU, HV, U, V, OM]])) == 0.0
    assert np.array_equal(out[:, H], Hnew)
    Tnew = (h * Tn) * (1.0 / Hnew)
    np.testing.assert_allclose(out[:, T], Tnew, rtol=1e-15)
    np.testing.assert_allclose(out[:, HT], Hnew * Tnew, rtol=1e-15)
    Snew = (h * Sn) * (1.0 / Hnew)
    rho = 1025.0 * (1.0 - 2e-4 * (Tnew - 10.0) + 7.6e-4 * (Snew - 35.0))
    np.testing.assert_allclose(out[:, RHO], rho, rtol=1e-14)


@pytest.mark.parametrize("nsub", [1, 5, 20])
def test_barotropic_rotation_and_stale_transport(orc, nsub):
    order, E = 2, 2
    nd = 27
    H0, u0, v0, f0 = 80.0, 0.2, -0.1, 1e-4
    F = _full(order, E, {H: H0, U: u0, V: v0, HU: H0 * u0, HV: H0 * v0, T: 15.0, S: 35.0,
                         HT: H0 * 15.0, HS: H0 * 35.0})
    h = np.full(E * nd, H0)
    dt3 = 1800.0
    # g = 0: with the large dt used here, gravity waves would amplify the
    # O(eps) residues of D acting on constants (CFL >> 1 in reference units)
    out = orc.mode_split_step(order, dt3, nsub, F.reshape(-1).copy(), h,
                              np.full(E * nd, f0), orc.physics(g=0.0)).reshape(E, NF, nd)
    z = -1j * f0 * dt3 / nsub
    w = (1 + z + z * z / 2) ** nsub * (H0 * u0 + 1j * H0 * v0)
    # depth-averaged velocity after recombine = rotated transport / H
    np.testing.assert_allclose(out[:, U], w.real / H0, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out[:, V], w.imag / H0, rtol=1e-9, atol=1e-12)
    # Hu, Hv are NOT rotated: H_new * (U_bar + u') with the original U_bar
    np.testing.assert_allclose(out[:, HU], H0 * u0, rtol=1e-12)
    np.testing.assert_allclose(out[:, HV], H0 * v0, rtol=1e-12)
    assert np.max(np.abs(out[:, ETA])) < 1e-10
    assert abs(w.real / H0 - u0) > 1e-3  # the rotation is visible at this dt
