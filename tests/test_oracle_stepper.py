"""Pins the oracle's Runge-Kutta steps (A9) and dt control (A10).

A uniform flow with g = nu = kappa = 0 makes the reference RHS exactly the
linear Coriolis rotation (a deep column, H = 1e6, keeps u = Hu/H tiny so that
advection cannot amplify rounding noise within a step) dHu/dt = f Hv, dHv/dt = -f Hu (all derivatives act on
constants), i.e. w = Hu + i Hv obeys dw/dt = -i f w.  One step of a scheme is
then its stability polynomial P(z), z = -i f dt, applied to w -- a closed form
independent of the implementation (reference tests:
tests/unit/test_time_stepper.cpp:17-195).
"""
import math

import numpy as np
import pytest

F0 = 1e-4


def _uniform_state(order, E=2, Hu=3.0, Hv=-1.5, H=1.0e6):
    nd = (order + 1) ** 3
    blk = np.concatenate([np.full(nd, Hu), np.full(nd, Hv), np.zeros(nd), np.full(nd, 15 * H),
                          np.full(nd, 35 * H)])
    U = np.tile(blk, E)
    h = np.full(E * nd, H)
    f = np.full(E * nd, F0)
    return U, h, f


def _w(U, order):
    nd = (order + 1) ** 3
    B = U.reshape(-1, 5, nd)
    return B[:, 0, :] + 1j * B[:, 1, :]


POLY = {
    0: lambda z: 1 + z,
    1: lambda z: 1 + z + z ** 2 / 2,
    2: lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6,
    3: lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24,
}


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_stability_polynomial(orc, scheme):
    order = 2
    U, h, f = _uniform_state(order)
    phys = orc.physics(g=0.0, nu_h=0.0, nu_v=0.0, kappa_v_T=0.0, kappa_v_S=0.0)
    dt = 2000.0
    V = orc.rk_step(scheme, order, 0.0, dt, U, h, f, phys)
    w0, w1 = _w(U, order), _w(V, order)
    expect = POLY[scheme](-1j * F0 * dt) * w0
    # derivatives of constants leave O(eps |D| dt) residues; a wrong stage
    # coefficient shows at the O(z^p / p!) >= 1e-5 level (z = 0.2)
    assert np.max(np.abs(w1 - expect)) < 1e-11 * np.max(np.abs(w0))


@pytest.mark.parametrize("scheme,p", [(4, 4), (5, 5), (3, 4), (2, 3)])
def test_local_error_order(orc, scheme, p):
    order = 2
    U, h, f = _uniform_state(order)
    phys = orc.physics(g=0.0, nu_h=0.0, nu_v=0.0, kappa_v_T=0.0, kappa_v_S=0.0)
    errs = []
    for dt in (4000.0, 2000.0):
        V = orc.rk_step(scheme, order, 0.0, dt, U, h, f, phys)
        exact = np.exp(-1j * F0 * dt) * _w(U, order)
        errs.append(np.max(np.abs(_w(V, order) - exact)))
    ratio = errs[0] / errs[1]
    assert abs(math.log2(ratio) - (p + 1)) < 0.35, (ratio, p)


def test_rk4_one_step_decay_closed_form(orc):
    """du/dt = -u analogue: P(z) for real z from the reference's RK4 test."""
    z = -0.3
    assert abs(POLY[3](z) - (1 - 0.3 + 0.09 / 2 - 0.027 / 6 + 0.0081 / 24)) < 1e-15


def test_compute_dt_clamps(orc):
    # tests/unit/test_time_stepper.cpp:216-233
    assert orc.compute_dt(0.0, 0.5, 0.5, 1e-6, 3600.0) == 3600.0
    assert orc.compute_dt(1e-15, 0.5, 0.5, 1e-6, 3600.0) == 3600.0
    assert orc.compute_dt(1e12, 0.5, 0.5, 1e-6, 3600.0) == 1e-6
    assert orc.compute_dt(10.0, 0.5, 0.5, 1e-6, 3600.0) == 0.5 * 0.5 / 10.0


def test_max_wave_speed_is_per_element(orc):
    """max_e [sqrt(9.81 max H_e) + max(max|u_e|, max|v_e|)] -- not the
    combination of global maxima (simulation_driver.cpp:153-168)."""
    nd = 8
    H = np.concatenate([np.full(nd, 100.0), np.full(nd, 1.0)])
    u = np.concatenate([np.zeros(nd), np.full(nd, 5.0)])
    v = np.concatenate([np.full(nd, -0.5), np.zeros(nd)])
    c = orc.max_wave_speed(1, H, u, v)
    brute = max(math.sqrt(9.81 * 100.0) + 0.5, math.sqrt(9.81 * 1.0) + 5.0)
    assert c == brute
    assert c != math.sqrt(9.81 * 100.0) + 5.0
