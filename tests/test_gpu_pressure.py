"""Hydrostatic pressure and pressure gradient (pressure.hip) against the
oracle (oracle/pressure.c, the algorithm of
include/physics/pressure_gradient.hpp:46-367) on random fields: every method
and part, orders 1-7, within 1e-13 relative (the line contractions add the
dense Kronecker rows' non-zero terms in their order, contraction off)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(order, E, seed):
    rng = np.random.default_rng(seed)
    nd = (order + 1) ** 3
    r = lambda lo, hi: rng.uniform(lo, hi, E * nd)  # noqa: E731
    return dict(eta=r(-0.1, 0.1), rho=r(1020.0, 1030.0), H=r(20.0, 60.0), sigma=r(-1.0, 0.0),
                deta_dx=r(-1e-3, 1e-3), deta_dy=r(-1e-3, 1e-3), dh_dx=r(-1e-2, 1e-2),
                dh_dy=r(-1e-2, 1e-2))


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 7])
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("parts", [1, 2, 3])
def test_pressure_gradient_against_oracle(orc, order, method, parts):
    import torch

    from drifter_amd import pressure

    E = 37
    F = _fields(order, E, 10 * order + method)
    T = {k: torch.from_numpy(v).cuda() for k, v in F.items()}
    pg = pressure.PressureGradientSolver(order, method)
    gx, gy = pg.compute(parts=parts, **T)
    torch.cuda.synchronize()
    wx, wy = orc.pressure_gradient(order, method, parts, F["eta"], F["rho"], F["H"], F["sigma"],
                                   F["deta_dx"], F["deta_dy"], F["dh_dx"], F["dh_dy"])
    for got, want in ((gx, wx), (gy, wy)):
        g = got.cpu().numpy()
        assert np.abs(g - want).max() <= 1e-13 * np.abs(want).max()
    # the metric inputs are optional (0 when absent)
    if method == 0:
        gx, _ = pg.compute(T["eta"], T["rho"], T["H"], T["sigma"], parts=parts)
        wx, _ = orc.pressure_gradient(order, method, parts, F["eta"], F["rho"], F["H"],
                                      F["sigma"])
        assert np.abs(gx.cpu().numpy() - wx).max() <= 1e-13 * np.abs(wx).max()


@pytest.mark.parametrize("order", [2, 4])
def test_hydrostatic_pressure_against_oracle(orc, order):
    import torch

    from drifter_amd import pressure

    F = _fields(order, 5000, 3)
    pg = pressure.PressureGradientSolver(order)
    p = pg.hydrostatic_pressure(torch.from_numpy(F["rho"]).cuda(), torch.from_numpy(F["H"]).cuda())
    want = orc.hydrostatic_pressure(order, F["rho"], F["H"])
    assert np.array_equal(p.cpu().numpy(), want)
