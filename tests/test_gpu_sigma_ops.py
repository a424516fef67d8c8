"""GPU parity of SigmaCoordinateOperators (src/dg/operators_3d.cpp:513-662)
against the oracle, bit-exact: the kernels contract element lines in the
order of the reference's dense Kronecker products, with contraction off."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 8])
def test_sigma_operators_bit_exact(torch, orc, order):
    from drifter_amd.sigma import SigmaCoordinateOperators

    nd = (order + 1) ** 3
    rng = np.random.default_rng(order)
    E = 301
    eta = rng.uniform(-0.5, 0.5, (E, nd))
    h = rng.uniform(5, 50, (E, nd))
    h[2, 7] = -eta[2, 7]  # a dry node
    ex, ey, hx, hy, et = rng.standard_normal((5, E, nd)) * 0.01
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    print(f"[sigma ops] order={order} E={E}", flush=True)
    ops = SigmaCoordinateOperators(order)
    ops.update_sigma_metrics(d(eta), d(h), d(ex), d(ey), d(hx), d(hy), d(et))
    m = orc.sigma_ops_metrics(order, eta, h, ex, ey, hx, hy, deta_dt=et)
    for k, t in (("H", ops.H), ("Hinv", ops.H_inv), ("sigma", ops.sigma),
                 ("dsdx", ops.dsigma_dx), ("dsdy", ops.dsigma_dy), ("dsdt", ops.dsigma_dt)):
        np.testing.assert_array_equal(t.cpu().numpy(), m[k], err_msg=k)
    U, a, b, c, f, g = rng.standard_normal((6, E, nd))
    gx, gy = ops.horizontal_gradient(d(U))
    wx, wy = orc.sigma_ops_apply(order, orc.SIGMA_OP_HGRAD, U, dsdx=m["dsdx"], dsdy=m["dsdy"])
    np.testing.assert_array_equal(gx.cpu().numpy(), wx)
    np.testing.assert_array_equal(gy.cpu().numpy(), wy)
    np.testing.assert_array_equal(ops.vertical_gradient(d(U)).cpu().numpy(),
                                  orc.sigma_ops_apply(order, orc.SIGMA_OP_VGRAD, U,
                                                      Hinv=m["Hinv"]))
    np.testing.assert_array_equal(ops.sigma_divergence(d(U), d(a), d(b)).cpu().numpy(),
                                  orc.sigma_ops_apply(order, orc.SIGMA_OP_DIV, U, a=a, b=b))
    np.testing.assert_array_equal(ops.ale_correction(d(U), d(a)).cpu().numpy(),
                                  orc.sigma_ops_apply(order, orc.SIGMA_OP_ALE, U, a=a,
                                                      Hinv=m["Hinv"]))
    np.testing.assert_array_equal(
        ops.material_derivative(d(U), d(a), d(b), d(c), d(f), d(g)).cpu().numpy(),
        orc.sigma_ops_apply(order, orc.SIGMA_OP_MATDER, U, a=a, b=b, c=c, f=f, g=g,
                            dsdx=m["dsdx"], dsdy=m["dsdy"], Hinv=m["Hinv"]))


def test_steady_metrics_and_strided_fields(torch, orc):
    """update_sigma_metrics without deta/dt (dsigma/dt = 0), and fields that
    are one variable of a packed [e][var][ndof] state."""
    from drifter_amd.sigma import SigmaCoordinateOperators

    order, nd = 3, 64
    rng = np.random.default_rng(11)
    full = rng.standard_normal((50, 8, nd))
    full[:, 1] = np.abs(full[:, 1]) + 10.0  # h
    dv = torch.from_numpy(full).cuda()
    f = [dv[:, k, :] for k in range(8)]
    print("[sigma ops] steady, strided", flush=True)
    ops = SigmaCoordinateOperators(order)
    ops.update_sigma_metrics(*f[:6])
    m = orc.sigma_ops_metrics(order, *[full[:, k].copy() for k in range(6)])
    np.testing.assert_array_equal(ops.dsigma_dx.cpu().numpy(), m["dsdx"])
    assert np.all(ops.dsigma_dt.cpu().numpy() == 0)
    gx, _ = ops.horizontal_gradient(f[6])
    assert gx.stride(0) == 8 * nd
    wx, _ = orc.sigma_ops_apply(order, orc.SIGMA_OP_HGRAD, full[:, 6].copy(), dsdx=m["dsdx"],
                                dsdy=m["dsdy"])
    np.testing.assert_array_equal(gx.cpu().numpy(), wx)
    with pytest.raises(ValueError):  # mixed strides are refused before any device work
        ops.sigma_divergence(f[6], f[7].contiguous(), f[5])
