"""GPU parity of the vertical velocity diagnosis against the oracle,
bit-exact (FP64, the reference's sequence of operations):
VerticalVelocityDiagnosis and ColumnVerticalVelocity
(src/physics/vertical_velocity.cpp:55-275)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _d(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("order", range(1, 9))
def test_diagnose_omega_bit_exact(torch, orc, order):
    from drifter_amd.vertical import VerticalVelocityDiagnosis

    nd = (order + 1) ** 3
    rng = np.random.default_rng(order)
    E = 517
    div = rng.standard_normal((E, nd))
    H = rng.uniform(0.5, 40, size=(E, nd))
    H[3, : order + 1] = 0.0  # a dry column
    print(f"[omega] order={order} E={E}", flush=True)
    vv = VerticalVelocityDiagnosis(order)
    got = vv.diagnose_omega(_d(torch, div), _d(torch, H)).cpu().numpy()
    np.testing.assert_array_equal(got, orc.diagnose_omega(order, div, H))
    a, b = rng.standard_normal((2, E, nd))
    got2 = vv.diagnose_omega_from_velocity(_d(torch, a), _d(torch, b), _d(torch, H)).cpu().numpy()
    np.testing.assert_array_equal(got2, orc.diagnose_omega_from_velocity(order, a, b, H))
    r = vv.continuity_residual(_d(torch, div), _d(torch, a)).cpu().numpy()
    np.testing.assert_array_equal(r, orc.continuity_residual(order, div, a))


def test_strided_field_view(torch, orc):
    """One field of a packed [e][field][ndof] state (element stride 5 ndof)."""
    from drifter_amd.vertical import VerticalVelocityDiagnosis

    order, nd = 3, 64
    rng = np.random.default_rng(9)
    full = rng.standard_normal((40, 5, nd))
    Hf = rng.uniform(1, 9, size=(40, 5, nd))
    print("[omega] strided", flush=True)
    vv = VerticalVelocityDiagnosis(order)
    got = vv.diagnose_omega(_d(torch, full)[:, 1, :], _d(torch, Hf)[:, 1, :])
    assert got.stride(0) == 5 * nd
    want = orc.diagnose_omega(order, full[:, 1].copy(), Hf[:, 1].copy())
    np.testing.assert_array_equal(got.cpu().numpy(), want)
    # fields of different element strides are refused before any device work
    with pytest.raises(ValueError):
        vv.diagnose_omega(_d(torch, full)[:, 1, :], _d(torch, Hf[:, 1].copy()))


def test_w_conversions_bit_exact(torch, orc):
    from drifter_amd.vertical import VerticalVelocityDiagnosis as V

    rng = np.random.default_rng(3)
    m = 100_003
    a = rng.standard_normal((9, m))
    a[5] = rng.uniform(-1, 0, m)  # sigma
    print("[omega] w conversions", flush=True)
    w = V.omega_to_w_simple(*[_d(torch, x) for x in a]).cpu().numpy()
    np.testing.assert_array_equal(w, orc.omega_to_w_simple(*a))
    w2 = V.omega_to_physical_w(*[_d(torch, x) for x in a[:7]]).cpu().numpy()
    np.testing.assert_array_equal(w2, orc.omega_to_w(*a[:7]))
    assert V.surface_omega(0.1, 1.0, 2.0, 0.01, 0.02, 5.0) == (0.1 + 1.0 * 0.01 + 2.0 * 0.02) / 5.0
    assert V.surface_omega(0.1, 1.0, 2.0, 0.01, 0.02, 0.0) == 0.0


@pytest.mark.parametrize("order", range(1, 9))
def test_column_bit_exact(torch, orc, order):
    from drifter_amd.vertical import ColumnVerticalVelocity

    n = order + 1
    rng = np.random.default_rng(order + 40)
    ncol = 10_001
    div = rng.standard_normal((ncol, n))
    H = rng.uniform(0.5, 30, size=(ncol, n))
    H[7] = 0.0
    ob = rng.standard_normal(ncol)
    cv = ColumnVerticalVelocity(order)
    np.testing.assert_array_equal(cv.sigma_nodes, orc.column_sigma_nodes(order))
    print(f"[omega] column order={order}", flush=True)
    got = cv.diagnose_column(_d(torch, div), _d(torch, H), _d(torch, ob)).cpu().numpy()
    np.testing.assert_array_equal(got, orc.diagnose_column(order, div, H, ob))
    got0 = cv.diagnose_column(_d(torch, div), _d(torch, H)).cpu().numpy()
    np.testing.assert_array_equal(got0, orc.diagnose_column(order, div, H))
