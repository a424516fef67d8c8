"""Host parts of the refinement-side library calls (no GPU needed):
characteristic lengths and the modal Vandermonde against the oracle."""
import numpy as np
import pytest

from drifter_amd import adapt


def test_char_length_matches_cbrt_of_volume():
    rng = np.random.default_rng(0)
    lo = rng.uniform(-5, 5, size=(100, 3))
    sz = rng.uniform(0.01, 3, size=(100, 3))
    b = np.stack([lo[:, 0], lo[:, 0] + sz[:, 0], lo[:, 1], lo[:, 1] + sz[:, 1],
                  lo[:, 2], lo[:, 2] + sz[:, 2]], axis=1)
    h = adapt.element_char_length(b)
    ref = np.cbrt((b[:, 1] - b[:, 0]) * (b[:, 3] - b[:, 2]) * (b[:, 5] - b[:, 4]))
    np.testing.assert_allclose(h, ref, rtol=4.5e-16, atol=0)  # numpy and libm cbrt: 1 ulp


@pytest.mark.parametrize("order", range(1, 9))
def test_modal_matrices_match_oracle(orc, order):
    V, Vi = adapt.modal_matrices(order)
    oV, oVi = orc.modal_matrices(order)
    np.testing.assert_allclose(V, oV, rtol=0, atol=1e-14)
    np.testing.assert_allclose(Vi, oVi, rtol=0, atol=1e-13)
    np.testing.assert_allclose(Vi @ V, np.eye(order + 1), atol=1e-13)


def test_estimator_entry_points_reject_bad_arguments():
    import ctypes as C

    from drifter_amd import _lib

    lib = _lib.load()
    assert lib.drf_estimate_dev(None, 10, None, 64, None, None, None, None, None, None, None,
                                None, None, None) == _lib.DRF_ERR_ARG
    assert lib.drf_modal_coeffs_dev(3, 10, None, 63, None, None) == _lib.DRF_ERR_ARG
    assert lib.drf_modal_coeffs_dev(9, 10, None, 1000, None, None) != _lib.DRF_OK
    assert lib.drf_modal_features_bf16_dev(3, 10, None, 64, 1.0, None, None, 100, 40,
                                           None) == _lib.DRF_ERR_ARG  # 40 + 64 > 100
    assert lib.drf_modal_coeffs_dev(3, 0, None, 64, None, None) == _lib.DRF_OK
