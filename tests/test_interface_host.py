"""Host parts of the face components in libdrifter_hip (no GPU needed):
LGL face maps and the mortar operators built by drf_mortar_create, against
the oracle (index maps bit-exact, operators to rounding)."""
import numpy as np
import pytest

from drifter_amd import interface as itf


@pytest.mark.parametrize("order", range(1, 9))
def test_face_map_bit_exact(orc, order):
    for face in range(6):
        np.testing.assert_array_equal(itf.face_map(order, face), orc.face_map(order, face))


def test_face_map_rejects_bad_face():
    from drifter_amd import DrfError

    with pytest.raises(DrfError):
        itf.face_map(3, 6)


@pytest.mark.parametrize("ctype", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("order", [1, 2, 3, 4, 6])
def test_mortar_operators_match_oracle(orc, order, ctype):
    ms = itf.MortarSpace(order, ctype, device=None)
    M, Mi, Pc, Lc, Pf, Lf = ms.operators()
    oM, oMi, oPc, oLc, oPf, oLf = orc.mortar_operators(order, ctype)
    assert ms.num_fine_elements() == len(oPf) and ms.num_dofs() == (order + 1) ** 2
    # same quadrature and basis arithmetic: mass and lift matrices identical
    np.testing.assert_array_equal(M, oM)
    np.testing.assert_array_equal(Lc, oLc)
    np.testing.assert_array_equal(Lf, oLf)
    # Cholesky here, Gauss-Jordan in the oracle: equal to rounding
    scale = np.abs(oMi).max()
    np.testing.assert_allclose(Mi, oMi, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(Pc, oPc, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Pf, oPf, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Mi @ M, np.eye(M.shape[0]), atol=1e-11)


def test_mortar_rejects_boundary_type():
    from drifter_amd import DrfError

    with pytest.raises(DrfError):
        itf.MortarSpace(3, itf.CONN_BOUNDARY, device=None)


def test_flux_entry_points_reject_bad_arguments():
    """argument errors are reported before any device work"""
    import ctypes as C

    from drifter_amd import _lib

    lib = _lib.load()
    p = np.array([9.81, 0, 0, 0])
    pd = p.ctypes.data_as(_lib._PD)
    assert lib.drf_flux_points(99, 10, None, None, None, pd, None, None) == _lib.DRF_ERR_ARG
    assert lib.drf_flux_points(0, 0, None, None, None, pd, None, None) == _lib.DRF_OK
    assert lib.drf_conforming_flux_batch(3, 0, pd, 5, None, None, None, None, None,
                                         None) == _lib.DRF_ERR_ARG  # not a scalar flux
    assert lib.drf_gather_face_traces(3, 10, 8, 1, None, None, None, None,
                                      None) == _lib.DRF_ERR_ARG  # stride < (N+1)^3
    m = C.c_void_p()
    assert lib.drf_mortar_create(3, 5, -1, C.byref(m)) == _lib.DRF_OK
    assert lib.drf_mortar_flux_batch(m, 5, pd, 1, None, None, None, None, None,
                                     None) == _lib.DRF_ERR_ARG  # null buffers
    one = np.zeros(64)
    vp = one.ctypes.data_as(C.c_void_p)
    assert lib.drf_mortar_flux_batch(m, 5, pd, 1, vp, vp, vp, vp, vp,
                                     None) == _lib.DRF_ERR_STATE  # host-only mortar
    lib.drf_mortar_destroy(m)


@pytest.mark.parametrize("ctype", [0, 1, 3, 5])
def test_verify_conservation_host_bit_exact(orc, ctype):
    """drf_verify_mortar_conservation against the oracle, same summation order"""
    ms = itf.MortarSpace(3, ctype, device=None)
    nm, nfine = ms.num_dofs(), ms.num_fine_elements()
    rng = np.random.default_rng(ctype)
    for _ in range(5):
        rc, rf = rng.normal(size=nm), rng.normal(size=(nfine, nm))
        wc, wf = rng.uniform(0, 1, nm), rng.uniform(0, 1, (nfine, nm))
        assert ms.verify_conservation(rc, rf, wc, wf) == orc.verify_mortar_conservation(
            rc, rf, wc, wf)
    assert ms.verify_conservation(np.zeros(nm), np.zeros((nfine, nm)), wc, wf) == 0.0
    with pytest.raises(ValueError):
        ms.verify_conservation(rc[:-1], rf, wc, wf)
