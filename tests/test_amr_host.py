"""Host parts of the solution transfer (no GPU needed): the leaf map onto a
refined mesh (drf_refine_transfer_map) built from the octree's own bounds,
argument checks of the device entry points, and the parents check of the
Python wrapper."""
import numpy as np
import pytest

from drifter_amd import _lib, amr
from drifter_amd.mesh import OctreeAdapter


def _mesh(nx=2, ny=2, nz=2):
    m = OctreeAdapter(0.0, 1.0, 0.0, 3.0, -1.0, 0.0)
    m.build_uniform(nx, ny, nz)
    return m


def _check_boxes(old_b, new_b, src, boxes):
    """Each new leaf's bounds are its box mapped through its source leaf."""
    for k in range(new_b.shape[0]):
        ob = old_b[src[k]]
        for a in range(3):
            lo, hi = ob[2 * a], ob[2 * a + 1]
            m = lo + 0.5 * (boxes[k, 2 * a:2 * a + 2] + 1.0) * (hi - lo)
            np.testing.assert_allclose(m, new_b[k, 2 * a:2 * a + 2], rtol=0, atol=1e-13)


def test_map_after_xyz_refinement_matches_project_to_children():
    m = _mesh()
    old_b, _, _ = m.elements()
    m.refine([3], [amr.MASK_XYZ])
    new_b, _, _ = m.elements()
    src, boxes = amr.refine_transfer_map(old_b, new_b)
    assert new_b.shape[0] == 8 + 7
    _check_boxes(old_b, new_b, src, boxes)
    # depth-first numbering: the children take the parent's place
    np.testing.assert_array_equal(src, [0, 1, 2] + [3] * 8 + [4, 5, 6, 7])
    full = np.array([-1, 1, -1, 1, -1, 1.0])
    for k in (0, 1, 2, 11, 12, 13, 14):
        np.testing.assert_array_equal(boxes[k], full)  # exact: copies
    # children in creation order x fastest, exactly the project_to_children boxes
    for c in range(8):
        h = [(c >> a) & 1 for a in range(3)]
        want = np.array([v for a in range(3) for v in (-1.0 + h[a], float(h[a]))])
        np.testing.assert_array_equal(boxes[3 + c], want)


@pytest.mark.parametrize("mask", [1, 2, 3, 4, 5, 6])
def test_map_with_anisotropic_masks(mask):
    m = _mesh(2, 1, 1)
    old_b, _, _ = m.elements()
    m.refine([0], [mask])
    new_b, _, _ = m.elements()
    src, boxes = amr.refine_transfer_map(old_b, new_b)
    _check_boxes(old_b, new_b, src, boxes)
    nch = amr.num_children(mask)
    assert (src == 0).sum() == nch
    # the children of element 0 tile its cube: box volumes add up to 8
    vol = np.prod(boxes[src == 0][:, 1::2] - boxes[src == 0][:, ::2], axis=1)
    assert vol.sum() == 8.0


def test_map_through_balance_and_two_levels():
    m = _mesh(3, 3, 1)
    old_b, _, _ = m.elements()
    m.refine([4], [amr.MASK_XYZ])
    E1 = m.num_elements()
    m.refine([4], [amr.MASK_XYZ])  # a child of the first refinement
    new_b, _, _ = m.elements()
    assert m.num_elements() >= E1 + 7
    src, boxes = amr.refine_transfer_map(old_b, new_b)
    _check_boxes(old_b, new_b, src, boxes)
    assert np.all(np.diff(src) >= 0)
    # quarter-width boxes appear (two levels below an old leaf)
    w = boxes[:, 1] - boxes[:, 0]
    assert np.isclose(w, 0.5).any()


def test_map_rejects_a_non_refinement():
    m = _mesh()
    old_b, _, _ = m.elements()
    shuffled = old_b[::-1].copy()
    with pytest.raises(_lib.DrfError):
        amr.refine_transfer_map(old_b, shuffled)


def test_entry_points_reject_bad_arguments():
    lib = _lib.load()
    assert lib.drf_project_to_children_dev(3, 0, 10, None, 64, None, 64, None) == _lib.DRF_ERR_ARG
    assert lib.drf_project_to_children_dev(3, 8, 10, None, 64, None, 64, None) == _lib.DRF_ERR_ARG
    assert lib.drf_project_to_children_dev(3, 7, 10, None, 63, None, 64, None) == _lib.DRF_ERR_ARG
    assert lib.drf_project_to_children_dev(9, 7, 10, None, 1000, None, 1000,
                                           None) == _lib.DRF_ERR_UNSUPPORTED
    assert lib.drf_project_from_children_dev(3, 7, 10, None, 64, None, 32,
                                             None) == _lib.DRF_ERR_ARG
    assert lib.drf_transfer_solution_dev(3, 10, None, 64, 11, None, None, 64,
                                         None) == _lib.DRF_ERR_ARG  # more parents than elements
    assert lib.drf_transfer_solution_dev(3, 10, None, 64, 2, None, None, 64,
                                         None) == _lib.DRF_ERR_ARG  # null pointers
    assert lib.drf_interpolate_boxes_dev(3, 5, 10, None, None, None, 64, None, 64,
                                         None) == _lib.DRF_ERR_ARG
    # empty batches are no-ops before any device work
    assert lib.drf_project_to_children_dev(3, 7, 0, None, 64, None, 64, None) == _lib.DRF_OK
    assert lib.drf_transfer_solution_dev(3, 0, None, 64, 0, None, None, 64, None) == _lib.DRF_OK


@pytest.mark.parametrize("parents", [[2, 1], [1, 1], [-1], [10]])
def test_transfer_wrapper_rejects_bad_parents(parents):
    torch = pytest.importorskip("torch")
    U = torch.zeros((10, 8), dtype=torch.float64)
    with pytest.raises(ValueError):
        amr.transfer_solution(1, U, parents)
