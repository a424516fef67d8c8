"""Connectivity parity (A17): the product's octree (drf_mesh_*, host C++)
against the oracle's restatement of OctreeAdapter, bit-exact on element
bounds, levels, Morton codes and every FaceConnection field, of the
reference's connectivity and of the 2:1 connectivity the face plans use.
Host-only code: runs without a GPU."""
import numpy as np
import pytest

from drifter_amd import mesh


def _compare(prod, orc_mesh):
    b1, l1, m1 = prod.elements()
    b2, l2, m2 = orc_mesh.elements()
    assert b1.shape == b2.shape
    assert np.array_equal(b1, b2)  # bit-exact FP64 bounds
    assert np.array_equal(l1, l2)
    assert np.array_equal(m1, m2)
    for two in (False, True):
        c1 = prod.build_face_connections(two_to_one=two)
        c2 = orc_mesh.connections(two_to_one=two)
        for a, b in zip(c1, c2):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 2, 1), (8, 8, 4), (5, 3, 2), (64, 64, 8)])
def test_uniform(orc, dims):
    p = mesh.OctreeAdapter(0.0, 10000.0, 0.0, 10000.0, -1.0, 0.0)
    p.build_uniform(*dims)
    o = orc.Mesh(0.0, 10000.0, 0.0, 10000.0, -1.0, 0.0)
    o.build_uniform(*dims)
    _compare(p, o)


def test_refine_balance_sequence(orc):
    rng = np.random.default_rng(7)
    p = mesh.OctreeAdapter(0.0, 1.0, 0.0, 3.0, -1.0, 0.0)  # non-dyadic aspect
    o = orc.Mesh(0.0, 1.0, 0.0, 3.0, -1.0, 0.0)
    p.build_uniform(4, 4, 2)
    o.build_uniform(4, 4, 2)
    for _ in range(4):
        E = p.num_elements()
        sel = np.unique(rng.integers(0, E, size=max(1, E // 8)))
        masks = rng.integers(1, 8, size=sel.size)
        p.refine(sel, masks)
        o.refine(sel, masks)
        _compare(p, o)


def test_coastline_mesh_config3(orc):
    """Base 32x32x4, three XY refinement rounds within the coastline band."""
    p = mesh.coastline_mesh(band=7000.0)
    o = orc.Mesh(0.0, 100000.0, 0.0, 100000.0, -1.0, 0.0)
    o.build_uniform(32, 32, 4)
    L = 100000.0
    for _ in range(3):
        b, _, _ = o.elements()
        cx = 0.5 * (b[:, 0] + b[:, 1])
        cy = 0.5 * (b[:, 2] + b[:, 3])
        yc = 0.8 * L + 0.05 * L * np.sin(2.0 * np.pi * cx / (0.25 * L))
        sel = np.nonzero(np.abs(cy - yc) < 7000.0)[0]
        o.refine(sel, np.full(sel.size, 3))
    _compare(p, o)
    assert p.num_elements() == 39280


def test_out_of_range_element():
    from drifter_amd import _lib

    p = mesh.OctreeAdapter(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    p.build_uniform(2, 2, 2)
    with pytest.raises(_lib.DrfError) as ei:
        p.element_bounds(8)
    assert ei.value.status == _lib.DRF_ERR_RANGE


def test_refine_balance_sequence_large(orc):
    """Random anisotropic refinement of a 32x32x4 base over three rounds
    (beyond 4,096 leaves, where balance filters and connectivity run on
    several threads), with the balance cascades that mixed masks cause."""
    rng = np.random.default_rng(11)
    p = mesh.OctreeAdapter(0.0, 2.0, 0.0, 1.0, -1.0, 0.0)
    o = orc.Mesh(0.0, 2.0, 0.0, 1.0, -1.0, 0.0)
    p.build_uniform(32, 32, 4)
    o.build_uniform(32, 32, 4)
    for _ in range(3):
        E = p.num_elements()
        sel = np.unique(rng.integers(0, E, size=E // 6))
        masks = rng.integers(1, 8, size=sel.size)
        p.refine(sel, masks)
        o.refine(sel, masks)
    assert p.num_elements() > 2 * 4096
    _compare(p, o)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_deep_local_refinement_cascades(orc, seed):
    """Refining the leaves near one point again and again opens level gaps of
    several levels at once, so balance needs several passes and leaves that
    split a neighbour must be checked again in the next pass."""
    rng = np.random.default_rng(seed)
    p = mesh.OctreeAdapter(0.0, 1.0, 0.0, 1.0, -1.0, 0.0)
    o = orc.Mesh(0.0, 1.0, 0.0, 1.0, -1.0, 0.0)
    p.build_uniform(32, 32, 4)
    o.build_uniform(32, 32, 4)
    c = rng.uniform(0.2, 0.8, 2)
    for r in range(4):
        b, _, _ = p.elements()
        cx, cy = 0.5 * (b[:, 0] + b[:, 1]), 0.5 * (b[:, 2] + b[:, 3])
        sel = np.nonzero((np.abs(cx - c[0]) < 0.02) & (np.abs(cy - c[1]) < 0.02))[0]
        masks = rng.integers(1, 8, size=sel.size)
        p.refine(sel, masks)
        o.refine(sel, masks)
    assert p.num_elements() > 4096
    _compare(p, o)


@pytest.mark.parametrize("seed", [4, 5])
def test_refinement_on_a_domain_with_inexact_bounds(orc, seed):
    """Bounds that are not dyadic fractions (a child's upper bound computed
    from its lower bound and width can differ from its parent's by rounding,
    and probe points can fall on such bounds): the neighbour descent that
    picks one child per level from the first child's bounds, the balance
    filter that skips faces between siblings and the 2:1 search that starts
    below the root must give what the oracle's scans over every child give."""
    rng = np.random.default_rng(seed)
    dom = (0.1, 12345.678, -3.3, 777.7, -1.7, -0.013)
    p = mesh.OctreeAdapter(*dom)
    o = orc.Mesh(*dom)
    p.build_uniform(12, 10, 3)
    o.build_uniform(12, 10, 3)
    for _ in range(4):
        E = p.num_elements()
        sel = np.unique(rng.integers(0, E, size=max(1, E // 5)))
        masks = rng.integers(1, 8, size=sel.size)
        p.refine(sel, masks)
        o.refine(sel, masks)
        _compare(p, o)
    p.balance()
    o.balance()
    _compare(p, o)
