"""Pins the oracle's solution transfer (oracle/projection.c, restating
SolutionProjection::project_to_children / project_from_children /
interpolate_3d, src/amr/refinement.cpp:217-371, and
AdaptiveMeshRefinement::transfer_solution :533-567) by closed forms:
polynomials of degree <= N per axis are reproduced exactly at the
children's nodes, children -> parent inverts parent -> children, child
counts per mask, and the output order of the transfer."""
import numpy as np
import pytest

import oracle as orc

MASKS = range(1, 8)


def _poly(rng, N):
    """A non-separable polynomial of degree <= N in each variable."""
    a = rng.standard_normal((2, 3, N + 1))

    def f(x, y, z):
        t = 0.0
        for s in range(2):
            t = t + (np.polyval(a[s, 0], x) * np.polyval(a[s, 1], y) * np.polyval(a[s, 2], z))
        return t

    return f


def _nodal(order, f, box=((-1, 1), (-1, 1), (-1, 1))):
    """f at the LGL nodes of a sub-box, node index i + n (j + n k)."""
    x, _ = orc.lgl(order + 1)
    m = [lo + 0.5 * (x + 1.0) * (hi - lo) for lo, hi in box]
    Z, Y, X = np.meshgrid(m[2], m[1], m[0], indexing="ij")
    return f(X, Y, Z).reshape(-1)


def _child_boxes(mask):
    nx, ny, nz = (2 if mask & b else 1 for b in (1, 2, 4))
    out = []
    for cz in range(nz):
        for cy in range(ny):
            for cx in range(nx):
                box = []
                for ref, c in ((mask & 1, cx), (mask & 2, cy), (mask & 4, cz)):
                    box.append((-1.0 + c, float(c)) if ref else (-1.0, 1.0))
                out.append(tuple(box))
    return out


@pytest.mark.parametrize("order", [1, 2, 3, 4, 7])
@pytest.mark.parametrize("mask", MASKS)
def test_children_reproduce_polynomials(order, mask):
    rng = np.random.default_rng(100 * order + mask)
    f = _poly(rng, order)
    U = _nodal(order, f)
    ch = orc.project_to_children(order, mask, U)
    assert ch.shape[0] == 1 << bin(mask).count("1")
    scale = np.abs(U).max()
    for c, box in enumerate(_child_boxes(mask)):
        np.testing.assert_allclose(ch[c], _nodal(order, f, box), rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("order", [2, 3])
def test_children_do_not_reproduce_higher_degree(order):
    # sensitivity: a degree N+1 term is not representable, so the check above
    # is not vacuous
    f = lambda x, y, z: x ** (order + 1)  # noqa: E731
    ch = orc.project_to_children(order, 1, _nodal(order, f))
    assert np.abs(ch[1] - _nodal(order, f, _child_boxes(1)[1])).max() > 1e-3


@pytest.mark.parametrize("order", [1, 3, 4, 8])
@pytest.mark.parametrize("mask", MASKS)
def test_from_children_closed_form(order, mask):
    rng = np.random.default_rng(7 + order * 10 + mask)
    boxes = _child_boxes(mask)
    fs = [_poly(rng, order) for _ in boxes]
    # child c holds g_c in its own reference coordinates
    children = np.stack([_nodal(order, g) for g in fs])
    got = orc.project_from_children(order, mask, children)
    x, _ = orc.lgl(order + 1)
    n = order + 1
    want = np.zeros(n ** 3)
    nx, ny = (2 if mask & 1 else 1), (2 if mask & 2 else 1)
    for k in range(n):
        for j in range(n):
            for i in range(n):
                h = [(1 if x[p] >= 0 else 0) if mask & b else 0
                     for p, b in ((i, 1), (j, 2), (k, 4))]
                c = h[0] + nx * (h[1] + ny * h[2])
                loc = []
                for p, b, hh in ((i, 1, h[0]), (j, 2, h[1]), (k, 4, h[2])):
                    lo, hi = ((-1.0 + hh, float(hh)) if mask & b else (-1.0, 1.0))
                    loc.append(np.clip(2 * (x[p] - lo) / (hi - lo) - 1, -1, 1))
                want[i + n * (j + n * k)] = fs[c](*loc)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("order", [1, 2, 4, 6])
@pytest.mark.parametrize("mask", MASKS)
def test_round_trip_is_identity(order, mask):
    U = np.random.default_rng(order + 31 * mask).standard_normal((order + 1) ** 3)
    back = orc.project_from_children(order, mask, orc.project_to_children(order, mask, U))
    np.testing.assert_allclose(back, U, rtol=0, atol=1e-13 * np.abs(U).max())


def test_transfer_order():
    order = 2
    nd = (order + 1) ** 3
    U = np.random.default_rng(5).standard_normal((6, nd))
    out = orc.transfer_solution(order, U, [1, 4])
    assert out.shape == (6 - 2 + 16, nd)
    np.testing.assert_array_equal(out[:4], U[[0, 2, 3, 5]])
    np.testing.assert_array_equal(out[4:12], orc.project_to_children(order, 7, U[1]))
    np.testing.assert_array_equal(out[12:20], orc.project_to_children(order, 7, U[4]))
    # no parents: a copy; every element a parent: children only
    np.testing.assert_array_equal(orc.transfer_solution(order, U, []), U)
    assert orc.transfer_solution(order, U, np.arange(6)).shape == (48, nd)
