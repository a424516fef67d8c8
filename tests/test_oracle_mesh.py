"""Pins the oracle's octree and face connectivity (A17) independently of its
code: brute-force geometric adjacency, reference count tests
(tests/unit/test_amr.cpp, tests/integration/test_mesh.cpp:29-62), 2:1 balance,
depth-first leaf order and Morton bit interleaving."""
import numpy as np
import pytest


def _mesh(orc, nx, ny, nz, L=10000.0):
    m = orc.Mesh(0.0, L, 0.0, L, -1.0, 0.0)
    m.build_uniform(nx, ny, nz)
    return m


def test_counts(orc):
    # test_mesh.cpp: 2x2x2 = 8 elements; 2x2x1 has boundary and interior faces
    assert _mesh(orc, 2, 2, 2).num_elements == 8
    m = _mesh(orc, 2, 2, 1)
    t, c, el, _, _ = m.connections()
    assert (t == orc.CONN_BOUNDARY).any() and (t == orc.CONN_SAME).any()
    assert _mesh(orc, 3, 1, 1).num_elements == 4  # counts round up to powers of two


def test_leaf_order_matches_depth_first_definition(orc):
    from drifter_amd import synth

    m = _mesh(orc, 8, 4, 2)
    b, lv, _ = m.elements()
    ref = synth.uniform_mesh(8, 4, 2, 0.0, 10000.0, 0.0, 10000.0, -1.0, 0.0)
    assert np.array_equal(b, ref)
    assert (lv == [3, 2, 1]).all()


def _brute_neighbours(b, e, f, tol=1e-9):
    """elements sharing face f of element e (geometric search)."""
    ax, hi = f // 2, f % 2
    plane = b[e, 2 * ax + hi]
    out = []
    for o in range(b.shape[0]):
        if o == e:
            continue
        if abs(b[o, 2 * ax + (1 - hi)] - plane) > tol:
            continue
        ok = True
        for a in range(3):
            if a == ax:
                continue
            lo = max(b[e, 2 * a], b[o, 2 * a])
            up = min(b[e, 2 * a + 1], b[o, 2 * a + 1])
            if up - lo <= tol:
                ok = False
        if ok:
            out.append(o)
    return out


@pytest.mark.parametrize("dims", [(4, 4, 2), (2, 8, 1), (4, 2, 4)])
def test_uniform_adjacency_brute_force(orc, dims):
    m = _mesh(orc, *dims)
    b, _, _ = m.elements()
    t, c, el, fa, su = m.connections()
    for e in range(b.shape[0]):
        for f in range(6):
            bf = _brute_neighbours(b, e, f)
            if not bf:
                assert t[e, f] == orc.CONN_BOUNDARY and c[e, f] == 0
                continue
            assert t[e, f] == orc.CONN_SAME and c[e, f] == 1
            assert [el[e, f, 0]] == bf
            assert fa[e, f, 0] == (f + 1 if f % 2 == 0 else f - 1)
            o = el[e, f, 0]
            assert el[o, fa[e, f, 0], 0] == e  # points back


def _morton_interleave(ix, iy, iz, bits):
    code = 0
    for k in range(bits):
        code |= ((ix >> k) & 1) << (3 * k)
        code |= ((iy >> k) & 1) << (3 * k + 1)
        code |= ((iz >> k) & 1) << (3 * k + 2)
    return code


def test_morton_codes_interleave(orc):
    m = _mesh(orc, 4, 4, 4)
    b, _, mo = m.elements()
    h = 10000.0 / 4
    for e in range(b.shape[0]):
        ix = int(round(b[e, 0] / h))
        iy = int(round(b[e, 2] / h))
        iz = int(round((b[e, 4] + 1.0) / 0.25))
        assert int(mo[e]) == _morton_interleave(ix, iy, iz, 2)


def _face_pairs(b):
    E = b.shape[0]
    for e in range(E):
        for f in range(6):
            for o in _brute_neighbours(b, e, f):
                yield e, o


def test_refine_then_balance_is_two_to_one(orc):
    m = _mesh(orc, 4, 4, 2)
    # refine one element three times in XY: the balance pass must propagate
    for _ in range(3):
        b, lv, _ = m.elements()
        tgt = int(np.argmin(np.abs(b[:, 0] - 5000.0) + np.abs(b[:, 2] - 5000.0) + b[:, 4]))
        m.refine([tgt], [3])
    b, lv, _ = m.elements()
    assert b.shape[0] > 32
    for e, o in _face_pairs(b):
        assert np.all(np.abs(lv[e] - lv[o]) <= 1), (e, o, lv[e], lv[o])
    # refine conserves volume exactly (dyadic bounds)
    vol = np.prod(b[:, 1::2] - b[:, ::2], axis=1).sum()
    assert vol == 10000.0 * 10000.0 * 1.0


def test_coarse_to_fine_face_reports_the_leaf_at_the_centre(orc):
    """get_neighbor descends to the leaf containing the centre of the would-be
    same-size neighbour (octree_adapter.cpp:523-580) and reports SameLevel."""
    m = _mesh(orc, 2, 2, 1)
    m.refine([1], [3])  # element 1 (x in [5000, 10000], y in [0, 5000]) split in XY
    b, lv, _ = m.elements()
    t, c, el, fa, su = m.connections()
    # element 0 looks across x+ at the refined region; centre (7500, 2500, -0.5)
    e0 = 0
    assert t[e0, 1] == orc.CONN_SAME and c[e0, 1] == 1
    nb = el[e0, 1, 0]
    p = np.array([7500.0, 2500.0, -0.5])
    inside = (b[:, 0] <= p[0]) & (p[0] <= b[:, 1]) & (b[:, 2] <= p[1]) & (p[1] <= b[:, 3])
    assert inside[nb]
    assert lv[nb][0] == lv[e0][0] + 1
