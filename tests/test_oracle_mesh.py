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


def _probe(bb, f):
    """Centre of the same-size box across face f (octree_adapter.cpp:530-558)."""
    p = 0.5 * (bb[0::2] + bb[1::2])
    a = f // 2
    lo, hi = bb[2 * a], bb[2 * a + 1]
    p[a] = hi + 0.5 * (hi - lo) if f % 2 else lo - 0.5 * (hi - lo)
    return p


def _leaf_at(b, p):
    """The leaf the reference's descent reaches: among the leaves whose
    bounds hold p (inclusive), the first in depth-first order, i.e. the
    smallest element number (at every level the descent takes the first
    child in creation order that holds p, and elements are numbered depth
    first with children in creation order)."""
    ins = np.all((b[:, 0::2] <= p) & (p <= b[:, 1::2]), axis=1)
    return int(np.nonzero(ins)[0][0])


def _random_refined(orc, seed, rounds=3):
    rng = np.random.default_rng(seed)
    m = orc.Mesh(0.0, 1.0, 0.0, 2.0, -1.0, 0.0)
    m.build_uniform(4, 4, 2)
    for _ in range(rounds):
        E = m.num_elements
        sel = np.unique(rng.integers(0, E, size=max(1, E // 5)))
        m.refine(sel, rng.integers(1, 8, size=sel.size))
    return m


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_refined_connectivity_brute_force(orc, seed):
    """After random anisotropic refine + balance, every FaceConnection of the
    reference's connectivity is the leaf holding the probe point found by a
    linear scan (SameLevel, one neighbour, opposite face), and every face of
    the 2:1 connectivity lists exactly the leaves sharing a piece of the face
    (geometric search), one covering leaf as SameLevel."""
    m = _random_refined(orc, seed)
    b, lv, _ = m.elements()
    dom = np.array([0.0, 1.0, 0.0, 2.0, -1.0, 0.0])
    t, c, el, fa, _ = m.connections()
    t2, c2, el2, fa2, _ = m.connections(two_to_one=True)
    nmortar = 0
    for e in range(b.shape[0]):
        for f in range(6):
            a = f // 2
            on = b[e, 2 * a + 1] >= dom[2 * a + 1] - 1e-12 if f % 2 else \
                b[e, 2 * a] <= dom[2 * a] + 1e-12
            if on:
                assert t[e, f] == orc.CONN_BOUNDARY and t2[e, f] == orc.CONN_BOUNDARY
                continue
            opp = f + 1 if f % 2 == 0 else f - 1
            assert t[e, f] == orc.CONN_SAME and c[e, f] == 1 and fa[e, f, 0] == opp
            assert el[e, f, 0] == _leaf_at(b, _probe(b[e], f))
            br = sorted(_brute_neighbours(b, e, f))
            assert br
            k = c2[e, f]
            assert sorted(el2[e, f, :k].tolist()) == br[:4]
            assert (fa2[e, f, :k] == opp).all()
            if len(br) == 1:
                assert t2[e, f] == orc.CONN_SAME
            else:
                nmortar += 1
                assert t2[e, f] != orc.CONN_SAME
    assert nmortar > 0


def _forcing_nodes(b, lv, x, mask, tol=1e-12):
    """Nodes that would make balance split node x with `mask`: outside x, with
    the probe point of one of their faces (the centre of the same-size box
    across it) inside x's bounds (inclusive), and a level gap to x above one
    on exactly the axes of mask (octree_adapter.cpp:138-182)."""
    bx = b[x]
    within = np.all((b[:, 0::2] >= bx[0::2] - tol) & (b[:, 1::2] <= bx[1::2] + tol), axis=1)
    need = ((lv - lv[x]) > 1).astype(np.int64) @ np.array([1, 2, 4])
    cand = (need == mask) & ~within
    c = 0.5 * (b[:, 0::2] + b[:, 1::2])
    s = b[:, 1::2] - b[:, 0::2]
    found = []
    for f in range(6):
        a = f // 2
        p = c.copy()
        p[:, a] = b[:, 2 * a + 1] + 0.5 * s[:, a] if f % 2 else b[:, 2 * a] - 0.5 * s[:, a]
        hit = cand & np.all((bx[0::2] <= p) & (p <= bx[1::2]), axis=1)
        found.extend(np.nonzero(hit)[0].tolist())
    return found


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 5])
def test_balance_splits_only_what_a_level_gap_forces(orc, seed):
    """Minimality of balance: every node that the balance pass split (not one
    the caller refined) has a node across one of its faces that is more than
    one level finer on exactly the axes of the split, and on no other axis.
    The tree only gets finer, so the node that forced a split is still in the
    final tree, found here by a linear scan over all nodes.  The same check
    rejects most of the caller's own (unforced) splits, and a split with one
    axis more than recorded, so it is not vacuous."""
    rng = np.random.default_rng(100 + seed)
    m = orc.Mesh(0.0, 1.0, 0.0, 2.0, -1.0, 0.0)
    m.build_uniform(4, 4, 2)
    nforced = nfree = nfree_unforced = nwider = 0
    for _ in range(3):
        _, _, _, nc0, lf0 = m.nodes()
        sel = np.unique(rng.integers(0, m.num_elements, size=max(1, m.num_elements // 5)))
        m.refine(sel, rng.integers(1, 8, size=sel.size))
        b, lv, mk, nc, _ = m.nodes()
        node_of_leaf = {int(l): i for i, l in enumerate(lf0) if l >= 0}
        refined = {node_of_leaf[int(e)] for e in sel}
        split_before = np.zeros(b.shape[0], bool)
        split_before[:nc0.size] = nc0 > 0
        for x in np.nonzero(nc > 0)[0]:
            if split_before[x]:
                continue
            if x in refined:
                nfree += 1
                nfree_unforced += not _forcing_nodes(b, lv, x, mk[x])
                continue
            nforced += 1
            assert _forcing_nodes(b, lv, x, mk[x]), (x, mk[x], lv[x])
            if mk[x] != 7:
                wider = mk[x] | (1 << int(np.nonzero([(mk[x] >> a) & 1 == 0 for a in range(3)])[0][0]))
                nwider += not _forcing_nodes(b, lv, x, wider)
    assert nforced > 0 and nwider > 0
    assert nfree_unforced > nfree // 2


def test_two_to_one_connectivity_a_coarse_face_lists_its_fine_leaves(orc):
    """The case the reference's descent cannot reach: a face whose same-size
    neighbour is refined lists the fine leaves across it (Fine2x2 here), and
    the fine leaves point back at the coarse element as SameLevel."""
    m = _mesh(orc, 2, 2, 1)
    m.refine([1], [3])  # element 1 split in XY: its -x face now holds two leaves
    b, lv, _ = m.elements()
    t, c, el, fa, su = m.connections(two_to_one=True)
    # element 0 (x in [0, 5000], y in [0, 5000]) across +x: two leaves, split in y only
    assert c[0, 1] == 2 and t[0, 1] == orc.CONN_FINE2X1
    fine = el[0, 1, :2]
    assert set(fine) == set(_brute_neighbours(b, 0, 1))
    for x in fine:
        assert t[x, 0] == orc.CONN_SAME and el[x, 0, 0] == 0
