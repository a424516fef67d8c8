"""Pins the oracle's face term of the DG right-hand side (oracle/faces.c) and
the interface lists it runs over, independently of its code:

  - the lists against a geometric search: conforming pairs share identical
    faces, a mortar's fine faces tile its coarse face on the sub-faces of
    MortarSpace::compute_subface_bounds (src/dg/mortar.cpp:27-78), and every
    interior (element, face) belongs to exactly one interface;
  - the face term in closed form for constant states: a conforming face adds
    M^-1 of w_q v.n u at the first N^2 face nodes (ConformingInterface's
    N-point rule, src/dg/mortar.cpp:297-343) and nothing elsewhere; a mortar's
    coarse side adds M^-1 of v.n u at every face node (L_c M^-1 F = F when
    the mortar basis is the coarse face's, src/dg/mortar.cpp:212-258);
  - the mass-weighted conforming contributions cancel over the mesh;
  - planted mistakes (the interface oriented the other way, the inverse mass
    left out) are caught by the closed forms.
"""
import numpy as np
import pytest

UPWIND, LF_ADV = 5, 6


def _mesh(orc, nx, ny, nz, refine=()):
    m = orc.Mesh(0.0, 1.0, 0.0, 1.0, -1.0, 0.0)
    m.build_uniform(nx, ny, nz)
    for sel, mask in refine:
        m.refine(sel, [mask] * len(sel))
    return m


def _band_mesh(orc, mask=3, axis=0):
    """8x8x2 base, two rounds of refinement (mask 1 = X, 3 = XY, 4 = Z) of
    the elements whose centre lies in a band across `axis`: conforming faces
    and 2:1 mortars (Fine2x1, Fine1x2 or, across z, Fine2x2)."""
    m = orc.Mesh(0.0, 1.0, 0.0, 1.0, -1.0, 0.0)
    m.build_uniform(8, 8, 2)
    for _ in range(2):
        b, _, _ = m.elements()
        c = 0.5 * (b[:, 2 * axis] + b[:, 2 * axis + 1])
        sel = np.nonzero(np.abs(c - 0.5) < 0.15)[0] if axis < 2 else np.nonzero(c > -0.5)[0]
        m.refine(sel, [mask] * sel.size)
    return m


def _tangents(f):
    a = f // 2
    return (1, 2) if a == 0 else ((0, 2) if a == 1 else (0, 1))


@pytest.mark.parametrize("mask,axis,kind", [(3, 0, 1), (4, 0, 2), (1, 1, 1), (3, 2, 5)])
def test_lists_against_geometry(orc, mask, axis, kind):
    m = _band_mesh(orc, mask, axis)
    b, lv, _ = m.elements()
    conn = m.connections(two_to_one=True)
    conf, mort, groups = orc.face_lists(b, lv, conn)
    assert mort.shape[0] > 0 and conf.shape[0] > 0
    assert kind in groups[:, 1]
    E = b.shape[0]
    used = np.zeros((E, 6), dtype=int)
    for el, fl, er, fr in conf:
        assert fr == fl ^ 1 and el < er
        a = fl // 2
        assert b[el, 2 * a + fl % 2] == b[er, 2 * a + fr % 2]  # the same plane
        for t in _tangents(fl):
            assert b[el, 2 * t] == b[er, 2 * t] and b[el, 2 * t + 1] == b[er, 2 * t + 1]
        used[el, fl] += 1
        used[er, fr] += 1
    kinds = {}
    for g in groups:
        for i in range(g[3], g[3] + g[4]):
            if g[1] != orc.CONN_SAME:
                kinds[i] = int(g[1])
    for i, row in enumerate(mort):
        ec, fc = row[0], int(row[1])
        fine = [x for x in row[2:] if x >= 0]
        k = kinds[i]
        t1, t2 = _tangents(fc)
        s1 = k in (orc.CONN_FINE2X1, orc.CONN_FINE2X2)
        s2 = k in (orc.CONN_FINE1X2, orc.CONN_FINE2X2)
        assert len(fine) == (4 if k == orc.CONN_FINE2X2 else 2)
        for s, fe in enumerate(fine):
            # sub-face s: t1 halves vary fastest (compute_subface_bounds)
            h1 = s % 2 if s1 else 0
            h2 = (s // 2 if s1 else s) if s2 else 0
            for t, split, h in ((t1, s1, h1), (t2, s2, h2)):
                lo, hi = b[ec, 2 * t], b[ec, 2 * t + 1]
                want = (lo + h * (hi - lo) / 2, lo + (h + 1) * (hi - lo) / 2) if split else (lo, hi)
                assert np.allclose(b[fe, 2 * t:2 * t + 2], want, rtol=0, atol=1e-12)
            used[fe, fc ^ 1] += 1
        used[ec, fc] += 1
    # every interior face in exactly one interface, boundary faces in none
    t = conn[0]
    assert ((used == 1) == (t != orc.CONN_BOUNDARY)).all()
    # groups in (axis, kind, level) order, lists in element order within a group
    keys = [tuple(g[:3]) for g in groups]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    for g in groups:
        rows = conf if g[1] == orc.CONN_SAME else mort
        sel = rows[g[3]:g[3] + g[4]]
        assert (np.diff(sel[:, 0] * 6 + sel[:, 1]) > 0).all()
        assert (sel[:, 1] // 2 == g[0]).all()
        assert all(max(lv[e]) == g[2] for e in sel[:, 0])


def _lgl(npts):
    """npts-point Legendre-Gauss-Lobatto rule from numpy's Legendre series:
    the nodes are +-1 and the roots of P_N' (N = npts - 1), the weights
    2 / (N (N+1) P_N(x)^2)."""
    N = npts - 1
    c = np.zeros(N + 1)
    c[-1] = 1.0
    inner = np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c))
    xs = np.r_[-1.0, np.sort(inner), 1.0]
    return xs, 2.0 / (N * (N + 1) * np.polynomial.legendre.legval(xs, c) ** 2)


def _weights(order):
    return _lgl(order + 1)


def _face_nodes(n, f):
    """(volume node, it, jt) of face f's DOFs it + n jt, from the geometry of
    the index triple: the normal index at 0 or N, tangents in axis order."""
    a = f // 2
    out = []
    for jt in range(n):
        for it in range(n):
            idx = [0, 0, 0]
            idx[a] = n - 1 if f % 2 else 0
            t1, t2 = _tangents(f)
            idx[t1], idx[t2] = it, jt
            out.append((idx[0] + n * (idx[1] + n * idx[2]), it, jt))
    return out


@pytest.mark.parametrize("order", [2, 3, 4])
def test_conforming_constant_states_closed_form(orc, order):
    """Two elements side by side in x, each with its own constant state; the
    upwind flux with v.n = 0.3 > 0 on element 0's +x face takes element 0's
    value: element 0 gains M^-1 w_q 0.3 u0 and element 1 loses it, at the
    face DOFs q < nq1^2 of the N-point rule only."""
    n, nd = order + 1, (order + 1) ** 3
    m = _mesh(orc, 2, 1, 1)
    b, lv, _ = m.elements()
    conn = m.connections(two_to_one=True)
    conf, mort, _ = orc.face_lists(b, lv, conn)
    assert conf.tolist() == [[0, 1, 1, 0]] and mort.shape[0] == 0
    u0 = np.array([1.5, -0.7, 0.25, 12.0, 33.0])
    u1 = np.array([0.5, 0.9, -0.05, 14.0, 31.0])
    U = np.concatenate([np.repeat(u0[:, None], nd, axis=1).reshape(-1),
                        np.repeat(u1[:, None], nd, axis=1).reshape(-1)])
    v = (0.3, 0.2, 0.1)
    dU = orc.face_term(order, 2, conf, mort, [], UPWIND, v, U).reshape(2, 5, nd)
    _, w = _weights(order)
    nq1 = order if order >= 2 else 2  # FaceQuadrature(face, order): `order` points
    _, rw = _lgl(nq1)
    W = np.array([w[d % n] * w[(d // n) % n] * w[d // (n * n)] for d in range(nd)])

    def closed(vn, upwind_state):
        out = np.zeros((2, 5, nd))
        for e, f, sgn in ((0, 1, 1.0), (1, 0, -1.0)):
            for q, (d, it, jt) in enumerate(_face_nodes(n, f)):
                if q < nq1 * nq1:
                    wq = rw[q % nq1] * rw[q // nq1]
                    out[e, :, d] = sgn * wq * vn * upwind_state / W[d]
        return out

    want = closed(0.3, u0)
    np.testing.assert_allclose(dU, want, rtol=1e-13, atol=1e-13)
    # planted mistakes, each caught by the closed form: the normal's sign
    # flipped (the same as v reversed for these fluxes: the upwind side and the
    # sign change), and the inverse mass left out
    flipped = orc.face_term(order, 2, conf, mort, [], UPWIND, (-0.3, -0.2, -0.1), U)
    assert not np.allclose(flipped.reshape(2, 5, nd), want, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(flipped.reshape(2, 5, nd), closed(-0.3, u1), rtol=1e-13,
                               atol=1e-13)
    assert not np.allclose(dU, want * W, rtol=1e-6, atol=1e-9)
    # the interface seen from its other side is the same interface
    other = orc.face_term(order, 2, [[1, 0, 0, 1]], mort, [], UPWIND, v, U).reshape(2, 5, nd)
    np.testing.assert_allclose(other, want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("order", [2, 3])
def test_mortar_coarse_side_constant_state_closed_form(orc, order):
    """Element 0 next to element 1 refined in XY: a Fine2x1 mortar on 0's +x
    face.  With v.n > 0 the upwind flux takes the coarse value, F = v.n u,
    and L_c M^-1 F = F, so the coarse face nodes gain (v.n u) / (w_i w_j w_k)."""
    n, nd = order + 1, (order + 1) ** 3
    m = _mesh(orc, 2, 2, 1, refine=[([1], 3)])
    b, lv, _ = m.elements()
    conn = m.connections(two_to_one=True)
    conf, mort, groups = orc.face_lists(b, lv, conn)
    mi = [i for i, r in enumerate(mort) if r[0] == 0 and r[1] == 1]
    assert len(mi) == 1
    mtype = [0] * mort.shape[0]
    for g in groups:
        if g[1] != orc.CONN_SAME:
            for i in range(g[3], g[3] + g[4]):
                mtype[i] = int(g[1])
    E = b.shape[0]
    u = np.array([1.5, -0.7, 0.25, 12.0, 33.0])
    U = np.repeat(np.tile(u, E)[:, None], nd, axis=1).reshape(-1)
    one = mort[mi]
    dU = orc.face_term(order, E, np.zeros((0, 4), np.int64), one, [mtype[mi[0]]], UPWIND,
                       (0.3, 0.2, 0.1), U).reshape(E, 5, nd)
    _, w = _weights(order)
    want = np.zeros((5, nd))
    for d, it, jt in _face_nodes(n, 1):
        i, j, k = d % n, (d // n) % n, d // (n * n)
        want[:, d] = 0.3 * u / (w[i] * w[j] * w[k])
    np.testing.assert_allclose(dU[0], want, rtol=1e-11, atol=1e-12)
    # the fine elements receive the opposite-signed share on their -x faces only
    for fe in one[0, 2:4]:
        nz = np.nonzero(np.abs(dU[fe]).sum(axis=0))[0]
        assert set(nz) <= {d for d, _, _ in _face_nodes(n, 0)}


def test_conforming_contributions_cancel_in_the_mass_weighted_sum(orc):
    order = 3
    n, nd = order + 1, (order + 1) ** 3
    m = _mesh(orc, 4, 3, 2)
    b, lv, _ = m.elements()
    conf, mort, _ = orc.face_lists(b, lv, m.connections(two_to_one=True))
    assert mort.shape[0] == 0
    E = b.shape[0]
    U = np.random.default_rng(3).standard_normal(E * 5 * nd)
    dU = orc.face_term(order, E, conf, mort, [], LF_ADV, (0.3, -0.2, 0.1, 0.6), U)
    _, w = _weights(order)
    W = np.array([w[d % n] * w[(d // n) % n] * w[d // (n * n)] for d in range(nd)])
    MdU = dU.reshape(E, 5, nd) * W
    tot = MdU.sum(axis=(0, 2))
    scale = np.abs(MdU).sum(axis=(0, 2))
    assert (np.abs(tot) <= 1e-13 * scale).all(), (tot, scale)
    assert (scale > 0).all()
