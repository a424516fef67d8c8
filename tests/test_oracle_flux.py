"""Pins the oracle's fluxes, conforming interface and mortars (A14-A16, A18)
by properties and closed forms computed independently in numpy.

Reference tests restated: tests/unit/test_numerical_flux.cpp:17-234
(consistency, LF antisymmetry, F(UL,UR,n) + F(UR,UL,-n) = 0 with
n = (0.6, 0.8, 0), upwind, dry states) and tests/unit/test_mortar.cpp:42-169
(mortar DOFs, mass symmetric positive definite, (2N+1)^2 quadrature)."""
import numpy as np
import pytest

G = 9.81
N06 = np.array([0.6, 0.8, 0.0])


def phys_sw(U, n, g=G):
    h, hu, hv = U
    u, v = hu / h, hv / h
    un = u * n[0] + v * n[1]
    return np.array([h * un, hu * un + 0.5 * g * h * h * n[0], hv * un + 0.5 * g * h * h * n[1]])


def phys_pe(U, n, g=G):
    H = U[2] + 10.0
    u, v = U[0] / H, U[1] / H
    un = u * n[0] + v * n[1]
    return np.array([U[0] * un + 0.5 * g * H * H * n[0], U[1] * un + 0.5 * g * H * H * n[1],
                     U[0] * n[0] + U[1] * n[1], U[3] * un, U[4] * un])


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_sw_consistency_and_conservation(orc, kind):
    rng = np.random.default_rng(kind)
    p = [G, 0.1]
    for _ in range(50):
        UL = np.array([rng.uniform(0.5, 5), rng.normal(), rng.normal()])
        UR = np.array([rng.uniform(0.5, 5), rng.normal(), rng.normal()])
        n = rng.normal(size=3)
        n[2] = 0.0
        n /= np.linalg.norm(n)
        F = orc.flux(kind, UL, UL, n, p)
        np.testing.assert_allclose(F, phys_sw(UL, n), rtol=1e-12, atol=1e-12)
        a = orc.flux(kind, UL, UR, N06, p)
        b = orc.flux(kind, UR, UL, -N06, p)
        np.testing.assert_allclose(a + b, 0.0, atol=1e-11 * (1 + np.abs(a).max()))


def test_hllc_supersonic_takes_upwind_flux(orc):
    # S_L >= 0 for flow much faster than the wave speed: F* = F(U_L)
    UL = np.array([1.0, 20.0, 0.0])
    UR = np.array([1.2, 24.0, 0.0])
    n = np.array([1.0, 0.0, 0.0])
    np.testing.assert_allclose(orc.flux(1, UL, UR, n, [G]), phys_sw(UL, n), rtol=1e-14)
    np.testing.assert_allclose(orc.flux(1, UR, UL, -n, [G]), -phys_sw(UL, n), rtol=1e-13)


def test_lf_penalty_closed_form(orc):
    UL = np.array([2.0, 1.0, -0.5])
    UR = np.array([1.5, -0.3, 0.2])
    n = N06
    lam = max(abs((UL[1] * n[0] + UL[2] * n[1]) / UL[0]) + np.sqrt(G * UL[0]),
              abs((UR[1] * n[0] + UR[2] * n[1]) / UR[0]) + np.sqrt(G * UR[0]))
    ref = 0.5 * (phys_sw(UL, n) + phys_sw(UR, n)) - 0.5 * lam * (UR - UL)
    np.testing.assert_allclose(orc.flux(0, UL, UR, n, [G]), ref, rtol=1e-13)


def test_dry_states_finite(orc):
    dry = np.array([0.0, 0.0, 0.0])
    wet = np.array([1.0, 0.3, 0.1])
    for kind in range(4):
        assert np.all(np.isfinite(orc.flux(kind, dry, wet, N06, [G, 0.1])))


def test_pe_lf5_consistency(orc):
    U = np.array([3.0, -1.0, 0.2, 150.0, 340.0])
    np.testing.assert_allclose(orc.flux(4, U, U, N06, [G]), phys_pe(U, N06), rtol=1e-13)
    UR = np.array([2.0, 0.5, -0.1, 140.0, 330.0])
    a = orc.flux(4, U, UR, N06, [G])
    b = orc.flux(4, UR, U, -N06, [G])
    np.testing.assert_allclose(a + b, 0.0, atol=1e-10)


def test_upwind_direction(orc):
    v = [1.0, 2.0, 0.0]
    assert orc.flux(5, [3.0], [7.0], [1.0, 0.0, 0.0], v)[0] == 3.0
    assert orc.flux(5, [3.0], [7.0], [-1.0, 0.0, 0.0], v)[0] == -7.0


@pytest.mark.parametrize("order", [2, 3, 4])
def test_face_map_extracts_face_nodes(orc, order):
    n = order + 1
    x, _ = orc.lgl(n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    pts = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], axis=1)
    tang = {0: (1, 2), 1: (1, 2), 2: (0, 2), 3: (0, 2), 4: (0, 1), 5: (0, 1)}
    for f in range(6):
        m = orc.face_map(order, f)
        ax, sign = f // 2, (1.0 if f % 2 else -1.0)
        t1, t2 = tang[f]
        for jt in range(n):
            for it in range(n):
                p = pts[m[it + n * jt]]
                assert p[ax] == sign and p[t1] == x[it] and p[t2] == x[jt]


def _lgl_numpy(n):
    N = n - 1
    P = np.polynomial.legendre.Legendre.basis(N)
    inner = np.sort(P.deriv().roots().real) if n > 2 else np.array([])
    x = np.r_[-1.0, inner, 1.0]
    w = 2.0 / (n * (n - 1) * P(x) ** 2)
    return x, w


@pytest.mark.parametrize("order", [2, 3, 4])
def test_conforming_uses_N_point_rule(orc, order):
    """FaceQuadrature(face, N, LGL) has N points per direction: only the first
    N^2 face values enter, with N-point LGL weights (mortar.cpp:306)."""
    nf = (order + 1) ** 2
    rng = np.random.default_rng(order)
    UL, UR = rng.normal(size=nf), rng.normal(size=nf)
    v = [0.7, -0.2, 0.4]
    n = np.array([0.0, 0.0, 1.0])
    rl, rr = orc.conforming_flux(order, 5, v, UL, UR, n)
    xq, wq = _lgl_numpy(order)
    W = np.outer(wq, wq).reshape(-1)  # index i + N j
    vn = 0.4
    F = vn * UL[: order * order]
    np.testing.assert_allclose(rl[: order * order], W * F, rtol=1e-13)
    np.testing.assert_allclose(rr[: order * order], -W * F, rtol=1e-13)
    assert np.all(rl[order * order:] == 0) and np.all(rr[order * order:] == 0)


def _lagrange_mass_1d(x):
    """exact int_{-1}^{1} l_i l_j for the Lagrange basis on nodes x"""
    n = len(x)
    polys = []
    for i in range(n):
        others = np.delete(x, i)
        p = np.polynomial.Polynomial.fromroots(others)
        polys.append(p / p(x[i]))
    M = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            q = (polys[i] * polys[j]).integ()
            M[i, j] = q(1.0) - q(-1.0)
    return M, polys


@pytest.mark.parametrize("order", [2, 3, 4])
def test_mortar_mass_and_coarse_projection(orc, order):
    M, Mi, Pc, Lc, Pf, Lf = orc.mortar_operators(order, orc.CONN_FINE2X2)
    nm = (order + 1) ** 2
    assert M.shape == (nm, nm) and len(Pf) == 4
    x, _ = orc.lgl(order + 1)
    M1, _ = _lagrange_mass_1d(x)
    np.testing.assert_allclose(M, np.kron(M1, M1), rtol=0, atol=1e-13)  # index a + np b
    assert np.allclose(M, M.T, atol=1e-15) and np.all(np.linalg.eigvalsh(M) > 0)
    np.testing.assert_allclose(Mi @ M, np.eye(nm), atol=1e-11)
    np.testing.assert_allclose(Pc, np.eye(nm), atol=1e-12)  # same basis: identity
    np.testing.assert_allclose(Lc, M.T, atol=1e-13)          # L_c = B_c^T = M^T


def _midline_extra(order, ctype, g):
    """sum_f B_f u_f - int Phi g: the GL nodes on a split line (exactly 0 for
    odd point counts) are counted by both sub-faces (mortar.cpp:184-185)."""
    nq = 2 * order + 1
    gx, gw = np.polynomial.legendre.leggauss(nq)
    x, _ = _lgl_numpy(order + 1)
    _, polys = _lagrange_mass_1d(x)
    np_ = order + 1
    extra = np.zeros(np_ * np_)
    for j in range(nq):
        for i in range(nq):
            mult = 1
            if ctype in (1, 5) and gx[i] == 0.0:
                mult *= 2
            if ctype in (2, 5) and gx[j] == 0.0:
                mult *= 2
            if mult == 1:
                continue
            phi = np.array([polys[a](gx[i]) * polys[b](gx[j]) for b in range(np_) for a in range(np_)])
            extra += (mult - 1) * gw[i] * gw[j] * phi * g(gx[i], gx[j])
    return extra


@pytest.mark.parametrize("ctype,nfine", [(1, 2), (2, 2), (5, 4)])
@pytest.mark.parametrize("order", [2, 3])
def test_mortar_fine_projection_closed_form(orc, ctype, nfine, order):
    M, Mi, Pc, Lc, Pf, Lf = orc.mortar_operators(order, ctype)
    assert len(Pf) == nfine
    np_ = order + 1
    x, _ = orc.lgl(np_)
    g = lambda s, t: 1.0 + 0.3 * s - 0.2 * t + 0.1 * s * t + 0.05 * s * s
    bounds = {1: [(-1, -1, 0, 1), (0, -1, 1, 1)], 2: [(-1, -1, 1, 0), (-1, 0, 1, 1)],
              5: [(-1, -1, 0, 0), (0, -1, 1, 0), (-1, 0, 0, 1), (0, 0, 1, 1)]}[ctype]
    total = np.zeros(np_ * np_)
    for f, (a0, b0, a1, b1) in enumerate(bounds):
        s = a0 + (x + 1) * (a1 - a0) / 2
        t = b0 + (x + 1) * (b1 - b0) / 2
        uf = np.array([g(s[i], t[j]) for j in range(np_) for i in range(np_)])
        B = M @ Pf[f]  # B_f = M P_f
        total += B @ uf
    gc = np.array([g(x[i], x[j]) for j in range(np_) for i in range(np_)])
    exact = M @ gc  # quadratic g lies in the mortar space: int Phi g = M g
    np.testing.assert_allclose(total - exact, _midline_extra(order, ctype, g), atol=1e-12)
    for f in range(nfine):
        np.testing.assert_allclose(Lf[f], (M @ Pf[f]).T, atol=1e-12)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_mortar_flux_same_level_is_antisymmetric(orc, order):
    nm = (order + 1) ** 2
    rng = np.random.default_rng(3)
    Uc, Uf = rng.normal(size=nm), rng.normal(size=(1, nm))
    rc, rf = orc.mortar_flux(order, 0, 6, [0.5, 0.1, -0.3, 0.8], Uc, Uf, [0.0, 1.0, 0.0])
    np.testing.assert_allclose(rf[0], -rc, atol=1e-13)
    # lifted flux integrates to the sum of nodal fluxes (partition of unity)
    F = 0.5 * 0.1 * (Uc + Uf[0]) - 0.5 * 0.8 * (Uf[0] - Uc)
    assert abs(rc.sum() - F.sum()) < 1e-12 * (1 + np.abs(F).sum())


# verify_mortar_conservation (src/dg/mortar.cpp:386-413).  No reference test
# calls it; these closed forms pin the oracle: exact dyadic cancellation,
# per-fine-element weights (a dropped fine index or the coarse weights used
# on the fine side change the ratio), the 1e-15 floor and scale invariance.
def test_conservation_exact_cancellation(orc):
    nm, nfine = 9, 4
    rc = np.arange(1, nm + 1, dtype=float)
    wc = np.full(nm, 0.5)
    rf = np.tile(-rc / 8.0, (nfine, 1))
    wf = np.ones((nfine, nm))
    assert orc.verify_mortar_conservation(rc, rf, wc, wf) == 0.0


def test_conservation_closed_form_ratio(orc):
    nm, nfine = 16, 4
    rc, wc = np.ones(nm), np.ones(nm)
    rf = -np.ones((nfine, nm))
    wf = np.array([(f + 1) / 8.0 for f in range(nfine)])[:, None] * np.ones((nfine, nm))
    # coarse nm, fine -nm (1+2+3+4)/8 = -1.25 nm: |-0.25 nm| / (2.25 nm) = 1/9
    assert abs(orc.verify_mortar_conservation(rc, rf, wc, wf) - 1.0 / 9.0) < 1e-16
    # same signs on both sides: the ratio is one
    assert orc.verify_mortar_conservation(rc, -rf, wc, wf) == 1.0


def test_conservation_norm_floor(orc):
    nm, nfine = 4, 2
    tiny = np.full(nm, 1e-17)
    assert orc.verify_mortar_conservation(tiny, np.zeros((nfine, nm)), np.ones(nm),
                                          np.ones((nfine, nm))) == 0.0
    assert orc.verify_mortar_conservation(np.zeros(nm), np.zeros((nfine, nm)), np.ones(nm),
                                          np.ones((nfine, nm))) == 0.0
    # just above the floor the ratio is reported (one-sided flux: 1)
    above = np.full(nm, 1e-15)
    assert orc.verify_mortar_conservation(above, np.zeros((nfine, nm)), np.ones(nm),
                                          np.ones((nfine, nm))) == 1.0


def test_conservation_scale_invariant(orc):
    rng = np.random.default_rng(5)
    nm, nfine = 9, 3
    rc, rf = rng.normal(size=nm), rng.normal(size=(nfine, nm))
    wc, wf = rng.uniform(0.1, 1, nm), rng.uniform(0.1, 1, (nfine, nm))
    r = orc.verify_mortar_conservation(rc, rf, wc, wf)
    assert 0.0 <= r <= 1.0
    assert orc.verify_mortar_conservation(rc * 1024.0, rf * 1024.0, wc, wf) == r


@pytest.mark.parametrize("order", [2, 3, 4])
def test_conservation_same_level_mortar_flux(orc, order):
    """a same-level mortar lifts +F to one side and -F to the other, so the
    unit-weight integrals cancel to rounding"""
    nm = (order + 1) ** 2
    rng = np.random.default_rng(order)
    rc, rf = orc.mortar_flux(order, 0, 6, [0.5, 0.1, -0.3, 0.8], rng.normal(size=nm),
                             rng.normal(size=(1, nm)), [0.0, 1.0, 0.0])
    assert orc.verify_mortar_conservation(rc, rf, np.ones(nm), np.ones((1, nm))) < 1e-14
