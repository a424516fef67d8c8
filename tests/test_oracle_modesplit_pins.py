"""Pins of the oracle's mode-split step (A11/A12) on the paths that move
water: the barotropic pressure gradient -g H D eta, the free-surface
divergence D_x2d HU_bar + D_y2d HV_bar, the Coriolis coupling, the substep
average eta_avg and the omega profile of update_baroclinic
(src/physics/mode_splitting.cpp:129-221, 293-332, 359-453, 487-600).

1. The reference's component tests (tests/unit/test_mode_splitting.cpp:52-186)
   restated on the oracle's element-level pieces, plus closed forms of each
   piece on linear fields, where the LGL derivative matrix is exact.
2. Closed forms through the whole step:
   Time steps are short in the reference's units (dt_3d = 0.05 with
   sqrt(g H) = 31 per unit length): the element-local operators are
   nilpotent, so rounding residues grow like (dt sqrt(g H) |D|)^N and a long
   step would swamp the closed forms; the Coriolis parameter is 2 per unit
   time, so the rotation is visible within such a step.
   * gravity: eta linear in the 2D fast digit, h = 100 - eta (H = 100), no
     transport.  w = HU + i HV obeys dw/dt = -i f w + F with F = -g H alpha,
     so one Heun substep maps w -> P w + dt_b F (1 + z/2), P = 1 + z + z^2/2,
     z = -i f dt_b; u, v after recombine are w / H; eta is re-laid from the
     2D index to the column layout; Hu stays 0 (not refreshed).
   * divergence: H u linear in the column's fast digit, so HU_bar(c) = beta
     x_(c mod n) and D_x2d HU_bar = beta.  With Coriolis, w(c) = beta
     x_(c mod n) P^k and eta_(k+1) = eta_k - (dt_b beta / 2) Re((2 + z) P^k);
     eta_avg is their mean, H_new = eta_avg + h(c), Hu = H_new u (not
     refreshed), omega(c n + m) = -(m / (n - 1)) beta / 100 with Hinv taken
     from H(c n) (h varies with the slow digit of c, so another column gives
     another value).
3. Each closed form has teeth: the same checks run against copies of the
   oracle compiled on the CPU with one planted mistake each (swapped
   D_x2d / D_y2d, a flipped sign, the wrong Hinv column, the final eta
   instead of eta_avg) and every copy fails at least one of them.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

NF = 12
HU, HV, ETA, HT, HS, H, U, V, T, S, OM, RHO = range(12)
G = 9.81
ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
_PD = C.POINTER(C.c_double)


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PD)


class Lib:
    """The oracle's mode-split entry points on a given liboracle build."""

    def __init__(self, path):
        import oracle

        self.l = C.CDLL(path)
        self.l.orc_mode_split_step.restype = C.c_int
        self.l.orc_mode_split_step.argtypes = [C.c_int, C.c_int64, C.c_double, C.c_int, _PD,
                                               _PD, _PD, C.POINTER(oracle.OrcPhysics)]
        for name, nargs in (("orc_ms_depth_average", 2), ("orc_ms_depth_integrate", 3),
                            ("orc_ms_deviation", 3), ("orc_ms_divergence_2d", 3)):
            fn = getattr(self.l, name)
            fn.restype = C.c_int
            fn.argtypes = [C.c_int] + [_PD] * nargs
        self.l.orc_ms_barotropic_rhs.restype = C.c_int
        self.l.orc_ms_barotropic_rhs.argtypes = [C.c_int] + [_PD] * 5 + [C.c_double] + [_PD] * 3
        self.l.orc_default_physics.argtypes = [C.POINTER(oracle.OrcPhysics)]
        self.OrcPhysics = oracle.OrcPhysics

    def physics(self, **kw):
        p = self.OrcPhysics()
        self.l.orc_default_physics(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def step(self, order, dt3, nsub, full, h, f, phys):
        V = np.array(full, dtype=np.float64, copy=True)
        nd = (order + 1) ** 3
        E = V.size // (NF * nd)
        assert self.l.orc_mode_split_step(order, E, dt3, nsub, _p(V), _p(h), _p(f),
                                          C.byref(phys)) == 0
        return V.reshape(E, NF, nd)

    def depth_average(self, order, u3):
        n = order + 1
        out = np.zeros(n * n)
        assert self.l.orc_ms_depth_average(order, _p(np.ascontiguousarray(u3)), _p(out)) == 0
        return out

    def depth_integrate(self, order, u3, Hh):
        n = order + 1
        out = np.zeros(n * n)
        assert self.l.orc_ms_depth_integrate(order, _p(np.ascontiguousarray(u3)),
                                             _p(np.ascontiguousarray(Hh)), _p(out)) == 0
        return out

    def deviation(self, order, u3, ub):
        out = np.zeros_like(u3)
        assert self.l.orc_ms_deviation(order, _p(np.ascontiguousarray(u3)),
                                       _p(np.ascontiguousarray(ub)), _p(out)) == 0
        return out

    def divergence(self, order, HUb, HVb):
        out = np.zeros_like(HUb)
        assert self.l.orc_ms_divergence_2d(order, _p(np.ascontiguousarray(HUb)),
                                           _p(np.ascontiguousarray(HVb)), _p(out)) == 0
        return out

    def barotropic_rhs(self, order, eta, HUb, HVb, Hh, f, g):
        de, dU, dV = (np.zeros_like(eta) for _ in range(3))
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (eta, HUb, HVb, Hh, f)]
        assert self.l.orc_ms_barotropic_rhs(order, *[_p(a) for a in args], g, _p(de), _p(dU),
                                            _p(dV)) == 0
        return de, dU, dV


def _nodes(order):
    import oracle

    return oracle.lgl(order + 1)[0]


def _digits(order):
    """i (xi), j (eta), k (zeta) index of every DOF d = i + n (j + n k)."""
    n = order + 1
    d = np.arange(n ** 3)
    return d % n, (d // n) % n, d // (n * n)


# ---------------------------------------------------------------------------
# 1. components (test_mode_splitting.cpp:52-186) and their closed forms
# ---------------------------------------------------------------------------

def check_components(L, order):
    n = order + 1
    nh, nd = n * n, n ** 3
    x = _nodes(order)
    rng = np.random.default_rng(order)
    # DepthAverageConstant :52-63
    np.testing.assert_allclose(L.depth_average(order, np.full(nd, 1.5)), 1.5, atol=1e-14)
    # DepthAverageLinear :66-83
    u1, u2 = rng.random(nd), rng.random(nd)
    np.testing.assert_allclose(L.depth_average(order, 2.5 * u1 - 1.3 * u2),
                               2.5 * L.depth_average(order, u1) - 1.3 * L.depth_average(order, u2),
                               atol=1e-13)
    # DepthIntegrate :86-102: constant u H integrates to u H (sum of 0.5 w is 1)
    np.testing.assert_allclose(L.depth_integrate(order, np.full(nd, 0.5), np.full(nd, 100.0)),
                               50.0, rtol=1e-14)
    # ComputeDeviation :105-123 and DeviationAveragesZero :126-142
    u = rng.random(nd)
    ub = L.depth_average(order, u)
    up = L.deviation(order, u, ub)
    np.testing.assert_allclose(up, u - np.repeat(ub, n), atol=1e-15)
    np.testing.assert_allclose(L.depth_average(order, up), 0.0, atol=1e-14)
    # depth average of a column = LGL quadrature (weights 0.5 w): a column
    # that is a polynomial of degree <= 2N - 1 in its fast index averages to
    # its exact mean over [-1, 1]
    col = np.tile(x ** 2, nh)
    np.testing.assert_allclose(L.depth_average(order, col), 1.0 / 3.0, rtol=1e-13)
    # BarotropicRHSQuiescent :145-168
    z = np.zeros(nh)
    de, dU, dV = L.barotropic_rhs(order, z, z, z, np.full(nh, 100.0), np.full(nh, 1e-4), G)
    assert np.max(np.abs(np.r_[de, dU, dV])) == 0.0
    # Divergence2D :171-186: constant transport -> 0
    np.testing.assert_allclose(L.divergence(order, np.full(nh, 100.0), np.full(nh, 50.0)), 0.0,
                               atol=1e-11)
    # closed forms: the 2D index is a + n b (a fast); D_x2d differentiates
    # along a, D_y2d along b, exactly for linear fields
    a, b = np.arange(nh) % n, np.arange(nh) // n
    np.testing.assert_allclose(L.divergence(order, x[a], z), 1.0, rtol=1e-12)
    np.testing.assert_allclose(L.divergence(order, z, x[b]), 1.0, rtol=1e-12)
    np.testing.assert_allclose(L.divergence(order, x[b], x[a]), 0.0, atol=1e-12)
    Hh = 80.0 + rng.random(nh)
    f = np.full(nh, 1e-4)
    de, dU, dV = L.barotropic_rhs(order, 0.02 * x[a], z, z, Hh, f, G)
    np.testing.assert_allclose(dU, -G * Hh * 0.02, rtol=1e-12)
    np.testing.assert_allclose(dV, 0.0, atol=1e-12)
    de, dU, dV = L.barotropic_rhs(order, 0.02 * x[b], z, z, Hh, f, G)
    np.testing.assert_allclose(dV, -G * Hh * 0.02, rtol=1e-12)
    np.testing.assert_allclose(dU, 0.0, atol=1e-12)
    # deta = -(D_x2d HU + D_y2d HV); Coriolis +f HV on HU, -f HU on HV (:204-205)
    de, dU, dV = L.barotropic_rhs(order, z, 3.0 * x[a], 5.0 * x[b], Hh, f, G)
    np.testing.assert_allclose(de, -8.0, rtol=1e-12)
    np.testing.assert_allclose(dU, 1e-4 * 5.0 * x[b], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(dV, -1e-4 * 3.0 * x[a], rtol=1e-12, atol=1e-18)


# ---------------------------------------------------------------------------
# 2. closed forms through the whole step
# ---------------------------------------------------------------------------

def _heun(f, dtb):
    z = -1j * f * dtb
    return z, 1 + z + z * z / 2


def check_gravity(L, order, f0, nsub=10, dt3=0.05, alpha=0.01):
    n = order + 1
    nd = n ** 3
    x = _nodes(order)
    i, j, k = _digits(order)
    E = 2
    eta = np.tile(alpha * x[i], (E, 1))
    h = np.tile(100.0 - alpha * x[i], (E, 1))
    Hf = eta + h
    F = np.zeros((E, NF, nd))
    F[:, ETA], F[:, H] = eta, Hf
    F[:, T], F[:, S] = 15.0, 35.0
    F[:, HT], F[:, HS] = Hf * 15.0, Hf * 35.0
    out = L.step(order, dt3, nsub, F.reshape(-1), h.reshape(-1), np.full(E * nd, f0),
                 L.physics(g=G))
    dtb = dt3 / nsub
    z, P = _heun(f0, dtb)
    Fc = -G * 100.0 * alpha
    w = 0j
    for _ in range(nsub):
        w = P * w + dtb * Fc * (1 + z / 2)
    # rounding residues (D applied to a constant is zero only to rounding)
    # are held to 1e-9 of the velocity the gravity term produces
    tol = 1e-9 * abs(w) / 100.0
    np.testing.assert_allclose(out[:, U], w.real / 100.0, rtol=1e-9, atol=tol)
    np.testing.assert_allclose(out[:, V], w.imag / 100.0, rtol=1e-9, atol=tol)
    # the final 2D eta at index c = d / n, whose fast digit is j(d)
    np.testing.assert_allclose(out[:, ETA], np.tile(alpha * x[j], (E, 1)), rtol=1e-9,
                               atol=1e-13)
    np.testing.assert_allclose(out[:, H], 100.0, rtol=1e-13)
    # update_baroclinic's Hu = H (U_bar + u') with the original (zero) u;
    # recombine does not refresh it
    assert np.max(np.abs(out[:, [HU, HV, OM]])) == 0.0
    assert abs(w.real) > 1e-3 and (f0 == 0.0 or abs(w.imag) > 1e-2 * abs(w.real))


def check_divergence(L, order, f0, nsub=10, dt3=0.05, beta=0.2, delta=3.0):
    n = order + 1
    nd = n ** 3
    x = _nodes(order)
    i, j, k = _digits(order)
    E = 2
    u = np.tile(beta * x[j] / 100.0, (E, 1))
    h = np.tile(100.0 + delta * j, (E, 1)).astype(np.float64)
    F = np.zeros((E, NF, nd))
    F[:, H] = 100.0
    F[:, U], F[:, HU] = u, 100.0 * u
    F[:, T], F[:, S] = 12.0, 34.0
    F[:, HT], F[:, HS] = 1200.0, 3400.0
    out = L.step(order, dt3, nsub, F.reshape(-1), h.reshape(-1), np.full(E * nd, f0),
                 L.physics(g=G))
    dtb = dt3 / nsub
    z, P = _heun(f0, dtb)
    e, eavg, W = 0.0, 0.0, 1.0 + 0j
    for _ in range(nsub):
        e -= 0.5 * dtb * beta * ((2 + z) * W).real
        W = P * W
        eavg += e
    eavg /= nsub
    c = np.arange(nd) // n  # column (2D index) of each DOF
    hc = 100.0 + delta * (c // n)  # h(c) = h at DOF c, c < n^2: 100 + delta j(c)
    np.testing.assert_allclose(out[:, ETA], e, rtol=1e-9)
    Hnew = eavg + hc
    np.testing.assert_allclose(out[:, H], np.tile(Hnew, (E, 1)), rtol=1e-12)
    # Hu = H_new * u_orig: pins eta_avg; Hv = 0
    np.testing.assert_allclose(out[:, HU], np.tile(Hnew * beta * x[j] / 100.0, (E, 1)),
                               rtol=1e-11, atol=1e-14)
    assert np.max(np.abs(out[:, HV])) < 1e-13
    # omega(c n + m) = -(m / (n - 1)) div(c) / H(c n), div(c) = H_new(c) beta / 100
    m = np.arange(nd) % n
    np.testing.assert_allclose(out[:, OM], np.tile(-(m / (n - 1)) * beta / 100.0, (E, 1)),
                               rtol=1e-9, atol=1e-15)
    # recombine: u = HU_final / H_final + u' with u' = 0
    Hfin = e + hc
    np.testing.assert_allclose(out[:, U], np.tile(beta * x[c % n] * W.real / Hfin, (E, 1)),
                               rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(out[:, V], np.tile(beta * x[c % n] * W.imag / Hfin, (E, 1)),
                               rtol=1e-9, atol=1e-15)
    assert abs(e) > 1e-3 and abs(eavg - e) > 1e-4
    assert f0 == 0.0 or abs(W.imag) > 1e-2


def check_eta_conserved(L, order):
    """BarotropicStepConservation (test_mode_splitting.cpp:189-213) through the
    whole step: a flat, resting surface keeps its total eta."""
    n = order + 1
    nd = n ** 3
    F = np.zeros((1, NF, nd))
    F[:, ETA], F[:, H] = 0.1, 100.1
    F[:, T], F[:, S], F[:, HT], F[:, HS] = 10.0, 35.0, 1001.0, 3503.5
    out = L.step(order, 0.1, 3, F.reshape(-1), np.full(nd, 100.0), np.full(nd, 1e-4),
                 L.physics(g=G))
    assert abs(out[0, ETA].sum() - F[0, ETA].sum()) < 1e-10


CHECKS = [
    ("components", lambda L, o: check_components(L, o)),
    ("gravity f=0", lambda L, o: check_gravity(L, o, 0.0)),
    ("gravity f=2", lambda L, o: check_gravity(L, o, 2.0)),
    ("divergence f=0", lambda L, o: check_divergence(L, o, 0.0)),
    ("divergence f=2", lambda L, o: check_divergence(L, o, 2.0)),
    ("eta conserved", lambda L, o: check_eta_conserved(L, o)),
]


@pytest.fixture(scope="module")
def lib():
    import oracle

    oracle.load()
    return Lib(oracle.LIB_PATH)


@pytest.mark.parametrize("order", [2, 3, 4])
@pytest.mark.parametrize("name", [c[0] for c in CHECKS])
def test_closed_forms(lib, order, name):
    dict(CHECKS)[name](lib, order)


# ---------------------------------------------------------------------------
# 3. planted mistakes (each in a copy of the oracle built on the CPU)
# ---------------------------------------------------------------------------

MUTANTS = {
    "pressure gradient along the other axis": [
        ("mvh(o, o->Dx, s->eta, ex);", "mvh(o, o->Dy, s->eta, ex);"),
        ("mvh(o, o->Dy, s->eta, ey);", "mvh(o, o->Dx, s->eta, ey);")],
    "free-surface divergence along the other axes": [
        ("mvh(o, o->Dx, s->HU, a);", "mvh(o, o->Dy, s->HU, a);"),
        ("mvh(o, o->Dy, s->HV, b);", "mvh(o, o->Dx, s->HV, b);")],
    "pressure gradient sign": [
        ("const double pgx = -g * s->H[i] * ex[i];", "const double pgx = g * s->H[i] * ex[i];"),
        ("const double pgy = -g * s->H[i] * ey[i];", "const double pgy = g * s->H[i] * ey[i];")],
    "free-surface divergence sign": [("de[i] = -(a[i] + b[i]);", "de[i] = (a[i] + b[i]);")],
    "Coriolis sign": [("const double cx = f[i] * s->HV[i];", "const double cx = -f[i] * s->HV[i];"),
                      ("const double cy = -f[i] * s->HU[i];", "const double cy = f[i] * s->HU[i];")],
    "omega sign": [("om[ih * n + k] = -sf * div * Hinv;", "om[ih * n + k] = sf * div * Hinv;")],
    "omega Hinv from H(c) instead of H(c n)": [("const double Hc = H[ih * n];",
                                                 "const double Hc = H[ih];")],
    "baroclinic divergence along the other axes": [
        ("mvh(o, o->Dx, HU, a);", "mvh(o, o->Dy, HU, a);"),
        ("mvh(o, o->Dy, HV, b);", "mvh(o, o->Dx, HV, b);")],
    "final eta instead of eta_avg": [("eavg[i] /= (double)nsub;", "eavg[i] = b.eta[i];")],
    "Heun corrector without the predictor stage": [
        ("b.eta[i] += 0.5 * dtb * (de0[i] + de1[i]);", "b.eta[i] += dtb * de0[i];")],
}


def _build_mutant(tmp_path, name, edits):
    src = open(os.path.join(ORACLE_DIR, "modesplit.c")).read()
    for old, new in edits:
        assert src.count(old) == 1, (name, old)
        src = src.replace(old, new)
    d = tmp_path / "mut"
    d.mkdir()
    for fn in os.listdir(ORACLE_DIR):
        if fn.endswith((".c", ".h")) and fn != "modesplit.c":
            shutil.copy(os.path.join(ORACLE_DIR, fn), d / fn)
    (d / "modesplit.c").write_text(src)
    so = d / "libmut.so"
    srcs = sorted(str(p) for p in d.glob("*.c"))
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-ffp-contract=off", "-D_GNU_SOURCE",
                    "-shared", "-o", str(so)] + srcs + ["-lm"], check=True)
    return str(so)


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_every_planted_mistake_is_caught(tmp_path, name):
    L = Lib(_build_mutant(tmp_path, name, MUTANTS[name]))
    caught = []
    for cname, fn in CHECKS:
        for order in (2, 3):
            try:
                fn(L, order)
            except AssertionError:
                caught.append((cname, order))
    assert caught, f"no closed form detects: {name}"
