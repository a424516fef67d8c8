"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference's FP64 algorithm used to
check the HIP implementation.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

EULER, RK2, SSPRK3, RK4, LSRK45, DP45 = range(6)


class OrcPhysics(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("g", "nu_h", "nu_v", "kappa_v_T", "kappa_v_S", "rho_0", "alpha", "beta")]


_lib = None
_PD = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_default_physics": (None, [C.POINTER(OrcPhysics)]),
        "orc_lgl_nodes": (C.c_int, [C.c_int, _PD, _PD]),
        "orc_gl_nodes": (C.c_int, [C.c_int, _PD, _PD]),
        "orc_derivative_matrix": (None, [C.c_int, _PD, _PD]),
        "orc_kron_diff": (None, [C.c_int, _PD, _PD, _PD, _PD]),
        "orc_legendre": (None, [C.c_int, C.c_double, _PD, _PD]),
        "orc_rhs": (C.c_int, [C.c_int, C.c_int64, _PD, _PD, _PD, C.POINTER(OrcPhysics), _PD]),
        "orc_rhs_range": (C.c_int, [C.c_int, C.c_int64, C.c_int64, _PD, _PD, _PD,
                                    C.POINTER(OrcPhysics), _PD]),
        "orc_update_derived": (None, [C.c_int, C.c_int64, _PD, _PD, _PD, _PD, _PD, _PD, _PD]),
        "orc_rk_step": (C.c_int, [C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double, _PD, _PD,
                                  _PD, C.POINTER(OrcPhysics)]),
        "orc_max_wave_speed": (C.c_double, [C.c_int, C.c_int64, _PD, _PD, _PD, C.c_int64]),
        "orc_compute_dt": (C.c_double, [C.c_double] * 5),
        "orc_mode_split_step": (C.c_int, [C.c_int, C.c_int64, C.c_double, C.c_int, _PD, _PD, _PD,
                                          C.POINTER(OrcPhysics)]),
        "orc_ms_depth_average": (C.c_int, [C.c_int, _PD, _PD]),
        "orc_ms_depth_integrate": (C.c_int, [C.c_int, _PD, _PD, _PD]),
        "orc_ms_deviation": (C.c_int, [C.c_int, _PD, _PD, _PD]),
        "orc_ms_divergence_2d": (C.c_int, [C.c_int, _PD, _PD, _PD]),
        "orc_ms_barotropic_rhs": (C.c_int, [C.c_int] + [_PD] * 5 + [C.c_double] + [_PD] * 3),
        "orc_mesh_create": (C.c_void_p, [C.c_double] * 6),
        "orc_mesh_destroy": (None, [C.c_void_p]),
        "orc_mesh_build_uniform": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
        "orc_mesh_refine": (None, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
        "orc_mesh_balance": (None, [C.c_void_p]),
        "orc_mesh_num_elements": (C.c_int64, [C.c_void_p]),
        "orc_mesh_elements": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_mesh_num_nodes": (C.c_int64, [C.c_void_p]),
        "orc_mesh_nodes": (None, [C.c_void_p] * 6),
        "orc_mesh_connections": (None, [C.c_void_p] * 6),
        "orc_mesh_connections_2to1": (None, [C.c_void_p] * 6),
        "orc_face_lists": (C.c_int, [C.c_int64] + [C.c_void_p] * 11),
        "orc_hydrostatic_pressure": (C.c_int, [C.c_int, C.c_int64, C.c_double, _PD, _PD, _PD]),
        "orc_pressure_gradient": (C.c_int, [C.c_int, C.c_int64, C.c_int, C.c_int, C.c_double,
                                            C.c_double] + [C.c_void_p] * 10),
        "orc_face_term": (C.c_int, [C.c_int, C.c_int64, C.c_int64, C.c_void_p, C.c_int64,
                                    C.c_void_p, C.c_void_p, C.c_int, _PD, _PD, _PD]),
        "orc_flux": (C.c_int, [C.c_int, _PD, _PD, _PD, _PD, _PD]),
        "orc_face_map": (None, [C.c_int, C.c_int, C.c_void_p]),
        "orc_conforming_flux": (C.c_int, [C.c_int, C.c_int, _PD, _PD, _PD, _PD, _PD, _PD]),
        "orc_mortar_operators": (C.c_int, [C.c_int, C.c_int] + [_PD] * 6),
        "orc_mortar_flux": (C.c_int, [C.c_int, C.c_int, C.c_int, _PD, _PD, _PD, _PD, _PD, _PD]),
        "orc_estimate": (None, [C.c_int, C.c_int64, _PD, C.c_int64, _PD, C.c_double, C.c_double,
                                C.c_double, _PD, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_mark_elements": (None, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
        "orc_modal_matrices": (C.c_int, [C.c_int, _PD, _PD]),
        "orc_modal_coeffs": (C.c_int, [C.c_int, C.c_int64, _PD, C.c_int64, _PD]),
        "orc_sigma_metrics": (None, [C.c_int64] + [_PD] * 13),
        "orc_bf16_round": (C.c_double, [C.c_double]),
        "orc_verify_mortar_conservation": (C.c_double, [C.c_int, C.c_int] + [_PD] * 4),
        "orc_project_to_children": (C.c_int, [C.c_int, C.c_int, _PD, _PD]),
        "orc_diagnose_omega": (C.c_int, [C.c_int, C.c_int64, _PD, _PD, C.c_int64, _PD]),
        "orc_diagnose_omega_from_velocity": (C.c_int, [C.c_int, C.c_int64, _PD, _PD, _PD,
                                                       C.c_int64, _PD]),
        "orc_omega_to_w": (None, [C.c_int64] + [_PD] * 8),
        "orc_omega_to_w_simple": (None, [C.c_int64] + [_PD] * 10),
        "orc_column_sigma_nodes": (C.c_int, [C.c_int, _PD]),
        "orc_diagnose_column": (C.c_int, [C.c_int, C.c_int64, _PD, _PD, C.c_void_p, _PD]),
        "orc_continuity_residual": (None, [C.c_int, C.c_int64, _PD, _PD, C.c_int64, _PD]),
        "orc_sigma_ops_metrics": (C.c_int, [C.c_int, C.c_int64] + [C.c_void_p] * 7 +
                                  [C.c_int64] + [_PD] * 6),
        "orc_sigma_ops_apply": (C.c_int, [C.c_int, C.c_int, C.c_int64] + [C.c_void_p] * 9 +
                                [C.c_int64, C.c_void_p, C.c_void_p]),
        "orc_project_from_children": (C.c_int, [C.c_int, C.c_int, _PD, _PD]),
        "orc_transfer_solution": (C.c_int64, [C.c_int, C.c_int64, _PD, C.c_int64, C.c_void_p,
                                              _PD]),
        "orc_mlp_loss_grad": (C.c_double, [_PD, C.c_int64, _PD, _PD, C.c_int, _PD, _PD]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PD)


def physics(**kw):
    p = OrcPhysics()
    load().orc_default_physics(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def lgl(n):
    x = np.zeros(n)
    w = np.zeros(n)
    assert load().orc_lgl_nodes(n, _p(x), _p(w)) == 0
    return x, w


def gl(n):
    x = np.zeros(n)
    w = np.zeros(n)
    assert load().orc_gl_nodes(n, _p(x), _p(w)) == 0
    return x, w


def dmat(nodes):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    n = len(nodes)
    D = np.zeros(n * n)
    load().orc_derivative_matrix(n, _p(nodes), _p(D))
    return D.reshape(n, n)


def kron(order, D1):
    n = order + 1
    nd = n ** 3
    D1 = np.ascontiguousarray(D1, dtype=np.float64)
    a = np.zeros(nd * nd)
    b = np.zeros(nd * nd)
    c = np.zeros(nd * nd)
    load().orc_kron_diff(order, _p(D1), _p(a), _p(b), _p(c))
    return a.reshape(nd, nd), b.reshape(nd, nd), c.reshape(nd, nd)


def rhs(order, U, h, f, phys=None, e0=None, e1=None):
    n = order + 1
    nd = n ** 3
    E = U.size // (5 * nd)
    out = np.zeros_like(U)
    p = phys or physics()
    if e0 is None:
        rc = load().orc_rhs(order, E, _p(U), _p(h), _p(f), C.byref(p), _p(out))
    else:
        rc = load().orc_rhs_range(order, e0, e1, _p(U), _p(h), _p(f), C.byref(p), _p(out))
    assert rc == 0
    return out


def rk_step(scheme, order, t, dt, U, h, f, phys=None):
    n = order + 1
    E = U.size // (5 * n ** 3)
    V = U.copy()
    p = phys or physics()
    assert load().orc_rk_step(scheme, order, E, t, dt, _p(V), _p(h), _p(f), C.byref(p)) == 0
    return V


def derived(order, U, h):
    n = order + 1
    nd = n ** 3
    E = U.size // (5 * nd)
    arrs = [np.zeros(E * nd) for _ in range(5)]
    load().orc_update_derived(order, E, _p(U), _p(h), *[_p(a) for a in arrs])
    return arrs  # H, u, v, T, S


def max_wave_speed(order, H, u, v, stride=None):
    n = order + 1
    nd = n ** 3
    stride = stride or nd
    E = H.size // stride
    return load().orc_max_wave_speed(order, E, _p(H), _p(u), _p(v), stride)


def compute_dt(c, dx_min, cfl, dt_min, dt_max):
    return load().orc_compute_dt(c, dx_min, cfl, dt_min, dt_max)


def mode_split_step(order, dt3, nsub, full, h, f, phys=None):
    n = order + 1
    E = full.size // (12 * n ** 3)
    V = full.copy()
    p = phys or physics()
    assert load().orc_mode_split_step(order, E, dt3, nsub, _p(V), _p(h), _p(f), C.byref(p)) == 0
    return V


CONN_SAME, CONN_FINE2X1, CONN_FINE1X2, CONN_FINE3_2P1, CONN_FINE3_1P2, CONN_FINE2X2, CONN_BOUNDARY = range(7)


class Mesh:
    """OctreeAdapter restatement (oracle/mesh.c)."""

    def __init__(self, xmin, xmax, ymin, ymax, zmin, zmax):
        self._l = load()
        self.h = self._l.orc_mesh_create(xmin, xmax, ymin, ymax, zmin, zmax)

    def __del__(self):
        try:
            self._l.orc_mesh_destroy(self.h)
        except Exception:
            pass

    def build_uniform(self, nx, ny, nz):
        self._l.orc_mesh_build_uniform(self.h, nx, ny, nz)

    def refine(self, elems, masks):
        e = np.ascontiguousarray(elems, dtype=np.int64)
        m = np.ascontiguousarray(masks, dtype=np.int32)
        self._l.orc_mesh_refine(self.h, e.size, e.ctypes.data, m.ctypes.data)

    def balance(self):
        self._l.orc_mesh_balance(self.h)

    @property
    def num_elements(self):
        return self._l.orc_mesh_num_elements(self.h)

    def elements(self):
        E = self.num_elements
        b = np.zeros((E, 6))
        lv = np.zeros((E, 3), dtype=np.int32)
        mo = np.zeros(E, dtype=np.uint64)
        self._l.orc_mesh_elements(self.h, b.ctypes.data, lv.ctypes.data, mo.ctypes.data)
        return b, lv, mo

    def nodes(self):
        """Every tree node in creation order: bounds (n, 6), levels (n, 3),
        split mask, child count, element number (-1 for interior nodes)."""
        n = self._l.orc_mesh_num_nodes(self.h)
        b = np.zeros((n, 6))
        lv = np.zeros((n, 3), dtype=np.int32)
        mk = np.zeros(n, dtype=np.int32)
        nc = np.zeros(n, dtype=np.int32)
        lf = np.zeros(n, dtype=np.int64)
        self._l.orc_mesh_nodes(self.h, b.ctypes.data, lv.ctypes.data, mk.ctypes.data,
                               nc.ctypes.data, lf.ctypes.data)
        return b, lv, mk, nc, lf

    def connections(self, two_to_one=False):
        E = self.num_elements
        t = np.zeros((E, 6), dtype=np.uint8)
        c = np.zeros((E, 6), dtype=np.int8)
        el = np.zeros((E, 6, 4), dtype=np.int64)
        fa = np.zeros((E, 6, 4), dtype=np.int8)
        su = np.zeros((E, 6, 4), dtype=np.int8)
        fn = self._l.orc_mesh_connections_2to1 if two_to_one else self._l.orc_mesh_connections
        fn(self.h, t.ctypes.data, c.ctypes.data, el.ctypes.data, fa.ctypes.data, su.ctypes.data)
        return t, c, el, fa, su


def face_lists(bounds, levels, conn):
    """Interface lists of a connectivity (oracle/faces.c): conforming (nc, 4),
    mortar (nm, 6), groups (ng, 5)."""
    t, c, el = (np.ascontiguousarray(conn[0], dtype=np.uint8),
                np.ascontiguousarray(conn[1], dtype=np.int8),
                np.ascontiguousarray(conn[2], dtype=np.int64))
    b = np.ascontiguousarray(bounds, dtype=np.float64)
    lv = np.ascontiguousarray(levels, dtype=np.int32)
    E = b.shape[0]
    cf = np.zeros((6 * E, 4), dtype=np.int64)
    mo = np.zeros((6 * E, 6), dtype=np.int64)
    gr = np.zeros((6 * E, 5), dtype=np.int64)
    nc, nm, ng = C.c_int64(), C.c_int64(), C.c_int64()
    rc = load().orc_face_lists(E, b.ctypes.data, lv.ctypes.data, t.ctypes.data, c.ctypes.data,
                               el.ctypes.data, cf.ctypes.data, C.byref(nc), mo.ctypes.data,
                               C.byref(nm), gr.ctypes.data, C.byref(ng))
    if rc:
        raise ValueError("a connection without a 2:1 mortar")
    return cf[:nc.value].copy(), mo[:nm.value].copy(), gr[:ng.value].copy()


def face_term(order, E, conf, mort, mort_type, kind, p, U, dU=None):
    """dU + the face term (packed layout); dU = 0 when None."""
    out = np.zeros(E * 5 * (order + 1) ** 3) if dU is None else np.array(dU, dtype=np.float64)
    conf = np.ascontiguousarray(conf, dtype=np.int64).reshape(-1, 4)
    mort = np.ascontiguousarray(mort, dtype=np.int64).reshape(-1, 6)
    mt = np.ascontiguousarray(mort_type, dtype=np.int32)
    pp = np.zeros(4)
    pp[: len(p)] = p
    rc = load().orc_face_term(order, E, conf.shape[0], conf.ctypes.data, mort.shape[0],
                              mort.ctypes.data, mt.ctypes.data, kind, _p(pp), _p(_v(U)), _p(out))
    assert rc == 0
    return out


SW_LF, SW_HLLC, SW_ROE, SW_CENTRAL, PE_LF5, UPWIND, LF_ADV = range(7)


def _v(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def flux(kind, UL, UR, n, p):
    out = np.zeros(5)
    nv = load().orc_flux(kind, _p(_v(UL)), _p(_v(UR)), _p(_v(n)), _p(_v(p)), _p(out))
    assert nv > 0
    return out[:nv]


def face_map(order, face):
    m = np.zeros((order + 1) ** 2, dtype=np.int32)
    load().orc_face_map(order, face, m.ctypes.data)
    return m


def conforming_flux(order, kind, p, UL, UR, n):
    nf = (order + 1) ** 2
    rl = np.zeros(nf)
    rr = np.zeros(nf)
    load().orc_conforming_flux(order, kind, _p(_v(p)), _p(_v(UL)), _p(_v(UR)), _p(_v(n)),
                               _p(rl), _p(rr))
    return rl, rr


def mortar_operators(order, ctype):
    nm = (order + 1) ** 2
    M = np.zeros((nm, nm))
    Mi = np.zeros((nm, nm))
    Pc = np.zeros((nm, nm))
    Lc = np.zeros((nm, nm))
    Pf = np.zeros((4, nm, nm))
    Lf = np.zeros((4, nm, nm))
    nfine = load().orc_mortar_operators(order, ctype, _p(M), _p(Mi), _p(Pc), _p(Lc), _p(Pf), _p(Lf))
    return M, Mi, Pc, Lc, Pf[:nfine], Lf[:nfine]


def mortar_flux(order, ctype, kind, p, Uc, Uf, n):
    nm = (order + 1) ** 2
    rc = np.zeros(nm)
    rf = np.zeros(4 * nm)
    nfine = load().orc_mortar_flux(order, ctype, kind, _p(_v(p)), _p(_v(Uc)), _p(_v(Uf).reshape(-1)),
                                   _p(_v(n)), _p(rc), _p(rf))
    return rc, rf[:nfine * nm].reshape(nfine, nm)


def estimate(order, field, bounds, refine_threshold=0.1, coarsen_threshold=0.01,
             smoothness_factor=0.5):
    """GradientErrorEstimator::estimate; field (E, ndof), bounds (E, 6)."""
    field = np.ascontiguousarray(field, dtype=np.float64)
    bounds = np.ascontiguousarray(bounds, dtype=np.float64)
    E = field.shape[0]
    value = np.zeros(E)
    mask, refine, coarsen = (np.zeros(E, dtype=np.uint8) for _ in range(3))
    load().orc_estimate(order, E, _p(field), field.shape[1], _p(bounds), refine_threshold,
                        coarsen_threshold, smoothness_factor, _p(value), mask.ctypes.data,
                        refine.ctypes.data, coarsen.ctypes.data)
    return value, mask, refine, coarsen


def mark_elements(refine, coarsen, mask, levels, max_level=8, min_level=0):
    E = len(refine)
    refine, coarsen, mask = (np.ascontiguousarray(a, dtype=np.uint8) for a in (refine, coarsen, mask))
    levels = np.ascontiguousarray(levels, dtype=np.int32)
    action, mout = np.zeros(E, dtype=np.uint8), np.zeros(E, dtype=np.uint8)
    load().orc_mark_elements(E, refine.ctypes.data, coarsen.ctypes.data, mask.ctypes.data,
                             levels.ctypes.data, max_level, min_level, action.ctypes.data,
                             mout.ctypes.data)
    return action, mout


def modal_matrices(order):
    n = order + 1
    V, Vi = np.zeros((n, n)), np.zeros((n, n))
    assert load().orc_modal_matrices(order, _p(V), _p(Vi)) == 0
    return V, Vi


def modal_coeffs(order, u):
    """u (E, ndof) nodal -> (E, ndof) orthonormal Legendre coefficients"""
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u)
    assert load().orc_modal_coeffs(order, u.shape[0], _p(u), u.shape[1], _p(out)) == 0
    return out


def sigma_metrics(sigma, eta, h, deta_dx, deta_dy, dh_dx, dh_dy, deta_dt=None):
    """SigmaMetrics::update -> (H, Hinv, dsigma_dx, dsigma_dy, dsigma_dt)"""
    a = [_v(x) for x in (sigma, eta, h, deta_dx, deta_dy, dh_dx, dh_dy)]
    n = a[0].size
    et = None if deta_dt is None else _v(deta_dt)
    outs = [np.zeros(n) for _ in range(5)]
    load().orc_sigma_metrics(n, *[_p(x) for x in a], None if et is None else _p(et),
                             *[_p(o) for o in outs])
    return tuple(outs)


MLP_IN, MLP_HIDDEN, MLP_NPARAM = 128, 64, 12481


def mlp_loss_grad(params, X, y, quantize=True, want_grad=True):
    """-> (loss, pred, grad) of the MSE loss of the indicator MLP"""
    params, X, y = _v(params), _v(X), _v(y)
    B = X.shape[0]
    pred = np.zeros(B)
    grad = np.zeros(MLP_NPARAM) if want_grad else None
    loss = load().orc_mlp_loss_grad(_p(params), B, _p(X), _p(y), int(quantize), _p(pred),
                                    None if grad is None else _p(grad))
    return loss, pred, grad


def project_to_children(order, mask, U):
    nd = (order + 1) ** 3
    out = np.zeros(8 * nd)
    nc = load().orc_project_to_children(order, mask, _p(_v(U)), _p(out))
    return out[: nc * nd].reshape(nc, nd)


def project_from_children(order, mask, children):
    nd = (order + 1) ** 3
    out = np.zeros(nd)
    load().orc_project_from_children(order, mask, _p(_v(children).reshape(-1)), _p(out))
    return out


def transfer_solution(order, U, parents):
    """U (E, ndof); parents ascending -> (E - n + 8 n, ndof)"""
    U = _v(U)
    E, nd = U.shape
    parents = np.ascontiguousarray(parents, dtype=np.int64)
    out = np.zeros((E + 8 * len(parents), nd))
    k = load().orc_transfer_solution(order, E, _p(U), len(parents), parents.ctypes.data, _p(out))
    return out[:k]


def diagnose_omega(order, div, H):
    """div, H (E, ndof) -> omega (E, ndof)"""
    div, H = _v(div), _v(H)
    out = np.zeros_like(div)
    load().orc_diagnose_omega(order, div.shape[0], _p(div), _p(H), div.shape[1], _p(out))
    return out


def diagnose_omega_from_velocity(order, dHu_dx, dHv_dy, H):
    a, b, H = _v(dHu_dx), _v(dHv_dy), _v(H)
    out = np.zeros_like(a)
    load().orc_diagnose_omega_from_velocity(order, a.shape[0], _p(a), _p(b), _p(H), a.shape[1],
                                            _p(out))
    return out


def omega_to_w(omega, u, v, H, dz_dx, dz_dy, dz_dt):
    a = [_v(x).reshape(-1) for x in (omega, u, v, H, dz_dx, dz_dy, dz_dt)]
    out = np.zeros(a[0].size)
    load().orc_omega_to_w(a[0].size, *[_p(x) for x in a], _p(out))
    return out


def omega_to_w_simple(omega, u, v, eta, h, sigma, deta_dx, deta_dy, deta_dt):
    a = [_v(x).reshape(-1) for x in (omega, u, v, eta, h, sigma, deta_dx, deta_dy, deta_dt)]
    out = np.zeros(a[0].size)
    load().orc_omega_to_w_simple(a[0].size, *[_p(x) for x in a], _p(out))
    return out


def column_sigma_nodes(order):
    out = np.zeros(order + 1)
    load().orc_column_sigma_nodes(order, _p(out))
    return out


def diagnose_column(order, div, H, omega_bottom=None):
    """div, H (ncol, order+1) -> omega (ncol, order+1)"""
    div, H = _v(div), _v(H)
    ob = None if omega_bottom is None else np.ascontiguousarray(omega_bottom, dtype=np.float64)
    out = np.zeros_like(div)
    load().orc_diagnose_column(order, div.shape[0], _p(div), _p(H),
                               None if ob is None else ob.ctypes.data, _p(out))
    return out


def continuity_residual(order, div, dHomega):
    """div, dHomega (E, ndof) -> RMS residual per element (E,)"""
    div, dh = _v(div), _v(dHomega)
    out = np.zeros(div.shape[0])
    load().orc_continuity_residual(order, div.shape[0], _p(div), _p(dh), div.shape[1], _p(out))
    return out


def _vp(a):
    return None if a is None else a.ctypes.data


def sigma_ops_metrics(order, eta, h, deta_dx, deta_dy, dh_dx, dh_dy, deta_dt=None):
    """(E, ndof) arrays -> dict H, Hinv, sigma, dsdx, dsdy, dsdt"""
    ins = [_v(a) for a in (eta, h, deta_dx, deta_dy, dh_dx, dh_dy)]
    dt = None if deta_dt is None else _v(deta_dt)
    E, nd = ins[0].shape
    outs = {k: np.zeros((E, nd)) for k in ("H", "Hinv", "sigma", "dsdx", "dsdy", "dsdt")}
    load().orc_sigma_ops_metrics(order, E, *[_vp(a) for a in ins], _vp(dt), nd,
                                 *[_p(outs[k]) for k in ("H", "Hinv", "sigma", "dsdx", "dsdy",
                                                         "dsdt")])
    return outs


SIGMA_OP_HGRAD, SIGMA_OP_VGRAD, SIGMA_OP_DIV, SIGMA_OP_ALE, SIGMA_OP_MATDER = range(5)


def sigma_ops_apply(order, op, U, a=None, b=None, c=None, f=None, g=None, dsdx=None, dsdy=None,
                    Hinv=None):
    U = _v(U)
    E, nd = U.shape
    conv = [None if x is None else _v(x) for x in (a, b, c, f, g, dsdx, dsdy, Hinv)]
    o1 = np.zeros((E, nd))
    o2 = np.zeros((E, nd)) if op == SIGMA_OP_HGRAD else None
    load().orc_sigma_ops_apply(order, op, E, _vp(U), *[_vp(x) for x in conv], nd, _vp(o1),
                               _vp(o2))
    return (o1, o2) if op == SIGMA_OP_HGRAD else o1


def verify_mortar_conservation(rc, rf, wc, wf):
    """rc (nm,), rf (nfine, nm), wc (nm,), wf (nfine, nm) -> relative error"""
    rc, rf, wc, wf = (_v(a) for a in (rc, rf, wc, wf))
    return load().orc_verify_mortar_conservation(rc.size, rf.shape[0], _p(rc), _p(rf), _p(wc),
                                                 _p(wf))


def hydrostatic_pressure(order, rho, H, g=9.81):
    rho, H = _v(rho), _v(H)
    nd = (order + 1) ** 3
    p = np.zeros_like(rho)
    assert load().orc_hydrostatic_pressure(order, rho.size // nd, g, _p(rho), _p(H), _p(p)) == 0
    return p


def pressure_gradient(order, method, parts, eta, rho, H, sigma, deta_dx=None, deta_dy=None,
                      dh_dx=None, dh_dy=None, g=9.81, rho0=1025.0):
    """(pg_x, pg_y) element-major; method 0 standard, 1 density Jacobian;
    parts 1 barotropic, 2 baroclinic, 3 both."""
    arrs = [_v(a) for a in (eta, rho, H, sigma)]
    opt = [None if a is None else _v(a) for a in (deta_dx, deta_dy, dh_dx, dh_dy)]
    nd = (order + 1) ** 3
    px = np.zeros_like(arrs[0])
    py = np.zeros_like(arrs[0])
    ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
    rc = load().orc_pressure_gradient(order, arrs[0].size // nd, method, parts, g, rho0,
                                      *(a.ctypes.data for a in arrs), *(ptr(a) for a in opt),
                                      px.ctypes.data, py.ctypes.data)
    assert rc == 0
    return px, py
