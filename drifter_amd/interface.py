"""Python mirror of the reference's face components, running on the HIP
library through the C ABI.

  NumericalFlux::flux          include/flux/numerical_flux.hpp:31-53
                               (one call evaluates a batch of face points)
  LGL face extraction          src/dg/basis_hexahedron.cpp:431-522
  ConformingInterface          include/dg/mortar.hpp:187-215
  MortarSpace                  include/dg/mortar.hpp:32-136

Device arrays are torch tensors (float64, contiguous) on the GPU; sizes and
layouts are those of include/drifter_hip/drf.h.  Errors raise DrfError.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .solver import _ptr, _sptr

SW_LF, SW_HLLC, SW_ROE, SW_CENTRAL, PE_LF5, UPWIND, LF_ADV = range(7)
NUM_VARS = {SW_LF: 3, SW_HLLC: 3, SW_ROE: 3, SW_CENTRAL: 3, PE_LF5: 5, UPWIND: 1, LF_ADV: 1}

(CONN_SAME_LEVEL, CONN_FINE_2X1, CONN_FINE_1X2, CONN_FINE_3_2PLUS1, CONN_FINE_3_1PLUS2,
 CONN_FINE_2X2, CONN_BOUNDARY) = range(7)


def _params(p):
    a = np.zeros(4, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64).reshape(-1)
    a[: len(p)] = p
    return a


def _dev(t, n, name):
    if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == 8
            and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous float64 device tensor")
    if t.numel() < n:
        raise ValueError(f"{name} holds {t.numel()} values, needs {n}")
    return _ptr(t)


def flux_points(kind, UL, UR, normals, params, out=None, stream=None):
    """F*(UL, UR, n) at npts points: UL, UR [nvar][npts], normals [3][npts]."""
    import torch

    nv = NUM_VARS[kind]
    npts = UL.shape[-1]
    if out is None:
        out = torch.empty((nv, npts), dtype=torch.float64, device=UL.device)
    p = _params(params)
    check(_lib.load().drf_flux_points(
        kind, npts, _dev(UL, nv * npts, "UL"), _dev(UR, nv * npts, "UR"),
        _dev(normals, 3 * npts, "normals"), p.ctypes.data_as(_lib._PD),
        _dev(out, nv * npts, "out"), _sptr(stream)))
    return out


def face_map(order, face):
    """Volume node of each face DOF it + (N+1) jt of face 0..5."""
    n = order + 1
    m = np.zeros(n * n, dtype=np.int32)
    check(_lib.load().drf_face_map(order, face, m.ctypes.data_as(C.c_void_p)))
    return m


def gather_face_traces(order, field, elems, faces, n_elem, elem_stride=None, stream=None):
    """out[k] = face values of field for element elems[k], face faces[k]."""
    import torch

    n = order + 1
    nf = n * n
    es = n ** 3 if elem_stride is None else elem_stride
    if elems.dtype != torch.int64 or faces.dtype != torch.int8:
        raise ValueError("elems must be int64 and faces int8 device tensors")
    k = elems.numel()
    out = torch.empty((k, nf), dtype=torch.float64, device=field.device)
    check(_lib.load().drf_gather_face_traces(
        order, n_elem, es, k, _ptr(elems), _ptr(faces),
        _dev(field, (n_elem - 1) * es + n ** 3 if n_elem else 0, "field"), _ptr(out),
        _sptr(stream)))
    return out


class ConformingInterface:
    """Same-level face pairs; the reference's direct per-point flux with an
    N-point LGL face rule (src/dg/mortar.cpp:297-343)."""

    def __init__(self, order):
        self.order = order
        self.nf = (order + 1) ** 2

    def compute_flux(self, kind, params, UL, UR, normals, stream=None):
        import torch

        nfaces = UL.shape[0]
        rl = torch.empty((nfaces, self.nf), dtype=torch.float64, device=UL.device)
        rr = torch.empty_like(rl)
        p = _params(params)
        check(_lib.load().drf_conforming_flux_batch(
            self.order, kind, p.ctypes.data_as(_lib._PD), nfaces,
            _dev(UL, nfaces * self.nf, "UL"), _dev(UR, nfaces * self.nf, "UR"),
            _dev(normals, 3 * nfaces, "normals"), _ptr(rl), _ptr(rr), _sptr(stream)))
        return rl, rr


class MortarSpace:
    """Projection / lift operators of one connection type and the batched
    mortar flux.  device=None builds the operators on the host only."""

    def __init__(self, order, conn_type, device=0):
        self._l = _lib.load()
        h = C.c_void_p()
        check(self._l.drf_mortar_create(order, conn_type, -1 if device is None else device,
                                        C.byref(h)))
        self._m = h
        self.order = order
        self.conn_type = conn_type

    def __del__(self):
        if getattr(self, "_m", None):
            self._l.drf_mortar_destroy(self._m)
            self._m = None

    def num_dofs(self):
        return self._l.drf_mortar_num_dofs(self._m)

    def num_fine_elements(self):
        return self._l.drf_mortar_num_fine(self._m)

    def operators(self):
        """(M, Minv, Pc, Lc, Pf[nfine], Lf[nfine]) as numpy arrays."""
        nm, nfine = self.num_dofs(), self.num_fine_elements()
        M, Mi, Pc, Lc = (np.zeros((nm, nm)) for _ in range(4))
        Pf, Lf = np.zeros((nfine, nm, nm)), np.zeros((nfine, nm, nm))
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        check(self._l.drf_mortar_operators(self._m, vp(M), vp(Mi), vp(Pc), vp(Lc), vp(Pf),
                                           vp(Lf)))
        return M, Mi, Pc, Lc, Pf, Lf

    def compute_mortar_flux(self, kind, params, Uc, Uf, normals, stream=None):
        """Uc [n][nm], Uf [n][nfine][nm], normals [n][3] -> (rhs_c, rhs_f)."""
        import torch

        nm, nfine = self.num_dofs(), self.num_fine_elements()
        n = Uc.shape[0]
        rc = torch.empty((n, nm), dtype=torch.float64, device=Uc.device)
        rf = torch.empty((n, nfine, nm), dtype=torch.float64, device=Uc.device)
        p = _params(params)
        check(self._l.drf_mortar_flux_batch(
            self._m, kind, p.ctypes.data_as(_lib._PD), n, _dev(Uc, n * nm, "Uc"),
            _dev(Uf, n * nfine * nm, "Uf"), _dev(normals, 3 * n, "normals"), _ptr(rc),
            _ptr(rf), _sptr(stream)))
        return rc, rf

    def verify_conservation(self, rhs_coarse, rhs_fine, w_coarse, w_fine):
        """verify_mortar_conservation (src/dg/mortar.cpp:386-413) for one
        interface (host numpy arrays): relative imbalance of the weighted
        coarse and fine integrals, 0 when both are below 1e-15 together."""
        nm, nfine = self.num_dofs(), self.num_fine_elements()
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                for a in (rhs_coarse, rhs_fine, w_coarse, w_fine)]
        for a, n, name in zip(arrs, (nm, nfine * nm, nm, nfine * nm),
                              ("rhs_coarse", "rhs_fine", "w_coarse", "w_fine")):
            if a.size != n:
                raise ValueError(f"{name}: expected {n} values, got {a.size}")
        out = C.c_double()
        check(self._l.drf_verify_mortar_conservation(
            self._m, *(a.ctypes.data_as(_lib._PD) for a in arrs), C.byref(out)))
        return out.value

    def verify_conservation_batch(self, rhs_coarse, rhs_fine, w_coarse, w_fine, stream=None):
        """The same for a batch on the device: rhs_coarse [n][nm], rhs_fine
        [n][nfine][nm] (as compute_mortar_flux returns them), weights shared by
        every interface -> relative errors [n]."""
        import torch

        nm, nfine = self.num_dofs(), self.num_fine_elements()
        n = rhs_coarse.shape[0]
        out = torch.empty(n, dtype=torch.float64, device=rhs_coarse.device)
        check(self._l.drf_mortar_conservation_dev(
            self._m, n, _dev(rhs_coarse, n * nm, "rhs_coarse"),
            _dev(rhs_fine, n * nfine * nm, "rhs_fine"), _dev(w_coarse, nm, "w_coarse"),
            _dev(w_fine, nfine * nm, "w_fine"), _ptr(out), _sptr(stream)))
        return out
