"""Face term of the DG right-hand side on an octree mesh, running on the HIP
library through the C ABI (drf_face_plan_*).

Mirrors the face loop of DG3DIntegration::compute_rhs
(src/dg/operators_3d.cpp:255-281): conforming interfaces as
compute_interface_flux_conforming (:320-343) computes them through
ConformingInterface, 2:1 mortars as compute_interface_flux_nonconforming
(:345-391) computes them through MortarSpace, and the inverse mass matrix
applied to what the faces add.  The interfaces are grouped by normal axis,
kind (conforming / mortar type) and level; each (axis, kind) runs as one
persistent kernel.

The reference's own connectivity (OctreeAdapter::build_face_connections)
never reports a refined neighbour, so a plan built from it holds conforming
interfaces only; `FacePlan.from_mesh(..., two_to_one=True)` takes the
connectivity whose neighbour search stops at the element's own size, which
gives the 2:1 mortar faces.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .solver import _dev_array, _sptr

UPWIND, LF_ADV = 5, 6
BY_INTERFACE, BY_ELEMENT, IN_VOLUME = 0, 1, 2
KIND_CONFORMING = 0


class FacePlan:
    """Interface lists of one mesh for one solver context (its order,
    element count, precision and device)."""

    def __init__(self, solver, bounds, levels, conn, kind=LF_ADV, params=(0.3, 0.2, 0.1, 0.5)):
        t, c, el, fa, _ = conn
        E = solver.num_elements
        bounds = np.ascontiguousarray(bounds, dtype=np.float64)
        levels = np.ascontiguousarray(levels, dtype=np.int32)
        arrs = [np.ascontiguousarray(t, dtype=np.uint8), np.ascontiguousarray(c, dtype=np.int8),
                np.ascontiguousarray(el, dtype=np.int64), np.ascontiguousarray(fa, dtype=np.int8)]
        if bounds.size != 6 * E or levels.size != 3 * E or arrs[0].size != 6 * E \
                or arrs[2].size != 24 * E:
            raise ValueError("bounds, levels and the connectivity must cover the solver's "
                             f"{E} elements")
        p = np.zeros(4)
        p[: len(params)] = params
        self._l = _lib.load()
        self.solver = solver
        self.kind = kind
        self.params = p
        h = C.c_void_p()
        check(self._l.drf_face_plan_create(
            solver.ctx, E, bounds.ctypes.data, levels.ctypes.data, arrs[0].ctypes.data,
            arrs[1].ctypes.data, arrs[2].ctypes.data, arrs[3].ctypes.data, kind,
            p.ctypes.data_as(_lib._PD), C.byref(h)))
        self._h = h

    @classmethod
    def from_mesh(cls, solver, mesh, two_to_one=True, **kw):
        b, lv, _ = mesh.elements()
        return cls(solver, b, lv, mesh.build_face_connections(two_to_one=two_to_one), **kw)

    def close(self):
        if self._h:
            self._l.drf_face_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        g, nc, nm = C.c_int(), C.c_int64(), C.c_int64()
        check(self._l.drf_face_plan_info(self._h, C.byref(g), C.byref(nc), C.byref(nm)))
        return g.value, nc.value, nm.value

    def groups(self):
        """(ngroups, 5) int64: axis, kind, level, first, count."""
        g, _, _ = self.info()
        d = np.zeros((g, 5), dtype=np.int64)
        check(self._l.drf_face_plan_groups(self._h, d.ctypes.data))
        return d

    def lists(self):
        """(conforming (nc, 4), mortar (nm, 6)) int64 in group order."""
        _, nc, nm = self.info()
        cf = np.zeros((nc, 4), dtype=np.int64)
        mo = np.zeros((nm, 6), dtype=np.int64)
        check(self._l.drf_face_plan_lists(self._h, cf.ctypes.data, mo.ctypes.data))
        return cf, mo

    def set_conforming_mode(self, mode=None):
        """How the conforming interfaces run: IN_VOLUME (the default: inside
        the volume kernel of an attached right-hand side, element by element
        in face_term), BY_ELEMENT (one element-ordered launch) or
        BY_INTERFACE (interface by interface, one launch per normal axis)."""
        check(self._l.drf_face_plan_set_conforming_mode(
            self._h, IN_VOLUME if mode is None else int(mode)))

    def face_term(self, U, dU, layout=_lib.LAYOUT_LINES, stream=None):
        """dU += the face term of state U (device tensors of the context's precision)."""
        s = self.solver
        check(self._l.drf_face_term_dev(self._h, s._dev_state(U, "U", layout),
                                        s._dev_state(dU, "dU", layout),
                                        layout, _sptr(stream)))

    def group_fluxes(self, group, U, layout=_lib.LAYOUT_LINES, stream=None):
        """Face vectors of one group before lifting: conforming -> (rhs_l, rhs_r)
        [n][5][nh]; mortar -> (rhs_c [n][5][nh], rhs_f [n][5][nfine][nh])."""
        import torch

        s = self.solver
        g = self.groups()[group]
        nh = s.n * s.n
        cnt, kind = int(g[4]), int(g[1])
        nf = {1: 2, 2: 2, 5: 4}.get(kind, 1)
        dt = s._state_dtype()
        a = torch.empty((cnt, 5, nh), dtype=dt, device=U.device)
        b = torch.empty((cnt, 5, nh) if kind == KIND_CONFORMING else (cnt, 5, nf, nh), dtype=dt,
                        device=U.device)
        check(self._l.drf_face_group_fluxes_dev(
            self._h, group, s._dev_state(U, "U", layout), layout,
            _dev_array(a, a.numel(), dt, "out_a"),
            _dev_array(b, b.numel(), dt, "out_b"), _sptr(stream)))
        return a, b

    def attach(self):
        """Add this plan's face term to every right-hand side of its solver."""
        check(self._l.drf_set_face_plan(self.solver.ctx, self._h))

    def detach(self):
        check(self._l.drf_set_face_plan(self.solver.ctx, None))
