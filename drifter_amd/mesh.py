"""Octree mesh and face connectivity through the C ABI (drf_mesh_*), mirroring
OctreeAdapter (include/mesh/octree_adapter.hpp:103-273) and FaceConnection
(include/dg/face_connection.hpp:51-104)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

SAME_LEVEL, FINE_2X1, FINE_1X2, FINE_3_2PLUS1, FINE_3_1PLUS2, FINE_2X2, BOUNDARY = range(7)
MASK_X, MASK_Y, MASK_Z = 1, 2, 4
MASK_XY = MASK_X | MASK_Y
MASK_XYZ = 7


class OctreeAdapter:
    def __init__(self, xmin, xmax, ymin, ymax, zmin, zmax):
        self._l = _lib.load()
        h = C.c_void_p()
        check(self._l.drf_mesh_create(xmin, xmax, ymin, ymax, zmin, zmax, C.byref(h)))
        self._h = h

    def __del__(self):
        try:
            self._l.drf_mesh_destroy(self._h)
        except Exception:
            pass

    def build_uniform(self, nx, ny, nz):
        check(self._l.drf_mesh_build_uniform(self._h, nx, ny, nz))

    def refine(self, elements, masks):
        e = np.ascontiguousarray(elements, dtype=np.int64)
        m = np.ascontiguousarray(masks, dtype=np.uint8)
        check(self._l.drf_mesh_refine(self._h, e.size, e.ctypes.data, m.ctypes.data))

    def balance(self):
        check(self._l.drf_mesh_balance(self._h))

    def num_elements(self):
        return self._l.drf_mesh_num_elements(self._h)

    def elements(self):
        E = self.num_elements()
        b = np.zeros((E, 6))
        lv = np.zeros((E, 3), dtype=np.int32)
        mo = np.zeros(E, dtype=np.uint64)
        check(self._l.drf_mesh_elements(self._h, b.ctypes.data, lv.ctypes.data, mo.ctypes.data))
        return b, lv, mo

    def element_bounds(self, e):
        b = np.zeros(6)
        check(self._l.drf_mesh_element_bounds(self._h, e, _lib.dptr(b)))
        return b

    def build_face_connections(self, two_to_one=False):
        """OctreeAdapter::build_face_connections (octree_adapter.cpp:369-389)
        as flat arrays (type, count, elems, faces, subs).  two_to_one: the
        neighbour search stops at the element's own size, so a refined
        neighbour gives its fine leaves and a 2:1 connection type (the
        reference's search always reaches a leaf and reports SameLevel);
        the face plans' mortars come from this mode."""
        E = self.num_elements()
        t = np.zeros((E, 6), dtype=np.uint8)
        c = np.zeros((E, 6), dtype=np.int8)
        el = np.zeros((E, 6, 4), dtype=np.int64)
        fa = np.zeros((E, 6, 4), dtype=np.int8)
        su = np.zeros((E, 6, 4), dtype=np.int8)
        fn = (self._l.drf_mesh_face_connections_2to1 if two_to_one
              else self._l.drf_mesh_face_connections)
        check(fn(self._h, t.ctypes.data, c.ctypes.data, el.ctypes.data, fa.ctypes.data,
                 su.ctypes.data))
        return t, c, el, fa, su


def coastline_mesh(base=(32, 32, 4), L=100000.0, rounds=3, band=5000.0, amp=0.05, wavelength=0.25):
    """Adaptive mesh near a sinusoidal coastline y_c(x) = 0.8 L + amp L sin(2 pi x / (wavelength L)):
    `rounds` times, every leaf whose centre lies within `band` of the curve
    (vertically) is refined in XY, each refine followed by the 2:1 balance."""
    m = OctreeAdapter(0.0, L, 0.0, L, -1.0, 0.0)
    m.build_uniform(*base)
    for _ in range(rounds):
        b, _, _ = m.elements()
        cx = 0.5 * (b[:, 0] + b[:, 1])
        cy = 0.5 * (b[:, 2] + b[:, 3])
        yc = 0.8 * L + amp * L * np.sin(2.0 * np.pi * cx / (wavelength * L))
        sel = np.nonzero(np.abs(cy - yc) < band)[0]
        m.refine(sel, np.full(sel.size, MASK_XY))
    return m
