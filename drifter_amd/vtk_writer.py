"""VTK output through the C ABI, mirroring VTKWriter, OceanVTKWriter and
HighOrderVTKWriter (include/io/vtk_writer.hpp, src/io/vtk_writer.cpp).  The
files are written on the host; fields may come from device tensors."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .solver import _ptr, _sptr

VTU, PVTU = 0, 1                # VTKFormat
ASCII, BINARY, BASE64 = 0, 1, 2  # VTKEncoding


class VTKWriter:
    def __init__(self, basename, fmt=VTU, encoding=BINARY):
        self._l = _lib.load()
        h = C.c_void_p()
        check(self._l.drf_vtk_create(str(basename).encode(), fmt, encoding, C.byref(h)))
        self._w = h

    def __del__(self):
        if getattr(self, "_w", None):
            self._l.drf_vtk_destroy(self._w)
            self._w = None

    def set_polynomial_order(self, order):
        check(self._l.drf_vtk_set_polynomial_order(self._w, order))

    def set_mesh(self, mesh):
        """mesh: an OctreeAdapter (drifter_amd.mesh) or leaf bounds (E, 6)."""
        b = mesh.elements()[0] if hasattr(mesh, "elements") else mesh
        b = np.ascontiguousarray(b, dtype=np.float64)
        check(self._l.drf_vtk_set_mesh(self._w, b.shape[0], _lib.dptr(b)))

    def add_point_data(self, name, num_components=1):
        check(self._l.drf_vtk_add_field(self._w, name.encode(), num_components, 0))

    def add_cell_data(self, name, num_components=1):
        check(self._l.drf_vtk_add_field(self._w, name.encode(), num_components, 1))

    def _set(self, name, data, cells):
        if hasattr(data, "is_cuda") and data.is_cuda:
            if cells or data.dim() != 2:
                raise ValueError("device fields: point data of shape (E, ndof)")
            check(self._l.drf_vtk_set_point_field_dev(self._w, name.encode(), data.shape[0],
                                                      _ptr(data), data.stride(0), data.shape[1],
                                                      _sptr(None)))
            return
        if isinstance(data, (list, tuple)):  # vector<VecX> per element
            data = np.concatenate([np.asarray(d, dtype=np.float64).reshape(-1) for d in data]) \
                if len(data) else np.zeros(0)
        d = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
        check(self._l.drf_vtk_set_field(self._w, name.encode(), int(cells), d.size,
                                        _lib.dptr(d)))

    def set_point_data(self, name, data):
        self._set(name, data, False)

    def set_cell_data(self, name, data):
        self._set(name, data, True)

    def get_filename(self, time_idx):
        buf = C.create_string_buffer(4096)
        check(self._l.drf_vtk_filename(self._w, time_idx, buf, 4096))
        return buf.value.decode()

    def write(self, time_idx, time):
        check(self._l.drf_vtk_write(self._w, time_idx, time))

    def write_timestep(self, time):
        check(self._l.drf_vtk_write_timestep(self._w, time))

    def finalize(self):
        check(self._l.drf_vtk_finalize(self._w))


class OceanVTKWriter:
    """OceanVTKWriter (:487-535): eta, velocity (3), temperature, salinity
    point fields, binary (base64) encoding.  As in the reference, write()
    leaves the velocity array empty (its combined data is never stored,
    :511-527)."""

    def __init__(self, basename, mesh, polynomial_order):
        self.writer = VTKWriter(basename, VTU, BINARY)
        self.writer.set_polynomial_order(polynomial_order)
        self.writer.set_mesh(mesh)
        for name, nc in (("eta", 1), ("velocity", 3), ("temperature", 1), ("salinity", 1)):
            self.writer.add_point_data(name, nc)
        self._idx = 0

    def write(self, time, eta, u, v, w, temperature, salinity):
        self.writer.set_point_data("eta", eta)
        self.writer.set_point_data("temperature", temperature)
        self.writer.set_point_data("salinity", salinity)
        self.writer.write(self._idx, time)
        self._idx += 1

    def finalize(self):
        self.writer.finalize()


class HighOrderVTKWriter:
    """HighOrderVTKWriter (:541-600): binary encoding; the DG -> VTK node
    reordering of the reference is the identity."""

    def __init__(self, basename, polynomial_order):
        self.writer = VTKWriter(basename, VTU, BINARY)
        self.writer.set_polynomial_order(polynomial_order)

    def set_mesh(self, mesh, basis=None):
        self.writer.set_mesh(mesh)

    def add_scalar_field(self, name):
        self.writer.add_point_data(name, 1)

    def add_vector_field(self, name):
        self.writer.add_point_data(name, 3)

    def set_field(self, name, element_data):
        self.writer.set_point_data(name, element_data)

    def write_timestep(self, time):
        self.writer.write_timestep(time)

    def finalize(self):
        self.writer.finalize()
