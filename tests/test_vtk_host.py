"""VTK output (VTKWriter, src/io/vtk_writer.cpp:47-475), host side.

The reference's own tests are restated first (tests/unit/test_io.cpp:40-177);
then the file is parsed and checked against closed forms: equispaced node
positions (and the order-1 placement, where the reference coordinate is the
node index), connectivity, offsets and cell types, ASCII values with 15
significant digits, base64 arrays (8-byte byte count, then the doubles),
fields in name order, the .pvd collection and the wrapper writers."""
import base64
import os
import re
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from drifter_amd import _lib
from drifter_amd import vtk_writer as V
from drifter_amd.mesh import OctreeAdapter


def _mesh(lo, hi, n):
    m = OctreeAdapter(lo, hi, lo, hi, lo, hi)
    m.build_uniform(n, n, n)
    return m


def _read(path):
    with open(path) as f:
        return f.read()


# ---- the reference's tests (test_io.cpp) ----

def test_vtu_format(tmp_path):  # :40-76
    base = str(tmp_path / "test_vtu")
    w = V.VTKWriter(base, V.VTU, V.ASCII)
    w.set_polynomial_order(1)
    w.set_mesh(_mesh(0.0, 1.0, 1))
    w.add_point_data("temperature", 1)
    w.set_point_data("temperature", np.arange(8.0))
    w.add_cell_data("cell_id", 1)
    w.set_cell_data("cell_id", [42.0])
    w.write(0, 0.0)
    fn = base + "_000000.vtu"
    assert os.path.exists(fn)
    c = _read(fn)
    assert "<?xml version" in c and "<VTKFile" in c and "UnstructuredGrid" in c


def test_vtu_point_count(tmp_path):  # :78-97
    base = str(tmp_path / "test_vtu_points")
    w = V.VTKWriter(base, V.VTU, V.ASCII)
    w.set_polynomial_order(1)
    w.set_mesh(_mesh(0.0, 2.0, 2))
    w.write(0, 0.0)
    assert 'NumberOfPoints="64"' in _read(base + "_000000.vtu")


def test_vtu_cell_data(tmp_path):  # :99-126
    base = str(tmp_path / "test_vtu_cells")
    w = V.VTKWriter(base, V.VTU, V.ASCII)
    w.set_polynomial_order(1)
    w.set_mesh(_mesh(0.0, 1.0, 1))
    w.add_cell_data("cell_id", 1)
    w.set_cell_data("cell_id", [42.0])
    w.write(0, 0.0)
    c = _read(base + "_000000.vtu")
    assert "<CellData>" in c and 'Name="cell_id"' in c


def test_construction_order_fields_filename(tmp_path):  # :131-177
    base = str(tmp_path / "output")
    V.VTKWriter(base + "_vtu", V.VTU)
    V.VTKWriter(base + "_ascii", V.VTU, V.ASCII)
    V.VTKWriter(base + "_pvtu", V.PVTU)
    w = V.VTKWriter(base)
    for p in (1, 2, 4):
        w.set_polynomial_order(p)
    w.add_point_data("temperature", 1)
    w.add_point_data("velocity", 3)
    w.add_cell_data("element_id", 1)
    w.add_cell_data("stress", 6)
    assert ".vtu" in w.get_filename(0) and "_000000" in w.get_filename(0)
    assert "_000123" in w.get_filename(123)
    wp = V.VTKWriter(base, V.PVTU)
    assert ".vtu" in wp.get_filename(0)


# ---- closed forms of the file contents ----

def _arrays(path):
    root = ET.parse(path).getroot()
    return root, {a.get("Name"): a for a in root.iter("DataArray")}


def _decode(arr):
    raw = base64.b64decode(arr.text)
    nbytes = struct.unpack("<Q", raw[:8])[0]
    assert nbytes == len(raw) - 8
    return np.frombuffer(raw[8:], dtype="<f8")


@pytest.mark.parametrize("order", [2, 3])
def test_high_order_geometry_and_ascii_values(tmp_path, order):
    m = OctreeAdapter(0.0, 2.0, 0.0, 1.0, -1.0, 0.0)
    m.build_uniform(2, 1, 1)
    b, _, _ = m.elements()
    n = order + 1
    rng = np.random.default_rng(order)
    vals = rng.standard_normal(2 * n ** 3)
    base = str(tmp_path / "ho")
    w = V.VTKWriter(base, V.VTU, V.ASCII)
    w.set_polynomial_order(order)
    w.set_mesh(m)
    w.add_point_data("f", 1)
    w.set_point_data("f", vals)
    w.write(3, 1.0 / 3.0)
    root, arr = _arrays(base + "_000003.vtu")
    assert arr["TIME"].text.strip() == "0.333333333333333"  # 15 significant digits
    pts = np.array(arr["Points"].text.split(), dtype=float).reshape(-1, 3)
    r = -1.0 + 2.0 * np.arange(n) / order  # equispaced, not LGL
    want = []
    for e in range(2):
        for k in range(n):
            for j in range(n):
                for i in range(n):
                    want.append([b[e, 2 * a] + 0.5 * (r[x] + 1.0) * (b[e, 2 * a + 1] - b[e, 2 * a])
                                 for a, x in enumerate((i, j, k))])
    np.testing.assert_allclose(pts, np.array(want), rtol=1e-14, atol=1e-14)
    conn = np.array(arr["connectivity"].text.split(), dtype=int)
    np.testing.assert_array_equal(conn, np.arange(2 * n ** 3))
    np.testing.assert_array_equal(np.array(arr["offsets"].text.split(), dtype=int),
                                  [n ** 3, 2 * n ** 3])
    assert arr["types"].text.split() == ["72", "72"]
    got = np.array(arr["f"].text.split(), dtype=float)
    np.testing.assert_allclose(got, vals, rtol=1e-14)
    assert all(len(re.sub(r"[-.e+]", "", t).lstrip("0")) <= 15 for t in arr["f"].text.split())


def test_order_one_placement_and_corner_order(tmp_path):
    base = str(tmp_path / "lin")
    w = V.VTKWriter(base, V.VTU, V.ASCII)
    w.set_polynomial_order(1)
    w.set_mesh(_mesh(0.0, 4.0, 1))
    w.write(0, 0.0)
    _, arr = _arrays(base + "_000000.vtu")
    pts = np.array(arr["Points"].text.split(), dtype=float).reshape(-1, 3)
    # reference coordinate = node index 0 / 1: the centre and the max face
    assert set(np.unique(pts)) == {2.0, 4.0}
    assert arr["connectivity"].text.split() == ["0", "1", "3", "2", "4", "5", "7", "6"]
    assert arr["types"].text.split() == ["12"]


@pytest.mark.parametrize("enc", [V.BINARY, V.BASE64])
def test_base64_arrays_and_field_order(tmp_path, enc):
    base = str(tmp_path / "b64")
    w = V.VTKWriter(base, V.VTU, enc)
    w.set_polynomial_order(2)
    w.set_mesh(_mesh(0.0, 1.0, 1))
    vals = {"zeta": np.linspace(0, 1, 27), "alpha": np.linspace(2, 3, 27),
            "mid": np.linspace(-1, 0, 27)}
    for k in ("zeta", "alpha", "mid"):  # registration order differs from name order
        w.add_point_data(k, 1)
        w.set_point_data(k, vals[k])
    w.add_point_data("empty", 3)  # registered but never set: a zero-length array
    w.write(0, 2.5)
    txt = _read(base + "_000000.vtu")
    assert ('compressor="vtkZLibDataCompressor"' in txt) == (enc == V.BASE64)
    root, arr = _arrays(base + "_000000.vtu")
    names = [a.get("Name") for a in root.find(".//PointData")]
    assert names == sorted(names) == ["alpha", "empty", "mid", "zeta"]
    for k, v in vals.items():
        np.testing.assert_array_equal(_decode(arr[k]), v)
    assert _decode(arr["empty"]).size == 0
    assert _decode(arr["Points"]).size == 27 * 3


def test_timesteps_pvd_and_errors(tmp_path):
    base = str(tmp_path / "run" / "ocean")  # the directory is created
    w = V.VTKWriter(base)
    w.set_mesh(_mesh(0.0, 1.0, 1))
    w.write_timestep(0.0)
    w.write_timestep(0.1)
    w.write(7, 12.25)
    w.finalize()
    pvd = ET.parse(base + ".pvd").getroot()
    ds = pvd.findall(".//DataSet")
    assert [d.get("timestep") for d in ds] == ["0", "0.1", "12.25"]
    assert [os.path.basename(d.get("file")) for d in ds] == [
        "ocean_000000.vtu", "ocean_000001.vtu", "ocean_000007.vtu"]
    with pytest.raises(_lib.DrfError, match="Unknown point field: nope"):
        w.set_point_data("nope", [1.0])
    with pytest.raises(_lib.DrfError, match="Unknown cell field: nope"):
        w.set_cell_data("nope", [1.0])


def test_ocean_and_high_order_writers(tmp_path):
    m = _mesh(0.0, 1.0, 1)
    n3 = 27
    f = [np.full(n3, 1.5)]
    ow = V.OceanVTKWriter(str(tmp_path / "oc"), m, 2)
    ow.write(0.5, f, f, f, f, [np.full(n3, 10.0)], [np.full(n3, 35.0)])
    ow.finalize()
    root, arr = _arrays(str(tmp_path / "oc_000000.vtu"))
    assert [a.get("Name") for a in root.find(".//PointData")] == [
        "eta", "salinity", "temperature", "velocity"]
    assert _decode(arr["velocity"]).size == 0  # never stored by the reference's write()
    np.testing.assert_array_equal(_decode(arr["temperature"]), np.full(n3, 10.0))
    hw = V.HighOrderVTKWriter(str(tmp_path / "ho"), 2)
    hw.set_mesh(m)
    hw.add_scalar_field("T")
    hw.set_field("T", [np.arange(27.0)])
    hw.write_timestep(0.0)
    _, arr = _arrays(str(tmp_path / "ho_000000.vtu"))
    np.testing.assert_array_equal(_decode(arr["T"]), np.arange(27.0))  # identity ordering
