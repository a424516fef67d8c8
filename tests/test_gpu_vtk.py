"""VTK output from device memory: a point field set from a device tensor (a
strided view of a packed state) writes the same file, byte for byte, as the
same values set from host memory (src/io/vtk_writer.cpp:173-192)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("enc", [0, 1])
def test_device_field_matches_host(tmp_path, enc):
    import torch

    from drifter_amd import vtk_writer as V
    from drifter_amd.mesh import OctreeAdapter

    m = OctreeAdapter(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    m.build_uniform(3, 2, 2)
    E, nd = m.num_elements(), 64
    full = np.random.default_rng(enc).standard_normal((E, 5, nd))
    print(f"[vtk] device field, encoding {enc}", flush=True)
    files = []
    for src in ("host", "device"):
        base = str(tmp_path / src)
        w = V.VTKWriter(base, V.VTU, enc)
        w.set_polynomial_order(3)
        w.set_mesh(m)
        w.add_point_data("eta", 1)
        if src == "host":
            w.set_point_data("eta", full[:, 2].copy())
        else:
            dev = torch.from_numpy(full).cuda()
            w.set_point_data("eta", dev[:, 2, :])  # element stride 5 * 64
        w.write(0, 0.25)
        files.append(open(base + "_000000.vtu", "rb").read())
    assert files[0] == files[1]
