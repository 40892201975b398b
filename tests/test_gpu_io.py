"""Solution output: pp_io, the counterpart of PencilDcmp::IO (src/pencilDcmp.cpp:1827-1832).

The written field must be the solution bit for bit (the reference stores Q as native double,
src/phdf5.cpp:1780-1783), in z, y, x order, and the XDMF index must describe a cell-centred
attribute "Q" on the (nx+1)(ny+1)(nz+1) corner mesh of the box (src/phdf5.cpp:1809-2070).
Default format: the reference's HDF5 file (datasets X, Y, Z, Q); PP_IO_RAW: a raw binary field.
"""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import pittpack_amd as pa
import _h5

pytestmark = pytest.mark.gpu


def _solve(nx, ny, nz, bc, X, seed):
    rng = np.random.default_rng(seed)
    M = pa.PoissonHIP(nx, ny, nz)
    M.assignBoundary(bc)
    M.setBox(X)
    M.setCoords(2)
    M.assignRhs(rng.standard_normal((nz, ny, nx)))
    M.pittPack()
    return M


@pytest.mark.parametrize("shape", [(32, 16, 24), (64, 64, 64)])
def test_io_writes_solution_and_index(tmp_path, shape):
    nx, ny, nz = shape
    X = [0.0, 2.0, -1.0, 0.5, 0.25, 3.25]
    print(f"io {shape}", flush=True)
    M = _solve(nx, ny, nz, "DDNDDD", X, 5)
    u = M.getSolution()
    M.IO(7, str(tmp_path), format=pa.IO_RAW)
    xmf, binf = tmp_path / "Pxdmf3d7.xmf", tmp_path / "Pxdmf3d7.bin"
    assert os.path.getsize(binf) == nx * ny * nz * 8
    q = np.fromfile(binf, dtype="<f8").reshape(nz, ny, nx)
    assert np.array_equal(q, u)

    root = ET.parse(xmf).getroot()
    assert root.tag == "Xdmf" and root.get("Version") == "2.0"
    grid = root.find("Domain/Grid")
    topo = grid.find("Topology")
    assert topo.get("TopologyType") == "3DCoRectMesh"
    assert [int(v) for v in topo.get("Dimensions").split()] == [nz + 1, ny + 1, nx + 1]
    geo = grid.find("Geometry")
    assert geo.get("GeometryType") == "ORIGIN_DXDYDZ"
    items = {d.get("Name"): [float(v) for v in d.text.split()] for d in geo.findall("DataItem")}
    assert items["Origin"] == [X[4], X[2], X[0]]
    np.testing.assert_allclose(items["Spacing"], [(X[5] - X[4]) / nz, (X[3] - X[2]) / ny, (X[1] - X[0]) / nx],
                               rtol=0, atol=1e-15)
    att = grid.find("Attribute")
    assert att.get("Name") == "Q" and att.get("Center") == "Cell"
    d = att.find("DataItem")
    assert [int(v) for v in d.get("Dimensions").split()] == [nz, ny, nx]
    assert d.get("Format") == "Binary" and d.get("Endian") == "Little" and d.get("Precision") == "8"
    assert d.text.strip() == "Pxdmf3d7.bin"


@pytest.mark.skipif(_h5.lib() is None, reason="no HDF5 library in this image")
def test_io_hdf5_default(tmp_path):
    nx, ny, nz = 24, 20, 18
    X = [-0.3, 0.7, 0.0, 2.0, 0.5, 1.0]
    print("io hdf5 (24, 20, 18)", flush=True)
    M = _solve(nx, ny, nz, "NNNNDD", X, 9)
    u = M.getSolution()
    M.IO(2, str(tmp_path))
    q, chunk, _ = _h5.read(str(tmp_path / "Pxdmf3d2.h5"), "/Q")
    assert q.shape == (nz, ny, nx, 1) and np.array_equal(q[..., 0], u)
    z, _, _ = _h5.read(str(tmp_path / "Pxdmf3d2.h5"), "/Z")
    assert z.shape == (nz + 1, ny + 1, nx + 1, 1)
    assert z[0, 0, 0, 0] == X[4] and abs(z[-1, 0, 0, 0] - X[5]) < 1e-15
    root = ET.parse(tmp_path / "Pxdmf3d2.xmf").getroot()
    assert root.find("Domain/Grid/Grid/Topology").get("TopologyType") == "3DSMesh"


def test_io_errors(tmp_path):
    M = _solve(16, 16, 16, "NNNNDD", [0, 1, 0, 1, 0, 1], 1)
    with pytest.raises(pa.PittPackError) as e:
        M.IO(0, str(tmp_path / "missing"))
    assert e.value.code == pa.IO_FAIL
    with pytest.raises(pa.PittPackError) as e:
        M.IO(-1, str(tmp_path))
    assert e.value.code == pa.INPUT_ARGS
    M.IO(0, str(tmp_path))  # the handle stays usable
    M.assignRhs(np.ones((16, 16, 16)))
    M.pittPack()
    M.IO(1, str(tmp_path), format=pa.IO_RAW)
    assert (tmp_path / "Pxdmf3d1.xmf").exists() and (tmp_path / "Pxdmf3d1.bin").exists()
    with pytest.raises(pa.PittPackError) as e:
        M.set_option(pa.OPT_IO_FORMAT, 5)
    assert e.value.code == pa.INPUT_ARGS
