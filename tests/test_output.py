"""Solution files on the CPU: pp_write_solution, the writer behind pp_io (PencilDcmp::IO,
src/pencilDcmp.cpp:1827-1832 -> Phdf5::writeMultiBlockCellCenter, src/phdf5.cpp:1299-1808).

HDF5: datasets X, Y, Z (node coordinates, (nz+1) x (ny+1) x (nx+1) x 1 -- getXcoord /
getYcoord / getZcoord with aligndir 0, src/phdf5.cpp:2238-2420) and Q (nz x ny x nx x 1, the
field in z, y, x order, getQ, src/phdf5.cpp:2427-2455), native double, one chunk per block;
the XDMF index points its 3DSMesh geometry and cell attribute at them
(src/phdf5.cpp:1809-2070)."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import pittpack_amd as pa
import _h5

pytestmark = pytest.mark.skipif(_h5.lib() is None, reason="no HDF5 library in this image")

H5T_FLOAT = 1


def test_hdf5_loads():
    assert pa.hdf5_available()


@pytest.mark.parametrize("shape,X", [((5, 4, 3), [0.0, 1.0, 0.0, 1.0, 0.0, 1.0]),
                                     ((7, 6, 9), [-0.3, 0.7, 2.0, 5.0, 0.25, 0.5])])
def test_hdf5_layout(tmp_path, shape, X):
    nz, ny, nx = shape
    u = np.random.default_rng(nx).standard_normal(shape)
    pa.write_solution(str(tmp_path), 3, u, X)
    h5 = str(tmp_path / "Pxdmf3d3.h5")
    q, qchunk, qt = _h5.read(h5, "/Q")
    assert q.shape == (nz, ny, nx, 1) and qchunk == (nz, ny, nx, 1) and qt == (H5T_FLOAT, 8)
    assert np.array_equal(q[..., 0], u)
    # node coordinates a + (b - a) / n * index, x fastest
    h = [(X[1] - X[0]) / nx, (X[3] - X[2]) / ny, (X[5] - X[4]) / nz]
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    for name, idx, a, hh in (("/X", i, X[0], h[0]), ("/Y", j, X[2], h[1]), ("/Z", k, X[4], h[2])):
        c, chunk, t = _h5.read(h5, name)
        assert c.shape == (nz + 1, ny + 1, nx + 1, 1) and chunk == c.shape and t == (H5T_FLOAT, 8)
        assert np.array_equal(c[..., 0], a + hh * idx.astype(np.float64)), name
    # index: a spatial collection with one 3DSMesh block; geometry and Q are hyperslabs of block 0
    root = ET.parse(tmp_path / "Pxdmf3d3.xmf").getroot()
    assert root.tag == "Xdmf" and root.get("Version") == "2.0"
    coll = root.find("Domain/Grid")
    assert coll.get("GridType") == "Collection" and coll.get("CollectionType") == "Spatial"
    grid = coll.find("Grid")
    topo = grid.find("Topology")
    assert topo.get("TopologyType") == "3DSMesh"
    assert [int(v) for v in topo.get("NumberOfElements").split()] == [nz + 1, ny + 1, nx + 1]
    geo = grid.find("Geometry")
    assert geo.get("GeometryType") == "X_Y_Z"
    slabs = geo.findall("DataItem")
    assert len(slabs) == 3
    for s, name in zip(slabs, "XYZ"):
        sel, data = s.findall("DataItem")
        assert [int(v) for v in sel.text.split()] == [0, 0, 0, 0, 1, 1, 1, 1, nz + 1, ny + 1, nx + 1, 1]
        assert data.get("Format") == "HDF" and data.get("Precision") == "8"
        assert data.text.strip() == f"Pxdmf3d3.h5:/{name}"
    att = grid.find("Attribute")
    assert att.get("Name") == "Q" and att.get("Center") == "Cell"
    sel, data = att.find("DataItem").findall("DataItem")
    assert [int(v) for v in sel.text.split()] == [0, 0, 0, 0, 1, 1, 1, 1, nz, ny, nx, 1]
    assert [int(v) for v in data.get("Dimensions").split()] == [nz, ny, nx, 1]
    assert data.text.strip() == "Pxdmf3d3.h5:/Q"


def test_raw_layout(tmp_path):
    u = np.random.default_rng(1).standard_normal((4, 5, 6))
    X = [0.0, 2.0, -1.0, 0.5, 0.25, 3.25]
    pa.write_solution(str(tmp_path), 0, u, X, format=pa.IO_RAW)
    q = np.fromfile(tmp_path / "Pxdmf3d0.bin", dtype="<f8").reshape(u.shape)
    assert np.array_equal(q, u)
    topo = ET.parse(tmp_path / "Pxdmf3d0.xmf").getroot().find("Domain/Grid/Topology")
    assert topo.get("TopologyType") == "3DCoRectMesh"


def test_write_errors(tmp_path):
    u = np.zeros((4, 4, 4))
    with pytest.raises(pa.PittPackError) as e:
        pa.write_solution(str(tmp_path / "missing"), 0, u, [0, 1, 0, 1, 0, 1])
    assert e.value.code == pa.IO_FAIL
    with pytest.raises(pa.PittPackError) as e:
        pa.write_solution(str(tmp_path), 0, u, [0, 1, 0, 1, 1, 0])
    assert e.value.code == pa.INPUT_ARGS
    with pytest.raises(pa.PittPackError) as e:
        pa.write_solution(str(tmp_path), 0, u, [0, 1, 0, 1, 0, 1], format=7)
    assert e.value.code == pa.INPUT_ARGS
    assert not os.listdir(tmp_path)
