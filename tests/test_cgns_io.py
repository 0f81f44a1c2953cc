"""CGNS/HDF5 writer and reader of the host driver (lbm-c_amd/host/cgns_io.c),
checked with h5dump, an HDF5 reader independent of this code: the node
tree, attributes, data shapes (CGNS dimensions reversed) and values of the
output LBM-C's CGNSOutputHandler writes (cgns_output_handler.cu:62-205,
264-293), and field input by name from zone 1 / solution 1
(cgns_input_handler.cu:57-103).  Against libcgns itself: parity unpinned
(no CGNS library in the image)."""
import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "lbm-c_amd", "libcgnsio.so")
H5DUMP = "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not (os.path.exists(LIB) and os.path.exists(H5DUMP)),
                                reason="libcgnsio.so or h5dump not available")

D = ct.POINTER(ct.c_double)


def lib():
    L = ct.CDLL(LIB)
    L.cgio_last_error.restype = ct.c_char_p
    L.cgio_quiet()
    return L


def append(L, fn, it, fields):
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in fields.values()]
    data = (D * len(arrs))(*[a.ctypes.data_as(D) for a in arrs])
    labels = (ct.c_char_p * len(arrs))(*[k.encode() for k in fields])
    assert L.cgio_append_solution(fn.encode(), it, len(arrs), data, labels) == 0, L.cgio_last_error()


def dump_data(fn, path):
    out = subprocess.run([H5DUMP, "-w", "0", "-m", "%.17g", "-d", path, fn], capture_output=True, text=True,
                         check=True).stdout
    dims = re.search(r"DATASPACE\s+SIMPLE \{ \( ([0-9, ]+) \)", out).group(1)
    shape = tuple(int(v) for v in dims.split(","))
    body = out[out.index("DATA {") + 6:]
    body = re.sub(r"\(\d+(?:,\d+)*\):", " ", body)
    vals = [float(v) for v in re.findall(r"-?\d+(?:\.\d+)?(?:e[-+]?\d+)?", body.split("}")[0])]
    return shape, np.array(vals)


def attr(fn, path, name):
    out = subprocess.run([H5DUMP, "-a", f"{path}/{name}", fn], capture_output=True, text=True, check=True).stdout
    return re.search(r'\(0\): "([^"]*)"', out).group(1)


def test_output_tree_2d(tmp_path):
    L = lib()
    fn = str(tmp_path / "out.cgns")
    assert L.cgio_write_grid(fn.encode(), 4, 3, 1) == 0
    u = np.arange(12.0)
    rho = 1.0 + 0.01 * np.arange(12.0)
    append(L, fn, 100, {"VelocityX": u, "Density": rho})
    append(L, fn, 200, {"VelocityX": 2 * u, "Density": rho})
    z = "/Base/LBM-C Output Data"
    assert attr(fn, "/", "name") == "HDF5 MotherNode"
    assert attr(fn, "/", "label") == "Root Node of HDF5 File"
    assert attr(fn, "/Base", "label") == "CGNSBase_t" and attr(fn, "/Base", "type") == "I4"
    assert tuple(dump_data(fn, "/Base/ data")[1]) == (2, 2)
    assert attr(fn, z, "label") == "Zone_t"
    shape, isize = dump_data(fn, f"{z}/ data")
    assert shape == (3, 2) and list(isize) == [5, 4, 4, 3, 0, 0]
    assert attr(fn, f"{z}/ZoneType", "label") == "ZoneType_t"
    shape, cx = dump_data(fn, f"{z}/GridCoordinates/CoordinateX/ data")
    assert shape == (4, 5) and np.array_equal(cx.reshape(4, 5), np.tile(np.arange(5.0), (4, 1)))
    shape, cy = dump_data(fn, f"{z}/GridCoordinates/CoordinateY/ data")
    assert np.array_equal(cy.reshape(4, 5), np.tile(np.arange(4.0)[:, None], (1, 5)))
    s1 = f"{z}/Solution @ Time = 100"
    assert attr(fn, s1, "label") == "FlowSolution_t"
    assert attr(fn, f"{s1}/VelocityX", "type") == "R8"
    shape, v = dump_data(fn, f"{s1}/VelocityX/ data")
    assert shape == (3, 4) and np.array_equal(v, u)
    shape, v = dump_data(fn, f"{z}/Solution @ Time = 200/VelocityX/ data")
    assert np.array_equal(v, 2 * u)
    assert attr(fn, "/Base/TimeIterValues", "label") == "BaseIterativeData_t"
    assert list(dump_data(fn, "/Base/TimeIterValues/IterationValues/ data")[1]) == [100, 200]
    shape, _ = dump_data(fn, f"{z}/ZoneIterativeData_t/FlowSolutionPointers/ data")
    assert shape == (2, 32)
    assert attr(fn, "/Base/SimulationType", "label") == "SimulationType_t"


def test_output_tree_3d(tmp_path):
    L = lib()
    fn = str(tmp_path / "out3.cgns")
    assert L.cgio_write_grid(fn.encode(), 3, 2, 2) == 0
    append(L, fn, 7, {"VelocityZ": np.arange(12.0)})
    z = "/Base/LBM-C Output Data"
    assert tuple(dump_data(fn, "/Base/ data")[1]) == (3, 3)
    assert list(dump_data(fn, f"{z}/ data")[1]) == [4, 3, 3, 3, 2, 2, 0, 0, 0]
    shape, cz = dump_data(fn, f"{z}/GridCoordinates/CoordinateZ/ data")
    assert shape == (3, 3, 4) and np.array_equal(cz.reshape(3, 3, 4)[:, 0, 0], [0, 1, 2])
    shape, v = dump_data(fn, f"{z}/Solution @ Time = 7/VelocityZ/ data")
    assert shape == (2, 2, 3) and np.array_equal(v, np.arange(12.0))


def test_input_fields_by_name(tmp_path):
    L = lib()
    fn = str(tmp_path / "in.cgns")
    n = 6 * 5
    geo = (np.arange(n) % 7 == 0).astype(np.float64)
    micro = (np.arange(n) % 3).astype(np.float64)
    assert L.cgio_write_grid(fn.encode(), 6, 5, 1) == 0
    append(L, fn, 0, {"Porosity": geo, "MicroBC": micro})
    append(L, fn, 1, {"Porosity": 1 - geo})  # solution 2 is not read
    out = np.zeros(n)
    assert L.cgio_read_field(fn.encode(), b"Porosity", out.ctypes.data_as(D), ct.c_longlong(n)) == 0
    assert np.array_equal(out, geo)
    iv = np.zeros(n, dtype=np.int32)
    assert L.cgio_read_field_int(fn.encode(), b"MicroBC", iv.ctypes.data_as(ct.POINTER(ct.c_int)),
                                 ct.c_longlong(n)) == 0
    assert np.array_equal(iv, micro.astype(np.int32))
    assert L.cgio_read_field(fn.encode(), b"ForceX", out.ctypes.data_as(D), ct.c_longlong(n)) == 1
    assert L.cgio_read_field(fn.encode(), b"Porosity", out.ctypes.data_as(D), ct.c_longlong(n + 1)) == -1
    assert b"values" in L.cgio_last_error()
