"""VTK rectilinear-grid snapshots and ParaView collections (no GPU): the layout of
src/io/vtk_writer.jl — node coordinates 0, dx, …, N·dx, fields as cell data, x fastest,
`snapshot_NNNNNNN` names, `.pvd` time series — written and read back."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from kraken_mi355x import vtk


def test_2d_round_trip(tmp_path):
    r = np.random.default_rng(1)
    rho = np.asfortranarray(r.uniform(0.9, 1.1, (7, 5)))
    ux = np.asfortranarray(r.standard_normal((7, 5)).astype(np.float32))
    p = vtk.write_vtk(str(tmp_path / "a"), 7, 5, 0.5, {"rho": rho, "ux": ux})
    assert p.endswith("a.vtr") and os.path.isfile(p)
    back = vtk.read_vtr(p)
    assert back["cells"] == (7, 5)
    assert np.array_equal(back["rho"], rho) and back["ux"].dtype == np.float32
    assert np.array_equal(back["ux"], ux)
    x, y, z = back["coords"]
    assert np.array_equal(x, 0.5 * np.arange(8)) and np.array_equal(y, 0.5 * np.arange(6))
    assert np.array_equal(z, [0.0])
    head = open(p, "rb").read().split(b"<AppendedData")[0] + b"</VTKFile>"
    root = ET.fromstring(head)
    assert root.get("type") == "RectilinearGrid" and root.get("header_type") == "UInt64"
    assert root.find("RectilinearGrid").get("WholeExtent") == "0 7 0 5 0 0"


def test_3d_vector_field_and_snapshots(tmp_path):
    r = np.random.default_rng(2)
    shp = (4, 3, 5)
    u = tuple(np.asfortranarray(r.standard_normal(shp)) for _ in range(3))
    T = np.asfortranarray(r.standard_normal(shp))
    pvd = vtk.create_pvd(str(tmp_path / "series"))
    paths = [vtk.write_snapshot_3d(str(tmp_path), s, *shp, 1.0, {"velocity": u, "Temp": T}, pvd=pvd,
                                   time=0.1 * s) for s in (0, 50, 100)]
    assert [os.path.basename(p) for p in paths] == ["snapshot_0000000.vtr", "snapshot_0000050.vtr",
                                                    "snapshot_0000100.vtr"]
    back = vtk.read_vtr(paths[-1])
    assert back["cells"] == shp
    for a, b in zip(back["velocity"], u):
        assert np.array_equal(a, b)
    assert np.array_equal(back["Temp"], T)
    entries = vtk.read_pvd(pvd.path)
    assert [t for t, _ in entries] == [0.0, 5.0, 10.0]
    assert [f for _, f in entries] == [os.path.abspath(p) for p in paths]
    p2 = vtk.write_snapshot_2d(str(tmp_path), 7, 3, 2, 0.25, {"rho": np.ones((3, 2))})
    assert os.path.basename(p2) == "snapshot_0000007.vtr"


def test_shape_errors(tmp_path):
    with pytest.raises(ValueError, match="shape"):
        vtk.write_vtk(str(tmp_path / "b"), 4, 4, 1.0, {"rho": np.ones((4, 5))})
    with pytest.raises(ValueError, match="three components"):
        vtk.write_vtk(str(tmp_path / "c"), 2, 2, 2, 1.0, {"u": (np.ones((2, 2, 2)),) * 2})
    with pytest.raises(TypeError):
        vtk.write_vtk(str(tmp_path / "d"), 4, {"rho": np.ones(4)})


def _decode_vtr_from_spec(path):
    """An independent decoder of the file, from the VTK XML format's definition of appended
    raw data (not the library's reader): every DataArray's `offset` counts bytes from the
    first byte after the `_` that opens the AppendedData section; each block is a UInt64
    byte count (header_type) followed by that many bytes, little-endian."""
    raw = open(path, "rb").read()
    head, _, tail = raw.partition(b"<AppendedData")
    blob = tail[tail.index(b"_") + 1:]
    root = ET.fromstring(head + b"</VTKFile>")
    assert root.get("byte_order") == "LittleEndian" and root.get("header_type") == "UInt64"
    types = {"Float64": "<f8", "Float32": "<f4", "Int32": "<i4", "UInt8": "u1"}
    out = {}
    for da in root.iter("DataArray"):
        assert da.get("format") == "appended"
        at = int(da.get("offset"))
        n = int.from_bytes(blob[at:at + 8], "little")
        out[da.get("Name")] = (np.frombuffer(blob[at + 8:at + 8 + n], dtype=types[da.get("type")]),
                               int(da.get("NumberOfComponents", "1")))
    return root, out


def test_file_decoded_from_the_format_definition(tmp_path):
    """Cell data in x-fastest order, vectors interleaved per cell, node coordinates 0 … N·dx
    (write_snapshot_3d!, vtk_writer.jl:182-208, through WriteVTK's rectilinear grid), checked
    with a decoder written from the format's definition.  The reference's own test
    (test/test_vtk_3d.jl:3-38) checks only that the files exist, and the reference holds no
    written file to compare bytes with."""
    r = np.random.default_rng(9)
    shp = (5, 3, 4)
    ux, uy, uz = (np.asfortranarray(r.standard_normal(shp)) for _ in range(3))
    rho = np.asfortranarray(r.uniform(0.9, 1.1, shp).astype(np.float32))
    p = vtk.write_snapshot_3d(str(tmp_path), 7, *shp, 0.25, {"rho": rho, "velocity": (ux, uy, uz)})
    root, arrays = _decode_vtr_from_spec(p)
    grid = root.find("RectilinearGrid")
    assert grid.get("WholeExtent") == "0 5 0 3 0 4" and grid.find("Piece").get("Extent") == "0 5 0 3 0 4"
    cell_names = [da.get("Name") for da in grid.find("Piece").find("CellData")]
    assert cell_names == ["rho", "velocity"]
    v, nc = arrays["rho"]
    assert nc == 1 and v.dtype == np.float32 and np.array_equal(v, rho.ravel(order="F"))
    v, nc = arrays["velocity"]
    assert nc == 3 and np.array_equal(v.reshape(-1, 3), np.stack([a.ravel(order="F") for a in (ux, uy, uz)], axis=1))
    for axis, n in zip("xyz", shp):
        c, nc = arrays[axis]
        assert nc == 1 and np.array_equal(c, 0.25 * np.arange(n + 1))
