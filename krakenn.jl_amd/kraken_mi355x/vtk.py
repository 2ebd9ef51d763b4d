"""VTK output of lattice fields: rectilinear-grid ``.vtr`` snapshots and ParaView ``.pvd``
time series, the on-disk formats of Kraken.jl's src/io/vtk_writer.jl, fed from the device
through ``Lattice.moments`` / ``kmi_moments``.

Grid convention (vtk_writer.jl:9-35): node coordinates ``0, dx, …, N·dx`` on each axis, so
an ``(Nx, Ny[, Nz])`` field is cell data, ``x`` fastest (the reference's column-major
layout).  Files are XML VTK 1.0 with raw appended binary data and 64-bit block headers,
which ParaView and the VTK readers open directly.  Snapshots are named
``snapshot_NNNNNNN.vtr`` (vtk_writer.jl:186-226).
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET
from typing import Mapping

import numpy as np

_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
              np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64",
              np.dtype(np.uint8): "UInt8"}


def _vtr_path(filename: str) -> str:
    return filename if filename.endswith(".vtr") else filename + ".vtr"


def _as_cell_array(name, data, dims):
    """(bytes, vtk type, components) of a scalar array or a 3-tuple of component arrays."""
    if isinstance(data, tuple):
        if len(data) != 3:
            raise ValueError(f"vector field {name!r} needs three components")
        comps = [np.asarray(c) for c in data]
        for c in comps:
            if c.shape != dims:
                raise ValueError(f"field {name!r}: component shape {c.shape} != grid {dims}")
        dt = np.result_type(*comps)
        flat = np.stack([c.astype(dt).ravel(order="F") for c in comps], axis=1)
        return flat.tobytes(order="C"), dt, 3
    a = np.asarray(data)
    if a.shape != dims:
        raise ValueError(f"field {name!r}: shape {a.shape} != grid {dims}")
    return a.ravel(order="F").tobytes(), a.dtype, 1


def write_vtk(filename: str, *args) -> str:
    """write_vtk(filename, Nx, Ny, dx, fields) or write_vtk(filename, Nx, Ny, Nz, dx, fields).

    ``fields`` maps names to arrays of the grid shape, or to ``(ux, uy, uz)`` tuples for a
    3-component field.  Returns the path written (``filename`` + ``.vtr``)."""
    if len(args) == 4:
        Nx, Ny, dx, fields = args
        dims = (int(Nx), int(Ny))
    elif len(args) == 5:
        Nx, Ny, Nz, dx, fields = args
        dims = (int(Nx), int(Ny), int(Nz))
    else:
        raise TypeError("write_vtk(filename, Nx, Ny[, Nz], dx, fields)")
    return _write_vtr(_vtr_path(filename), dims, float(dx), fields)


def _write_vtr(path, dims, dx, fields: Mapping):
    n3 = tuple(dims) + (0,) * (3 - len(dims))  # cell counts; 0 = flat axis
    coords = [np.arange(n + 1, dtype=np.float64) * dx if n > 0 else np.zeros(1) for n in n3]
    extent = " ".join(f"0 {n}" for n in n3)
    blocks, header = [], []
    offset = 0

    def add(payload: bytes):
        nonlocal offset
        blocks.append(np.uint64(len(payload)).tobytes() + payload)
        at = offset
        offset += 8 + len(payload)
        return at

    cell_lines = []
    for name in fields:  # insertion order of the mapping
        payload, dt, ncomp = _as_cell_array(name, fields[name], tuple(dims))
        vt = _VTK_TYPES.get(np.dtype(dt))
        if vt is None:
            raise ValueError(f"field {name!r}: element type {dt} has no VTK equivalent")
        at = add(payload)
        comp = f' NumberOfComponents="{ncomp}"' if ncomp > 1 else ""
        cell_lines.append(f'        <DataArray type="{vt}" Name="{name}"{comp} format="appended" offset="{at}"/>')
    coord_lines = []
    for axis, c in zip("xyz", coords):
        at = add(c.tobytes())
        coord_lines.append(f'        <DataArray type="Float64" Name="{axis}" format="appended" offset="{at}"/>')
    header.append('<?xml version="1.0"?>')
    header.append('<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">')
    header.append(f'  <RectilinearGrid WholeExtent="{extent}">')
    header.append(f'    <Piece Extent="{extent}">')
    header.append("      <PointData>\n      </PointData>")
    header.append("      <CellData>")
    header.extend(cell_lines)
    header.append("      </CellData>")
    header.append("      <Coordinates>")
    header.extend(coord_lines)
    header.append("      </Coordinates>")
    header.append("    </Piece>")
    header.append("  </RectilinearGrid>")
    header.append('  <AppendedData encoding="raw">')
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(("\n".join(header) + "\n   _").encode())
        for b in blocks:
            fh.write(b)
        fh.write(b"\n  </AppendedData>\n</VTKFile>\n")
    os.replace(tmp, path)
    return path


class PVDCollection:
    """ParaView ``.pvd`` time series (create_pvd / paraview_collection, vtk_writer.jl:80-82).
    Each ``add`` rewrites the collection file, so it is valid after every snapshot."""

    def __init__(self, filename: str):
        self.path = filename if filename.endswith(".pvd") else filename + ".pvd"
        self.entries: list[tuple[float, str]] = []
        self.save()

    def __setitem__(self, time: float, vtr_path: str):
        self.add(time, vtr_path)

    def add(self, time: float, vtr_path: str):
        rel = os.path.relpath(os.path.abspath(vtr_path), os.path.dirname(os.path.abspath(self.path)))
        self.entries.append((float(time), rel))
        self.save()

    def save(self):
        lines = ['<?xml version="1.0"?>',
                 '<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">',
                 "  <Collection>"]
        for t, f in self.entries:
            lines.append(f'    <DataSet timestep="{t!r}" part="0" file="{f}"/>')
        lines += ["  </Collection>", "</VTKFile>", ""]
        tmp = self.path + ".tmp"
        with open(tmp, "w") as fh:
            fh.write("\n".join(lines))
        os.replace(tmp, self.path)
        return self.path

    close = save


def create_pvd(filename: str) -> PVDCollection:
    return PVDCollection(filename)


def write_vtk_to_pvd(pvd: PVDCollection, filename: str, *args, time: float = 0.0) -> str:
    """write_vtk, then register the file in `pvd` at `time` (vtk_writer.jl:42-76)."""
    path = write_vtk(filename, *args)
    pvd.add(time, path)
    return path


def setup_output_dir(path: str) -> str:
    os.makedirs(path, exist_ok=True)
    return os.path.abspath(path)


def _snapshot_name(output_dir, step):
    return os.path.join(output_dir, f"snapshot_{int(step):07d}")


def write_snapshot_2d(output_dir, step, Nx, Ny, dx, fields, pvd=None, time=0.0) -> str:
    """write_snapshot_2d! (vtk_writer.jl:167-180)."""
    name = _snapshot_name(output_dir, step)
    if pvd is not None:
        return write_vtk_to_pvd(pvd, name, Nx, Ny, float(dx), fields, time=float(time))
    return write_vtk(name, Nx, Ny, float(dx), fields)


def write_snapshot_3d(output_dir, step, Nx, Ny, Nz, dx, fields, pvd=None, time=0.0) -> str:
    """write_snapshot_3d! (vtk_writer.jl:195-226): scalar arrays or (ux, uy, uz) tuples."""
    name = _snapshot_name(output_dir, step)
    if pvd is not None:
        return write_vtk_to_pvd(pvd, name, Nx, Ny, Nz, float(dx), fields, time=float(time))
    return write_vtk(name, Nx, Ny, Nz, float(dx), fields)


def lattice_fields(lat, temperature=None, velocity_vector=False):
    """Moments of a lattice as VTK fields: rho, ux, uy(, uz)(, Temp); with
    ``velocity_vector`` the 3D velocity is one (ux, uy, uz) vector field."""
    thermal = bool(lat.desc.thermal) if temperature is None else bool(temperature)
    m = lat.moments(temperature=thermal)
    out = {"rho": m["rho"]}
    if velocity_vector and "uz" in m:
        out["velocity"] = (m["ux"], m["uy"], m["uz"])
    else:
        out["ux"], out["uy"] = m["ux"], m["uy"]
        if "uz" in m:
            out["uz"] = m["uz"]
    if thermal:
        out["Temp"] = m["Temp"]
    return out


def write_lattice_snapshot(lat, output_dir, pvd=None, dx=1.0, time=None, **kw) -> str:
    """Snapshot of a lattice's moments at its current step (zero-padded step number)."""
    fields = lattice_fields(lat, **kw)
    step = lat.steps
    t = float(step) if time is None else float(time)
    if len(lat.shape) == 2:
        return write_snapshot_2d(output_dir, step, *lat.shape, dx, fields, pvd=pvd, time=t)
    return write_snapshot_3d(output_dir, step, *lat.shape, dx, fields, pvd=pvd, time=t)


def read_vtr(path: str) -> dict:
    """Read a file written by write_vtk: {"cells": dims, "coords": (x, y, z), name: array}."""
    with open(path, "rb") as fh:
        raw = fh.read()
    start = raw.index(b'<AppendedData encoding="raw">')
    us = raw.index(b"_", start) + 1
    end = raw.rindex(b"</AppendedData>")
    data = raw[us:end]
    xml = raw[:start] + b"</VTKFile>"
    root = ET.fromstring(xml)
    piece = root.find("RectilinearGrid/Piece")
    ext = [int(v) for v in piece.get("Extent").split()]
    n3 = (ext[1] - ext[0], ext[3] - ext[2], ext[5] - ext[4])
    dims = tuple(n for n in n3 if n > 0)
    inv = {v: k for k, v in _VTK_TYPES.items()}

    def block(el):
        off = int(el.get("offset"))
        nbytes = int(np.frombuffer(data[off:off + 8], dtype=np.uint64)[0])
        return np.frombuffer(data[off + 8:off + 8 + nbytes], dtype=inv[el.get("type")])

    out = {"cells": dims}
    out["coords"] = tuple(block(el) for el in piece.find("Coordinates"))
    for el in piece.find("CellData"):
        a = block(el)
        nc = int(el.get("NumberOfComponents", "1"))
        if nc == 1:
            out[el.get("Name")] = a.reshape(dims, order="F")
        else:
            a = a.reshape(-1, nc)
            out[el.get("Name")] = tuple(a[:, k].reshape(dims, order="F") for k in range(nc))
    return out


def read_pvd(path: str) -> list[tuple[float, str]]:
    root = ET.parse(path).getroot()
    base = os.path.dirname(os.path.abspath(path))
    return [(float(d.get("timestep")), os.path.join(base, d.get("file")))
            for d in root.find("Collection")]
