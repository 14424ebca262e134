"""Reading and writing the simulation's output files.

Writers go through the C-ABI (``pf_write_vtk``; a simulation's own fields
through ``Simulation.write``).  The readers parse what those write: raw
little-endian FP32 ``.bin`` dumps of nx*ny*nz values (x fastest), and VTK
XML ImageData (``.vti``) or RectilinearGrid (``.vtr``) files with raw
appended point arrays, as ParaView reads them.
"""
from __future__ import annotations

import ctypes
import xml.etree.ElementTree as ET

import numpy as np

from . import _abi

FORMATS = {"vti": _abi.OUT_VTI, "vtr": _abi.OUT_VTR}


def write_vtk(path: str, arrays: dict, fmt: str = "vtr") -> None:
    """Write float32 arrays of shape (nz, ny, nx) as one .vti or .vtr file."""
    names = list(arrays)
    data = [np.ascontiguousarray(arrays[n], dtype=np.float32) for n in names]
    nz, ny, nx = data[0].shape
    if any(a.shape != (nz, ny, nx) for a in data):
        raise ValueError("arrays differ in shape")
    cnames = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    cptrs = (ctypes.c_void_p * len(names))(*[a.ctypes.data for a in data])
    rc = _abi.load().pf_write_vtk(str(path).encode(), FORMATS[fmt], nx, ny, nz, cnames, cptrs, len(names))
    if rc:
        raise OSError(f"pf_write_vtk({path}) failed: {_abi.STATUS.get(rc, rc)}")


def read_bin(path: str, shape) -> np.ndarray:
    """A raw FP32 field dump as an array of shape (nz, ny, nx)."""
    a = np.fromfile(path, dtype="<f4")
    return a.reshape(tuple(shape))


def read_vtk(path: str):
    """(meta, arrays) of a .vti/.vtr file: meta holds the grid type, extent,
    header type and the XML tree; arrays maps point-array names to
    (nz, ny, nx) float32 arrays and, for a RectilinearGrid, 'x', 'y', 'z'
    to its coordinates."""
    raw = open(path, "rb").read()
    start = raw.index(b"<AppendedData")
    mark = raw.index(b"_", raw.index(b">", start)) + 1  # appended data begins after '_'
    if not raw.rstrip().endswith(b"</VTKFile>"):
        raise ValueError(f"{path}: truncated file")
    root = ET.fromstring(raw[:start].decode() + "</VTKFile>")
    kind = root.get("type")
    grid = root.find(kind)
    header = root.get("header_type", "UInt32")
    hsize, hfmt = (8, "<u8") if header == "UInt64" else (4, "<u4")
    ext = [int(v) for v in grid.get("WholeExtent").split()]
    shape = (ext[5] - ext[4] + 1, ext[3] - ext[2] + 1, ext[1] - ext[0] + 1)

    def block(da):
        off = mark + int(da.get("offset"))
        nbytes = int(np.frombuffer(raw, dtype=hfmt, count=1, offset=off)[0])
        return np.frombuffer(raw, dtype="<f4", count=nbytes // 4, offset=off + hsize).copy()

    out = {da.get("Name"): block(da).reshape(shape) for da in grid.findall("Piece/PointData/DataArray")}
    for da in grid.findall("Piece/Coordinates/DataArray"):
        out[da.get("Name")] = block(da)
    meta = {"type": kind, "extent": ext, "shape": shape, "header_type": header, "xml": root}
    return meta, out
