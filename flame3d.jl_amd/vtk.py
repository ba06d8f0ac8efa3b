"""Plot files in the VTK XML structured-grid format (.vts), the fields the
reference's plot output writes (solver.jl:394-434): interior node
coordinates, rho, velocity (3 components), p, T, the shock sensor phi, each
species' partial density under "Y" + species name, tag, and the time as
field data.  Point data are Float32 as there (tag Int64), in the reference's
column-major [i, j, k] order; the binary payload is appended raw with a
UInt64 byte count per array (VTK XML "appended" encoding), uncompressed.

`read_vts` reads such a file back into numpy arrays (used by the tests and
for post-processing without VTK installed).
"""
from __future__ import annotations

import struct
import xml.etree.ElementTree as ET
from typing import Dict, Sequence

import numpy as np

NG = 4
_VTK_TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64",
              np.dtype(np.int64): "Int64", np.dtype(np.int32): "Int32"}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


def _interior(t, ng: int = NG):
    """Device or host array (..., Nz+8, Ny+8, Nx+8) -> host numpy interior."""
    a = t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
    return a[..., ng:-ng, ng:-ng, ng:-ng]


def write_vts(path: str, x, y, z, Q, rhoi, species: Sequence[str], phi=None, tag=None,
              time: float = 0.0):
    """Write one plot file.  x, y, z: padded node coordinates; Q: (7, ...)
    primitive state (rho, u, v, w, p, T, e); rhoi: (Ns, ...) partial
    densities; phi: shock sensor (optional); tag: int64 (optional)."""
    q = _interior(Q)
    r = _interior(rhoi)
    nz, ny, nx = q.shape[1:]
    pts = np.stack([_interior(x), _interior(y), _interior(z)], axis=-1).astype(np.float32)
    fields: Dict[str, np.ndarray] = {
        "rho": q[0].astype(np.float32),
        "velocity": np.stack([q[1], q[2], q[3]], axis=-1).astype(np.float32),
        "p": q[4].astype(np.float32),
        "T": q[5].astype(np.float32),
    }
    if phi is not None:
        fields["phi"] = _interior(phi).astype(np.float32)
    for k, name in enumerate(species):
        fields["Y" + name] = r[k].astype(np.float32)
    if tag is not None:
        fields["tag"] = _interior(tag).astype(np.int64)
    blocks, offset, arrays = [], 0, []
    for name, a in fields.items():
        nc = a.shape[-1] if a.ndim == 4 else 1
        comp = f' NumberOfComponents="{nc}"' if nc > 1 else ""
        arrays.append(f'        <DataArray type="{_VTK_TYPES[a.dtype]}" Name="{name}"{comp} '
                      f'format="appended" offset="{offset}"/>')
        raw = np.ascontiguousarray(a).tobytes()  # [k][j][i](comp): i fastest
        blocks.append(struct.pack("<Q", len(raw)) + raw)
        offset += 8 + len(raw)
    raw = np.ascontiguousarray(pts).tobytes()
    pts_xml = (f'        <DataArray type="Float32" NumberOfComponents="3" format="appended" '
               f'offset="{offset}"/>')
    blocks.append(struct.pack("<Q", len(raw)) + raw)
    ext = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    head = "\n".join([
        '<?xml version="1.0"?>',
        '<VTKFile type="StructuredGrid" version="1.0" byte_order="LittleEndian" '
        'header_type="UInt64">',
        f'  <StructuredGrid WholeExtent="{ext}">',
        '    <FieldData>',
        f'      <DataArray type="Float64" Name="Time" NumberOfTuples="1" format="ascii">'
        f'{time!r}</DataArray>',
        '    </FieldData>',
        f'    <Piece Extent="{ext}">',
        '      <PointData>',
        *arrays,
        '      </PointData>',
        '      <Points>',
        pts_xml,
        '      </Points>',
        '    </Piece>',
        '  </StructuredGrid>',
        '  <AppendedData encoding="raw">',
        '_'])
    with open(path, "wb") as fh:
        fh.write(head.encode())
        for b in blocks:
            fh.write(b)
        fh.write(b"\n  </AppendedData>\n</VTKFile>\n")


def read_vts(path: str):
    """(fields dict name -> array [nz, ny, nx(, comp)], points [nz, ny, nx, 3],
    time) from a file written by write_vts."""
    data = open(path, "rb").read()
    start = data.index(b'<AppendedData encoding="raw">')
    us = data.index(b"_", start) + 1
    end = data.rindex(b"</AppendedData>")
    xml = data[:us - 1] + b"\n  </AppendedData>\n</VTKFile>\n"
    root = ET.fromstring(xml)
    payload = data[us:end]
    sg = root.find("StructuredGrid")
    e = [int(v) for v in sg.get("WholeExtent").split()]
    shape = (e[5] - e[4] + 1, e[3] - e[2] + 1, e[1] - e[0] + 1)

    def arr(el):
        off = int(el.get("offset"))
        (nbytes,) = struct.unpack_from("<Q", payload, off)
        dt = np.dtype(_NP_TYPES[el.get("type")])
        a = np.frombuffer(payload, dtype=dt, count=nbytes // dt.itemsize, offset=off + 8).copy()
        nc = int(el.get("NumberOfComponents", "1"))
        return a.reshape(shape + ((nc,) if nc > 1 else ()))

    piece = sg.find("Piece")
    fields = {el.get("Name"): arr(el) for el in piece.find("PointData")}
    pts = arr(piece.find("Points").find("DataArray"))
    time = float(sg.find("FieldData").find("DataArray").text)
    return fields, pts, time
