"""VTU snapshots of the density field.

``export_mesh_with_info`` keeps the call surface of
``core/visualization.py:22-84`` (points, one cell block, named point and
cell arrays, written to ``filepath``) and ``export_submesh`` that of
``mesh/visualization.py:11-37`` (the elements kept at ``rho > threshold``
plus every non-design element, with compacted node numbering).  The
reference writes through meshio; this module writes the VTK XML
UnstructuredGrid format itself (inline base64 binary arrays, UInt64 block
headers), readable by ParaView, VTK and meshio.  Cell node order is the
mesh's own corner order, which for hex8 is VTK_HEXAHEDRON's and for tet4
VTK_TETRA's.  ``read_vtu`` reads files of this writer back.
"""
from __future__ import annotations

import base64
import os
import xml.etree.ElementTree as ET

import numpy as np

_VTK_TYPE = {"hex8": 12, "tet4": 10}
_NP2VTK = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
           np.dtype(np.int64): "Int64", np.dtype(np.int32): "Int32",
           np.dtype(np.uint8): "UInt8"}
_VTK2NP = {v: k for k, v in _NP2VTK.items()}


def _encode(a: np.ndarray) -> str:
    raw = np.ascontiguousarray(a).tobytes()
    return base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode("ascii")


def _array_xml(name, a, ncomp=1):
    a = np.asarray(a)
    if a.dtype not in _NP2VTK:
        a = a.astype(np.float64)
    attrs = f'type="{_NP2VTK[a.dtype]}" Name="{name}" format="binary"'
    if ncomp > 1:
        attrs += f' NumberOfComponents="{ncomp}"'
    return f"<DataArray {attrs}>{_encode(a)}</DataArray>"


def write_vtu(filepath, points, cells, kind, point_data=None, cell_data=None):
    """points (3, n_nodes), cells (nen, n_el) int, kind 'hex8' / 'tet4'."""
    if kind not in _VTK_TYPE:
        raise ValueError(f"Unsupported mesh type: {kind}")
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).T)
    conn = np.ascontiguousarray(np.asarray(cells, dtype=np.int64).T)
    n_nodes, n_el, nen = pts.shape[0], conn.shape[0], conn.shape[1]
    offsets = np.arange(1, n_el + 1, dtype=np.int64) * nen
    types = np.full(n_el, _VTK_TYPE[kind], dtype=np.uint8)
    parts = ['<?xml version="1.0"?>',
             '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
             'header_type="UInt64">',
             "<UnstructuredGrid>",
             f'<Piece NumberOfPoints="{n_nodes}" NumberOfCells="{n_el}">',
             "<Points>", _array_xml("Points", pts, 3), "</Points>",
             "<Cells>", _array_xml("connectivity", conn.ravel()),
             _array_xml("offsets", offsets), _array_xml("types", types), "</Cells>"]
    if point_data:
        parts.append("<PointData>")
        parts += [_array_xml(k, v) for k, v in point_data.items()]
        parts.append("</PointData>")
    if cell_data:
        parts.append("<CellData>")
        parts += [_array_xml(k, v) for k, v in cell_data.items()]
        parts.append("</CellData>")
    parts += ["</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    d = os.path.dirname(filepath)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(filepath, "w") as f:
        f.write("\n".join(parts))


def read_vtu(filepath):
    """(points (3, n), cells (nen, n_el), vtk types, point_data, cell_data)."""
    root = ET.parse(filepath).getroot()
    piece = root.find("UnstructuredGrid/Piece")

    def arr(node):
        raw = base64.b64decode(node.text.strip())
        n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
        a = np.frombuffer(raw[8:8 + n], dtype=_VTK2NP[node.get("type")])
        nc = int(node.get("NumberOfComponents", "1"))
        return a.reshape(-1, nc) if nc > 1 else a

    pts = arr(piece.find("Points/DataArray"))
    cells = {c.get("Name"): arr(c) for c in piece.findall("Cells/DataArray")}
    nen = int(cells["offsets"][0])
    conn = cells["connectivity"].reshape(-1, nen).T
    pd = {c.get("Name"): arr(c) for c in piece.findall("PointData/DataArray")}
    cd = {c.get("Name"): arr(c) for c in piece.findall("CellData/DataArray")}
    return pts.T, conn, cells["types"], pd, cd


def export_mesh_with_info(mesh, point_data_values=None, point_data_names=None,
                          cell_data_values=None, cell_data_names=None,
                          filepath="output.vtu"):
    """``core/visualization.py:22-84``."""
    pd = dict(zip(point_data_names, point_data_values)) \
        if point_data_values and point_data_names else None
    cd = dict(zip(cell_data_names, cell_data_values)) \
        if cell_data_values and cell_data_names else None
    write_vtu(filepath, mesh.p, mesh.t, mesh.kind, pd, cd)


def export_submesh(tsk, rho_projected, threshold, dst_path):
    """``mesh/visualization.py:11-37``: drop the design elements with
    ``rho <= threshold``, keep the rest, renumber the used nodes in
    ascending order."""
    mesh = tsk.mesh
    rho = np.asarray(rho_projected)
    design = np.asarray(tsk.design_elements)
    remove = design[rho[design] <= threshold]
    keep = np.setdiff1d(np.arange(mesh.nelements), remove)
    kept_t = mesh.t[:, keep]
    used = np.unique(kept_t)
    new_t = np.searchsorted(used, kept_t)
    write_vtu(dst_path, mesh.p[:, used], new_t, mesh.kind)
    return mesh.p[:, used], new_t
