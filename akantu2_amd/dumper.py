"""ParaView output of the model's fields (the role of the reference's
DumperParaview, src/io/dumper/dumper_iohelper_paraview.cc, and
Model::addDumpField / dump): one VTK XML UnstructuredGrid file (.vtu) per
dump, and a .pvd collection that lists them with their times.

Nodal fields are read back from the device at dump time; element fields
(grad_u, stress) are written per element with nqp x 9 components, the
quadrature-point values of the element in order.  Arrays are raw binary,
base64-encoded with a UInt64 byte-count header (VTK XML "binary" format).
"""
from __future__ import annotations

import base64
import os
import struct
from xml.sax.saxutils import quoteattr

import numpy as np

from .mesh import HEX8, TET4

VTK_CELL = {TET4: 10, HEX8: 12}  # VTK_TETRA, VTK_HEXAHEDRON (same node order)
NODAL = ("displacement", "velocity", "acceleration", "internal_force",
         "external_force", "blocked_dofs", "mass")
ELEMENTAL = ("grad_u", "stress")


def _data_array(name, arr):
    a = np.ascontiguousarray(arr)
    if a.dtype == np.float64:
        vt = "Float64"
    elif a.dtype == np.int64:
        vt = "Int64"
    elif a.dtype == np.int32:
        vt = "Int32"
    elif a.dtype == np.uint8:
        vt = "UInt8"
    else:
        raise TypeError(f"unsupported dtype {a.dtype}")
    comps = 1 if a.ndim == 1 else a.shape[1]
    raw = a.tobytes()
    payload = base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")
    return (f'<DataArray type="{vt}" Name={quoteattr(name)} '
            f'NumberOfComponents="{comps}" format="binary">{payload}</DataArray>')


def write_vtu(path, nodes, conn, etype, point_data=None, cell_data=None):
    """One unstructured-grid file: points, cells and the given arrays."""
    nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, np.int64)
    nel, nnpe = conn.shape
    offsets = np.arange(1, nel + 1, dtype=np.int64) * nnpe
    types = np.full(nel, VTK_CELL[etype], dtype=np.uint8)
    parts = ['<?xml version="1.0"?>',
             '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
             'header_type="UInt64">', "<UnstructuredGrid>",
             f'<Piece NumberOfPoints="{nodes.shape[0]}" NumberOfCells="{nel}">',
             "<Points>", _data_array("positions", nodes), "</Points>", "<Cells>",
             _data_array("connectivity", conn.ravel()), _data_array("offsets", offsets),
             _data_array("types", types), "</Cells>"]
    if point_data:
        parts.append("<PointData>")
        parts += [_data_array(k, v) for k, v in point_data.items()]
        parts.append("</PointData>")
    if cell_data:
        parts.append("<CellData>")
        parts += [_data_array(k, v) for k, v in cell_data.items()]
        parts.append("</CellData>")
    parts += ["</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        fh.write("\n".join(parts))
    os.replace(tmp, path)


def write_pvd(path, entries):
    """Collection file: entries = [(time, file name relative to the .pvd)]."""
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">',
             "<Collection>"]
    lines += [f'<DataSet timestep="{t!r}" group="" part="0" file={quoteattr(f)}/>'
              for t, f in entries]
    lines += ["</Collection>", "</VTKFile>"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines))


class DumperParaview:
    """model.setBaseName / addDumpField / dump, for one model."""

    def __init__(self, model, base_name="solid_mechanics_model", directory="paraview"):
        self.model, self.base_name, self.directory = model, base_name, directory
        self.fields = []
        self.entries = []

    def addDumpField(self, name):
        if name not in NODAL + ELEMENTAL:
            raise ValueError(f"unknown dump field {name!r}")
        if name not in self.fields:
            self.fields.append(name)

    def dump(self, time=None, step=None):
        m = self.model
        step = len(self.entries) if step is None else int(step)
        time = float(step * m.getTimeStep()) if time is None else float(time)
        pd, cd = {}, {}
        getters = {"displacement": m.getDisplacement, "velocity": m.getVelocity,
                   "acceleration": m.getAcceleration, "internal_force": m.getInternalForce,
                   "external_force": m.getExternalForce, "blocked_dofs": m.getBlockedDOFs,
                   "mass": m.getMass}
        for f in self.fields:
            if f in getters:
                pd[f] = np.array(getters[f]())
            else:
                nq = m.nqp
                full = np.zeros((m.mesh.nb_elements, nq * 9))
                for mat in m.materials:
                    vals = mat.getGradU() if f == "grad_u" else mat.getStress()
                    ids = getattr(mat, "element_ids", None)
                    ids = np.arange(m.mesh.nb_elements) if ids is None else ids
                    full[ids] = vals.reshape(-1, nq * 9)
                cd[f] = full
        os.makedirs(self.directory, exist_ok=True)
        name = f"{self.base_name}_{step:04d}.vtu"
        write_vtu(os.path.join(self.directory, name), m.mesh.nodes, m.mesh.conn,
                  m.mesh.etype, pd, cd)
        self.entries.append((time, name))
        write_pvd(os.path.join(self.directory, self.base_name + ".pvd"), self.entries)
        return os.path.join(self.directory, name)
