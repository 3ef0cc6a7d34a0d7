"""The ParaView writer (akantu2_amd.dumper): the .vtu it writes decodes back
to the arrays it was given, cells carry VTK's tetra / hexahedron codes with
Akantu's node order (the same for these two types), and the .pvd lists the
dumps with their times."""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np

from akantu2_amd import mesh as M
from akantu2_amd.dumper import write_pvd, write_vtu

_NP = {"Float64": np.float64, "Int64": np.int64, "Int32": np.int32, "UInt8": np.uint8}


def _decode(da):
    raw = base64.b64decode(da.text.strip())
    n = struct.unpack("<Q", raw[:8])[0]
    a = np.frombuffer(raw[8:8 + n], dtype=_NP[da.get("type")])
    c = int(da.get("NumberOfComponents"))
    return a.reshape(-1, c) if c > 1 else a


def test_vtu_round_trip(tmp_path):
    for etype, code in (("_tetrahedron_4", 10), ("_hexahedron_8", 12)):
        mesh = M.box_mesh(etype, 3, 2, 2, jitter=0.1, seed=1)
        rng = np.random.default_rng(0)
        u = rng.normal(size=(mesh.nb_nodes, 3))
        bl = M.clamped_face_x0(mesh)
        st = rng.normal(size=(mesh.nb_elements, mesh.nqp * 9))
        path = str(tmp_path / f"m{code}.vtu")
        write_vtu(path, mesh.nodes, mesh.conn, mesh.etype,
                  {"displacement": u, "blocked_dofs": bl}, {"stress": st})
        root = ET.parse(path).getroot()
        piece = root.find("UnstructuredGrid/Piece")
        assert int(piece.get("NumberOfPoints")) == mesh.nb_nodes
        assert int(piece.get("NumberOfCells")) == mesh.nb_elements
        arrays = {da.get("Name"): _decode(da) for da in piece.iter("DataArray")}
        np.testing.assert_array_equal(arrays["positions"], mesh.nodes)
        nnpe = mesh.conn.shape[1]
        np.testing.assert_array_equal(arrays["connectivity"].reshape(-1, nnpe), mesh.conn)
        np.testing.assert_array_equal(arrays["offsets"],
                                      nnpe * np.arange(1, mesh.nb_elements + 1))
        assert (arrays["types"] == code).all()
        np.testing.assert_array_equal(arrays["displacement"], u)
        np.testing.assert_array_equal(arrays["blocked_dofs"], bl)
        np.testing.assert_array_equal(arrays["stress"], st)


def test_pvd_lists_the_dumps(tmp_path):
    p = str(tmp_path / "run.pvd")
    write_pvd(p, [(0.0, "run_0000.vtu"), (0.5, "run_0001.vtu")])
    ds = ET.parse(p).getroot().findall("Collection/DataSet")
    assert [(float(d.get("timestep")), d.get("file")) for d in ds] == \
        [(0.0, "run_0000.vtu"), (0.5, "run_0001.vtu")]
