"""Reader for gmsh ASCII ``.msh`` files (format 2.x), the subset Akantu's
``MeshIOMSH`` (src/io/mesh_io/mesh_io_msh.cc) needs for Tet4 and Hex8 meshes.

gmsh element type 4 is ``_tetrahedron_4`` and 5 is ``_hexahedron_8``; both
keep gmsh's node order (mesh_io_msh.cc:200-263 leaves them unpermuted).
Node tags are renumbered to 0-based indices in file order.  Boundary facets
(2 = ``_triangle_3``, 3 = ``_quadrangle_4``) become element groups named by
their physical name, or their physical tag, or (tag 0) their elementary tag,
as the reference's groups built from the mesh data "tag_0" / "tag_1".
"""
from __future__ import annotations

import numpy as np

from .mesh import HEX8, TET4, BoxMesh, add_group

_MSH_TYPES = {4: (TET4, 4), 5: (HEX8, 8)}
# boundary facets: gmsh 2 = _triangle_3, 3 = _quadrangle_4
_MSH_FACETS = {2: (2, 3), 3: (3, 4)}


def read_msh(path, etype=None):
    """Return a ``BoxMesh``-like object with the volume elements of the file.

    ``etype`` selects Tet4 or Hex8 when the file holds both; by default the
    first volume type found is used."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    it = iter(range(len(lines)))
    nodes, tag_to_idx, elems = [], {}, {}
    facets, names = {}, {}
    i = 0
    while i < len(lines):
        tok = lines[i].strip()
        if tok == "$MeshFormat":
            version = float(lines[i + 1].split()[0])
            if not (2.0 <= version < 3.0):
                raise ValueError(f"unsupported msh version {version}")
            i += 2
        elif tok == "$Nodes":
            n = int(lines[i + 1])
            for k in range(n):
                p = lines[i + 2 + k].split()
                tag_to_idx[int(p[0])] = k
                nodes.append([float(p[1]), float(p[2]), float(p[3])])
            i += 2 + n
        elif tok == "$PhysicalNames":
            n = int(lines[i + 1])
            for k in range(n):
                p = lines[i + 2 + k].split(maxsplit=2)
                names[int(p[1])] = p[2].strip().strip('"')
            i += 2 + n
        elif tok == "$Elements":
            n = int(lines[i + 1])
            for k in range(n):
                p = [int(x) for x in lines[i + 2 + k].split()]
                mtype, ntags = p[1], p[2]
                if mtype in _MSH_TYPES:
                    et, nn = _MSH_TYPES[mtype]
                    c = p[3 + ntags: 3 + ntags + nn]
                    elems.setdefault(et, []).append([tag_to_idx[t] for t in c])
                elif mtype in _MSH_FACETS and ntags >= 1:
                    ft, nn = _MSH_FACETS[mtype]
                    # physical tag, or the elementary one when it is 0
                    tag = p[3] if p[3] != 0 or ntags < 2 else p[4]
                    c = p[3 + ntags: 3 + ntags + nn]
                    facets.setdefault((tag, ft), []).append([tag_to_idx[t] for t in c])
            i += 2 + n
        else:
            i += 1
    del it
    if etype is None:
        if not elems:
            raise ValueError("no Tet4/Hex8 elements in " + str(path))
        etype = sorted(elems)[0]
    conn = np.asarray(elems[etype], dtype=np.int32)
    X = np.asarray(nodes, dtype=np.float64)
    mesh = BoxMesh(etype, X, conn, shape=(0, 0, 0), meta={"file": str(path)})
    for (tag, ft), fc in sorted(facets.items()):
        fc = np.asarray(fc, dtype=np.int32)
        add_group(mesh, names.get(tag, str(tag)),
                  {"facet_type": ft, "facets": fc, "nodes": np.unique(fc)})
    return mesh


def boundary_nodes(mesh) -> np.ndarray:
    """Nodes lying on faces that belong to exactly one element (the groups
    ``Mesh::createBoundaryGroupFromGeometry`` builds from the facets)."""
    if mesh.etype == TET4:
        faces = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    else:
        faces = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5),
                 (2, 3, 7, 6), (3, 0, 4, 7)]
    allf = np.concatenate([np.sort(mesh.conn[:, list(f)], axis=1)
                           for f in faces])
    uniq, counts = np.unique(allf, axis=0, return_counts=True)
    on = np.zeros(mesh.nb_nodes, dtype=bool)
    on[np.unique(uniq[counts == 1].ravel())] = True
    return on
