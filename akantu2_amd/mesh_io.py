"""Reader for gmsh ASCII ``.msh`` files, the formats Akantu's ``MeshIOMSH``
reads (src/io/mesh_io/mesh_io_msh.cc): version 1 (``$NOD`` / ``$ELM``),
2.x, 4.0 and 4.1.  Like the reference (mesh_io_msh.cc:698-713) it reads
ASCII files only and refuses binary ones.

gmsh element type 4 is ``_tetrahedron_4`` and 5 is ``_hexahedron_8``; both
keep gmsh's node order (the reference's read order is the identity for them,
mesh_io_msh.cc:213-263).  Node tags are renumbered to 0-based indices in file
order.  Boundary facets (2 = ``_triangle_3``, 3 = ``_quadrangle_4``) become
element groups:
- versions 1 and 2: named by their physical name, or their physical tag, or
  (physical tag 0) their elementary tag, as the groups built from the mesh
  data "tag_0" / "tag_1";
- version 4: the physical groups of the facet's entity, named by the
  physical name or ``msh_block_<tag>`` (mesh_io_msh.cc:501-549).
"""
from __future__ import annotations

import numpy as np

from .mesh import HEX8, TET4, BoxMesh, add_group

_MSH_TYPES = {4: (TET4, 4), 5: (HEX8, 8)}
# boundary facets: gmsh 2 = _triangle_3, 3 = _quadrangle_4
_MSH_FACETS = {2: (2, 3), 3: (3, 4)}


class _Lines:
    """Line cursor over the file (blank lines skipped, as the reference's
    trimmed getline / >> reads do)."""

    def __init__(self, text):
        self.lines = [ln.strip() for ln in text.split("\n")]
        self.i = 0

    def next(self):
        while self.i < len(self.lines) and not self.lines[self.i]:
            self.i += 1
        if self.i >= len(self.lines):
            raise ValueError("reached the end of the file inside a section")
        ln = self.lines[self.i]
        self.i += 1
        return ln

    def ints(self):
        return [int(x) for x in self.next().split()]

    def at_end(self):
        while self.i < len(self.lines) and not self.lines[self.i]:
            self.i += 1
        return self.i >= len(self.lines)


def _version(text: str) -> int:
    """major * 1000 + minor, as mesh_io_msh.cc:698-713 encodes it."""
    major, minor = (int(x) for x in f"{float(text):.1f}".split("."))
    return major * 1000 + minor


class _Reader:
    def __init__(self):
        self.version = 1000  # a file without $MeshFormat is version 1
        self.nodes = []
        self.tag_to_idx = {}
        self.elems = {}
        self.facets = {}       # (group name, facet type) -> [[nodes]]
        self.names = {}        # physical tag -> name
        self.entity_phys = {}  # (entity dim, entity tag) -> [physical tags]

    # ---------------------------------------------------------- sections
    def mesh_format(self, L):
        p = L.next().split()
        self.version = _version(p[0])
        if len(p) > 1 and int(p[1]) != 0:
            raise ValueError("This reader can only read ASCII files.")

    def physical_names(self, L):
        for _ in range(int(L.next())):
            p = L.next().split(maxsplit=2)
            self.names[int(p[1])] = p[2].strip().strip('"')

    def entities(self, L):
        counts = L.ints()
        for dim in range(4):
            for _ in range(counts[dim]):
                p = L.next().split()
                tag = int(p[0])
                # points of 4.1 have one coordinate triple, everything else a
                # bounding box (mesh_io_msh.cc:540-543)
                k = 4 if (dim == 0 and self.version >= 4001) else 7
                nphys = int(p[k])
                self.entity_phys[(dim, tag)] = [int(x) for x in p[k + 1:k + 1 + nphys]]

    def nodes_v2(self, L):
        for _ in range(int(L.next())):
            p = L.next().split()
            self.tag_to_idx[int(p[0])] = len(self.nodes)
            self.nodes.append([float(p[1]), float(p[2]), float(p[3])])

    def nodes_v4(self, L):
        head = L.ints()
        for _ in range(head[0]):
            b = L.ints()
            if self.version >= 4001:
                _dim, _tag, parametric, n = b
                if parametric:
                    raise ValueError("parametric nodes are not supported in msh files")
                # 4.1 lists all tags of the block first, then the coordinates
                tags = [int(L.next()) for _ in range(n)]
                for t in tags:
                    self.tag_to_idx[t] = len(self.nodes)
                    xyz = L.next().split()
                    self.nodes.append([float(xyz[0]), float(xyz[1]), float(xyz[2])])
            else:
                _tag, _dim, parametric, n = b
                if parametric:
                    raise ValueError("parametric nodes are not supported in msh files")
                for _ in range(n):
                    p = L.next().split()
                    self.tag_to_idx[int(p[0])] = len(self.nodes)
                    self.nodes.append([float(p[1]), float(p[2]), float(p[3])])

    def _add(self, mtype, conn_tags, group_names):
        if mtype in _MSH_TYPES:
            et, nn = _MSH_TYPES[mtype]
            self.elems.setdefault(et, []).append(
                [self.tag_to_idx[t] for t in conn_tags[:nn]])
        elif mtype in _MSH_FACETS:
            ft, nn = _MSH_FACETS[mtype]
            c = [self.tag_to_idx[t] for t in conn_tags[:nn]]
            for g in group_names:
                self.facets.setdefault((g, ft), []).append(c)

    def elements_v1(self, L):
        for _ in range(int(L.next())):
            p = [int(x) for x in L.next().split()]
            # index type reg-phys reg-elem nb-nodes nodes...
            mtype, phys, elem = p[1], p[2], p[3]
            tag = phys if phys != 0 else elem
            self._add(mtype, p[5:], [self.names.get(tag, str(tag))])

    def elements_v2(self, L):
        for _ in range(int(L.next())):
            p = [int(x) for x in L.next().split()]
            mtype, ntags = p[1], p[2]
            names = []
            if ntags >= 1:
                # physical tag, or the elementary one when it is 0
                tag = p[3] if p[3] != 0 or ntags < 2 else p[4]
                names = [self.names.get(tag, str(tag))]
            self._add(mtype, p[3 + ntags:], names)

    def elements_v4(self, L):
        head = L.ints()
        for _ in range(head[0]):
            b = L.ints()
            if self.version >= 4001:
                dim, tag, mtype, n = b
            else:
                tag, dim, mtype, n = b
            phys = self.entity_phys.get((dim, tag), [])
            names = [self.names.get(t, f"msh_block_{t}") for t in phys]
            for _ in range(n):
                p = [int(x) for x in L.next().split()]
                self._add(mtype, p[1:], names)

    # -------------------------------------------------------------- read
    def read(self, text):
        L = _Lines(text)
        while not L.at_end():
            tok = L.next()
            # section ends: $EndNodes ... (versions 2, 4), $ENDNOD, $ENDELM (1)
            if not tok.startswith("$") or tok.upper().startswith("$END"):
                continue
            if tok == "$MeshFormat":
                self.mesh_format(L)
            elif tok == "$PhysicalNames":
                self.physical_names(L)
            elif tok == "$Entities":
                self.entities(L)
            elif tok in ("$Nodes", "$NOD"):
                (self.nodes_v4 if self.version >= 4000 else self.nodes_v2)(L)
            elif tok in ("$Elements", "$ELM"):
                if self.version >= 4000:
                    self.elements_v4(L)
                elif self.version >= 2000:
                    self.elements_v2(L)
                else:
                    self.elements_v1(L)
            else:
                # a section this reader does not use: skip to its end
                end = "$End" + tok[1:]
                while L.next() != end:
                    pass


def read_msh(path, etype=None):
    """Return a ``BoxMesh``-like object with the volume elements of the file.

    ``etype`` selects Tet4 or Hex8 when the file holds both; by default the
    first volume type found is used."""
    with open(path, "rb") as fh:
        raw = fh.read()
    try:
        text = raw.decode("ascii")
    except UnicodeDecodeError:
        raise ValueError("This reader can only read ASCII files.") from None
    r = _Reader()
    r.read(text)
    if etype is None:
        if not r.elems:
            raise ValueError("no Tet4/Hex8 elements in " + str(path))
        etype = sorted(r.elems)[0]
    conn = np.asarray(r.elems[etype], dtype=np.int32)
    X = np.asarray(r.nodes, dtype=np.float64)
    mesh = BoxMesh(etype, X, conn, shape=(0, 0, 0),
                   meta={"file": str(path), "msh_version": r.version})
    for (name, ft), fc in sorted(r.facets.items()):
        fc = np.asarray(fc, dtype=np.int32)
        add_group(mesh, name, {"facet_type": ft, "facets": fc, "nodes": np.unique(fc)})
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
