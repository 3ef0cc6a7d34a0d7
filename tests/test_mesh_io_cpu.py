"""gmsh reader (akantu2_amd.mesh_io) for the formats the reference's
MeshIOMSH reads (mesh_io_msh.cc): versions 1, 2.x, 4.0 and 4.1 ASCII, and
the refusal of binary files.

- the reference's 4.1 mesh of its scalability example (a Tet4 cube
  [-1, 1]^3): node and element counts of the file, every tet positively
  oriented, total volume 8, the physical surfaces as facet groups whose
  nodes lie on their planes and whose areas add up to the faces';
- the reference's 4.0 phase-field mesh (one quadrangle, 2-D): its nodes are
  read, and it is refused as a volume mesh;
- the reference's 2.1 patch meshes written back in formats 1, 4.0 and 4.1
  (by a writer in this test) read to the same nodes, connectivity and
  groups as the 2.1 file.
"""
import os

import numpy as np
import pytest

from akantu2_amd import mesh_io

from tests.conftest import REFERENCE_FIXTURES


def _fx(name):
    return os.path.join(REFERENCE_FIXTURES, name)


def _tet_volumes(X, c):
    return np.einsum("ij,ij->i", np.cross(X[c[:, 1]] - X[c[:, 0]], X[c[:, 2]] - X[c[:, 0]]),
                     X[c[:, 3]] - X[c[:, 0]]) / 6.0


def _tri_areas(X, f):
    return 0.5 * np.linalg.norm(np.cross(X[f[:, 1]] - X[f[:, 0]], X[f[:, 2]] - X[f[:, 0]]),
                                axis=1)


def test_v41_scalability_cube():
    path = _fx("scalability_cube_v41.msh")
    text = open(path).read().split("\n")
    # counts from the section headers of the file itself
    i = text.index("$Nodes")
    nb_nodes = int(text[i + 1].split()[1])
    m = mesh_io.read_msh(path)
    assert m.meta["msh_version"] == 4001
    assert m.etype == mesh_io.TET4 and m.nb_nodes == nb_nodes == 141
    assert m.nb_elements == 373
    vol = _tet_volumes(m.nodes, m.conn)
    assert vol.min() > 0
    assert abs(vol.sum() - 8.0) < 1e-12
    assert np.allclose(m.nodes.min(axis=0), -1) and np.allclose(m.nodes.max(axis=0), 1)
    groups = m.meta["groups"]
    assert set(groups) == {"top", "bottom", "side left", "side right"}
    planes = {"top": (2, 1.0), "bottom": (2, -1.0), "side left": (0, -1.0),
              "side right": (0, 1.0)}
    for name, (axis, value) in planes.items():
        g = groups[name]
        assert g["facet_type"] == 2
        assert np.allclose(m.nodes[g["nodes"], axis], value), name
        # each physical surface covers one 2 x 2 face of the cube
        assert abs(_tri_areas(m.nodes, g["facets"]).sum() - 4.0) < 1e-12, name


def test_v40_two_dimensional_mesh():
    path = _fx("phase_field_one_element_v40.msh")
    r = mesh_io._Reader()
    r.read(open(path).read())
    assert r.version == 4000
    assert np.array_equal(np.asarray(r.nodes),
                          [[-0.5, -0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0], [-0.5, 0.5, 0]])
    assert r.entity_phys[(2, 1)] == [1] and r.names[1] == "plate"
    with pytest.raises(ValueError, match="no Tet4/Hex8 elements"):
        mesh_io.read_msh(path)


def test_binary_files_are_refused(tmp_path):
    p = tmp_path / "bin.msh"
    p.write_bytes(b"$MeshFormat\n4.1 1 8\n\x01\x00\x00\x00\n$EndMeshFormat\n")
    with pytest.raises(ValueError, match="only read ASCII"):
        mesh_io.read_msh(str(p))
    p.write_bytes(b"$MeshFormat\n2.2 1 8\n$EndMeshFormat\n")
    with pytest.raises(ValueError, match="only read ASCII"):
        mesh_io.read_msh(str(p))


# ------------------------------------------------------------------ writers
def _elements(m):
    """(gmsh type, physical tag, node tags) of the volume elements and the
    facets of a mesh read from a 2.x file, tags 1-based."""
    vol_type = 4 if m.etype == mesh_io.TET4 else 5
    out = [(vol_type, 1, c + 1) for c in m.conn]
    tags = {}
    for k, (name, g) in enumerate(sorted(m.meta.get("groups", {}).items())):
        tags[name] = 100 + k
        ft = 2 if g["facet_type"] == 2 else 3
        out += [(ft, tags[name], f + 1) for f in g["facets"]]
    return out, tags


def _write_v1(m, path):
    els, tags = _elements(m)
    lines = ["$NOD", str(m.nb_nodes)]
    lines += [f"{k + 1} {float(x)!r} {float(y)!r} {float(z)!r}" for k, (x, y, z) in enumerate(m.nodes)]
    lines += ["$ENDNOD", "$ELM", str(len(els))]
    for k, (t, phys, c) in enumerate(els):
        lines.append(f"{k + 1} {t} {phys} {phys} {len(c)} " + " ".join(map(str, c)))
    lines.append("$ENDELM")
    open(path, "w").write("\n".join(lines) + "\n")
    return {str(v): k for k, v in tags.items()}


def _write_v4(m, path, minor):
    els, tags = _elements(m)
    names = {v: k for k, v in tags.items()}
    names[1] = "volume"
    ents = {}
    for t, phys, _ in els:
        dim = 3 if t in (4, 5) else 2
        ents.setdefault((dim, phys), []).append(None)
    lines = ["$MeshFormat", f"4.{minor} 0 8", "$EndMeshFormat", "$PhysicalNames",
             str(len(names))]
    lines += [f"{3 if p == 1 else 2} {p} \"{n}\"" for p, n in sorted(names.items())]
    lines += ["$EndPhysicalNames", "$Entities"]
    surf = sorted(p for d, p in ents if d == 2)
    lines.append(f"0 0 {len(surf)} 1")
    for p in surf:  # entity tag = physical tag, one physical group each
        lines.append(f"{p} 0 0 0 1 1 1 1 {p} 0")
    lines.append("1 0 0 0 1 1 1 1 1 0")
    lines += ["$EndEntities", "$Nodes"]
    n = m.nb_nodes
    if minor == 1:
        lines += [f"1 {n} 1 {n}", f"3 1 0 {n}"]
        lines += [str(k + 1) for k in range(n)]
        lines += [f"{float(x)!r} {float(y)!r} {float(z)!r}" for x, y, z in m.nodes]
    else:
        lines += [f"1 {n}", f"1 3 0 {n}"]
        lines += [f"{k + 1} {float(x)!r} {float(y)!r} {float(z)!r}" for k, (x, y, z) in enumerate(m.nodes)]
    lines += ["$EndNodes", "$Elements"]
    blocks = {}
    for t, phys, c in els:
        blocks.setdefault((3 if t in (4, 5) else 2, phys, t), []).append(c)
    lines.append(f"{len(blocks)} {len(els)}")
    k = 1
    for (dim, phys, t), cs in blocks.items():
        lines.append(f"{dim} {phys} {t} {len(cs)}" if minor == 1 else
                     f"{phys} {dim} {t} {len(cs)}")
        for c in cs:
            lines.append(f"{k} " + " ".join(map(str, c)))
            k += 1
    lines.append("$EndElements")
    open(path, "w").write("\n".join(lines) + "\n")
    return {k: k for k in tags}  # the physical names carry the group names


@pytest.mark.parametrize("name", ["patch_tetrahedron_4.msh", "patch_hexahedron_8.msh"])
@pytest.mark.parametrize("fmt", ["v1", "v40", "v41"])
def test_formats_read_the_same_mesh(tmp_path, name, fmt):
    ref = mesh_io.read_msh(_fx(name))
    assert ref.meta["msh_version"] == 2001
    path = str(tmp_path / f"{fmt}.msh")
    if fmt == "v1":
        rename = _write_v1(ref, path)
        version = 1000
    else:
        rename = _write_v4(ref, path, 0 if fmt == "v40" else 1)
        version = 4000 if fmt == "v40" else 4001
    m = mesh_io.read_msh(path)
    assert m.meta["msh_version"] == version
    assert m.etype == ref.etype
    assert np.array_equal(m.nodes, ref.nodes)
    assert np.array_equal(m.conn, ref.conn)
    got = m.meta.get("groups", {})
    want = ref.meta.get("groups", {})
    assert len(got) == len(want)
    for gname, g in got.items():
        r = want[rename[gname]]
        assert g["facet_type"] == r["facet_type"]
        assert np.array_equal(g["facets"], r["facets"])
        assert np.array_equal(g["nodes"], r["nodes"])
