"""BC::Neumann::FromTraction in the oracle (boundary_condition_tmpl.hh
:100-160) pinned by closed forms: the total force is t times the loaded area,
the moment about the origin is t times the face's first moment of area, a
unit-square face of a structured Tet4 box gives each node t * (area of its
triangles) / 3, and the mesh helpers and the gmsh reader find the faces."""
import os

import numpy as np
import pytest

import oracle
from akantu2_amd import mesh as M
from akantu2_amd.mesh_io import read_msh
from tests.conftest import ROOT

T = np.array([0.3, -1.2, 0.7])


def _oracle(mesh):
    o = oracle.OracleModel(mesh.etype, mesh.nodes, mesh.conn, "elastic")
    o.init_full()
    return o


@pytest.mark.parametrize("etype", ["_tetrahedron_4", "_hexahedron_8"])
@pytest.mark.parametrize("axis,value,area", [(0, 0.0, 3 * 2), (1, 3.0, 4 * 2),
                                             (2, 2.0, 4 * 3)])
def test_total_force_and_moment(etype, axis, value, area):
    mesh = M.box_mesh(etype, 4, 3, 2, jitter=0.1, seed=3)
    g = M.face_group(mesh, axis, value)
    assert g["facet_type"] == (M.TRI3 if mesh.etype == M.TET4 else M.QUAD4)
    o = _oracle(mesh)
    o.apply_traction(g["facet_type"], g["facets"], T)
    np.testing.assert_allclose(o.f_ext.sum(axis=0), T * area, rtol=1e-13)
    # first moment: the face centroid times the area
    lo, hi = mesh.nodes.min(axis=0), mesh.nodes.max(axis=0)
    c = 0.5 * (lo + hi)
    c[axis] = value
    np.testing.assert_allclose(np.cross(mesh.nodes, o.f_ext).sum(axis=0),
                               np.cross(c, T) * area, rtol=1e-12, atol=1e-12)
    # only the face's nodes are loaded
    off = np.ones(mesh.nb_nodes, bool)
    off[g["nodes"]] = False
    assert not o.f_ext[off].any()


def test_tet4_nodal_shares():
    mesh = M.box_mesh("_tetrahedron_4", 2, 2, 2)
    g = M.face_group(mesh, 2, 0.0)
    o = _oracle(mesh)
    o.apply_traction(M.TRI3, g["facets"], T)
    share = np.zeros(mesh.nb_nodes)
    for f in g["facets"]:
        p = mesh.nodes[f]
        a = 0.5 * np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0]))
        share[f] += a / 3
    np.testing.assert_allclose(o.f_ext, share[:, None] * T[None, :], rtol=1e-14, atol=1e-16)


def test_traction_adds_to_existing_external_force():
    mesh = M.box_mesh("_hexahedron_8", 3, 2, 2)
    g = M.face_group(mesh, 0, 3.0)
    o = _oracle(mesh)
    o.f_ext[:] = 1.0
    o.apply_traction(M.QUAD4, g["facets"], T)
    o.apply_traction(M.QUAD4, g["facets"], -T)
    np.testing.assert_allclose(o.f_ext, 1.0, rtol=0, atol=1e-15)


def test_boundary_facets_of_a_box():
    for etype, per_face in (("_tetrahedron_4", 2), ("_hexahedron_8", 1)):
        mesh = M.box_mesh(etype, 4, 3, 2)
        f = M.boundary_facets(mesh)
        assert f.shape[0] == per_face * 2 * (4 * 3 + 4 * 2 + 3 * 2)


@pytest.mark.parametrize("name", ["patch_hexahedron_8.msh", "patch_tetrahedron_4.msh"])
def test_reference_mesh_groups_carry_the_loads(name):
    mesh = read_msh(os.path.join(ROOT, "tests", "golden", "reference", name))
    groups = mesh.meta["groups"]
    assert len(groups) == 6  # the six faces of the patch-test cube
    o = _oracle(mesh)
    total_area = 0.0
    for g in groups.values():
        o.f_ext[:] = 0.0
        o.apply_traction(g["facet_type"], g["facets"], T)
        p = mesh.nodes[g["nodes"]]
        ext = p.max(axis=0) - p.min(axis=0)
        area = np.prod(np.sort(ext)[1:])  # planar rectangular face
        np.testing.assert_allclose(o.f_ext.sum(axis=0), T * area, rtol=1e-12)
        total_area += area
    lo, hi = mesh.nodes.min(axis=0), mesh.nodes.max(axis=0)
    d = hi - lo
    assert total_area == pytest.approx(2 * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]))
