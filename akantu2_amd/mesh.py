"""Structured box meshes and seeded initial conditions for the explicit
benchmarks.

Element conventions are Akantu's:
- ``_tetrahedron_4``: nodes (n0, n1, n2, n3) with natural coordinates
  n0 = 0, n1 = eta, n2 = zeta, n3 = xi
  (element_class_tetrahedron_4_inline_impl.hh:61-77).
- ``_hexahedron_8``: the gmsh/Akantu ordering with nodes 0-3 on zeta = -1 and
  4-7 on zeta = +1 (element_class_hexahedron_8_inline_impl.hh:20-40).

Every element is positively oriented (det J > 0), so the reference's
``IntegratorGauss::checkJacobians`` (integrator_gauss_inline_impl.hh:176-195)
would accept the mesh.

Random numbers come from numpy's PCG64 ``default_rng(seed)``; the seed is
part of every workload description.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

TET4 = 0
HEX8 = 1
TRI3 = 2   # _triangle_3, boundary facet of Tet4
QUAD4 = 3  # _quadrangle_4, boundary facet of Hex8
ELEMENT_TYPES = {"_tetrahedron_4": TET4, "_hexahedron_8": HEX8}
NNPE = {TET4: 4, HEX8: 8}
NQP = {TET4: 1, HEX8: 8}


@dataclass
class BoxMesh:
    """Nodes [nn][3] (float64) and connectivity [nel][nnpe] (int32)."""

    etype: int
    nodes: np.ndarray
    conn: np.ndarray
    shape: tuple  # cells per direction (nx, ny, nz)
    h: tuple = (1.0, 1.0, 1.0)
    meta: dict = field(default_factory=dict)

    @property
    def nb_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def nb_elements(self) -> int:
        return int(self.conn.shape[0])

    @property
    def nqp(self) -> int:
        return NQP[self.etype]

    @property
    def length(self) -> float:
        return self.shape[0] * self.h[0]


def _grid_nodes(nx, ny, nz, h):
    i = np.arange(nx + 1, dtype=np.float64) * h[0]
    j = np.arange(ny + 1, dtype=np.float64) * h[1]
    k = np.arange(nz + 1, dtype=np.float64) * h[2]
    # node id = i + (nx+1) * (j + (ny+1) * k): x runs fastest
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    return np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)


def _cell_corners(nx, ny, nz):
    """Node ids of the 8 corners (bit order x, y, z) of every cell, cells
    numbered x fastest."""
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij")
    ci, cj, ck = (a.transpose(2, 1, 0).ravel() for a in (ci, cj, ck))
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    corners = np.empty((ci.size, 8), dtype=np.int64)
    for b in range(8):
        bx, by, bz = b & 1, (b >> 1) & 1, (b >> 2) & 1
        corners[:, b] = (ci + bx) + sx * (cj + by) + sy * (ck + bz)
    return corners


def _tet_patterns():
    """Kuhn split of the unit cube along the (0,0,0)-(1,1,1) diagonal into six
    tetrahedra, each oriented so that det J > 0 in Akantu's Tet4 convention."""
    pats = []
    for perm in itertools.permutations(range(3)):
        b0 = 0
        b1 = 1 << perm[0]
        b2 = b1 | (1 << perm[1])
        b3 = 7
        pts = [np.array([(b >> d) & 1 for d in range(3)], float)
               for b in (b0, b1, b2, b3)]
        # J rows: X3 - X0 (xi), X1 - X0 (eta), X2 - X0 (zeta)
        det = np.linalg.det(np.stack([pts[3] - pts[0], pts[1] - pts[0],
                                      pts[2] - pts[0]]))
        pat = [b0, b1, b2, b3] if det > 0 else [b0, b2, b1, b3]
        pats.append(pat)
    return np.array(pats, dtype=np.int64)


HEX8_CORNER_ORDER = np.array([0, 1, 3, 2, 4, 5, 7, 6], dtype=np.int64)


def box_mesh(etype, nx, ny, nz, h=(1.0, 1.0, 1.0), jitter=0.0, seed=0):
    """Structured box of nx*ny*nz cells (6 Tet4 or 1 Hex8 per cell).

    ``jitter`` moves every interior node by U(-jitter, jitter) * h per
    coordinate (seeded), boundary nodes stay on the box faces."""
    if isinstance(etype, str):
        etype = ELEMENT_TYPES[etype]
    h = tuple(float(x) for x in (h if np.ndim(h) else (h, h, h)))
    nodes = _grid_nodes(nx, ny, nz, h)
    corners = _cell_corners(nx, ny, nz)
    if etype == TET4:
        pats = _tet_patterns()
        conn = corners[:, pats].reshape(-1, 4)
    else:
        conn = corners[:, HEX8_CORNER_ORDER]
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, size=nodes.shape) * np.asarray(h)
        i = np.arange(nodes.shape[0])
        ix = i % (nx + 1)
        iy = (i // (nx + 1)) % (ny + 1)
        iz = i // ((nx + 1) * (ny + 1))
        interior = ((ix > 0) & (ix < nx) & (iy > 0) & (iy < ny) & (iz > 0)
                    & (iz < nz))
        nodes[interior] += d[interior]
    if conn.max(initial=0) >= 2**31 - 1:
        raise ValueError("mesh too large for int32 connectivity")
    return BoxMesh(etype, np.ascontiguousarray(nodes),
                   np.ascontiguousarray(conn.astype(np.int32)), (nx, ny, nz), h,
                   {"jitter": jitter, "seed": seed})


def cells_for(etype, nb_elements, aspect=1.0):
    """Cells (nx, ny, nz) of a box with about ``nb_elements`` elements and
    x : y : z = aspect : 1 : 1."""
    if isinstance(etype, str):
        etype = ELEMENT_TYPES[etype]
    cells = nb_elements / (6 if etype == TET4 else 1)
    ny = max(1, int(round((cells / aspect) ** (1.0 / 3.0))))
    nx = max(1, int(round(cells / (ny * ny))))
    return nx, ny, ny


# local faces of each element type, each in a cyclic order (facet node order)
_FACES = {TET4: [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)],
          HEX8: [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5),
                 (2, 3, 7, 6), (3, 0, 4, 7)]}


def boundary_facets(mesh: BoxMesh) -> np.ndarray:
    """Faces that belong to exactly one element, [n][3] (Tet4) or [n][4]
    (Hex8), in the element's local cyclic order."""
    faces = np.concatenate([mesh.conn[:, list(f)] for f in _FACES[mesh.etype]])
    key = np.sort(faces, axis=1)
    _, idx, counts = np.unique(key, axis=0, return_index=True, return_counts=True)
    return np.ascontiguousarray(faces[np.sort(idx[counts == 1])], dtype=np.int32)


def add_group(mesh: BoxMesh, name, group):
    """Register an element / node group (Mesh::createElementGroup)."""
    mesh.meta.setdefault("groups", {})[name] = group


def face_group(mesh: BoxMesh, axis: int, value: float, tol: float = 1e-12) -> dict:
    """The boundary facets lying on the plane X[axis] = value, as a group
    {"facet_type", "facets", "nodes"}."""
    f = boundary_facets(mesh)
    on = np.all(np.abs(mesh.nodes[f, axis] - value) <= tol, axis=1)
    f = f[on]
    return {"facet_type": TRI3 if mesh.etype == TET4 else QUAD4, "facets": f,
            "nodes": np.unique(f)}


def clamped_face_x0(mesh: BoxMesh) -> np.ndarray:
    """Blocked DOFs [nn][3] (uint8): the x = 0 face clamped in x, y and z."""
    blocked = np.zeros((mesh.nb_nodes, 3), dtype=np.uint8)
    blocked[mesh.nodes[:, 0] == 0.0] = 1
    return blocked


def initial_velocity(mesh: BoxMesh, amplitude=1e-3, seed=0) -> np.ndarray:
    """v0(node) = amplitude * sin(pi x / L) * r_node, r_node ~ U(-1, 1)^3 drawn
    in node order; zero on the clamped x = 0 face."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(-1.0, 1.0, size=(mesh.nb_nodes, 3))
    env = np.sin(np.pi * mesh.nodes[:, 0] / mesh.length)
    return amplitude * env[:, None] * r
