"""BC::Neumann::FromTraction and the Dirichlet functors through the model's
applyBC on the GPU, against the oracle: the assembled external force is
bit-identical, the run after it matches to 1e-12."""
import numpy as np
import pytest

import oracle
from akantu2_amd import BC, SolidMechanicsModel
from akantu2_amd import mesh as M
from tests.helpers import relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("etype,law", [("_tetrahedron_4", "elastic"),
                                       ("_hexahedron_8", "neohookean")])
def test_traction_loaded_bar(etype, law):
    mesh = M.box_mesh(etype, 8, 3, 3, jitter=0.1, seed=8)
    M.add_group(mesh, "left", M.face_group(mesh, 0, 0.0))
    M.add_group(mesh, "right", M.face_group(mesh, 0, 8.0))
    t = np.array([2e-4, -1e-4, 5e-5])
    g = SolidMechanicsModel(mesh)
    g.addMaterial(law, rho=1.0, E=1.0, nu=0.3)
    g.initFull()
    for ax in (BC._x, BC._y, BC._z):
        g.applyBC(BC.Dirichlet.FixedValue(0.0, ax), "left")
    g.applyBC(BC.Neumann.FromTraction(t), "right")
    dt = 0.5 * g.getStableTimeStep()
    g.setTimeStep(dt)

    o = oracle.OracleModel(mesh.etype, mesh.nodes, mesh.conn, law)
    o.init_full()
    o.blocked[M.face_group(mesh, 0, 0.0)["nodes"]] = 1
    right = M.face_group(mesh, 0, 8.0)
    o.apply_traction(right["facet_type"], right["facets"], t)
    o.set_time_step(dt)
    assert np.array_equal(g.getExternalForce(), o.f_ext)
    assert np.array_equal(g.getBlockedDOFs(), o.blocked)
    for _ in range(3):
        g.solveStep(100)
        o.solve_steps(100)
        eu = relerr(g.getDisplacement(), o.u)
        ef = relerr(g.getInternalForce(), o.f_int)
        print(f"{etype} {law} u={eu:.2e} f_int={ef:.2e}", flush=True)
        assert eu <= 1e-12 and ef <= 1e-12
    assert np.abs(o.u).max() > 0
    assert g.getEnergy("external work") == pytest.approx(o.energy("external work"),
                                                         rel=1e-12)
