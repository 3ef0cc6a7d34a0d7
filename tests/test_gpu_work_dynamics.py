"""The reference's explicit work / energy test on the GPU path
(test/test_model/test_solid_mechanics_model/test_energies/
test_solid_mechanics_model_work_dynamics.cc, WorkExplicit): a 10 x 1 x 1 bar,
elastic E = 1, nu = 0, rho = 1 (test_solid_mechanics_model_work_material.dat),
a traction of 0.5 along x on the x = max face (BC::Neumann::FromTraction),
the x = min face held at u_x = 0, 0.01 and -0.01 in turn
(BC::Dirichlet::FixedValue), dt = 0.1 dt_crit.  For each value: 25 steps to
settle, then 200 steps summing getEnergy("external work") after each; the
work must equal the change of kinetic + potential energy within 5e-2, the
reference's bound.

The bar meshes are gmsh-generated from bar_*.geo in the reference and are
not in it; structured boxes of the same size and spacing stand in.  The
oracle runs the same sequence and the GPU path must equal it (u to 1e-12,
energies to 1e-12 relative).
"""
import numpy as np
import pytest

import oracle
from akantu2_amd import bc as BC
from akantu2_amd import mesh as M
from tests.helpers import relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("etype", ["_hexahedron_8", "_tetrahedron_4"])
def test_work_equals_energy_change(etype):
    from akantu2_amd import SolidMechanicsModel
    mesh = M.box_mesh(etype, 17, 2, 2, h=(10.0 / 17, 0.5, 0.5))
    M.add_group(mesh, "el_apply_force", M.face_group(mesh, 0, 10.0, tol=1e-6))
    M.add_group(mesh, "el_fixed", M.face_group(mesh, 0, 0.0, tol=1e-6))
    groups = mesh.meta["groups"]
    g = SolidMechanicsModel(mesh)
    g.addMaterial("elastic", rho=1.0, E=1.0, nu=0.0)
    g.initFull()
    o = oracle.OracleModel(mesh.etype, mesh.nodes, mesh.conn, "elastic", rho=1.0, E=1.0,
                           nu=0.0)
    o.init_full()
    t = np.array([0.5, 0.0, 0.0])
    g.applyBC(BC.Neumann.FromTraction(t), "el_apply_force")
    ff = groups["el_apply_force"]
    o.apply_traction(ff["facet_type"], ff["facets"], t)
    dt = 0.1 * g.getStableTimeStep()
    assert 0.1 * o.stable_dt() == dt
    g.setTimeStep(dt)
    o.set_time_step(dt)
    fixed = groups["el_fixed"]["nodes"]
    for u in (0.0, 0.01, -0.01):
        g.applyBC(BC.Dirichlet.FixedValue(u, 0), "el_fixed")
        o.blocked[fixed, 0] = 1
        o.u[fixed, 0] = u
        for _ in range(25):
            g.solveStep()
        o.solve_steps(25)
        e0 = g.getEnergy("potential") + g.getEnergy("kinetic")
        e0_o = o.energy("potential") + o.energy("kinetic")
        W = W_o = 0.0
        for _ in range(200):
            g.solveStep()
            o.solve_steps(1)
            W += g.getEnergy("external work")
            W_o += o.energy("external work")
        de = g.getEnergy("potential") + g.getEnergy("kinetic") - e0
        de_o = o.energy("potential") + o.energy("kinetic") - e0_o
        print(etype, u, "W", W, "dE", de, "oracle", W_o, de_o, flush=True)
        assert abs(W - de) <= 5e-2
        assert abs(W) > 1e-3  # the traction does work
        assert relerr(g.getDisplacement(), o.u) <= 1e-12
        assert abs(W - W_o) <= 1e-12 * abs(W_o) and abs(de - de_o) <= 1e-12 * abs(W_o)
