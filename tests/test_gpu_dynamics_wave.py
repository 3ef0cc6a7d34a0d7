"""The reference's explicit dynamics acceptance test on the GPU path
(test/test_model/test_solid_mechanics_model/test_solid_mechanics_model_dynamics.cc,
TestSMMFixtureBarExplicit): a plane P wave and two S waves along x in a
10 x 1 x 1 bar, elastic E = 1, nu = 0, rho = 1
(test_solid_mechanics_model_dynamics_material.dat), dt = dt_crit / 10, for
5 time units.  Every step, before solveStep, the boundary nodes get the
analytical displacement and velocity and their x component is blocked
(SolutionFunctor: flags(0) = true); the error is
sqrt(sum_nodes |u - u_exact|^2) / nb_nodes, and the test requires its
maximum over the run to stay below 2e-3.  The analytical wave speeds use
nu = 0.3, as the reference test's constants do.

The bar meshes are gmsh-generated from bar_*.geo (h = 0.6, transfinite for
the hexahedra) and are not in the reference; structured boxes of the same
size and spacing stand in (17 x 2 x 2 cells, Hex8 or 6 Tet4 per cell).  The
oracle runs the same loop and the GPU path must equal it to 1e-12.
"""
import numpy as np
import pytest

import oracle
from akantu2_amd import mesh as M
from tests.helpers import relerr

pytestmark = pytest.mark.gpu

A = 1e-1
E_T, NU_T, RHO = 1.0, 0.3, 1.0
LAM_T = E_T * NU_T / (1 + NU_T) / (1 - 2 * NU_T)
MU_T = E_T / 2 / (1 + NU_T)
CP, CS = np.sqrt((LAM_T + 2 * MU_T) / RHO), np.sqrt(MU_T / RHO)
K = np.array([0.5, 0.0, 0.0])
KS1 = np.cross(K, [0.0, 0.0, 1.0])
KS2 = np.cross(K, [0.0, 1.0, 0.0])
KN = np.linalg.norm(K)


def exact(X, t):
    """Verification<3>::displacement / velocity of the reference test."""
    pp = X @ K - KN * CP * t
    ps = X @ K - KN * CS * t
    u = A * (np.outer(np.cos(pp), K) + np.outer(np.cos(ps), KS1 + KS2))
    v = A * (np.outer(KN * CP * np.cos(pp), K) + np.outer(KN * CS * np.cos(ps), KS1 + KS2))
    return u, v


@pytest.mark.parametrize("etype", ["_hexahedron_8", "_tetrahedron_4"])
def test_bar_plane_waves(etype):
    from akantu2_amd import SolidMechanicsModel
    mesh = M.box_mesh(etype, 17, 2, 2, h=(10.0 / 17, 0.5, 0.5))
    X = mesh.nodes
    bnd = ((X[:, 0] <= 1e-12) | (X[:, 0] >= 10 - 1e-12) | (X[:, 1] <= 1e-12) |
           (X[:, 1] >= 1 - 1e-12) | (X[:, 2] <= 1e-12) | (X[:, 2] >= 1 - 1e-12))
    g = SolidMechanicsModel(mesh)
    g.addMaterial("elastic", rho=RHO, E=1.0, nu=0.0)
    g.initFull()
    dt = g.getStableTimeStep() / 10.0
    g.setTimeStep(dt)
    o = oracle.OracleModel(mesh.etype, X, mesh.conn, "elastic", rho=RHO, E=1.0, nu=0.0)
    o.init_full()
    assert o.stable_dt() / 10.0 == dt
    o.set_time_step(dt)
    u0, v0 = exact(X, 0.0)
    g.getDisplacement()[:] = u0
    g.getVelocity()[:] = v0
    o.u[:] = u0
    o.v[:] = v0
    blocked = np.zeros((mesh.nb_nodes, 3), np.uint8)
    blocked[bnd, 0] = 1
    g.getBlockedDOFs()[:] = blocked
    o.blocked[:] = blocked
    max_steps = int(5.0 / dt)
    max_error, t = -1.0, 0.0
    for _ in range(max_steps):
        ub, vb = exact(X[bnd], t)
        g.getDisplacement()[bnd] = ub
        g.getVelocity()[bnd] = vb
        o.u[bnd] = ub
        o.v[bnd] = vb
        ue, _ = exact(X, t)
        err = np.sqrt(np.sum((g.getDisplacement() - ue) ** 2)) / mesh.nb_nodes
        max_error = max(max_error, err)
        g.solveStep()
        o.solve_steps(1)
        t += dt
    eu = relerr(g.getDisplacement(), o.u)
    ev = relerr(g.getVelocity(), o.v)
    print(etype, "steps", max_steps, "max error", max_error, "vs oracle", eu, ev, flush=True)
    assert max_steps >= 100
    assert max_error <= 2e-3
    assert eu <= 1e-12 and ev <= 1e-12
