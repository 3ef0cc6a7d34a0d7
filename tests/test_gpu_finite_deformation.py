"""The reference's finite-deformation objectivity test on the GPU path
(test/test_model/test_solid_mechanics_model/test_materials/
test_finite_deformation.cc, TestFiniteDeformation.NotUnit): one Tet4 (the
unit tetrahedron of 1_tetrahedron.msh), elastic with finite_deformation =
true, rho = 3000, E = 40e9, nu = 0.3 (material_finite_deformation.dat).  A
linear displacement u0 = a0 + A X is imposed and the internal force f0
formed; the deformed body is then rotated by 45 degrees about x,
u1 = R (X + u0) - X, and the internal force must rotate with it:
|f1 - R f0| <= 1e-5 at every node (forces of order 1e9).  The reference
runs a static step with every DOF blocked; here
SolidMechanicsModel::assembleInternalForces forms the same forces.  The same
holds for the neo-Hookean law, and both equal the oracle to 1e-12.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import relerr

pytestmark = pytest.mark.gpu

X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
CONN = np.array([[0, 1, 2, 3]], dtype=np.int32)
ALPHA = np.array([[0.00, 0.02, 0.03, 0.04], [0.00, 0.06, 0.07, 0.08],
                  [0.00, 0.10, 0.11, 0.12]])
TH = np.pi / 4
R = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(TH), -np.sin(TH)], [0.0, np.sin(TH), np.cos(TH)]])


@pytest.mark.parametrize("law", ["elastic_fd", "neohookean"])
def test_internal_force_rotates_with_the_body(law):
    from akantu2_amd import SolidMechanicsModel
    from akantu2_amd import mesh as M
    mesh = M.BoxMesh(M.TET4, X.copy(), CONN.copy(), (1, 1, 1), (1.0, 1.0, 1.0), {})
    g = SolidMechanicsModel(mesh)
    g.addMaterial(law, rho=3000.0, E=40e9, nu=0.3)
    g.initFull()
    o = oracle.OracleModel(oracle.TET4, X, CONN, law, rho=3000.0, E=40e9, nu=0.3)
    o.init_full()
    g.getBlockedDOFs()[:] = 1
    u0 = ALPHA[:, 0] + X @ ALPHA[:, 1:].T
    g.getDisplacement()[:] = u0
    g.assembleInternalForces()
    f0 = g.getInternalForce().copy()
    o.u[:] = u0
    o.internal_forces()
    assert relerr(f0, o.f_int) <= 1e-12
    u1 = (X + u0) @ R.T - X
    g.getDisplacement()[:] = u1
    g.assembleInternalForces()
    f1 = g.getInternalForce().copy()
    o.u[:] = u1
    o.internal_forces()
    assert relerr(f1, o.f_int) <= 1e-12
    err = np.linalg.norm(f1 - f0 @ R.T, axis=1)
    print(law, "max |f0|", np.abs(f0).max(), "max |f1 - R f0|", err.max(), flush=True)
    assert np.abs(f0).max() > 1e8
    assert err.max() <= 1e-5
