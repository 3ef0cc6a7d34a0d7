"""The Material::computeStress plugin on the GPU path (AKX_LAW_USER).

User materials are written as in the reference's Python example
(examples/python/solid_mechanics_model/custom-material/custom-material.py):
a subclass of Material overriding computeStress(el_type, ghost_type),
registered with MaterialFactory.getInstance().registerAllocator.

- A user linear-elastic law written in numpy in the same operation order as
  the built-in one gives the built-in trajectory bit for bit (Tet4, Hex8); a
  user Saint-Venant-Kirchhoff law (finite deformation, PK2) the built-in
  finite-strain elastic one.
- The user stress is what the force pass uses: twice the elastic stress
  equals the built-in law with twice E at the same time step.
- getGradU / getStress outside a step run the law once and return what it
  wrote; an exception raised in computeStress propagates from solveStep.
"""
import numpy as np
import pytest

import akantu2_amd as aka
from tests.helpers import make_gpu, relerr, workload

pytestmark = pytest.mark.gpu


class LocalElastic(aka.Material):
    """sigma = lambda tr(eps) I + 2 mu eps, in the reference's operation
    order (material_elastic_inline_impl.hh:33-49)."""

    def __init__(self, model, _id):
        super().__init__(model, _id)
        self.registerParamReal("E", aka._pat_readable | aka._pat_parsable, "Young's modulus")
        self.registerParamReal("nu", aka._pat_readable | aka._pat_parsable, "Poisson ratio")
        self.scale = 1.0
        self.calls = 0

    def initMaterial(self):
        E, nu = self.getReal("E"), self.getReal("nu")
        self.lame_lambda = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu))
        self.lame_mu = E / (2.0 * (1.0 + nu))

    def computeStress(self, el_type, ghost_type):
        self.calls += 1
        g = self.getGradU(el_type, ghost_type).reshape(-1, 3, 3)  # [q][j][i] = du_i/dX_j
        sigma = self.getStress(el_type, ghost_type).reshape(-1, 3, 3)
        tr = (g[:, 0, 0] + g[:, 1, 1]) + g[:, 2, 2]
        two_mu = 2.0 * self.lame_mu
        eps = 0.5 * (g + g.transpose(0, 2, 1))
        s = two_mu * eps
        for i in range(3):
            s[:, i, i] = two_mu * eps[:, i, i] + self.lame_lambda * tr
        sigma[:] = self.scale * s


class LocalSVK(LocalElastic):
    """Saint-Venant-Kirchhoff: S = lambda tr(E) I + 2 mu E, E = (g + g^T +
    g^T g) / 2 (the reference's finite-deformation elastic law)."""

    def computeStress(self, el_type, ghost_type):
        self.calls += 1
        gt = self.getGradU(el_type, ghost_type).reshape(-1, 3, 3)  # g transposed
        g = gt.transpose(0, 2, 1)                                  # g[q][i][j]
        S = self.getStress(el_type, ghost_type).reshape(-1, 3, 3)
        E = np.empty_like(g)
        for i in range(3):
            for j in range(3):
                p = g[:, 0, i] * g[:, 0, j]
                p = p + g[:, 1, i] * g[:, 1, j]
                p = p + g[:, 2, i] * g[:, 2, j]
                E[:, i, j] = (p + g[:, j, i] + g[:, i, j]) / 2.0
        tr = (E[:, 0, 0] + E[:, 1, 1]) + E[:, 2, 2]
        two_mu = 2.0 * self.lame_mu
        s = np.empty_like(E)
        for i in range(3):
            for j in range(3):
                eps = 0.5 * (E[:, j, i] + E[:, i, j])
                s[:, i, j] = two_mu * eps + self.lame_lambda * tr if i == j else two_mu * eps
        S[:] = s.transpose(0, 2, 1)  # back to column-major [q][j][i]


class Failing(LocalElastic):
    def computeStress(self, el_type, ghost_type):
        raise ValueError("law failed on purpose")


FACTORY = aka.MaterialFactory.getInstance()
FACTORY.registerAllocator("local_elastic", lambda dim, unused, model, _id: LocalElastic(model, _id))
FACTORY.registerAllocator("local_svk", lambda dim, unused, model, _id: LocalSVK(model, _id))
FACTORY.registerAllocator("failing", lambda dim, unused, model, _id: Failing(model, _id))


def _user_model(mesh, blocked, v0, name, dt, finite_deformation=False, **mat):
    g = aka.SolidMechanicsModel(mesh)
    law = g.addMaterial(name, finite_deformation=finite_deformation, **mat)
    g.initFull(_analysis_method=aka._explicit_lumped_mass)
    g.getBlockedDOFs()[:] = blocked
    g.getVelocity()[:] = v0
    g.setTimeStep(dt)
    return g, law


@pytest.mark.parametrize("etype,shape", [("_tetrahedron_4", (8, 3, 3)),
                                         ("_hexahedron_8", (6, 3, 2))])
def test_user_elastic_equals_builtin(etype, shape):
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=31)
    mat = dict(rho=1.0, E=1.3, nu=0.27)
    ref, dt = make_gpu(mesh, blocked, v0, "elastic", **mat)
    usr, law = _user_model(mesh, blocked, v0, "local_elastic", dt, **mat)
    assert usr.getStableTimeStep() == ref.getStableTimeStep()
    ref.solveSteps(200)
    usr.solveSteps(200)
    assert law.calls == 200
    for f in ("getDisplacement", "getVelocity", "getInternalForce"):
        a, b = getattr(usr, f)(), getattr(ref, f)()
        print(etype, f, relerr(a, b), flush=True)
        assert np.array_equal(a, b), f


def test_user_svk_equals_builtin_finite_strain():
    mesh, blocked, v0 = workload("_hexahedron_8", (5, 3, 3), jitter=0.1, seed=32, amp=2e-2)
    mat = dict(rho=1.0, E=0.9, nu=0.31)
    ref, dt = make_gpu(mesh, blocked, v0, "elastic_fd", **mat)
    usr, law = _user_model(mesh, blocked, v0, "local_svk", dt, finite_deformation=True, **mat)
    ref.solveSteps(150)
    usr.solveSteps(150)
    eu = relerr(usr.getDisplacement(), ref.getDisplacement())
    ef = relerr(usr.getInternalForce(), ref.getInternalForce())
    print("svk", eu, ef, flush=True)
    assert eu <= 1e-12 and ef <= 1e-12


def test_user_stress_drives_the_forces():
    mesh, blocked, v0 = workload("_tetrahedron_4", (8, 3, 3), jitter=0.1, seed=33)
    ref, dt = make_gpu(mesh, blocked, v0, "elastic", rho=1.0, E=2.0, nu=0.3)
    usr, law = _user_model(mesh, blocked, v0, "local_elastic", dt, rho=1.0, E=1.0, nu=0.3)
    law.scale = 2.0
    ref.solveSteps(120)
    usr.solveSteps(120)
    eu = relerr(usr.getDisplacement(), ref.getDisplacement())
    print("scaled", eu, flush=True)
    assert eu <= 1e-12


def test_internals_and_errors():
    mesh, blocked, v0 = workload("_hexahedron_8", (4, 2, 2), jitter=0.1, seed=34)
    mat = dict(rho=1.0, E=1.1, nu=0.25)
    ref, dt = make_gpu(mesh, blocked, v0, "elastic", **mat)
    usr, law = _user_model(mesh, blocked, v0, "local_elastic", dt, **mat)
    ref.solveSteps(30)
    usr.solveSteps(30)
    calls = law.calls
    assert np.array_equal(usr.getMaterial(0).getGradU(), ref.getMaterial(0).getGradU())
    assert np.array_equal(usr.getMaterial(0).getStress(), ref.getMaterial(0).getStress())
    assert law.calls == calls + 1  # computeAllStresses ran the law once
    bad, _ = _user_model(mesh, blocked, v0, "failing", dt, **mat)
    with pytest.raises(ValueError, match="on purpose"):
        bad.solveSteps(1)
