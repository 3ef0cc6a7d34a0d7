"""The drop-in boundary on the GPU: ``solveStep(solver_id)`` as the reference
declares it (model_solver.hh:78, py_model.cc:130-135), and getter arrays that
behave like the reference's live ``Array`` (solid_mechanics_model.hh:440-468)
when a caller holds them across steps.  Bar: 1e-12 relative against the
oracle (the plain path is bit-equal to it in practice)."""
import numpy as np
import pytest

from akantu2_amd import _native as N
from tests.helpers import make_gpu, make_oracle, relerr, workload

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _pair(etype="_tetrahedron_4", shape=(6, 3, 3), law="elastic", seed=1):
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=seed)
    o, dt = make_oracle(mesh, blocked, v0, law)
    g, _ = make_gpu(mesh, blocked, v0, law, dt=dt)
    return g, o


def test_solve_step_takes_the_reference_solver_ids():
    """The published scalability driver calls solveStep("explicit_lumped")
    (examples/c++/scalability_test/cohesive_extrinsic.cc:267)."""
    g, o = _pair()
    print("solveStep ids: '', 'explicit_lumped'", flush=True)
    for k in range(10):
        g.solveStep("explicit_lumped" if k % 2 else "")
    o.solve_steps(10)
    assert relerr(g.getDisplacement(), o.u) <= TOL
    assert relerr(g.getVelocity(), o.v) <= TOL
    assert relerr(g.getInternalForce(), o.f_int) <= TOL
    with pytest.raises(RuntimeError, match="does not exists"):
        g.solveStep("implicit")
    with pytest.raises(TypeError):
        g.solveStep(5)
    # the C ABI refuses an unknown solver id the same way, without stepping
    before = g.getDisplacement().copy()
    with pytest.raises(N.AkantuError, match="does not exists"):
        N.check(N.lib().akx_model_solve_step(g._h, b"static"))
    g.synchronize()
    assert np.array_equal(g.getDisplacement(), before)


def test_held_displacement_write_after_a_step():
    """u = getDisplacement(); solveStep(); u[k] += d; solveStep(): the edit
    lands on the current state and nothing else reverts."""
    g, o = _pair()
    print("held u write", flush=True)
    u = g.getDisplacement()
    g.solveStep()
    o.solve_steps(1)
    bump = np.array([2e-5, -1e-5, 3e-5])
    u[7] += bump
    o.u[7] += bump
    g.solveStep()
    o.solve_steps(1)
    assert relerr(u, o.u) <= TOL          # the held array is live
    assert relerr(g.getVelocity(), o.v) <= TOL
    assert relerr(g.getInternalForce(), o.f_int) <= TOL
    assert u is g.getDisplacement()


def test_held_views_stay_live_and_writable():
    """A view of a view (``v[:, 0]``) held alone keeps the field live; its
    writes are uploaded; read-only fields refuse writes."""
    g, o = _pair("_hexahedron_8", (5, 3, 3), "neohookean", seed=2)
    print("held views", flush=True)
    vx = g.getVelocity()[:, 0]
    a = g.getAcceleration()
    f = g.getInternalForce()
    g.solveSteps(7)
    o.solve_steps(7)
    assert relerr(vx, o.v[:, 0]) <= TOL
    assert relerr(a, o.a) <= TOL
    assert relerr(f, o.f_int) <= TOL
    vx[3] = 1e-3
    o.v[3, 0] = 1e-3
    for _ in range(5):
        g.solveStep()
    o.solve_steps(5)
    assert relerr(g.getDisplacement(), o.u) <= TOL
    assert relerr(vx, o.v[:, 0]) <= TOL
    with pytest.raises(ValueError):
        f[0, 0] = 1.0
    with pytest.raises(ValueError):
        g.getMass()[0, 0] = 1.0


def test_dropped_arrays_are_not_refreshed():
    """With no getter array alive a step refreshes nothing on the host."""
    g, o = _pair()
    g.getDisplacement()[:] = 0.0   # temporary: uploaded, then forgotten
    g.solveStep()
    assert not g._mirrors
    o.solve_steps(1)
    assert relerr(g.getDisplacement(), o.u) <= TOL


def test_plastic_finite_deformation_is_refused():
    mesh, blocked, v0 = workload("_tetrahedron_4", (2, 2, 2))
    from akantu2_amd import SolidMechanicsModel
    g = SolidMechanicsModel(mesh)
    with pytest.raises(N.AkantuError, match="finite_deformation") as ei:
        g.addMaterial("plastic_linear_isotropic_hardening", E=1.0, nu=0.3, sigma_y=1e-3,
                      h=0.01, finite_deformation=True)
    assert ei.value.code == -5
