"""Parity of the gfx950 plain FP64 path with the CPU oracle (same seeded
inputs).  Bar: 1e-12 relative (max norm) on u, v, a and f_int after N steps,
the tolerance the explicit path is held to.  The plain path keeps the
reference's operation order and in practice reproduces the oracle bit for
bit, except for the last bit of log() in the neo-Hookean law."""
import os

import numpy as np
import pytest

import oracle
from akantu2_amd import mesh as M
from tests.helpers import make_gpu, make_oracle, relerr, workload

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _compare(g, o, tag=""):
    eu = relerr(g.getDisplacement(), o.u)
    ev = relerr(g.getVelocity(), o.v)
    ef = relerr(g.getInternalForce(), o.f_int)
    ea = relerr(g.getAcceleration(), o.a)
    print(f"{tag} rel err u={eu:.3e} v={ev:.3e} f_int={ef:.3e} a={ea:.3e}", flush=True)
    assert eu <= TOL and ev <= TOL and ef <= TOL and ea <= TOL
    return eu, ev, ef


PRECISIONS = ["fp64"]


@pytest.mark.parametrize("precision", PRECISIONS)
def test_tet4_bar_plain_fp64_1000_steps(precision):
    """The 20x4x4 Kuhn Tet4 bar, E=1 nu=0.3 rho=1, clamped x=0, seeded
    sin-enveloped initial velocity, dt = 0.5 dt_crit, 1000 steps."""
    mesh, blocked, v0 = workload("_tetrahedron_4", (20, 4, 4), seed=0)
    o, dt = make_oracle(mesh, blocked, v0, "elastic")
    g, dtg = make_gpu(mesh, blocked, v0, "elastic", precision=precision)
    assert dtg == dt  # same stable time step to the bit
    np.testing.assert_array_equal(g.getMass(), o.mass)
    for k in range(10):
        o.solve_steps(100)
        g.solveSteps(100)
        _compare(g, o, f"Tet4 bar step {(k + 1) * 100}")
    if precision == "fp64":  # bit for bit
        assert np.array_equal(g.getDisplacement(), o.u)
        assert np.array_equal(g.getInternalForce(), o.f_int)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("etype,law", [
    ("_hexahedron_8", "neohookean"), ("_hexahedron_8", "elastic"),
    ("_hexahedron_8", "elastic_fd"), ("_tetrahedron_4", "neohookean"),
    ("_tetrahedron_4", "elastic_fd")])
def test_laws_and_elements_jittered(etype, law, precision):
    mesh, blocked, v0 = workload(etype, (8, 4, 3), jitter=0.1, seed=3)
    o, dt = make_oracle(mesh, blocked, v0, law, E=1.0, nu=0.3)
    g, dtg = make_gpu(mesh, blocked, v0, law, E=1.0, nu=0.3, precision=precision)
    assert dtg == dt
    for k in range(3):
        o.solve_steps(100)
        g.solveSteps(100)
        _compare(g, o, f"{etype} {law} step {(k + 1) * 100}")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_hex8_box_reduced_checked_every_100_steps(precision):
    """The jittered Hex8 neo-Hookean box at reduced size, checked against
    the oracle at every 100th step over 1000 steps."""
    mesh, blocked, v0 = workload("_hexahedron_8", (20, 10, 10), jitter=0.1, seed=1)
    o, dt = make_oracle(mesh, blocked, v0, "neohookean")
    g, _ = make_gpu(mesh, blocked, v0, "neohookean", dt=dt, precision=precision)
    for k in range(10):
        o.solve_steps(100)
        g.solveSteps(100)
        _compare(g, o, f"reduced Hex8 box step {(k + 1) * 100}")


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("etype", ["_tetrahedron_4", "_hexahedron_8"])
def test_plastic_linear_isotropic_hardening(etype, precision):
    mesh, blocked, v0 = workload(etype, (10, 3, 3), jitter=0.05, seed=5, amp=3e-2)
    kw = dict(E=1.0, nu=0.3, sigma_y=2e-3, h=0.05)
    o, dt = make_oracle(mesh, blocked, v0, "plastic_linear_isotropic_hardening", **kw)
    g, _ = make_gpu(mesh, blocked, v0, "plastic_linear_isotropic_hardening", dt=dt,
                    precision=precision, **kw)
    for k in range(3):
        o.solve_steps(100)
        g.solveSteps(100)
        _compare(g, o, f"plastic step {(k + 1) * 100}")
    assert o.iso.max() > 0  # the run does yield
    mat = g.getMaterial(0)
    assert relerr(mat.getIsoHardening(), o.iso) <= TOL
    assert relerr(mat.getStress(), o.stress) <= TOL
    assert relerr(mat.getInelasticStrain(), o.eps_p) <= TOL


@pytest.mark.parametrize("precision", PRECISIONS)
def test_per_element_parameters(precision):
    """The 60 M-element Tet4 box's per-element E (log-uniform [0.5, 2]) and nu (U[0.2, 0.35])."""
    mesh, blocked, v0 = workload("_tetrahedron_4", (10, 4, 4), jitter=0.1, seed=7)
    rng = np.random.default_rng(7)
    E = np.exp(rng.uniform(np.log(0.5), np.log(2.0), mesh.nb_elements))
    nu = rng.uniform(0.2, 0.35, mesh.nb_elements)
    o, dt = make_oracle(mesh, blocked, v0, "elastic", E_el=E, nu_el=nu)
    g, dtg = make_gpu(mesh, blocked, v0, "elastic", E_el=E, nu_el=nu,
                      precision=precision)
    assert dtg == dt
    o.solve_steps(200)
    g.solveSteps(200)
    _compare(g, o, "per-element params")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_stresses_and_gradients_match(precision):
    mesh, blocked, v0 = workload("_hexahedron_8", (6, 3, 3), jitter=0.1, seed=2)
    o, dt = make_oracle(mesh, blocked, v0, "neohookean")
    g, _ = make_gpu(mesh, blocked, v0, "neohookean", dt=dt, precision=precision)
    o.solve_steps(50)
    g.solveSteps(50)
    mat = g.getMaterial(0)
    assert relerr(mat.getGradU(), o.grad_u) <= TOL
    assert relerr(mat.getStress(), o.stress) <= TOL


def test_bit_reproducible_runs():
    mesh, blocked, v0 = workload("_hexahedron_8", (12, 6, 6), jitter=0.1, seed=4)
    out = []
    for _ in range(2):
        g, _ = make_gpu(mesh, blocked, v0, "neohookean")
        g.solveSteps(150)
        out.append((g.getDisplacement().copy(), g.getInternalForce().copy()))
    assert out[0][0].tobytes() == out[1][0].tobytes()
    assert out[0][1].tobytes() == out[1][1].tobytes()


def test_step_batching_is_invisible():
    """solveStep(n) (fused predictor) equals n single solveStep() calls."""
    mesh, blocked, v0 = workload("_tetrahedron_4", (6, 3, 3), jitter=0.1, seed=9)
    g1, dt = make_gpu(mesh, blocked, v0, "elastic")
    g2, _ = make_gpu(mesh, blocked, v0, "elastic", dt=dt)
    g1.solveSteps(37)
    for _ in range(37):
        g2.solveStep()
    assert np.array_equal(g1.getDisplacement(), g2.getDisplacement())
    assert np.array_equal(g1.getVelocity(), g2.getVelocity())
    # f_int is stored by the last step of a call only; both hold the same
    assert np.array_equal(g1.getInternalForce(), g2.getInternalForce())


def test_user_writes_through_getters():
    """Editing getDisplacement() between steps is seen by the next step, as
    in the reference where the getter returns the model's own Array."""
    mesh, blocked, v0 = workload("_tetrahedron_4", (6, 3, 3), seed=1)
    o, dt = make_oracle(mesh, blocked, v0, "elastic")
    g, _ = make_gpu(mesh, blocked, v0, "elastic", dt=dt)
    o.solve_steps(10)
    g.solveSteps(10)
    bump = np.zeros_like(o.u)
    bump[5] = [1e-4, -2e-4, 3e-4]
    o.u += bump
    g.getDisplacement()[:] += bump
    o.solve_steps(10)
    g.solveSteps(10)
    _compare(g, o, "after user write")


def test_external_force_and_assemble_internal_forces():
    mesh, blocked, v0 = workload("_hexahedron_8", (5, 2, 2), seed=3)
    o, dt = make_oracle(mesh, blocked, v0, "elastic")
    g, _ = make_gpu(mesh, blocked, v0, "elastic", dt=dt)
    fext = np.zeros_like(o.u)
    fext[mesh.nodes[:, 0] == mesh.nodes[:, 0].max(), 0] = 1e-4
    o.f_ext[:] = fext
    g.getExternalForce()[:] = fext
    o.solve_steps(40)
    g.solveSteps(40)
    _compare(g, o, "f_ext")
    rng = np.random.default_rng(0)
    o.u[:] = rng.uniform(-1e-3, 1e-3, o.u.shape)
    g.getDisplacement()[:] = o.u
    o.internal_forces()
    g.assembleInternalForces()
    assert relerr(g.getInternalForce(), o.f_int) <= TOL


def test_negative_jacobian_raises():
    from akantu2_amd import SolidMechanicsModel
    from akantu2_amd._native import JacobianError
    mesh = M.box_mesh("_tetrahedron_4", 2, 1, 1)
    mesh.conn[3, [1, 2]] = mesh.conn[3, [2, 1]]
    g = SolidMechanicsModel(mesh)
    g.addMaterial("elastic")
    with pytest.raises(JacobianError, match=r"Quadrature Point 0:3"):
        g.initFull()


def test_fixture_meshes_mass_and_dt(fixtures_dir):
    from akantu2_amd import mesh_io
    for name in ("fe_engine_tetrahedron_4.msh", "fe_engine_hexahedron_8.msh"):
        mesh = mesh_io.read_msh(os.path.join(fixtures_dir, name))
        blocked = np.zeros((mesh.nb_nodes, 3), np.uint8)
        v0 = np.zeros((mesh.nb_nodes, 3))
        o, dt = make_oracle(mesh, blocked, v0, "elastic", rho=7800.0, E=2.1e11)
        g, dtg = make_gpu(mesh, blocked, v0, "elastic", rho=7800.0, E=2.1e11)
        assert dtg == dt
        np.testing.assert_array_equal(g.getMass(), o.mass)


@pytest.mark.parametrize("etype,law,shape", [
    ("_hexahedron_8", "neohookean", (24, 6, 5)),
    ("_tetrahedron_4", "elastic", (20, 5, 4)),
])
def test_pipelined_schedule_is_invisible(etype, law, shape):
    """The pipelined launch schedule (element parts overlapped with the
    gather on a second stream) gives the plain schedule's bits and the
    oracle's values."""
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=9)
    o, dt = make_oracle(mesh, blocked, v0, law)
    plain, _ = make_gpu(mesh, blocked, v0, law, dt=dt)
    piped, _ = make_gpu(mesh, blocked, v0, law, dt=dt)
    assert piped.setPipeline(4)
    for n in (1, 2, 57):
        plain.solveSteps(n)
        piped.solveSteps(n)
        o.solve_steps(n)
        for get in ("getDisplacement", "getVelocity", "getAcceleration",
                    "getInternalForce"):
            assert np.array_equal(getattr(plain, get)(), getattr(piped, get)()), get
        _compare(piped, o, f"pipelined {etype} after {n}")
