"""FEEngine / DOFManager calls on the device (akx_fe_*, akx_dof_*).

- gradientOnIntegrationPoints: the gradient of the node coordinates is the
  identity, that of a linear field its matrix; the displacement gradient
  equals the material's grad_u bit for bit.
- integrate: element volumes summing to the box volume.
- assembleElementalArrayLocalArray: ones assemble to the node valences.
- computeBtD -> integrate -> assemble(scale -1) on the elastic stress is the
  reference's Material::assembleInternalForces (material.cc:467-510) and
  reproduces the internal force of the element pass + gather bit for bit.
- A filtered call returns the rows of the filtered elements.
"""
import numpy as np
import pytest

from tests.helpers import make_gpu, workload

pytestmark = pytest.mark.gpu

CASES = [("_tetrahedron_4", (5, 3, 2)), ("_hexahedron_8", (4, 3, 2))]


@pytest.mark.parametrize("etype,shape", CASES)
def test_gradient_closed_forms_and_volume(etype, shape):
    mesh, blocked, v0 = workload(etype, shape, jitter=0.0, seed=41)
    g, dt = make_gpu(mesh, blocked, v0, "elastic")
    fe = g.getFEEngine()
    gx = fe.gradientOnIntegrationPoints(mesh.nodes, 3)
    assert gx.shape == (mesh.nb_elements * mesh.nqp, 9)
    assert np.abs(gx - np.eye(3).reshape(-1, order="F")).max() <= 1e-13
    A = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [0.2, 0.1, -0.15]])
    u = mesh.nodes @ A.T + np.array([1.0, -2.0, 0.5])
    gu = fe.gradientOnIntegrationPoints(u, 3)
    assert np.abs(gu - A.reshape(-1, order="F")).max() <= 1e-13
    vol = fe.integrate(np.ones((mesh.nb_elements * mesh.nqp, 1)), 1)
    assert vol.shape == (mesh.nb_elements, 1)
    assert abs(vol.sum() - np.prod(shape)) <= 1e-12 * np.prod(shape)
    valence = g.getDOFManager().assembleElementalArrayLocalArray(
        np.ones((mesh.nb_elements, mesh.conn.shape[1])), np.zeros(mesh.nb_nodes))
    assert np.array_equal(valence, np.bincount(mesh.conn.ravel(), minlength=mesh.nb_nodes))


@pytest.mark.parametrize("etype,shape", CASES)
def test_chain_reproduces_internal_force(etype, shape):
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=42, amp=1e-2)
    g, dt = make_gpu(mesh, blocked, v0, "elastic", E=1.4, nu=0.28)
    g.solveSteps(60)
    mat = g.getMaterial(0)
    fe, dof = g.getFEEngine(), g.getDOFManager()
    grad = fe.gradientOnIntegrationPoints(g.getDisplacement(), 3)
    assert np.array_equal(grad, mat.getGradU())
    sigma = mat.getStress()
    btd = fe.computeBtD(sigma)
    nnpe = mesh.conn.shape[1]
    assert btd.shape == (mesh.nb_elements * mesh.nqp, 3 * nnpe)
    r_e = fe.integrate(btd, 3 * nnpe)
    f = dof.assembleElementalArrayLocalArray(r_e, np.zeros((mesh.nb_nodes, 3)),
                                             scale_factor=-1.0)
    g.assembleInternalForces()
    assert np.array_equal(f, g.getInternalForce())


def test_filtered_rows():
    mesh, blocked, v0 = workload("_hexahedron_8", (4, 3, 2), jitter=0.1, seed=43)
    g, dt = make_gpu(mesh, blocked, v0, "elastic")
    fe = g.getFEEngine()
    u = np.random.default_rng(3).normal(size=(mesh.nb_nodes, 2))
    full = fe.gradientOnIntegrationPoints(u, 2)
    filt = np.arange(1, mesh.nb_elements, 3, dtype=np.int32)
    part = fe.gradientOnIntegrationPoints(u, 2, filter_elements=filt)
    rows = (filt[:, None] * mesh.nqp + np.arange(mesh.nqp)).ravel()
    assert np.array_equal(part, full[rows])
    w = fe.integrate(np.ones((filt.size * mesh.nqp, 1)), 1, filter_elements=filt)
    wall = fe.integrate(np.ones((mesh.nb_elements * mesh.nqp, 1)), 1)
    assert np.array_equal(w, wall[filt])
