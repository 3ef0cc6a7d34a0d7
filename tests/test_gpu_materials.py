"""Several materials on one mesh (element filters, Material::assembleInternalForces
called per material in creation order, material.cc:98-120 through
solid_mechanics_model.cc:333-370).

- Two elastic materials on the even and the odd elements equal one material
  with per-element E, nu in the oracle: the same values per element, summed
  at the nodes in another order (material by material), so to 1e-12, with
  the same stable time step to the bit.
- The same with the plastic law (per-element yield stress and hardening).
- Two materials splitting the elements into a first and a second half, with
  the same parameters, give the single material's bits: the per-node order
  (material, then filter position) is then the element order.
- The FP32 form with two materials stays within 1e-5 of the oracle.
"""
import numpy as np
import pytest

from tests.helpers import make_gpu, make_oracle, relerr, workload

pytestmark = pytest.mark.gpu

PL = "plastic_linear_isotropic_hardening"


def _two_materials(mesh, blocked, v0, law, parts, dt, precision="fp64"):
    from akantu2_amd import SolidMechanicsModel
    g = SolidMechanicsModel(mesh)
    for elements, mat in parts:
        g.addMaterial(law, elements=elements, **mat)
    g.setPrecision(precision)
    g.initFull()
    g.getBlockedDOFs()[:] = blocked
    g.getVelocity()[:] = v0
    dtg = g.getStableTimeStep() * 0.5
    g.setTimeStep(dt)
    return g, dtg


def _per_element(nel, parts, key):
    out = np.empty(nel)
    for elements, mat in parts:
        out[elements] = mat[key]
    return out


@pytest.mark.parametrize("etype,shape", [("_tetrahedron_4", (8, 4, 3)),
                                         ("_hexahedron_8", (8, 4, 3))])
def test_even_odd_elastic_materials_equal_per_element_oracle(etype, shape):
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=51)
    ids = np.arange(mesh.nb_elements, dtype=np.int32)
    parts = [(ids[0::2], dict(E=1.3, nu=0.27)), (ids[1::2], dict(E=0.7, nu=0.33))]
    o, dt = make_oracle(mesh, blocked, v0, "elastic",
                        E_el=_per_element(mesh.nb_elements, parts, "E"),
                        nu_el=_per_element(mesh.nb_elements, parts, "nu"))
    g, dtg = _two_materials(mesh, blocked, v0, "elastic", parts, dt)
    assert dtg == dt
    for k in range(3):
        o.solve_steps(70)
        g.solveSteps(70)
        eu = relerr(g.getDisplacement(), o.u)
        ef = relerr(g.getInternalForce(), o.f_int)
        print(etype, k, f"u {eu:.2e} f_int {ef:.2e}", flush=True)
        assert eu <= 1e-12 and ef <= 1e-12


def test_two_plastic_materials_equal_per_element_oracle():
    mesh, blocked, v0 = workload("_tetrahedron_4", (10, 3, 3), jitter=0.05, seed=52, amp=3e-2)
    ids = np.arange(mesh.nb_elements, dtype=np.int32)
    parts = [(ids[0::2], dict(E=1.0, nu=0.3, sigma_y=2e-3, h=0.05)),
             (ids[1::2], dict(E=1.0, nu=0.3, sigma_y=3e-3, h=0.02))]
    o, dt = make_oracle(mesh, blocked, v0, PL, E=1.0, nu=0.3,
                        sigma_y_el=_per_element(mesh.nb_elements, parts, "sigma_y"),
                        h_el=_per_element(mesh.nb_elements, parts, "h"))
    g, _ = _two_materials(mesh, blocked, v0, PL, parts, dt)
    for k in range(3):
        o.solve_steps(100)
        g.solveSteps(100)
        eu = relerr(g.getDisplacement(), o.u)
        print("plastic", k, f"u {eu:.2e}", flush=True)
        assert eu <= 1e-12
    assert o.iso.max() > 0  # the run does yield
    iso = np.empty(mesh.nb_elements)
    for m, (elements, _) in zip(g.materials, parts):
        iso[elements] = m.getIsoHardening().ravel()
    assert relerr(iso, o.iso) <= 1e-12


@pytest.mark.parametrize("etype", ["_tetrahedron_4", "_hexahedron_8"])
def test_contiguous_halves_give_the_single_material_bits(etype):
    mesh, blocked, v0 = workload(etype, (8, 3, 3), jitter=0.1, seed=53)
    ids = np.arange(mesh.nb_elements, dtype=np.int32)
    half = mesh.nb_elements // 2
    mat = dict(E=1.1, nu=0.29)
    one, dt = make_gpu(mesh, blocked, v0, "neohookean", **mat)
    two, _ = _two_materials(mesh, blocked, v0, "neohookean",
                            [(ids[:half], mat), (ids[half:], mat)], dt)
    one.solveSteps(150)
    two.solveSteps(150)
    assert np.array_equal(one.getDisplacement(), two.getDisplacement())
    assert np.array_equal(one.getInternalForce(), two.getInternalForce())


def test_fp32_form_with_two_materials():
    mesh, blocked, v0 = workload("_hexahedron_8", (10, 4, 4), jitter=0.1, seed=54)
    ids = np.arange(mesh.nb_elements, dtype=np.int32)
    parts = [(ids[0::2], dict(E=1.3, nu=0.27)), (ids[1::2], dict(E=0.7, nu=0.33))]
    o, dt = make_oracle(mesh, blocked, v0, "elastic",
                        E_el=_per_element(mesh.nb_elements, parts, "E"),
                        nu_el=_per_element(mesh.nb_elements, parts, "nu"))
    g, _ = _two_materials(mesh, blocked, v0, "elastic", parts, dt, precision="fp32")
    o.solve_steps(300)
    g.solveSteps(300)
    eu = relerr(g.getDisplacement(), o.u)
    print("fp32 two materials", f"u {eu:.2e}", flush=True)
    assert eu <= 1e-5
