"""Energies on the device (akx_model_get_energy) against the oracle:
SolidMechanicsModel::getEnergy("kinetic" | "external work" | "potential" |
"plastic") and Material::getEnergy.  The plastic energy is the work
accumulated after every step (afterSolveStep -> updateEnergies,
material_plastic.cc:86-99), so it checks a12 along the whole run.  The
device sums are tree sums and the reference's are sequential or
Math::reduce: agreement is to rounding, bar 1e-12 relative."""
import numpy as np
import pytest

from tests.helpers import make_gpu, make_oracle, workload

pytestmark = pytest.mark.gpu

PL = "plastic_linear_isotropic_hardening"


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("etype,law,amp,mat", [
    ("_tetrahedron_4", "elastic", 1e-3, {}),
    ("_hexahedron_8", "neohookean", 1e-3, {}),
    ("_hexahedron_8", "elastic_fd", 1e-3, {}),
    ("_tetrahedron_4", PL, 3e-2, dict(sigma_y=2e-3, h=0.05)),
    ("_hexahedron_8", PL, 3e-2, dict(sigma_y=2e-3, h=0.05)),
])
def test_energies_match_oracle(etype, law, amp, mat):
    mesh, blocked, v0 = workload(etype, (8, 3, 3), jitter=0.1, seed=6, amp=amp)
    rng = np.random.default_rng(6)
    fext = 1e-5 * rng.uniform(-1, 1, (mesh.nb_nodes, 3))
    o, dt = make_oracle(mesh, blocked, v0, law, **mat)
    o.f_ext[:] = fext
    g, _ = make_gpu(mesh, blocked, v0, law, dt=dt, **mat)
    g.getExternalForce()[:] = fext
    for k in range(3):
        o.solve_steps(70)
        g.solveSteps(70)
        for eid in ("kinetic", "potential", "external work", "plastic", "dissipated"):
            eg, eo = g.getEnergy(eid), o.energy(eid)
            print(f"{etype} {law} step {(k + 1) * 70} {eid}: gpu {eg:.15e} oracle {eo:.15e}",
                  flush=True)
            if eo == 0.0:
                assert eg == 0.0
            else:
                assert _rel(eg, eo) <= 1e-12, (eid, eg, eo)
    assert g.getMaterial(0).getEnergy("potential") == g.getEnergy("potential")
    if law == PL:
        assert o.energy("plastic") > 0


def test_energies_are_reproducible_and_do_not_change_the_state():
    mesh, blocked, v0 = workload("_hexahedron_8", (6, 3, 3), jitter=0.1, seed=2, amp=3e-2)
    kw = dict(sigma_y=2e-3, h=0.05)
    g, dt = make_gpu(mesh, blocked, v0, PL, **kw)
    h, _ = make_gpu(mesh, blocked, v0, PL, dt=dt, **kw)
    g.solveSteps(40)
    h.solveSteps(40)
    first = [g.getEnergy(e) for e in ("kinetic", "potential", "plastic")]
    again = [g.getEnergy(e) for e in ("kinetic", "potential", "plastic")]
    assert first == again
    g.solveSteps(40)
    h.solveSteps(40)
    assert np.array_equal(g.getDisplacement(), h.getDisplacement())
    assert g.getEnergy("plastic") == h.getEnergy("plastic")
