"""akx_model_get_energy_noalloc (the entry point that allocates no device
memory) against akx_model_get_energy, through the C ABI: the same ids give
the same bits, on two calls in a row, the second on a smaller model (the
scratch array keeps the first call's partial sums, which must not leak into
the second)."""
import ctypes as C

import numpy as np
import pytest

from akantu2_amd import _native as N
from tests.helpers import make_gpu, workload

pytestmark = pytest.mark.gpu

PL = "plastic_linear_isotropic_hardening"
IDS = ("kinetic", "external work", "potential", "plastic", "dissipated")


def _energy(fn, g, eid, mat=-1):
    out = C.c_double(np.nan)
    N.check(fn(g._h, eid.encode(), mat, C.byref(out)))
    return out.value


def _both(g, eid, mat=-1):
    lib = N.lib()
    return (_energy(lib.akx_model_get_energy, g, eid, mat),
            _energy(lib.akx_model_get_energy_noalloc, g, eid, mat))


def _model(etype, law, shape, seed, amp, steps, **mat):
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=seed, amp=amp)
    g, _ = make_gpu(mesh, blocked, v0, law, **mat)
    rng = np.random.default_rng(seed)
    g.getExternalForce()[:] = 1e-5 * rng.uniform(-1, 1, (mesh.nb_nodes, 3))
    g.solveSteps(steps)
    return g


@pytest.mark.parametrize("etype,law,mat", [
    ("_tetrahedron_4", "elastic", {}),
    ("_hexahedron_8", "neohookean", {}),
    ("_tetrahedron_4", PL, dict(sigma_y=2e-3, h=0.05)),
    ("_hexahedron_8", PL, dict(sigma_y=2e-3, h=0.05)),
])
def test_noalloc_energy_equals_allocating_energy(etype, law, mat):
    amp = 3e-2 if law == PL else 1e-3
    # the first model spans several first-level blocks (> 2048 nodes and
    # elements), the second fits one
    big = _model(etype, law, (40, 12, 12), 3, amp, 60, **mat)
    small = _model(etype, law, (5, 2, 2), 4, amp, 30, **mat)
    for g in (big, small):
        # "potential" of the allocating entry point first: it allocates the
        # stress arrays the no-allocation one reuses
        for eid in IDS:
            a, b = _both(g, eid)
            print(f"{etype} {law} nodes {g.getDisplacement().shape[0]} {eid}: "
                  f"{a:.17e} {b:.17e}", flush=True)
            assert a == b, (eid, a, b)
        assert _both(g, "potential", 0)[0] == _both(g, "potential", 0)[1]
    if law == PL:
        assert _both(big, "plastic")[1] > 0


def test_noalloc_potential_needs_existing_stress_arrays():
    g = _model("_hexahedron_8", "elastic", (4, 2, 2), 5, 1e-3, 5)
    lib = N.lib()
    with pytest.raises(N.AkantuError, match="do not exist yet"):
        _energy(lib.akx_model_get_energy_noalloc, g, "potential")
    # kinetic needs nothing: it runs before any allocating call
    k = _energy(lib.akx_model_get_energy_noalloc, g, "kinetic")
    assert k == _energy(lib.akx_model_get_energy, g, "kinetic") and k > 0
    _energy(lib.akx_model_get_energy, g, "potential")
    a, b = _both(g, "potential")
    assert a == b


def test_noalloc_energy_leaves_the_state_alone():
    g = _model("_tetrahedron_4", "elastic", (6, 3, 3), 6, 1e-3, 10)
    h = _model("_tetrahedron_4", "elastic", (6, 3, 3), 6, 1e-3, 10)
    _both(g, "potential")
    for eid in IDS:
        _energy(N.lib().akx_model_get_energy_noalloc, g, eid)
    g.solveSteps(10)
    h.solveSteps(10)
    assert np.array_equal(g.getDisplacement(), h.getDisplacement())
    assert np.array_equal(g.getInternalForce(), h.getInternalForce())
