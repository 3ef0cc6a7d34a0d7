"""The opt-in arithmetic forms against the FP64 oracle:
- "fp64_fast": FP64 with FMA and the regrouped contraction
  r_e = sum_q (w detJ P J^{-1})_q dN/dxi_q.  Equal to the reference order up
  to rounding per evaluation; over many steps the difference grows where the
  constitutive law cancels (the neo-Hookean S = mu (I - C^{-1}) + ... loses
  about three digits at strains of 1e-3), so its bound is stated per law.
- "fp32": grad u, stress and B^T sigma in FP32, FP64 nodal update with
  compensated summation; the target is <= 1e-5 relative displacement error."""
import numpy as np
import pytest

from tests.helpers import make_gpu, make_oracle, relerr, workload

pytestmark = pytest.mark.gpu


def _errs(g, o):
    return (relerr(g.getDisplacement(), o.u), relerr(g.getVelocity(), o.v),
            relerr(g.getInternalForce(), o.f_int))


@pytest.mark.parametrize("etype,law,steps,tol", [
    ("_tetrahedron_4", "elastic", 1000, 1e-12),
    ("_hexahedron_8", "elastic", 500, 1e-12),
    ("_hexahedron_8", "neohookean", 500, 1e-10),
])
def test_fp64_fast(etype, law, steps, tol):
    mesh, blocked, v0 = workload(etype, (12, 4, 4), jitter=0.1, seed=3)
    o, dt = make_oracle(mesh, blocked, v0, law)
    g, _ = make_gpu(mesh, blocked, v0, law, dt=dt, precision="fp64_fast")
    o.solve_steps(steps)
    g.solveSteps(steps)
    eu, ev, ef = _errs(g, o)
    print(f"fp64_fast {etype} {law} {steps} steps: u {eu:.2e} v {ev:.2e} f {ef:.2e}")
    assert max(eu, ev, ef) <= tol


@pytest.mark.parametrize("etype,law,tol_u", [("_tetrahedron_4", "elastic", 1e-5),
                                             ("_hexahedron_8", "neohookean", 1e-5)])
def test_fp32_form_error_against_fp64_oracle(etype, law, tol_u):
    """The 60 M-element Tet4 box in miniature (jittered Tet4, per-element E and nu, linear
    elastic, 2000 steps) and the Hex8 neo-Hookean case (a serving-form request
    type): u within 1e-5 of the FP64 oracle.  Reports L2 and max relative error per field."""
    mesh, blocked, v0 = workload(etype, (20, 4, 4), jitter=0.1, seed=11)
    kw = {}
    if etype == "_tetrahedron_4":
        rng = np.random.default_rng(11)
        kw = dict(E_el=np.exp(rng.uniform(np.log(0.5), np.log(2.0), mesh.nb_elements)),
                  nu_el=rng.uniform(0.2, 0.35, mesh.nb_elements))
    o, dt = make_oracle(mesh, blocked, v0, law, **kw)
    g, _ = make_gpu(mesh, blocked, v0, law, dt=dt, precision="fp32", **kw)
    o.solve_steps(2000)
    g.solveSteps(2000)
    for name, a, b in (("u", g.getDisplacement(), o.u), ("v", g.getVelocity(), o.v),
                       ("f_int", g.getInternalForce(), o.f_int)):
        l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
        mx = relerr(a, b)
        print(f"fp32 {etype} {law} {name}: L2 {l2:.2e} max {mx:.2e}")
        if name == "u":
            assert l2 <= tol_u and mx <= tol_u
        else:
            assert l2 <= 1e-4 and mx <= 1e-4


PL = "plastic_linear_isotropic_hardening"


@pytest.mark.parametrize("etype,law,amp,mat", [
    ("_hexahedron_8", "elastic", 1e-3, dict(E=1.3, nu=0.27)),
    ("_hexahedron_8", "elastic_fd", 1e-3, dict(E=0.9, nu=0.31)),
    ("_hexahedron_8", PL, 3e-2, dict(E=1.2, nu=0.22, sigma_y=1.5e-3, h=0.02)),
    ("_tetrahedron_4", "neohookean", 1e-3, dict(E=1.1, nu=0.28)),
    ("_tetrahedron_4", PL, 3e-2, dict(E=0.8, nu=0.3, sigma_y=2e-3, h=0.05)),
])
def test_fp32_form_every_law(etype, law, amp, mat):
    """The FP32 element bodies of every law and element type (the Hex8 one is
    the sum-factorised pass) against the FP64 oracle after 600 steps: u within
    1e-5; the plastic cases yield."""
    mesh, blocked, v0 = workload(etype, (10, 3, 3), jitter=0.1, seed=21, amp=amp)
    o, dt = make_oracle(mesh, blocked, v0, law, **mat)
    g, _ = make_gpu(mesh, blocked, v0, law, dt=dt, precision="fp32", **mat)
    o.solve_steps(600)
    g.solveSteps(600)
    eu = relerr(g.getDisplacement(), o.u)
    ev = relerr(g.getVelocity(), o.v)
    ef = relerr(g.getInternalForce(), o.f_int)
    print(f"fp32 {etype} {law}: u {eu:.2e} v {ev:.2e} f_int {ef:.2e}", flush=True)
    assert eu <= 1e-5
    assert ev <= 1e-4 and ef <= 1e-4
    if law == PL:
        assert o.iso.max() > 0



def _fp32_model(mesh, blocked, v0, block_assembly, dt=None, law="elastic", **mat):
    from akantu2_amd import SolidMechanicsModel
    g = SolidMechanicsModel(mesh)
    g.addMaterial(law, rho=1.0, E=1.0, nu=0.3, **mat)
    g.setPrecision("fp32")
    g.setBlockAssembly(block_assembly)
    g.initFull()
    assert g.blockAssembly() == block_assembly
    g.getBlockedDOFs()[:] = blocked
    g.getVelocity()[:] = v0
    g.setTimeStep(dt if dt is not None else 0.5 * g.getStableTimeStep())
    return g


def test_fp32_tet4_staged_pass_equals_per_element_pass():
    """With one table entry per element corner, the FP32 Tet4 element pass of
    a whole contiguous material stages each 256-element block's node records
    in LDS; the pipelined schedule's partial launches read the records per
    element instead.  Same arithmetic, so the two give the same bits
    (per-element E and nu, several step counts)."""
    mesh, blocked, v0 = workload("_tetrahedron_4", (30, 7, 5), jitter=0.1, seed=4)
    rng = np.random.default_rng(4)
    E = np.exp(rng.uniform(np.log(0.5), np.log(2.0), mesh.nb_elements))
    nu = rng.uniform(0.2, 0.35, mesh.nb_elements)
    staged = _fp32_model(mesh, blocked, v0, False, E_el=E, nu_el=nu)
    parts = _fp32_model(mesh, blocked, v0, False, dt=staged.getTimeStep(), E_el=E, nu_el=nu)
    assert parts.setPipeline(3)
    for n in (1, 2, 40):
        staged.solveSteps(n)
        parts.solveSteps(n)
        for get in ("getDisplacement", "getVelocity", "getAcceleration", "getInternalForce"):
            assert np.array_equal(getattr(staged, get)(), getattr(parts, get)()), (get, n)


@pytest.mark.parametrize("law,amp,mat", [
    ("elastic", 1e-3, {}),
    ("neohookean", 1e-3, {}),
    ("plastic_linear_isotropic_hardening", 3e-2, dict(sigma_y=2e-3, h=0.05)),
])
def test_fp32_tet4_block_assembly(law, amp, mat):
    """Block-partial assembly (one table entry per (256-element block, node)
    instead of one per element corner) against the FP64 oracle after 1,000
    steps (bar 1e-5 on u, as the per-element assembly), against the
    per-element assembly, and bit-reproducible: two runs give the same bits.
    The internal forces still sum to zero over the nodes."""
    mesh, blocked, v0 = workload("_tetrahedron_4", (24, 9, 7), jitter=0.1, seed=9, amp=amp)
    if law == "elastic":
        rng = np.random.default_rng(9)
        mat = dict(E_el=np.exp(rng.uniform(np.log(0.5), np.log(2.0), mesh.nb_elements)),
                   nu_el=rng.uniform(0.2, 0.35, mesh.nb_elements))
    o, dt = make_oracle(mesh, blocked, v0, law, **mat)
    bpa = _fp32_model(mesh, blocked, v0, True, dt=dt, law=law, **mat)
    again = _fp32_model(mesh, blocked, v0, True, dt=dt, law=law, **mat)
    per_el = _fp32_model(mesh, blocked, v0, False, dt=dt, law=law, **mat)
    for n in (1, 99, 900):
        o.solve_steps(n)
        for g in (bpa, again, per_el):
            g.solveSteps(n)
        eu = relerr(bpa.getDisplacement(), o.u)
        eu_el = relerr(per_el.getDisplacement(), o.u)
        ef = relerr(bpa.getInternalForce(), o.f_int)
        f = bpa.getInternalForce()
        tr = float((np.abs(f.sum(axis=0)) / np.abs(f).sum(axis=0)).max())
        print(f"{law} step {n}: block u {eu:.2e} f_int {ef:.2e}, per-element u {eu_el:.2e}, "
              f"sum f_int {tr:.1e}", flush=True)
        assert eu <= 1e-5 and eu_el <= 1e-5
        assert tr <= 1e-5
        for get in ("getDisplacement", "getVelocity", "getInternalForce"):
            assert np.array_equal(getattr(bpa, get)(), getattr(again, get)()), (get, n)


@pytest.mark.parametrize("law,amp,mat", [
    ("elastic", 1e-3, {}),
    ("neohookean", 1e-3, {}),
    ("plastic_linear_isotropic_hardening", 3e-2, dict(sigma_y=2e-3, h=0.05)),
])
def test_fp32_hex8_block_assembly(law, amp, mat):
    """FP32 Hex8 block partials (192-element blocks, one table entry per
    (block, node) instead of one per element corner): within 1e-5 of the FP64
    oracle on u after 1,000 steps like the per-corner assembly, two runs give
    the same bits, and the internal forces still sum to zero."""
    mesh, blocked, v0 = workload("_hexahedron_8", (30, 7, 6), jitter=0.1, seed=19, amp=amp)
    if law == "elastic":
        rng = np.random.default_rng(19)
        mat = dict(E_el=np.exp(rng.uniform(np.log(0.5), np.log(2.0), mesh.nb_elements)),
                   nu_el=rng.uniform(0.2, 0.35, mesh.nb_elements))
    o, dt = make_oracle(mesh, blocked, v0, law, **mat)
    bpa = _fp32_model(mesh, blocked, v0, True, dt=dt, law=law, **mat)
    again = _fp32_model(mesh, blocked, v0, True, dt=dt, law=law, **mat)
    per_el = _fp32_model(mesh, blocked, v0, False, dt=dt, law=law, **mat)
    for n in (1, 99, 900):
        o.solve_steps(n)
        for g in (bpa, again, per_el):
            g.solveSteps(n)
        eu = relerr(bpa.getDisplacement(), o.u)
        eu_el = relerr(per_el.getDisplacement(), o.u)
        f = bpa.getInternalForce()
        tr = float((np.abs(f.sum(axis=0)) / np.abs(f).sum(axis=0)).max())
        print(f"hex8 {law} step {n}: block u {eu:.2e}, per-corner u {eu_el:.2e}, "
              f"sum f_int {tr:.1e}", flush=True)
        assert eu <= 1e-5 and eu_el <= 1e-5
        assert tr <= 1e-5
        for get in ("getDisplacement", "getVelocity", "getInternalForce"):
            assert np.array_equal(getattr(bpa, get)(), getattr(again, get)()), (get, n)


@pytest.mark.parametrize("etype,law", [("_tetrahedron_4", "elastic"),
                                       ("_hexahedron_8", "neohookean")])
def test_fp32_steps_split_over_calls_keep_bits(etype, law):
    """The FP32 form keeps the half-step velocity w = v + dt/2 a between the
    steps of one call and (v, a) between calls; the call boundary performs
    the same two compensated additions as a step inside a call, so a run of
    steps split over calls gives the bits of one call (with an external
    force, and blocked DOFs)."""
    mesh, blocked, v0 = workload(etype, (10, 4, 3), jitter=0.1, seed=21)
    fext = 1e-5 * np.random.default_rng(21).uniform(-1, 1, (mesh.nb_nodes, 3))
    one = _fp32_model(mesh, blocked, v0, True, law=law)
    split = _fp32_model(mesh, blocked, v0, True, dt=one.getTimeStep(), law=law)
    for g in (one, split):
        g.getExternalForce()[:] = fext
    one.solveSteps(47)
    for n in (30, 1, 2, 1, 13):
        split.solveSteps(n)
    for get in ("getDisplacement", "getVelocity", "getAcceleration", "getInternalForce"):
        assert np.array_equal(getattr(one, get)(), getattr(split, get)()), get
    # blocked components keep their velocity and acceleration
    fixed = np.asarray(blocked, dtype=bool)
    assert fixed.shape == (mesh.nb_nodes, 3) and fixed.any()
    assert np.all(one.getVelocity()[fixed] == np.asarray(v0)[fixed])
    assert np.all(one.getAcceleration()[fixed] == 0.0)
