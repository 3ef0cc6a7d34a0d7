"""Batched multi-instance engine (akx_batch) on the GPU.

- FP64 batch: every instance matches the CPU oracle run on its own mesh to
  1e-12 relative (the plain path's bar), across uneven instance sizes, both
  element types, all four laws, odd step splits (history ping-pong parity)
  and graph / no-graph launches.
- An instance gives the same bits alone and in the mix (FP64 and FP32), and
  after other instances were removed or admitted around it; a run of steps
  split over calls gives the bits of one call (FP64 and FP32).
- FP32 batch (the precision form): <= 1e-5 relative displacement error
  against the FP64 oracle on the plasticity workload of the batching form.
- The autotuner leaves the simulation state untouched.
- A bad instance (inverted element) is refused without disturbing the
  instances already resident; a full arena refuses with AKX_ERR_MEMORY.
"""
import numpy as np
import pytest

from akantu2_amd import _native as N
from akantu2_amd import mesh as M
from akantu2_amd.batch import Batch
from tests.helpers import make_oracle, relerr, workload

pytestmark = pytest.mark.gpu

PL = "plastic_linear_isotropic_hardening"
# (element type, cells, law, material, velocity amplitude, seed)
SPECS = [
    ("_tetrahedron_4", (6, 3, 3), "elastic", dict(E=1.0, nu=0.3), 1e-3, 11),
    ("_hexahedron_8", (5, 3, 2), "neohookean", dict(E=1.5, nu=0.25), 1e-3, 12),
    ("_tetrahedron_4", (7, 3, 2), PL, dict(E=0.8, nu=0.3, sigma_y=2e-3, h=0.05), 3e-2, 13),
    ("_hexahedron_8", (6, 2, 2), PL, dict(E=1.2, nu=0.22, sigma_y=1.5e-3, h=0.02), 3e-2, 14),
    ("_tetrahedron_4", (4, 4, 3), "elastic_fd", dict(E=0.6, nu=0.33), 1e-3, 15),
    ("_hexahedron_8", (3, 3, 3), "elastic", dict(E=2.0, nu=0.2), 1e-3, 16),
    ("_tetrahedron_4", (12, 5, 4), "neohookean", dict(E=1.1, nu=0.28), 2e-3, 17),
]
SPLITS = [37, 64, 1, 2]  # odd and even step counts: parity, graph and eager paths


def _case(spec):
    etype, shape, law, mat, amp, seed = spec
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=seed, amp=amp)
    return mesh, blocked, v0, law, mat


def _batch(precision, specs, plastic=True, launch=None):
    b = Batch(precision, nodes=1 << 16, tets=1 << 15, hexes=1 << 13, plastic=plastic)
    if launch:
        b.setLaunch(**launch)
    ids = []
    for spec in specs:
        mesh, blocked, v0, law, mat = _case(spec)
        ids.append(b.add(mesh, law, blocked=blocked, velocity=v0, dt_factor=0.5, **mat))
    return b, ids


def _fields(b, iid):
    return {f: b.get(iid, f) for f in ("displacement", "velocity", "acceleration",
                                       "internal_force")}


@pytest.mark.parametrize("launch", [None, dict(block=64, tile=64, blocks_per_cu=1),
                                    dict(block=128, tile=256, blocks_per_cu=2,
                                         graph=False)])
def test_fp64_batch_matches_oracle_per_instance(launch):
    b, ids = _batch("fp64", SPECS, launch=launch)
    oracles = []
    for spec, iid in zip(SPECS, ids):
        mesh, blocked, v0, law, mat = _case(spec)
        o, dt = make_oracle(mesh, blocked, v0, law, **mat)
        assert b.timeStep(iid)[0] == dt  # stable time step to the bit
        oracles.append(o)
    done = 0
    for n in SPLITS:
        b.solveStep(n)
        done += n
        for k, (o, iid) in enumerate(zip(oracles, ids)):
            o.solve_steps(n)
            g = _fields(b, iid)
            errs = dict(u=relerr(g["displacement"], o.u), v=relerr(g["velocity"], o.v),
                        a=relerr(g["acceleration"], o.a),
                        f=relerr(g["internal_force"], o.f_int))
            print(f"step {done} instance {k} {SPECS[k][0]} {SPECS[k][2]} "
                  + " ".join(f"{a}={e:.2e}" for a, e in errs.items()), flush=True)
            assert max(errs.values()) <= 1e-12, (k, errs)
            assert b.steps(iid) == done
    # the plastic instances did yield
    assert oracles[2].iso.max() > 0 and oracles[3].iso.max() > 0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_instance_same_alone_and_in_mix(precision):
    nsteps = 41
    mix, ids = _batch(precision, SPECS)
    mix.solveStep(nsteps)
    for k, spec in enumerate(SPECS):
        alone, (iid,) = _batch(precision, [spec])
        alone.solveStep(nsteps)
        a, m = _fields(alone, iid), _fields(mix, ids[k])
        for f in a:
            assert np.array_equal(a[f], m[f]), (k, f)
        alone.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_step_split_over_calls_keeps_bits(precision):
    """A run of steps split over several calls gives the bits of one call.
    The FP32 form keeps the half-step velocity between the steps of a call
    and the velocity between calls: the call boundary performs the same two
    compensated additions as a step inside a call."""
    one, ids1 = _batch(precision, SPECS)
    one.solveStep(44)
    split, ids2 = _batch(precision, SPECS)
    for n in (37, 1, 2, 1, 3):
        split.solveStep(n)
    for k in range(len(SPECS)):
        a, m = _fields(one, ids1[k]), _fields(split, ids2[k])
        for f in a:
            assert np.array_equal(a[f], m[f]), (k, f)
    assert np.abs(a["acceleration"]).max() > 0
    one.close()
    split.close()


def test_admission_and_retirement_leave_residents_unchanged():
    """Instances admitted and retired between chunks (the serving pattern):
    a resident instance's trajectory is the one it has alone."""
    nsteps = 30
    b, ids = _batch("fp64", SPECS[:3])
    b.solveStep(10)
    b.remove(ids[1])
    new = _case(SPECS[4])
    nid = b.add(new[0], new[3], blocked=new[1], velocity=new[2], **new[4])
    assert nid == ids[1]  # the freed id is reused
    b.solveStep(nsteps - 10)
    for k in (0, 2):
        alone, (iid,) = _batch("fp64", [SPECS[k]])
        alone.solveStep(nsteps)
        for f, arr in _fields(alone, iid).items():
            assert np.array_equal(arr, b.get(ids[k], f)), (k, f)
        alone.close()
    # the newcomer is a fresh simulation: 20 steps from its initial state
    alone, (iid,) = _batch("fp64", [SPECS[4]])
    alone.solveStep(nsteps - 10)
    assert b.steps(nid) == nsteps - 10
    for f, arr in _fields(alone, iid).items():
        assert np.array_equal(arr, b.get(nid, f)), f


def test_fp32_batch_plastic_error_bound():
    """Batching-form workload in miniature: Tet4 plasticity, per-instance
    E, nu, sigma_y, h, FP32 element pass with compensated FP64 update;
    displacement error <= 1e-5 relative against the FP64 oracle."""
    rng = np.random.default_rng(4)
    specs = []
    for k in range(4):
        shape = (int(rng.integers(5, 12)), int(rng.integers(2, 5)), int(rng.integers(2, 5)))
        mat = dict(E=float(rng.uniform(0.5, 2)), nu=float(rng.uniform(0.2, 0.35)),
                   sigma_y=float(rng.uniform(1e-3, 5e-3)), h=float(rng.uniform(0, 0.1)))
        specs.append(("_tetrahedron_4", shape, PL, mat, 2e-2, 100 + k))
    b, ids = _batch("fp32", specs)
    b.solveStep(300)
    yielded = 0
    for spec, iid in zip(specs, ids):
        mesh, blocked, v0, law, mat = _case(spec)
        o, dt = make_oracle(mesh, blocked, v0, law, **mat)
        o.solve_steps(300)
        eu = relerr(b.getDisplacement(iid), o.u)
        ev = relerr(b.getVelocity(iid), o.v)
        ef = relerr(b.getInternalForce(iid), o.f_int)
        print(f"fp32 batch {spec[1]} u={eu:.2e} v={ev:.2e} f_int={ef:.2e}", flush=True)
        assert eu <= 1e-5, eu
        yielded += int(o.iso.max() > 0)
    assert yielded >= 2


def test_fp32_batch_every_law_against_oracle():
    """FP32 batch of the seven-instance mix (both element types, all four
    laws, the Hex8 instances on the sum-factorised pass): each instance's
    displacement within 1e-5 of the FP64 oracle after 240 steps."""
    b, ids = _batch("fp32", SPECS)
    b.solveStep(240)
    for spec, iid in zip(SPECS, ids):
        mesh, blocked, v0, law, mat = _case(spec)
        o, dt = make_oracle(mesh, blocked, v0, law, **mat)
        o.solve_steps(240)
        eu = relerr(b.getDisplacement(iid), o.u)
        ef = relerr(b.getInternalForce(iid), o.f_int)
        print(f"fp32 batch {spec[0]} {law} u={eu:.2e} f_int={ef:.2e}", flush=True)
        assert eu <= 1e-5, (spec, eu)
        assert ef <= 1e-4, (spec, ef)


def test_autotune_keeps_state():
    ref, ids = _batch("fp64", SPECS)
    ref.solveStep(20)
    b, ids2 = _batch("fp64", SPECS)
    b.solveStep(7)
    best = b.autotune(repeats=1)
    print("autotune", best, flush=True)
    assert best["block"] in (64, 128, 256) and best["element_pass_ms"] > 0
    assert b.launch()["tile"] == best["tile"]
    b.solveStep(13)
    for i, j in zip(ids, ids2):
        for f, arr in _fields(ref, i).items():
            assert np.array_equal(arr, b.get(j, f)), f


def test_bad_instance_refused_and_residents_unharmed():
    b, ids = _batch("fp64", SPECS[:2])
    b.solveStep(5)
    before = _fields(b, ids[0])
    cap = b.capacityLeft()
    mesh = M.box_mesh("_tetrahedron_4", 3, 2, 2)
    bad = M.BoxMesh(mesh.etype, mesh.nodes, mesh.conn[:, [0, 2, 1, 3]].copy(), mesh.shape)
    with pytest.raises(N.JacobianError):
        b.add(bad, "elastic")
    assert b.capacityLeft() == cap  # nothing kept
    assert b.stats()["live"] == 2
    for f, arr in _fields(b, ids[0]).items():
        assert np.array_equal(arr, before[f])
    # arena full
    small = Batch("fp64", nodes=512, tets=64, hexes=0)
    big = M.box_mesh("_tetrahedron_4", 4, 4, 4)
    with pytest.raises(N.AkantuError) as ei:
        small.add(big, "elastic")
    assert ei.value.code == -4
    b.solveStep(5)  # residents keep stepping
    ref, rid = _batch("fp64", SPECS[:2])
    ref.solveStep(10)
    for f, arr in _fields(ref, rid[0]).items():
        assert np.array_equal(arr, b.get(ids[0], f))
