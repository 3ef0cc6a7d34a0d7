"""The batching workload at its own size: 512 independent Tet4 boxes of
5e3-4e5 elements (log-uniform, seed 4), aspect ratios U(1, 8), linear
isotropic hardening with per-instance E, nu, sigma_y, h, initial velocity of
amplitude 1e-2 (about 51.5 M elements in all), advanced 500 steps by one
persistent launch per substep inside a captured hipGraph, launch shape chosen
by the autotuner, in both arithmetic forms.

- FP64 batch (the plain path's arithmetic): the four smallest instances equal
  the CPU oracle run on their own meshes to 1e-12 relative in u, v, f_int.
- FP32 batch (the precision form): every instance's displacement is within
  1e-5 relative (max norm) of the FP64 batch's, which is pinned to the oracle.
- Four sampled instances, run alone in a batch of their own, give the same
  bits as in the mix (FP64 and FP32).
"""
import numpy as np
import pytest

import bench
from tests.helpers import make_oracle, relerr

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

PL = "plastic_linear_isotropic_hardening"
STEPS = 500
CHECKS = (20, STEPS)  # step counts at which results are read


@pytest.fixture(scope="module")
def specs():
    s = bench.batch_instances(512, scale=1.0)
    assert len(s) == 512
    return s


def _run(precision, specs, idx):
    b, ids = bench._fill_batch(precision, [specs[k] for k in idx])
    st = b.stats()
    best = b.autotune(repeats=1)
    launch = b.launch()
    assert launch["graph"], launch
    out = {}
    done = 0
    for n in CHECKS:
        b.solveStep(n - done)
        done = n
        out[n] = {k: {"u": b.getDisplacement(i), "v": b.getVelocity(i),
                      "f": b.getInternalForce(i)} for k, i in zip(idx, ids)}
        assert all(b.steps(i) == n for i in ids)
    b.close()
    return out, st, best


@pytest.fixture(scope="module")
def full(specs):
    idx = list(range(len(specs)))
    res = {}
    for prec in ("fp64", "fp32"):
        out, st, best = _run(prec, specs, idx)
        print(f"{prec} batch: {st['live']} instances, {st['elements']} Tet4, "
              f"{st['nodes']} nodes, launch {best}", flush=True)
        res[prec] = out
    assert st["live"] == 512 and st["elements"] > 4e7
    return res


def test_fp64_batch_equals_oracle_on_smallest(specs, full):
    small = np.argsort([np.prod(s["cells"]) for s in specs])[:4]
    for k in small:
        mesh, blocked, v0 = bench.instance_mesh(specs[k])
        o, _ = make_oracle(mesh, blocked, v0, PL, **specs[k]["mat"])
        done = 0
        for n in CHECKS:
            o.solve_steps(n - done)
            done = n
            g = full["fp64"][n][k]
            errs = dict(u=relerr(g["u"], o.u), v=relerr(g["v"], o.v),
                        f_int=relerr(g["f"], o.f_int))
            print(f"instance {k} ({mesh.nb_elements} Tet4) step {n}: {errs}", flush=True)
            assert max(errs.values()) <= 1e-12, (k, n, errs)
        # the instance yields: a trajectory with plastic flow, not an elastic one
        assert o.energy("plastic") > 0, k


def test_fp32_batch_within_1e5_of_fp64_on_every_instance(full):
    worst = {}
    for n in CHECKS:
        errs = np.array([relerr(full["fp32"][n][k]["u"], full["fp64"][n][k]["u"])
                         for k in range(512)])
        l2 = np.array([np.linalg.norm(full["fp32"][n][k]["u"] - full["fp64"][n][k]["u"])
                       / np.linalg.norm(full["fp64"][n][k]["u"]) for k in range(512)])
        worst[n] = (float(errs.max()), int(errs.argmax()), float(l2.max()))
        print(f"step {n}: FP32 vs FP64 u max rel error over 512 instances "
              f"{errs.max():.3e} (instance {errs.argmax()}), median {np.median(errs):.3e}, "
              f"L2 max {l2.max():.3e}", flush=True)
    for n, (mx, k, _) in worst.items():
        assert mx <= 1e-5, (n, k, mx)


def test_instances_alone_equal_in_mix(specs, full):
    rng = np.random.default_rng(44)
    sample = sorted(rng.choice(512, 4, replace=False).tolist())
    for prec in ("fp64", "fp32"):
        alone = {}
        for k in sample:
            out, _, _ = _run(prec, specs, [k])
            alone[k] = out
        for k in sample:
            for n in CHECKS:
                for f in ("u", "v", "f"):
                    assert np.array_equal(alone[k][n][k][f], full[prec][n][k][f]), \
                        (prec, k, n, f)
        print(f"{prec}: instances {sample} alone == in the mix, bit for bit", flush=True)

