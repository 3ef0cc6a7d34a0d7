"""Serving form on the GPU: continuous batching of solve requests.

- Each served result is bit-identical to the same request run alone
  (FP32 form), while requests arrive, overlap on the device and retire at
  different times.
- FP64 server: each result matches the CPU oracle to 1e-12; FP32 server:
  displacement within the precision form's bound of the FP64 oracle.
- A small HBM budget makes requests queue; all still complete, unchanged.
- Bad requests (inverted element, NaN velocity) fail alone.
"""
import numpy as np
import pytest

from akantu2_amd import serving as S
from tests.helpers import make_oracle, relerr

pytestmark = pytest.mark.gpu


class Tick:
    """Virtual clock (1 ms per reading) so that admission is deterministic."""

    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-3
        return self.t


def _requests(n, seed=3, rate=40.0, max_el=3e3, steps=(60, 240)):
    rs = S.synthetic_requests(n, rate=rate, seed=seed, min_el=3e2, max_el=max_el,
                              steps=steps)
    return [r.materialize() for r in rs]


def test_served_results_equal_alone_fp32():
    rs = _requests(10, rate=200.0)
    srv = S.Server("fp32", hbm_bytes=1 << 30, chunk_ms=5.0)
    rep = srv.run(rs, clock=Tick())
    print(f"served {len(rep.done)} in {rep.chunks} chunks, max live {rep.max_live}",
          flush=True)
    assert len(rep.done) == 10 and not rep.failed
    assert rep.max_live >= 3      # requests overlapped on the device
    assert rep.chunks > 10        # and joined / left between chunks
    for r in rep.done:
        alone = S.run_alone(r, "fp32")
        for f in ("displacement", "velocity", "internal_force"):
            assert np.array_equal(alone[f], r.result[f]), (r.rid, f)
    srv.close()


def _oracle_for(r):
    return make_oracle(r.mesh, r.blocked, r.velocity, r.law, **r.material)[0]


def test_fp64_server_matches_oracle():
    rs = _requests(5, seed=8)
    srv = S.Server("fp64", hbm_bytes=1 << 30, chunk_ms=5.0)
    rep = srv.run(rs, clock=Tick())
    assert len(rep.done) == 5
    for r in rep.done:
        o = _oracle_for(r)
        o.solve_steps(r.steps)
        eu = relerr(r.result["displacement"], o.u)
        ev = relerr(r.result["velocity"], o.v)
        ef = relerr(r.result["internal_force"], o.f_int)
        print(f"request {r.rid} {r.etype} {r.law} {r.mesh.nb_elements} el {r.steps} steps "
              f"u={eu:.2e} v={ev:.2e} f={ef:.2e}", flush=True)
        assert max(eu, ev, ef) <= 1e-12
    srv.close()


def test_fp32_server_error_vs_oracle():
    rs = _requests(6, seed=9, steps=(300, 600))
    srv = S.Server("fp32", hbm_bytes=1 << 30, chunk_ms=5.0)
    rep = srv.run(rs, clock=Tick())
    for r in rep.done:
        o = _oracle_for(r)
        o.solve_steps(r.steps)
        eu = relerr(r.result["displacement"], o.u)
        print(f"request {r.rid} {r.etype} {r.law} u={eu:.2e}", flush=True)
        assert eu <= 1e-5
    srv.close()


def test_small_budget_queues_requests():
    rs = _requests(8, seed=4, rate=1000.0)
    fps = [S.footprint("fp32", r.mesh.etype, r.law, r.mesh.nb_nodes, r.mesh.nb_elements)
           for r in rs]
    budget = 2 * max(fps)
    srv = S.Server("fp32", hbm_bytes=budget, chunk_ms=5.0)
    rep = srv.run(rs, clock=Tick())
    assert len(rep.done) == 8 and rep.max_live < 8
    waited = [r for r in rep.done if r.admitted_at > r.arrival + 0.01]
    assert waited  # some requests queued for memory
    for r in rep.done[:3]:
        alone = S.run_alone(r, "fp32")
        assert np.array_equal(alone["displacement"], r.result["displacement"])
    srv.close()


def test_bad_requests_do_not_poison_the_mix():
    rs = _requests(6, seed=6)
    rs[1].velocity = rs[1].velocity.copy()
    rs[1].velocity[5, 1] = np.nan
    m = rs[4].mesh
    m.conn = m.conn[:, [0, 2, 1, 3] if m.etype == 0 else [4, 5, 6, 7, 0, 1, 2, 3]].copy()
    srv = S.Server("fp32", hbm_bytes=1 << 30, chunk_ms=5.0)
    rep = srv.run(rs, clock=Tick())
    assert sorted(r.rid for r in rep.failed) == [1, 4]
    print([r.error for r in rep.failed], flush=True)
    assert "non-finite" in rep.failed[0].error or "non-finite" in rep.failed[1].error
    assert any("jacobian" in r.error.lower() for r in rep.failed)
    assert len(rep.done) == 4
    for r in rep.done:
        alone = S.run_alone(r, "fp32")
        assert np.array_equal(alone["displacement"], r.result["displacement"])
    srv.close()


def test_large_requests_among_small_ones():
    """The upper end of the serving workload's sizes: a 2e6-element Hex8
    neo-Hookean request of 1,000 steps and a 1.2e5-element Tet4 elastic
    request of 300 steps, served in a stream with small requests that arrive
    while they run.
    - both large results are bit-identical to the request run alone;
    - the 1.2e5-element result is within 1e-5 of the FP64 oracle (on the
      host's cores), the 2e6-element one within 1e-5 of the plain FP64 path
      (bit-pinned to the oracle at this size by the full-size tests)."""
    import oracle
    from akantu2_amd import mesh as M
    from tests.helpers import make_gpu
    from tests.test_gpu_full_size import _oracle, _threads

    big = S.Request(0, 0.0, "_hexahedron_8", M.cells_for("_hexahedron_8", 2e6, 2.0),
                    "neohookean", dict(rho=1.0, E=1.3, nu=0.27), 1000, 20_000)
    mid = S.Request(1, 0.0, "_tetrahedron_4", M.cells_for("_tetrahedron_4", 1.2e5, 3.0),
                    "elastic", dict(rho=1.0, E=0.7, nu=0.31), 300, 20_001)
    small = _requests(6, seed=12, rate=1000.0, steps=(200, 400))
    for k, r in enumerate(small):
        r.rid = 2 + k
    rs = [big.materialize(), mid.materialize()] + small
    assert big.mesh.nb_elements >= 1.9e6 and mid.mesh.nb_elements >= 1e5
    srv = S.Server("fp32", hbm_bytes=8 << 30, chunk_ms=5.0)
    rep = srv.run(rs, clock=Tick())
    srv.close()
    print(f"served {len(rep.done)} in {rep.chunks} chunks, max live {rep.max_live}",
          flush=True)
    assert len(rep.done) == len(rs) and not rep.failed and rep.max_live >= 3
    for r in (big, mid):
        alone = S.run_alone(r, "fp32")
        for f in ("displacement", "velocity", "internal_force"):
            assert np.array_equal(alone[f], r.result[f]), (r.rid, f)
    o = _oracle(mid.mesh, mid.blocked, mid.velocity, mid.law, **mid.material)
    o.solve_steps(mid.steps)
    oracle.set_threads(1)
    eu = relerr(mid.result["displacement"], o.u)
    print(f"request 1: {mid.mesh.nb_elements} Tet4, {mid.steps} steps, u vs oracle {eu:.2e}"
          f" ({_threads()} oracle threads)", flush=True)
    assert eu <= 1e-5
    g, _ = make_gpu(big.mesh, big.blocked, big.velocity, big.law, **big.material)
    g.solveSteps(big.steps)
    eu = relerr(big.result["displacement"], g.getDisplacement())
    ev = relerr(big.result["velocity"], g.getVelocity())
    print(f"request 0: {big.mesh.nb_elements} Hex8, {big.steps} steps, u vs plain FP64 "
          f"{eu:.2e}, v {ev:.2e}", flush=True)
    assert eu <= 1e-5
