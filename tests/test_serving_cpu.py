"""Host logic of the serving loop (akantu2_amd.serving.Server) on the CPU,
with a stand-in for the device batch that records what the server asks of
it: FIFO admission under the HBM budget and the slot limit, chunks that
never run a request past its step count, failures that stay with their
request, every admitted request finished and its memory given back."""
import numpy as np
import pytest

from akantu2_amd import _native as N
from akantu2_amd import serving as S


class FakeBatch:
    """Counts steps per instance; fields encode (instance steps, request)."""

    def __init__(self, refuse_memory_when_live=None):
        self.live = {}
        self.next = 0
        self.log = []
        self.refuse = refuse_memory_when_live

    def add(self, mesh, law, blocked=None, velocity=None, dt_factor=0.5, **mat):
        if not np.isfinite(velocity).all():
            raise N.AkantuError(-1, "non-finite initial velocity")
        if mesh.meta.get("inverted"):
            raise N.JacobianError(-3, "Negative jacobian")
        if self.refuse is not None and len(self.live) >= self.refuse:
            raise N.AkantuError(-4, "node arena full")
        iid = self.next
        self.next += 1
        self.live[iid] = [0, mesh.nb_nodes]
        self.log.append(("add", iid))
        return iid

    def remove(self, iid):
        del self.live[iid]
        self.log.append(("remove", iid))

    def solveStep(self, n):
        assert n > 0
        for v in self.live.values():
            v[0] += n
        self.log.append(("step", n, len(self.live)))

    def synchronize(self):
        pass

    def _f(self, iid):
        steps, nn = self.live[iid]
        return np.full((nn, 3), float(steps))

    getDisplacement = getVelocity = getInternalForce = _f


def fp(precision, etype, law, nn, nel):
    return 1000 * nel


class Tick:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-3
        return self.t


def _reqs(n=12, rate=200.0):
    rs = S.synthetic_requests(n, rate=rate, seed=1, scale=1e-3, min_el=1e3, max_el=2e4,
                              steps=(20, 90))
    return [r.materialize() for r in rs]


def test_all_requests_finish_with_their_own_step_count():
    b = FakeBatch()
    srv = S.Server(batch=b, footprint_fn=fp, hbm_bytes=1 << 40)
    rs = _reqs()
    rep = srv.run(rs, clock=Tick())
    assert len(rep.done) == len(rs) and not rep.failed
    for r in rep.done:
        # the fake field holds the steps the instance received: exactly its own
        assert (r.result["displacement"] == r.steps).all()
        assert r.finished_at >= r.arrival and r.admitted_at >= r.arrival
    assert srv.used == 0 and not b.live
    assert rep.max_live >= 2  # requests overlapped


def test_budget_and_slots_bound_admission():
    rs = _reqs(16)
    budget = 3 * max(fp(None, 0, 0, 0, r.mesh.nb_elements) for r in rs)
    b = FakeBatch()
    srv = S.Server(batch=b, footprint_fn=fp, hbm_bytes=budget, max_live=2)
    used_seen = []
    orig = b.solveStep

    def step(n):
        used_seen.append(srv.used)
        orig(n)
    b.solveStep = step
    rep = srv.run(rs, clock=Tick())
    assert len(rep.done) == len(rs)
    assert max(used_seen) <= budget
    assert rep.max_live <= 2
    # FIFO: admission order is arrival order
    adds = [e[1] for e in b.log if e[0] == "add"]
    assert adds == sorted(adds)
    order = sorted(rep.done, key=lambda r: r.admitted_at)
    assert [r.rid for r in order] == sorted(r.rid for r in rs)


def test_arena_refusal_waits_instead_of_failing():
    rs = _reqs(8)
    b = FakeBatch(refuse_memory_when_live=3)
    srv = S.Server(batch=b, footprint_fn=fp, hbm_bytes=1 << 40)
    rep = srv.run(rs, clock=Tick())
    assert len(rep.done) == 8 and not rep.failed
    assert rep.max_live == 3


def test_bad_requests_fail_alone():
    rs = _reqs(10)
    rs[3].velocity = rs[3].velocity.copy()
    rs[3].velocity[0, 0] = np.nan
    rs[6].mesh.meta["inverted"] = True
    b = FakeBatch()
    srv = S.Server(batch=b, footprint_fn=fp, hbm_bytes=1 << 40)
    rep = srv.run(rs, clock=Tick())
    assert sorted(r.rid for r in rep.failed) == [3, 6]
    assert all(r.error for r in rep.failed)
    assert len(rep.done) == 8
    for r in rep.done:
        assert (r.result["displacement"] == r.steps).all()


def test_oversized_request_fails_and_queue_moves_on():
    rs = _reqs(5)
    budget = max(fp(None, 0, 0, 0, r.mesh.nb_elements) for r in rs) - 1
    srv = S.Server(batch=FakeBatch(), footprint_fn=fp, hbm_bytes=budget)
    rep = srv.run(rs, clock=Tick())
    assert len(rep.failed) >= 1 and len(rep.done) + len(rep.failed) == 5


def test_poisson_stream_is_seeded():
    a = S.synthetic_requests(20, rate=3.0, seed=9)
    b = S.synthetic_requests(20, rate=3.0, seed=9)
    assert [(r.arrival, r.cells, r.steps, r.law) for r in a] == \
           [(r.arrival, r.cells, r.steps, r.law) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert (gaps > 0).all()
    assert all(200 <= r.steps <= 5000 for r in a)


def test_batch_refuses_without_gpu():
    from akantu2_amd.batch import Batch
    if N.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="no gfx950 device"):
        Batch("fp32", nodes=1024, tets=64)


def test_footprint_is_host_only_and_monotone():
    from akantu2_amd.batch import footprint
    a = footprint("fp32", 0, "elastic", 1000, 5000)
    b = footprint("fp32", 0, "elastic", 2000, 10000)
    c = footprint("fp64", 0, "elastic", 1000, 5000)
    d = footprint("fp32", 0, "plastic_linear_isotropic_hardening", 1000, 5000)
    assert 0 < a < b and a < c and a < d
