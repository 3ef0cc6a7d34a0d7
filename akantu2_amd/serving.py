"""Serving form: a request loop that runs many explicit simulations at once
on one gfx950 device by continuous batching.

A request is one simulation: a mesh, one material, initial conditions and
a number of steps.  The server keeps one ``Batch`` (device arenas) and loops:

1. requests that have arrived join a FIFO queue;
2. queued requests are admitted while their HBM footprint
   (``akx_batch_footprint``) fits the server's budget, a slot is free and the
   arenas have room; admission uploads the mesh and computes the lumped mass
   and the stable time step on the device, and the state then stays resident;
3. one chunk of steps advances every live request together (one persistent
   element launch per group and one gather launch per step, replayed from a
   hipGraph); the chunk ends when the nearest request finishes or after about
   ``chunk_ms`` of device time, so newly arrived requests wait at most one
   chunk;
4. finished requests are read back and retired, which frees their memory for
   the queue.

A request that cannot be admitted (inverted element, non-finite input, larger
than the whole budget) fails alone; the others are not affected, since every
instance's arithmetic is its own.  Results do not depend on the mix: a
request gives the same bits alone and served with others.

The reference has no serving layer; per request the work is the reference's
``SolidMechanicsModel::solveStep`` loop (model_solver.cc:267-280) on one
model.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import mesh as M
from .batch import Batch, footprint, law_id


@dataclass
class Request:
    rid: int
    arrival: float              # seconds after the server starts
    etype: str                  # "_tetrahedron_4" or "_hexahedron_8"
    cells: tuple                # (nx, ny, nz)
    law: str                    # "elastic" or "neohookean"
    material: dict              # rho, E, nu
    steps: int
    seed: int
    jitter: float = 0.1
    amplitude: float = 1e-3
    mesh: M.BoxMesh = None
    blocked: np.ndarray = None
    velocity: np.ndarray = None
    # filled by the server
    admitted_at: float = None
    finished_at: float = None
    done_steps: int = 0
    result: dict = None
    error: str = None
    iid: int = -1
    footprint: int = 0

    def materialize(self):
        """Build the mesh and initial conditions (the client's side)."""
        if self.mesh is None:
            self.mesh = M.box_mesh(self.etype, *self.cells, jitter=self.jitter,
                                   seed=self.seed)
            self.blocked = M.clamped_face_x0(self.mesh)
            self.velocity = M.initial_velocity(self.mesh, self.amplitude, self.seed)
        return self

    @property
    def work(self) -> int:
        """Element quadrature-point updates the request asks for."""
        return self.mesh.nb_elements * self.mesh.nqp * self.steps

    @property
    def time_to_result(self) -> float:
        return self.finished_at - self.arrival


@dataclass
class Report:
    done: list = field(default_factory=list)
    failed: list = field(default_factory=list)
    makespan: float = 0.0
    busy: float = 0.0          # seconds spent in chunks of steps
    chunks: int = 0
    max_live: int = 0

    @property
    def throughput(self) -> float:
        """Element quadrature-point updates per second over the makespan."""
        return sum(r.work for r in self.done) / self.makespan if self.makespan else 0.0


def synthetic_requests(n, rate, seed=5, scale=1.0, min_el=1e3, max_el=2e6,
                       steps=(200, 5000)):
    """A seeded Poisson stream: inter-arrival times ~ Exp(rate); Tet4 or Hex8
    boxes of log-uniform size in [min_el, max_el] * scale, aspect U(1, 4);
    elastic or neo-Hookean with E ~ U(0.5, 2), nu ~ U(0.2, 0.35); steps
    ~ U{steps}."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    out = []
    for k in range(n):
        etype = "_tetrahedron_4" if rng.random() < 0.5 else "_hexahedron_8"
        nel = np.exp(rng.uniform(np.log(min_el), np.log(max_el))) * scale
        cells = M.cells_for(etype, max(nel, 6.0), rng.uniform(1.0, 4.0))
        law = "elastic" if rng.random() < 0.5 else "neohookean"
        mat = dict(rho=1.0, E=float(rng.uniform(0.5, 2.0)), nu=float(rng.uniform(0.2, 0.35)))
        out.append(Request(k, float(t[k]), etype, cells, law, mat,
                           int(rng.integers(steps[0], steps[1] + 1)), 10_000 + k))
    return out


class Server:
    """Continuous-batching server on one device.

    hbm_bytes: device memory the live requests may take (admission budget);
    the arenas are sized for it.  max_live: slots.  chunk_ms: device time of
    one chunk of steps at most (admission latency)."""

    def __init__(self, precision="fp32", hbm_bytes=16 << 30, max_live=1024, chunk_ms=50.0,
                 device=0, tet_share=0.5, batch=None, footprint_fn=footprint):
        self.precision = precision
        self.budget = int(hbm_bytes)
        self.max_live = max_live
        self.chunk_ms = chunk_ms
        self.footprint = footprint_fn
        if batch is None:
            # arenas: the budget split between Tet4 and Hex8 requests by the
            # bytes one element of each takes (with its share of nodes)
            per_tet = footprint_fn(precision, M.TET4, 0, 1 << 20, 6 << 20) / (6 << 20)
            per_hex = footprint_fn(precision, M.HEX8, 0, 1 << 20, 1 << 20) / (1 << 20)
            tets = int(self.budget * tet_share / per_tet)
            hexes = int(self.budget * (1 - tet_share) / per_hex)
            # the arenas index with int32: at most 2^31 force-table entries
            # (one per element corner, 25 % slice padding)
            f = min(1.0, 1.6e9 / (1.25 * (4 * tets + 8 * hexes)))
            tets, hexes = int(tets * f), int(hexes * f)
            nodes = tets // 5 + hexes + 4096 * 256
            batch = Batch(precision, nodes=nodes, tets=tets, hexes=hexes, device=device)
        self.batch = batch
        self.used = 0

    def close(self):
        self.batch.close()

    # ---------------------------------------------------------------- admit
    def _admit(self, r: Request, now: float) -> bool:
        """True if admitted; False if it must wait; a request that cannot
        run at all is marked failed (and True is returned so it leaves the
        queue)."""
        r.materialize()
        r.footprint = self.footprint(self.precision, r.mesh.etype, r.law, r.mesh.nb_nodes,
                                     r.mesh.nb_elements)
        if r.footprint > self.budget:
            r.error = "request larger than the device budget"
            return True
        if self.used + r.footprint > self.budget:
            return False
        try:
            r.iid = self.batch.add(r.mesh, r.law, blocked=r.blocked, velocity=r.velocity,
                                   dt_factor=0.5, **r.material)
        except N.AkantuError as e:
            if e.code == -4:  # an arena has no contiguous room right now
                return False
            r.error = str(e)
            return True
        self.used += r.footprint
        r.admitted_at = now
        return True

    # ------------------------------------------------------------------ run
    def run(self, requests, clock=time.perf_counter) -> Report:
        rep = Report()
        pending = deque(sorted(requests, key=lambda q: q.arrival))
        waiting = deque()
        live = {}
        step_s = None  # device seconds per step of the current mix, measured
        t0 = clock()
        now = 0.0
        while pending or waiting or live:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            while waiting and len(live) < self.max_live:
                r = waiting[0]
                if not self._admit(r, now):
                    break
                waiting.popleft()
                if r.error is not None:
                    r.finished_at = clock() - t0
                    rep.failed.append(r)
                else:
                    live[r.iid] = r
                    step_s = None  # the mix changed
            if not live:
                if waiting:  # nothing live and the head still does not fit
                    r = waiting.popleft()
                    r.error = "could not be admitted into an empty server"
                    r.finished_at = clock() - t0
                    rep.failed.append(r)
                    continue
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - (clock() - t0), 0.05)))
                continue
            rep.max_live = max(rep.max_live, len(live))
            rem = min(r.steps - r.done_steps for r in live.values())
            if step_s is None:
                chunk = min(rem, 8)  # short chunk to measure the new mix
            else:
                chunk = min(rem, max(8, int(self.chunk_ms * 1e-3 / step_s)))
            c0 = clock()
            self.batch.solveStep(chunk)
            self.batch.synchronize()
            dtc = clock() - c0
            rep.busy += dtc
            rep.chunks += 1
            step_s = dtc / chunk
            for r in live.values():
                r.done_steps += chunk
            for iid in [i for i, r in live.items() if r.done_steps >= r.steps]:
                r = live.pop(iid)
                r.result = {"displacement": self.batch.getDisplacement(iid),
                            "velocity": self.batch.getVelocity(iid),
                            "internal_force": self.batch.getInternalForce(iid)}
                self.batch.remove(iid)
                self.used -= r.footprint
                r.finished_at = clock() - t0
                rep.done.append(r)
                step_s = None
        first = min((r.arrival for r in requests), default=0.0)
        rep.makespan = (clock() - t0) - first
        return rep


def run_alone(r: Request, precision="fp32") -> dict:
    """The request on its own (a fresh batch holding only it)."""
    r.materialize()
    b = Batch(precision, nodes=r.mesh.nb_nodes + 256,
              tets=r.mesh.nb_elements if r.mesh.etype == M.TET4 else 0,
              hexes=r.mesh.nb_elements if r.mesh.etype == M.HEX8 else 0)
    iid = b.add(r.mesh, r.law, blocked=r.blocked, velocity=r.velocity, dt_factor=0.5,
                **r.material)
    b.solveStep(r.steps)
    out = {"displacement": b.getDisplacement(iid), "velocity": b.getVelocity(iid),
           "internal_force": b.getInternalForce(iid)}
    b.close()
    return out


__all__ = ["Request", "Report", "Server", "synthetic_requests", "run_alone", "law_id"]
