"""Shared set-up of matching GPU and oracle models for the parity tests."""
from __future__ import annotations

import numpy as np

import oracle
from akantu2_amd import mesh as M

LAW_IDS = {"elastic": oracle.ELASTIC, "elastic_fd": oracle.ELASTIC_FD,
           "neohookean": oracle.NEOHOOKEAN,
           "plastic_linear_isotropic_hardening": oracle.PLASTIC_LIH}


def relerr(a, b):
    """max |a - b| / max |b| (0 when both vanish)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    return 0.0 if num == 0 else num / (den if den > 0 else 1.0)


def workload(etype, shape, jitter=0.0, seed=0, amp=1e-3):
    mesh = M.box_mesh(etype, *shape, jitter=jitter, seed=seed)
    blocked = M.clamped_face_x0(mesh)
    v0 = M.initial_velocity(mesh, amp, seed)
    return mesh, blocked, v0


def make_oracle(mesh, blocked, v0, law="elastic", rho=1.0, E=1.0, nu=0.3,
                sigma_y=0.0, h=0.0, dt_factor=0.5, **per_el):
    o = oracle.OracleModel(mesh.etype, mesh.nodes, mesh.conn, law, rho=rho, E=E,
                           nu=nu, sigma_y=sigma_y, h=h, **per_el)
    o.init_full()
    o.blocked[:] = blocked
    o.v[:] = v0
    dt = dt_factor * o.stable_dt()
    o.set_time_step(dt)
    return o, dt


def make_gpu(mesh, blocked, v0, law="elastic", rho=1.0, E=1.0, nu=0.3,
             sigma_y=0.0, h=0.0, dt=None, precision="fp64", dt_factor=0.5,
             **per_el):
    from akantu2_amd import SolidMechanicsModel
    g = SolidMechanicsModel(mesh)
    g.addMaterial(law, rho=rho, E=E, nu=nu, sigma_y=sigma_y, h=h, **per_el)
    g.setPrecision(precision)
    g.initFull()
    g.getBlockedDOFs()[:] = blocked
    g.getVelocity()[:] = v0
    if dt is None:
        dt = dt_factor * g.getStableTimeStep()
    g.setTimeStep(dt)
    return g, dt
