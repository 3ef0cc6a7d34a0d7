"""Many independent explicit simulations advanced together on one gfx950
device (the C ABI's ``akx_batch``).

Each instance is what the reference runs as its own ``SolidMechanicsModel``
with one material, ``_explicit_lumped_mass`` and zero external force
(solid_mechanics_model.cc:186-232, model_solver.cc:267-280).  All live
instances step together: one persistent element launch per (element type,
law) group and one persistent nodal gather launch per step, replayed from a
hipGraph of two steps.  An instance's results do not depend on which other
instances share the batch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .mesh import BoxMesh

_LAWS = {"elastic": N.LAW_ELASTIC, "elastic_fd": N.LAW_ELASTIC_FD,
         "neohookean": N.LAW_NEOHOOKEAN,
         "plastic_linear_isotropic_hardening": N.LAW_PLASTIC_LIH}
_PRECISIONS = {"fp64": N.PRECISION_FP64, "fp32": N.PRECISION_FP32}
_FIELDS = {"displacement": N.FIELD_DISPLACEMENT, "velocity": N.FIELD_VELOCITY,
           "acceleration": N.FIELD_ACCELERATION,
           "internal_force": N.FIELD_INTERNAL_FORCE, "nodes": N.FIELD_NODES}


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def law_id(law) -> int:
    return _LAWS[law] if isinstance(law, str) else int(law)


def footprint(precision, etype, law, nb_nodes, nb_elements) -> int:
    """Device bytes one instance takes in a batch."""
    out = C.c_int64()
    N.check(N.lib().akx_batch_footprint(_PRECISIONS[precision], int(etype), law_id(law),
                                        int(nb_nodes), int(nb_elements), C.byref(out)))
    return out.value


class Batch:
    """Device arenas for up to ``nodes`` nodes, ``tets`` Tet4 and ``hexes``
    Hex8 elements; ``plastic`` reserves the plastic history."""

    def __init__(self, precision="fp64", nodes=1 << 20, tets=0, hexes=0, plastic=False,
                 device=0):
        if N.device_count() == 0:
            raise RuntimeError("no gfx950 device visible: the MI355X path has no "
                               "CPU fallback")
        self.precision = precision
        h = C.c_void_p()
        N.check(N.lib().akx_batch_create(device, _PRECISIONS[precision], int(nodes),
                                         int(tets), int(hexes), int(bool(plastic)),
                                         C.byref(h)))
        self._h = h
        self._nn = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().akx_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def close(self):
        self.__del__()

    # ------------------------------------------------------------ instances
    def add(self, mesh: BoxMesh, law="elastic", rho=1.0, E=1.0, nu=0.3, sigma_y=0.0,
            h=0.0, blocked=None, velocity=None, dt_factor=0.5) -> int:
        """Admit one instance (initFull + getStableTimeStep + setTimeStep(
        dt_factor * stable dt)); returns its id."""
        params = np.array([rho, E, nu, sigma_y, h], np.float64)
        nodes = np.ascontiguousarray(mesh.nodes, np.float64)
        conn = np.ascontiguousarray(mesh.conn, np.int32)
        bl = None if blocked is None else np.ascontiguousarray(blocked, np.uint8)
        v = None if velocity is None else np.ascontiguousarray(velocity, np.float64)
        for a in (bl, v):
            if a is not None and a.shape != (nodes.shape[0], 3):
                raise ValueError("blocked / velocity must be [nb_nodes][3]")
        iid = C.c_int(-1)
        N.check(N.lib().akx_batch_add_instance(
            self._h, mesh.etype, law_id(law), params.ctypes.data_as(C.POINTER(C.c_double)),
            nodes.shape[0], _p(nodes), conn.shape[0], _p(conn), _p(bl), _p(v),
            float(dt_factor), C.byref(iid)))
        self._nn[iid.value] = nodes.shape[0]
        return iid.value

    def remove(self, iid: int):
        N.check(N.lib().akx_batch_remove_instance(self._h, int(iid)))
        self._nn.pop(iid, None)

    def timeStep(self, iid: int):
        dt, crit = C.c_double(), C.c_double()
        N.check(N.lib().akx_batch_get_time_step(self._h, int(iid), C.byref(dt),
                                                C.byref(crit)))
        return dt.value, crit.value

    def setTimeStep(self, iid: int, dt: float):
        N.check(N.lib().akx_batch_set_time_step(self._h, int(iid), float(dt)))

    def steps(self, iid: int) -> int:
        s = C.c_int64()
        N.check(N.lib().akx_batch_get_steps(self._h, int(iid), C.byref(s)))
        return s.value

    # ---------------------------------------------------------------- steps
    def solveStep(self, nb_steps: int = 1):
        N.check(N.lib().akx_batch_solve_step(self._h, int(nb_steps)))

    def synchronize(self):
        N.check(N.lib().akx_batch_synchronize(self._h))

    def get(self, iid: int, field: str) -> np.ndarray:
        out = np.empty((self._nn[iid], 3), np.float64)
        N.check(N.lib().akx_batch_get_field(self._h, int(iid), _FIELDS[field], _p(out)))
        return out

    def getDisplacement(self, iid):
        return self.get(iid, "displacement")

    def getVelocity(self, iid):
        return self.get(iid, "velocity")

    def getInternalForce(self, iid):
        return self.get(iid, "internal_force")

    # ---------------------------------------------------------- launch shape
    def setLaunch(self, block=256, tile=2048, blocks_per_cu=0, graph=True):
        """Threads per block, elements per tile and hipGraph on/off.  Every
        workgroup does one tile; blocks_per_cu (0..32) has no effect."""
        N.check(N.lib().akx_batch_set_launch(self._h, int(block), int(tile),
                                             int(blocks_per_cu), int(bool(graph))))

    def launch(self) -> dict:
        v = [C.c_int() for _ in range(4)]
        N.check(N.lib().akx_batch_get_launch(self._h, *[C.byref(x) for x in v]))
        return dict(block=v[0].value, tile=v[1].value, blocks_per_cu=v[2].value,
                    graph=bool(v[3].value))

    def autotune(self, repeats: int = 3) -> dict:
        blk, tile, bpc = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        N.check(N.lib().akx_batch_autotune(self._h, int(repeats), C.byref(blk),
                                           C.byref(tile), C.byref(bpc), C.byref(ms)))
        return dict(block=blk.value, tile=tile.value, blocks_per_cu=bpc.value,
                    element_pass_ms=ms.value)

    # ------------------------------------------------------------ inventory
    def stats(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        N.check(N.lib().akx_batch_get_stats(self._h, *[C.byref(x) for x in v]))
        return dict(live=v[0].value, elements=v[1].value,
                    quadrature_points=v[2].value, nodes=v[3].value)

    def capacityLeft(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        N.check(N.lib().akx_batch_capacity_left(self._h, *[C.byref(x) for x in v]))
        return dict(nodes=v[0].value, tets=v[1].value, hexes=v[2].value,
                    entries=v[3].value)

    def enableKernelTiming(self, enable=True):
        N.check(N.lib().akx_batch_enable_kernel_timing(self._h, int(enable)))

    def kernelTiming(self, kernel_class):
        """(total ms, launches) of kernel class 0 (element pass, all groups of
        one step) or 1 (gather + update) since timing was enabled."""
        ms, n = C.c_double(), C.c_int64()
        N.check(N.lib().akx_batch_get_kernel_timing(self._h, int(kernel_class),
                                                    C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernelBytes(self, kernel_class):
        b = C.c_double()
        N.check(N.lib().akx_batch_get_kernel_bytes(self._h, int(kernel_class), C.byref(b)))
        return b.value
