"""The C++ host API (include/akantu_mi355x.hpp), driven like an Akantu C++
driver by tests/cpp/drive_explicit.cc, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT
from tests.helpers import make_oracle, relerr, workload

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "tests", "cpp", "drive_explicit")


@pytest.mark.parametrize("etype,law,fd,shape,steps", [
    ("_tetrahedron_4", "elastic", 0, (20, 4, 4), 1000),
    ("_hexahedron_8", "neohookean", 0, (10, 5, 5), 300),
    ("_hexahedron_8", "elastic", 1, (8, 4, 4), 200),
    ("_hexahedron_8", "local_elastic", 0, (6, 4, 3), 150),  # a ConstitutiveLaw
    ("_tetrahedron_4", "local_elastic", 0, (10, 3, 3), 300),
    ("_hexahedron_8", "device_elastic", 0, (6, 4, 3), 150),  # Material::setDeviceStress
    ("_tetrahedron_4", "device_elastic", 0, (10, 3, 3), 300),
])
def test_cpp_driver_matches_oracle(tmp_path, etype, law, fd, shape, steps):
    assert os.path.exists(DRIVER), "build() compiles tests/cpp/drive_explicit"
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=2)
    d = str(tmp_path)
    mesh.nodes.astype(np.float64).tofile(os.path.join(d, "nodes.f64"))
    mesh.conn.astype(np.int32).tofile(os.path.join(d, "conn.i32"))
    blocked.astype(np.uint8).tofile(os.path.join(d, "blocked.u8"))
    v0.astype(np.float64).tofile(os.path.join(d, "v0.f64"))
    np.array([1.0, 1.0, 0.3, 0.0, 0.0, 0.5]).tofile(os.path.join(d, "params.f64"))
    r = subprocess.run([DRIVER, d, str(mesh.etype), law, str(fd), str(steps)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    oracle_law = "elastic_fd" if (law == "elastic" and fd) else law
    if law in ("local_elastic", "device_elastic"):  # the driver's own law
        oracle_law = "elastic"
    o, dt = make_oracle(mesh, blocked, v0, oracle_law)
    assert np.fromfile(os.path.join(d, "dt.f64"))[0] == dt
    o.solve_steps(steps)
    rd = lambda n: np.fromfile(os.path.join(d, n)).reshape(-1, 3)
    assert relerr(rd("u.f64"), o.u) <= 1e-12
    assert relerr(rd("v.f64"), o.v) <= 1e-12
    assert relerr(rd("fint.f64"), o.f_int) <= 1e-12
    import ctypes
    import oracle
    oracle.lib().orc_compute_all_stresses(ctypes.byref(o._s))
    stress = np.fromfile(os.path.join(d, "stress.f64")).reshape(-1, 9)
    assert relerr(stress, o.stress) <= 1e-12


@pytest.mark.parametrize("etype,law,fd,shape,steps", [
    ("_tetrahedron_4", "elastic", 0, (12, 4, 4), 301),
    ("_hexahedron_8", "neohookean", 0, (6, 4, 3), 200),
])
def test_cpp_batch_matches_oracle(tmp_path, etype, law, fd, shape, steps):
    """The C++ Batch class: one instance among two, against the oracle."""
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=4)
    d = str(tmp_path)
    mesh.nodes.astype(np.float64).tofile(os.path.join(d, "nodes.f64"))
    mesh.conn.astype(np.int32).tofile(os.path.join(d, "conn.i32"))
    blocked.astype(np.uint8).tofile(os.path.join(d, "blocked.u8"))
    v0.astype(np.float64).tofile(os.path.join(d, "v0.f64"))
    np.array([1.0, 1.0, 0.3, 0.0, 0.0, 0.5]).tofile(os.path.join(d, "params.f64"))
    r = subprocess.run([DRIVER, d, str(mesh.etype), law, str(fd), str(steps), "batch"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o, dt = make_oracle(mesh, blocked, v0, law)
    assert np.fromfile(os.path.join(d, "dt.f64"))[0] == dt
    o.solve_steps(steps)
    rd = lambda n: np.fromfile(os.path.join(d, n)).reshape(-1, 3)
    assert relerr(rd("u.f64"), o.u) <= 1e-12
    assert relerr(rd("v.f64"), o.v) <= 1e-12
    assert relerr(rd("fint.f64"), o.f_int) <= 1e-12


def test_cpp_held_reference_write(tmp_path):
    """`auto & u = model.getDisplacement()` held across the loop and edited
    after half the steps: the edit lands on the current state (live Array)."""
    mesh, blocked, v0 = workload("_tetrahedron_4", (8, 3, 3), jitter=0.1, seed=5)
    d = str(tmp_path)
    mesh.nodes.astype(np.float64).tofile(os.path.join(d, "nodes.f64"))
    mesh.conn.astype(np.int32).tofile(os.path.join(d, "conn.i32"))
    blocked.astype(np.uint8).tofile(os.path.join(d, "blocked.u8"))
    v0.astype(np.float64).tofile(os.path.join(d, "v0.f64"))
    np.array([1.0, 1.0, 0.3, 0.0, 0.0, 0.5]).tofile(os.path.join(d, "params.f64"))
    steps = 40
    r = subprocess.run([DRIVER, d, str(mesh.etype), "elastic", "0", str(steps), "held"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o, dt = make_oracle(mesh, blocked, v0, "elastic")
    o.solve_steps(steps // 2)
    o.u[7] += np.array([2e-5, -1e-5, 3e-5])
    o.solve_steps(steps - steps // 2)
    rd = lambda n: np.fromfile(os.path.join(d, n)).reshape(-1, 3)
    assert relerr(rd("u.f64"), o.u) <= 1e-12
    assert relerr(rd("v.f64"), o.v) <= 1e-12
    assert relerr(rd("fint.f64"), o.f_int) <= 1e-12
