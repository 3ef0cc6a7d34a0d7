"""The FP32 form's nodal update against the reference's, on the host.

The FP32 form keeps the half-step velocity w = v + dt/2 a between the steps
of one solve call and forms a = (f_ext + f_int) / M from the force alone;
the plain path runs the reference's corrector and predictor
(newmark-beta.cc:47-77, 164-197; non_linear_solver_lumped.cc:76-89).  Both
are the functions the kernels call (akx_math.hpp: reference_gather /
reference_predict, half_step_gather / half_step_predict), compiled for the
host and run in FP64 with FP32 compensation terms on a chain of springs with
lumped masses, an external force and blocked components.

- In exact arithmetic the two are the same central-difference scheme; after
  2,000 steps they agree to rounding (relative 1e-12; measured 3e-15 on u,
  2e-14 on v, 1e-13 on a, whose reference form a + (f - M a) / M cancels).
- A run split over calls (37, 1, 2, 1, ... steps) has the bits of one call.
- The blocked components and the system's motion are non-trivial (the far
  end moved)."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_half_step_equals_reference_update(tmp_path):
    exe = tmp_path / "half_step"
    src = os.path.join(ROOT, "tests", "cpp", "half_step.cc")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                        "-ffp-contract=off", src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    eu, ev, ea, bits, moved = out.stdout.split()
    print(out.stdout)
    assert float(eu) <= 1e-12 and float(ev) <= 1e-12 and float(ea) <= 1e-12
    assert bits == "1"
    assert float(moved) > 1e-6
