"""FP32 helpers of the precision form (akx_math.hpp) run on the host.

akx_log1p<float> takes the Taylor series of log1p to x^7 for |x| < 1/16
(in Horner form with FMAs) and the library log1pf beyond; it is checked
against log1p in double over [-0.1, 0.1] and at tiny arguments.  The series
must stay within 1.5 FP32 epsilons of the exact value (its truncation is
below 5e-10 relative); a wrong sign or coefficient of any term up to x^5
moves it by more than 4 epsilons at |x| = 1/16."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_log1p_series_against_double(tmp_path):
    exe = tmp_path / "log1p_series"
    src = os.path.join(ROOT, "tests", "cpp", "log1p_series.cc")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                        "-ffp-contract=fast", src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    series, lib, n = out.stdout.split()
    print(out.stdout)
    assert int(n) > 700000
    assert float(series) <= 1.5
    assert float(lib) <= 2.0
