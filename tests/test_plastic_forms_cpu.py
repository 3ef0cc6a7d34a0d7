"""The FP32 batch form's total-strain radial return (compact state: plastic
strain and hardening) against the reference's incremental radial return
(material_linear_isotropic_hardening_inline_impl.hh:31-86), both the device
header's code run on the host in FP64 along one seeded load / yield /
unload / reload path.  In exact arithmetic the two are the same map; in FP64
they agree to rounding, which pins the reformulation independently of the
FP32 rounding it is used with."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_total_form_equals_incremental_form(tmp_path):
    exe = tmp_path / "plastic_forms"
    src = os.path.join(ROOT, "tests", "cpp", "plastic_forms.cc")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                        "-ffp-contract=off", src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    ds, de, nyield, diso = out.stdout.split()
    print(out.stdout)
    assert int(nyield) > 50          # the path does yield, and for long
    assert float(ds) < 1e-11         # stress
    assert float(de) < 1e-11         # plastic strain
    assert float(diso) < 1e-12       # hardening
