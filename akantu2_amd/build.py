"""In-tree build of the gfx950 library (hipcc; no CMake needed).

``build_library()`` compiles ``csrc/*.hip`` into
``akantu2_amd/libakantu_mi355x.so`` for gfx950.  Flags:
- ``-ffp-contract=off``: the FP64 plain path keeps the reference's separate
  rounding of every multiply and add (the FP32 form opts back into
  contraction inside its own code);
- ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  instructions.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libakantu_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["akx_model.hip"]
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
         "-Wno-unused-variable"]


def build_library(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
