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

# source -> floating-point contraction of its translation unit
SOURCES = {"akx_model.hip": "off", "akx_fast.hip": "fast", "akx_batch.hip": "off"}
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
OBJDIR = os.path.join(HERE, "_build")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def build_library(verbose: bool = False) -> str:
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = []
    for src, contract in SOURCES.items():
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        jobs.append((obj, [HIPCC, *FLAGS, f"-ffp-contract={contract}", "-c",
                           os.path.join(CSRC, src), "-o", obj]))
    with ThreadPoolExecutor(len(jobs)) as ex:
        logs = list(ex.map(lambda j: _run(j[1]), jobs))
    tmp = LIB + ".tmp"
    logs.append(_run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp,
                      *[j[0] for j in jobs], "-lhiprtc"]))
    if verbose:
        print("".join(logs))
    os.replace(tmp, LIB)
    return LIB


CPP_DRIVER_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "drive_explicit.cc")
CPP_DRIVER = os.path.join(os.path.dirname(HERE), "tests", "cpp", "drive_explicit")


def build_cpp_driver() -> str:
    """The C++ API test driver (tests/cpp/drive_explicit.cc), linked against
    the in-tree library with an rpath so it runs from any checkout."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    _run(["g++", "-O2", "-std=c++17", f"-I{inc}", "-o", CPP_DRIVER, CPP_DRIVER_SRC,
          f"-L{HERE}", "-lakantu_mi355x", "-Wl,-rpath,$ORIGIN/../../akantu2_amd"])
    return CPP_DRIVER


if __name__ == "__main__":
    print(build_library(verbose=True))
