"""Build recipe for the oracle (test infrastructure).

* ``oracle/liboracle.so`` -- the C restatement in ``oracle.c`` (gcc).
* ``oracle/_ref/ref_layout`` -- a checker compiled from the reference's own host
  headers (``Complex.h``, ``DataInterface.h``, ``SimpleMath.h``) where they lie under
  /root/reference, driven by ``ref_layout.cpp``.
* ``oracle/_ref/dropin_check`` -- ``tests/cpp/dropin_check.cpp`` compiled against the
  reference's own ``FFT.h`` / ``DataInterface.h`` / ``Complex.h`` (its factory, FFT2D
  and data interfaces) and linked to ``estl_amd/libestl_hip.so``: the reference's
  interface driving the FFT_HIP backend.
Both are built only when that tree exists; the outputs stay under ``oracle/_ref/``
(git-ignored, carried to the GPU machine).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = "/root/reference"


def build_oracle(verbose: bool = True) -> str:
    out = os.path.join(HERE, "liboracle.so")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-o", out, os.path.join(HERE, "oracle.c"), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


def build_ref(verbose: bool = True) -> str | None:
    """Compile the reference-header checker, if the reference tree is present."""
    if not os.path.isdir(REFERENCE):
        if verbose:
            print("reference tree absent: oracle/_ref not rebuilt", flush=True)
        return None
    os.makedirs(os.path.join(HERE, "_ref"), exist_ok=True)
    out = os.path.join(HERE, "_ref", "ref_layout")
    cmd = ["g++", "-O1", "-w", "-I", REFERENCE, "-o", out, os.path.join(HERE, "ref_layout.cpp")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


def build_dropin(verbose: bool = True) -> str | None:
    """Compile the drop-in check against the reference's FFT interface headers (needs
    the reference tree and a built estl_amd/libestl_hip.so)."""
    repo = os.path.dirname(HERE)
    lib = os.path.join(repo, "estl_amd", "libestl_hip.so")
    if not os.path.isdir(REFERENCE) or not os.path.exists(lib):
        if verbose:
            print("reference tree or libestl_hip.so absent: oracle/_ref/dropin_check not rebuilt", flush=True)
        return None
    os.makedirs(os.path.join(HERE, "_ref"), exist_ok=True)
    out = os.path.join(HERE, "_ref", "dropin_check")
    cmd = ["g++", "-O1", "-std=c++17", "-w",
           "-I", REFERENCE,                                   # the reference's FFT.h first
           "-I", os.path.join(repo, "include"), "-I", os.path.join(repo, "include", "estl"),
           "-I", "/opt/rocm/include",
           os.path.join(repo, "tests", "cpp", "dropin_check.cpp"), "-o", out,
           "-L", os.path.join(repo, "estl_amd"), "-lestl_hip", "-Wl,-rpath,$ORIGIN/../../estl_amd",
           "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


if __name__ == "__main__":
    build_oracle()
    build_ref()
    build_dropin()
    sys.exit(0)
