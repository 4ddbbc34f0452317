"""Build recipe for the oracle (test infrastructure).

* ``oracle/liboracle.so`` -- the C restatement in ``oracle.c`` (gcc).
* ``oracle/_ref/ref_layout`` -- a checker compiled from the reference's own host
  headers (``Complex.h``, ``DataInterface.h``, ``SimpleMath.h``) where they lie under
  /root/reference, driven by ``ref_layout.cpp``.  Built only when that tree exists;
  the output stays under ``oracle/_ref/`` (git-ignored, carried to the GPU machine).
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


if __name__ == "__main__":
    build_oracle()
    build_ref()
    sys.exit(0)
