"""Build libestl_hip.so in-tree with hipcc for gfx950.

    python -m estl_amd.build      (or __graft_entry__.build())

Every HIP source is compiled with the packed-FP32 target feature switched off
(``-Xclang -target-feature -Xclang -packed-fp32-ops``): hipcc -O3 otherwise forms
v_pk_* FP32 instructions on its own.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libestl_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(HERE, "..", "include"), "-Wall", "-Wno-unused-function"]

# (source, extra defines, object name)
SOURCES = [("kernels.hip", [], "kernels"),
           ("pass.hip", ["-DESTL_PASS_DIR=-1"], "pass_fwd"),
           ("pass.hip", ["-DESTL_PASS_DIR=+1"], "pass_bwd"),
           ("workqueue.hip", [], "workqueue"),
           ("fp16.hip", [], "fp16"),
           ("conv.hip", [], "conv"),
           ("noalloc.hip", [], "noalloc"),
           ("estl_c.cpp", [], "estl_c"),
           ("tune.cpp", [], "tune"),
           ("fp16_c.cpp", [], "fp16_c"),
           ("ubatch_c.cpp", [], "ubatch_c"),
           ("conv_c.cpp", [], "conv_c"),
           ("server.cpp", [], "server"),
           ("profile.cpp", [], "profile"),
           ("plan2d.cpp", [], "plan2d")]


def _compile(item, verbose: bool) -> str:
    src, defines, name = item
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, name + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC, *COMMON, *defines, *lang, "-c", path, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{out}")
    if out and verbose:
        print(out, flush=True)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s[0]))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_cpp_check(verbose: bool = True) -> str:
    """tests/cpp/dropin_check.cpp against this repository's C++ headers (include/estl),
    linked to the library: the C++ drop-in layer exercised by the GPU test suite."""
    repo = os.path.dirname(HERE)
    out = os.path.join(repo, "tests", "cpp", "dropin_check_own")
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(repo, "include"),
           "-I", os.path.join(repo, "include", "estl"), "-I", "/opt/rocm/include",
           os.path.join(repo, "tests", "cpp", "dropin_check.cpp"), "-o", out,
           "-L", HERE, "-lestl_hip", "-Wl,-rpath,$ORIGIN/../../estl_amd", "-L/opt/rocm/lib",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


if __name__ == "__main__":
    build()
    build_cpp_check()
    sys.exit(0)
