"""C++ drop-in size limits on the host (no GPU): tests/cpp/limits_check.cpp built with
AddressSanitizer and UBSan against include/estl, linked to the library (only its
argument checks and the 64-bit batch query with an explicit budget are called)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "estl_amd")


@pytest.fixture(scope="module")
def limits_binary(tmp_path_factory):
    if not os.path.exists(os.path.join(LIB_DIR, "libestl_hip.so")):
        pytest.fail("estl_amd/libestl_hip.so is not built")
    out = str(tmp_path_factory.mktemp("limits") / "limits_check")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-Wall", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "include", "estl"),
           os.path.join(ROOT, "tests", "cpp", "limits_check.cpp"), "-o", out,
           "-L", LIB_DIR, "-lestl_hip", "-Wl,-rpath," + LIB_DIR, "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_dropin_size_limits_under_sanitizers(limits_binary):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", HIP_VISIBLE_DEVICES="")
    r = subprocess.run([limits_binary], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") >= 20
