"""The overlap-save window choice on the CPU (no GPU): conv_ols_log2 / conv_ols_both
(estl_amd/csrc/plan.h) compiled over the host stand-in of the HIP runtime and checked by
tests/host/ols_model_test.cpp against the window lengths that measured clearly faster on
an MI355X (256 CUs), and against the bounds of each path (taps, row length,
ESTL_CONV_FULL_LENGTH)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "host")


def test_window_choice(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    exe = str(tmp_path / "ols_model_test")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(HOST, "fake_hip"),
                        os.path.join(HOST, "ols_model_test.cpp"), "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and "FAIL" not in r.stdout and "OK: 0 failure(s)" in r.stdout, r.stdout
