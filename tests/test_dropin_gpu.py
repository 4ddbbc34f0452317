"""The C++ drop-in layer on the GPU: tests/cpp/dropin_check.cpp drives FFT_HIP<float>
through an FFT<D> factory, FFT2D and setData/getData exactly as the reference's harness
drives its backends (tests/Test.cpp:45-184): the known-answer test (|diff|^2 <= 1e-3 M),
the 2-D differential test against a naive-DFT factory (|diff|^2 <= 1e-3 Mx My) and a
batched inverse on every row.

Two builds of the same source: against this repository's headers (dropin_check_own)
and against the reference's own FFT.h / DataInterface.h / Complex.h (oracle/_ref,
built where the reference tree exists).
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARIES = {
    "own headers": os.path.join(ROOT, "tests", "cpp", "dropin_check_own"),
    "reference headers": os.path.join(ROOT, "oracle", "_ref", "dropin_check"),
}


@pytest.mark.parametrize("which", sorted(BINARIES))
def test_dropin_check(gpu, which):
    exe = BINARIES[which]
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (build() builds it; the reference-header variant needs the reference tree)")
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(res.stdout[-3000:])
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert "OK: 0 failure(s)" in res.stdout
    assert res.stdout.count("PASS") >= 40
