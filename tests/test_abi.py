"""C ABI: the library loads and exports every entry point include/*.h declares, and
argument validation that happens before any device work reports the documented codes
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

import estl_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(estl_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_declared_symbols():
    L = estl_amd.lib()
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version():
    assert estl_amd.lib().estl_version() >= 100


@pytest.mark.parametrize("n", [0, 3, 12, 4097, -8])
def test_non_power_of_two_rejected(n):
    h = ctypes.c_void_p()
    rc = estl_amd.lib().estl_plan_create(ctypes.byref(h), n, 1, -1, 0, 0)
    assert rc == estl_amd.ERR_INVALID
    assert h.value is None
    assert b"power of two" in estl_amd.lib().estl_last_error()


def test_bad_direction_and_batch_rejected():
    h = ctypes.c_void_p()
    assert estl_amd.lib().estl_plan_create(ctypes.byref(h), 64, 1, 0, 0, 0) == estl_amd.ERR_INVALID
    assert estl_amd.lib().estl_plan_create(ctypes.byref(h), 64, 0, -1, 0, 0) == estl_amd.ERR_INVALID
    assert estl_amd.lib().estl_plan_create(None, 64, 1, -1, 0, 0) == estl_amd.ERR_INVALID


def test_max_batch_budget_is_64bit():
    # 288 GB budget must not overflow (uOpenCL.h:323-337 truncated it to int)
    b = estl_amd.lib().estl_max_batch(4096, estl_amd.FP32, 288 * 10 ** 9)
    assert b == 288 * 10 ** 9 // (16 * 4096)
    assert estl_amd.lib().estl_max_batch(3, estl_amd.FP32, 10 ** 9) == estl_amd.ERR_INVALID


def test_server_rejects_bad_arguments_without_a_device():
    L = estl_amd.lib()
    h = ctypes.c_void_p()
    assert L.estl_server_create(None, 1 << 20, 1) == estl_amd.ERR_INVALID
    assert L.estl_server_create(ctypes.byref(h), 0, 1) == estl_amd.ERR_INVALID
    assert h.value is None
    rid = ctypes.c_longlong()
    assert L.estl_server_submit_fft(None, 64, 1, -1, 1, None, ctypes.byref(rid)) == estl_amd.ERR_INVALID
    assert L.estl_server_submit_conv(None, 64, 1, 1, 1, 4, 0, None, ctypes.byref(rid)) == estl_amd.ERR_INVALID
    assert L.estl_server_wait(None, 1, 0.0) == estl_amd.ERR_INVALID
    assert L.estl_server_release(None, 1) == estl_amd.ERR_INVALID
    assert L.estl_server_destroy(None) == estl_amd.OK


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message saying so, before
    any device work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU only" in (r.stdout + r.stderr)
