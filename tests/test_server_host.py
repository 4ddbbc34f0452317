"""The request server on the CPU (no GPU): estl_amd/csrc/server.cpp compiled unchanged
over a host-only stand-in of the HIP runtime (tests/host/fake_hip) and CPU stand-ins of
the launch functions (tests/host/fake_launch.cpp), driven through its C ABI by
tests/host/server_host_test.cpp, in three builds: plain, AddressSanitizer + UBSan, and
ThreadSanitizer (ROCm's clang, whose TSan runtime understands the condition-variable
waits libstdc++ makes on a steady clock).

Cases (see the driver's header): a seeded mixed stream with request groups equal to
each request alone, bit for bit, and to an FP64 reference; admission denial and
malformed requests; launch failures when issued and on the device, whose neighbours
finish bit-identically; a failed request's queued launches never write into the next
request's memory; an output copy queued before release still reads that output;
shutdown with work in flight."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "host")
SOURCES = [os.path.join(ROOT, "estl_amd", "csrc", "server.cpp"), os.path.join(HOST, "fake_launch.cpp"),
           os.path.join(HOST, "server_host_test.cpp")]
CLANG = "/opt/rocm/llvm/bin/clang++"

BUILDS = {
    "plain": (["g++"], ["-O2"]),
    "asan": (["g++"], ["-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]),
    "tsan": ([CLANG], ["-O1", "-fsanitize=thread"]),
}
CASES = ["mix", "denial", "launchfail", "asyncfail", "stale", "copyout", "shutdown"]


@pytest.fixture(scope="module")
def binaries(tmp_path_factory):
    out = {}
    d = tmp_path_factory.mktemp("server_host")
    for name, (cc, flags) in BUILDS.items():
        if not shutil.which(cc[0]) and not os.path.exists(cc[0]):
            continue
        exe = str(d / f"server_host_{name}")
        cmd = [*cc, "-std=c++17", "-g", "-Wall", *flags, "-I", os.path.join(HOST, "fake_hip"), *SOURCES, "-o", exe,
               "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
        out[name] = exe
    return out


@pytest.mark.parametrize("build", list(BUILDS))
@pytest.mark.parametrize("case", CASES)
def test_server_host(binaries, build, case):
    if build not in binaries:
        pytest.skip(f"no compiler for the {build} build")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1")
    r = subprocess.run([binaries[build], case, "4"], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-3000:])
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-6000:])
    assert "WARNING: ThreadSanitizer" not in r.stderr
    assert "FAIL" not in r.stdout
    assert "OK: 0 failure(s)" in r.stdout


@pytest.mark.parametrize("launchers", [1, 2])
@pytest.mark.parametrize("case", ["mix", "launchfail"])
def test_server_host_fewer_launchers(binaries, case, launchers):
    """ESTL_SERVER_LAUNCHERS: fewer launcher threads than streams (each launcher takes
    its streams in turn), under ThreadSanitizer."""
    build = "tsan" if "tsan" in binaries else "plain"
    env = dict(os.environ, ESTL_SERVER_LAUNCHERS=str(launchers), TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([binaries[build], case, "5"], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-6000:])
    assert "WARNING: ThreadSanitizer" not in r.stderr
    assert "OK: 0 failure(s)" in r.stdout
