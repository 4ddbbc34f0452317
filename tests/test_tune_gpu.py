"""GPU tests of the autotuner (estl_tune, ESTL_FLAG_AUTOTUNE, ESTL_TUNE_CACHE).

The tuner replaces the reference's one fixed plan per size (radix-16 passes with the
remainder last, FFT_OpenCL_Contiguous.h:55-60) by the fastest of a timed plan space.
Whatever it picks must compute the same transform, so every plan it chooses is checked
against the FP64 oracle at the workloads the benchmark tunes:

* batched rows N = 4096 x 16384 (seed 1), FP32 and FP16 storage, both directions;
* the single N = 2^26 transform (seed 2), FP32 and FP16 storage, over the whole
  spectrum against ``oracle.fft64`` (computed once for the module);
* a tune-cache file written by one process is what a fresh process builds from.

Bounds: FP32 relative L2 <= 1e-6 log2 N per row; FP16 storage <= 2^-11 sqrt(stores),
stores = 1 (pack) + the plan's launches (every launch stores binary16 once).
"""
import json
import math
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import estl_amd
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
U = 2.0 ** -11


def tol(n):
    return 1e-6 * math.log2(n)


def fp16_bound(launches):
    return U * math.sqrt(1 + launches)


def plan_shape(plan):
    d = plan.describe()
    if d["kind"] == "rows":
        return ("rows",)
    return ("passes",) + tuple((p["radix"], p["tile"]) for p in d["passes"])


def run_fp32(torch, plan, x):
    y = torch.empty_like(x)
    plan.execute(x, y)
    torch.cuda.synchronize()
    return y


def run_fp16(torch, plan, x, n, batch):
    x16 = torch.empty(batch * n * 2, dtype=torch.float16, device=x.device)
    y16 = torch.empty_like(x16)
    s_in = torch.empty(batch, dtype=torch.float32, device=x.device)
    s_out = torch.empty_like(s_in)
    estl_amd.pack_fp16(x, x16, s_in, n, batch)
    plan.execute_fp16(x16, s_in, y16, s_out)
    y = torch.empty_like(x)
    estl_amd.unpack_fp16(y16, s_out, y, n, batch)
    torch.cuda.synchronize()
    return y


def test_tune_report_lists_candidates_and_best_is_used(gpu):
    """estl_tune times the rows kernel and pass splits at (4096, 16384); the plan that
    FLAG_AUTOTUNE builds afterwards is the cached choice, and is the fastest candidate."""
    rep = estl_amd.tune(4096, 16384, estl_amd.FORWARD)
    cands = rep["candidates"]
    assert rep["n"] == 4096 and rep["batch"] == 16384
    assert any(c["plan"] == "rows" for c in cands)
    assert any(isinstance(c["plan"], list) for c in cands)
    best = min(cands, key=lambda c: c["us"])
    p = estl_amd.Plan(4096, 16384, estl_amd.FORWARD, flags=estl_amd.FLAG_AUTOTUNE)
    d = p.describe()
    if best["plan"] == "rows":
        assert d["kind"] == "rows"
    else:
        assert [q["radix"] for q in d["passes"]] == [1 << b for b in best["plan"]]


@pytest.mark.parametrize("direction", [estl_amd.FORWARD, estl_amd.BACKWARD])
def test_autotuned_4096x16384_fp32_matches_fp64(gpu, direction):
    import torch
    n, B = 4096, 16384
    x = torch.empty(B, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 1)
    p = estl_amd.Plan(n, B, direction, flags=estl_amd.FLAG_AUTOTUNE)
    y = run_fp32(torch, p, x)
    rows = [0, 1, 777, 4095, 8192, 16383]
    xs, ys = x[rows].cpu().numpy(), y[rows].cpu().numpy()
    ref = oracle.fft64(xs.astype(np.complex128), direction)
    for i in range(len(rows)):
        assert oracle.rel_l2(ys[i], ref[i]) <= tol(n), (rows[i], plan_shape(p))
    # every row: Parseval in FP64 on the device
    ex = (x.abs().double() ** 2).sum(dim=1)
    ey = (y.abs().double() ** 2).sum(dim=1)
    assert torch.max(torch.abs(ey / (n * ex) - 1)).item() < 1e-5


def test_autotuned_4096x16384_fp16_matches_fp64(gpu):
    import torch
    n, B = 4096, 16384
    x = torch.empty(B, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 1)
    p = estl_amd.Plan(n, B, estl_amd.FORWARD, estl_amd.FP16, estl_amd.FLAG_AUTOTUNE)
    y = run_fp16(torch, p, x, n, B)
    rows = [0, 5, 9999, 16383]
    xs, ys = x[rows].cpu().numpy(), y[rows].cpu().numpy()
    ref = oracle.fft64(xs.astype(np.complex128), -1)
    bound = fp16_bound(p.launches())
    for i in range(len(rows)):
        assert oracle.rel_l2(ys[i], ref[i]) <= bound, (rows[i], plan_shape(p))


@pytest.fixture(scope="module")
def large_2_26(gpu):
    """The N = 2^26 workload (seed 2) and its FP64 spectrum from the oracle (the
    oracle's FP64 restatement takes tens of seconds on one core at this size)."""
    import torch
    n = 1 << 26
    x = torch.empty(1, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 2)
    xh = x[0].cpu().numpy().astype(np.complex128)
    print(f"[large_2_26] computing the FP64 oracle spectrum of 2^26 values", flush=True)
    ref = oracle.fft64(xh, -1)
    return x, ref


@pytest.mark.timeout(600)
def test_autotuned_2_26_fp32_whole_spectrum(gpu, large_2_26):
    import torch
    x, ref = large_2_26
    n = x.shape[1]
    p = estl_amd.Plan(n, 1, estl_amd.FORWARD, flags=estl_amd.FLAG_AUTOTUNE)
    assert p.describe()["kind"] == "passes"
    y = run_fp32(torch, p, x)
    e = oracle.rel_l2(y[0].cpu().numpy(), ref)
    print(f"2^26 FP32 tuned {plan_shape(p)}: rel-L2 {e:.3e}", flush=True)
    assert e <= tol(n), (e, plan_shape(p))


@pytest.mark.timeout(600)
def test_autotuned_2_26_fp16_storage_whole_spectrum(gpu, large_2_26):
    """The FP16-storage variant of the 2^26 workload with the tuner's plan, and with the
    default plan, over the whole spectrum."""
    import torch
    x, ref = large_2_26
    n = x.shape[1]
    for flags in (estl_amd.FLAG_AUTOTUNE, 0):
        p = estl_amd.Plan(n, 1, estl_amd.FORWARD, estl_amd.FP16, flags)
        y = run_fp16(torch, p, x, n, 1)
        yh = y[0].cpu().numpy()
        assert np.all(np.isfinite(yh))
        e = oracle.rel_l2(yh, ref)
        bound = fp16_bound(p.launches())
        print(f"2^26 FP16 flags={flags} {plan_shape(p)}: rel-L2 {e:.3e} (bound {bound:.3e})", flush=True)
        assert e <= bound, (flags, e, bound, plan_shape(p))


_CHILD = textwrap.dedent("""
    import json, sys
    sys.path.insert(0, {root!r})
    import estl_amd
    estl_amd.init(0)
    n, batch, prec, mode = {n}, {batch}, {prec}, {mode!r}
    if mode == "tune":
        estl_amd.tune(n, batch, estl_amd.FORWARD, prec)
    p = estl_amd.Plan(n, batch, estl_amd.FORWARD, prec, estl_amd.FLAG_AUTOTUNE)
    print("PLAN " + json.dumps(p.describe()), flush=True)
""")


def _child(n, batch, prec, mode, cache):
    code = _CHILD.format(root=ROOT, n=n, batch=batch, prec=prec, mode=mode)
    env = dict(os.environ, ESTL_TUNE_CACHE=cache)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("PLAN ")][-1]
    return json.loads(line[5:])


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", [estl_amd.FP32, estl_amd.FP16])
def test_tune_cache_reload_in_fresh_process(gpu, tmp_path, prec):
    """A process tunes (2^20, 4) and writes the cache file; a fresh process builds the
    same plan from the file.  Then the file's entry is replaced by another valid plan
    that the tuner never proposes for 2^20 (split [4, 8, 8]: its three-pass candidates
    use radices 2^5..2^8 near the balanced split; tiles of 2^4, 2^3, 2^3 columns) and a
    third process builds exactly that one -- it read the file rather than re-tuning --
    and the plan it names computes the transform within the bound."""
    import torch
    n, batch = 1 << 20, 4
    cache = str(tmp_path / "tune_cache.txt")
    first = _child(n, batch, prec, "tune", cache)
    assert os.path.exists(cache)
    lines = open(cache).read().split("\n")
    lines = [l for l in lines if l.strip()]
    assert len(lines) == 1, lines
    second = _child(n, batch, prec, "load", cache)
    assert first == second
    # replace the cached choice by a different valid plan
    f = lines[0].split()
    arch = f[0]
    forced = f"{arch} 20 {batch} -1 passes 3 4 8 8 4 3 3 1.0\n"
    open(cache, "w").write(forced)
    third = _child(n, batch, prec, "load", cache)
    assert [q["radix"] for q in third["passes"]] == [16, 256, 256]
    assert [q["tile"] for q in third["passes"]] == [16, 8, 8]
    # the forced plan, built here from the same split and tiles, is correct
    x = torch.empty(batch, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 44)
    p = estl_amd.Plan(n, batch, estl_amd.FORWARD, prec, split=[4, 8, 8], tiles=[4, 3, 3])
    assert plan_shape(p) == ("passes", (16, 16), (256, 8), (256, 8))
    y = run_fp32(torch, p, x) if prec == estl_amd.FP32 else run_fp16(torch, p, x, n, batch)
    ref = oracle.fft64(x.cpu().numpy().astype(np.complex128), -1)
    bound = tol(n) if prec == estl_amd.FP32 else fp16_bound(p.launches())
    ys = y.cpu().numpy()
    for r in range(batch):
        assert oracle.rel_l2(ys[r], ref[r]) <= bound, r
