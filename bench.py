"""Benchmark of the FFT hot path on one MI355X.

Workload (one step): batched C2C FP32 forward + backward round trip, N = 4096,
batch 16384 rows (512 MiB per direction), re/im uniform[-1,1) from seed 1, resident in
HBM before the timed region.  Metric: whole-step algorithmic HBM throughput in GB/s
(16 bytes per complex value per direction: one read + one write), beside:

  * roofline of the dominant kernel (k_fft_rows, N=4096): algorithmic bytes per launch
    / mean launch duration from HIP events on the launch stream, against 8 TB/s;
  * the per-call form (one launch per transform, same kernels, same stream) and the
    batched/per-call throughput ratio;
  * p50/p99 step latency, and p50/p99 latency of a single N = 65536 transform;
  * relative L2 error of sampled rows against the FP64 oracle;
  * a CPU baseline: the oracle's FP32 restatement of the reference algorithm, one core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N = 4096
BATCH = 16384
SEED = 1
METRIC = "batched C2C fp32 GB/s & % HBM peak vs per-call baseline; p50/p99 latency"


_ROCTX = []


def leg(name: str):
    """A named range around one leg of the benchmark, recorded by a profiler run with
    `rocprofv3 --marker-trace` through rocprofiler-sdk's roctx library (nothing is
    recorded, and nothing changes, when that library is absent)."""
    import contextlib
    import ctypes
    if not _ROCTX:
        try:
            lib = ctypes.CDLL("/opt/rocm/lib/librocprofiler-sdk-roctx.so")
            lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
            _ROCTX.append(lib)
        except OSError:
            _ROCTX.append(None)

    @contextlib.contextmanager
    def rng():
        lib = _ROCTX[0]
        if lib is not None:
            lib.roctxRangePushA(name.encode())
        try:
            yield
        finally:
            if lib is not None:
                lib.roctxRangePop()
    return rng()


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * p / 100.0
    lo, hi = int(math.floor(k)), int(math.ceil(k))
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def load_traffic(leg: str = "rows"):
    """Calibrated HBM traffic of a leg from the committed rocprofv3 counter summaries
    (profiles/traffic.json, written from `bench.py --counters LEG` runs under
    `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`), if present."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(leg)


def cpu_baseline(rows: int, reps: int):
    import numpy as np
    import oracle
    x = oracle.uniform_complex(rows, N, SEED)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        y = oracle.fft32(x, oracle.FORWARD)
        oracle.fft32(y, oracle.BACKWARD)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    bytes_ = 2 * 16.0 * N * rows
    return {"value": round(bytes_ / t / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP32 restatement of FFT_OpenCL_Contiguous (radix-16 Stockham passes), "
                      f"{rows} rows x N={N}, forward + backward, median of {reps} runs, 1 thread"}


def cpu_library(rows: int, reps: int = 3):
    """FFTW3 is not installed on this image; the all-cores CPU library figure is
    scipy.fft (pocketfft) on the same workload, labelled as such (a baseline, not a
    target)."""
    try:
        import numpy as np
        import scipy.fft as sf
    except ImportError:
        return None
    import oracle
    workers = min(16, os.cpu_count() or 1)
    x = oracle.uniform_complex(rows, N, SEED)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        y = sf.fft(x, axis=1, workers=workers)
        sf.ifft(y, axis=1, norm="forward", workers=workers)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    return {"value": round(2 * 16.0 * N * rows / t / 1e9, 3), "unit": "GB/s", "cores": workers,
            "library": "scipy.fft (pocketfft); FFTW3 unavailable on this image",
            "sample": f"{rows} rows x N={N} complex64, forward + backward, median of {reps}"}


def bench_uneven(torch, estl_amd, stream, reps: int = 10):
    """8192 transforms, N = 2^U{8..18} (seed 3), 10 % unit tones: one graph-captured
    work-queue launch vs one plan launch per transform (same stream, device-resident)."""
    import numpy as np
    import estl_amd.workloads as wl
    b = wl.uneven_batch()
    x = torch.empty(b.total, dtype=torch.complex64, device="cuda:0")
    y = torch.empty_like(x)
    with torch.cuda.stream(stream):
        wl.fill_uneven(b, x, torch)
    stream.synchronize()
    ub = estl_amd.UnevenBatch(b.sizes.tolist())
    for _ in range(2):
        ub.execute(x, y, stream, graph=True)
    stream.synchronize()
    times = []
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        ub.execute(x, y, stream, graph=True)
        e.record(stream)
        times.append((s, e))
    stream.synchronize()
    ms = sorted(s.elapsed_time(e) for s, e in times)
    t_batched = ms[len(ms) // 2] * 1e-3
    # tones: exact spectrum N at bin f
    worst = 0.0
    for i in np.nonzero(b.tone)[0][:200]:
        n, off, f = int(b.sizes[i]), int(b.offsets[i]), int(b.freq[i])
        yi = y[off:off + n]
        peak = yi[f].item()
        err2 = (yi.abs().double() ** 2).sum().item() - abs(peak) ** 2 + abs(peak - n) ** 2
        worst = max(worst, math.sqrt(max(err2, 0.0)) / n)
    # per-call form
    plans = {}
    for n in sorted(set(b.sizes.tolist())):
        plans[n] = estl_amd.Plan(n, 1, estl_amd.FORWARD)
    xb, yb = x.data_ptr(), y.data_ptr()
    launches = sum(plans[n].launches() for n in b.sizes.tolist())

    def percall():
        for n, off in zip(b.sizes.tolist(), b.offsets.tolist()):
            plans[n].execute(xb + 8 * off, yb + 8 * off, stream)

    percall()
    stream.synchronize()
    t0 = time.perf_counter()
    percall()
    stream.synchronize()
    t_percall = time.perf_counter() - t0
    bytes_ = 16.0 * b.total
    return {"workload": "8192 C2C fp32 transforms, N = 2^U{8..18} (seed 3), 10% unit tones, packed",
            "elements": b.total, "batched_ms_p50": round(t_batched * 1e3, 4),
            "batched_ms_min": round(ms[0], 4),
            "batched_gbs": round(bytes_ / t_batched / 1e9, 1), "launches_batched": 1,
            "per_call_ms": round(t_percall * 1e3, 2), "per_call_gbs": round(bytes_ / t_percall / 1e9, 2),
            "launches_per_call": launches, "batched_over_per_call": round(t_percall / t_batched, 1),
            "tone_rel_l2_max": worst, "work_items": ub.info["items"], "workgroups": ub.info["workgroups"],
            "traffic": load_traffic("uneven")}


def bench_large(torch, estl_amd, stream, reps: int = 10, check: bool = True):
    """One N = 2^26 transform (seed 2): FP32 multi-pass plan (default split and the
    autotuned one) and the FP16-storage plan, each timed as the median of `reps`
    launches; relative L2 error of each result against the FP64 oracle."""
    import numpy as np
    import oracle
    n = 1 << 26
    x = torch.empty(n, dtype=torch.complex64, device="cuda:0")
    y = torch.empty_like(x)
    estl_amd.fill_uniform(x, 2, stream=stream)
    stream.synchronize()

    def timed(fn):
        fn()
        stream.synchronize()
        ev = []
        for _ in range(reps):
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record(stream)
            fn()
            e.record(stream)
            ev.append((s, e))
        stream.synchronize()
        return sorted(s.elapsed_time(e) for s, e in ev)[reps // 2]

    out = {"workload": "single C2C forward, N=2^26, uniform[-1,1) seed 2, HBM-resident"}
    ref = None
    if check:
        ref = oracle.fft64(x.cpu().numpy().astype(np.complex128), -1)
    for name, flags in (("fp32_default", 0), ("fp32_tuned", estl_amd.FLAG_AUTOTUNE)):
        p = estl_amd.Plan(n, 1, estl_amd.FORWARD, estl_amd.FP32, flags)
        ms = timed(lambda: p.execute(x, y, stream))
        out[name] = {"ms": round(ms, 4), "gbs": round(16.0 * n / (ms * 1e-3) / 1e9, 1),
                     "passes": [[s["radix"], s["tile"]] for s in p.describe().get("passes", [])],
                     "rel_l2_vs_fp64": oracle.rel_l2(y.cpu().numpy(), ref) if check else None}
    x16 = torch.empty(2 * n, dtype=torch.float16, device="cuda:0")
    y16 = torch.empty_like(x16)
    s_in = torch.empty(1, dtype=torch.float32, device="cuda:0")
    s_out = torch.empty_like(s_in)
    estl_amd.pack_fp16(x, x16, s_in, n, 1, stream)
    p16 = estl_amd.Plan(n, 1, estl_amd.FORWARD, estl_amd.FP16, estl_amd.FLAG_AUTOTUNE)
    ms = timed(lambda: p16.execute_fp16(x16, s_in, y16, s_out, stream))
    estl_amd.unpack_fp16(y16, s_out, y, n, 1, stream)
    stream.synchronize()
    out["fp16_storage"] = {"ms": round(ms, 4), "gbs": round(8.0 * n / (ms * 1e-3) / 1e9, 1),
                           "fp32_equivalent_gbs": round(16.0 * n / (ms * 1e-3) / 1e9, 1),
                           "scaling": "per-transform max-abs input scale, 1/sqrt(radix) per pass",
                           "passes": [[q["radix"], q["tile"]] for q in p16.describe().get("passes", [])],
                           "rel_l2_vs_fp64": oracle.rel_l2(y.cpu().numpy(), ref) if check else None}
    out["bound_fp32"] = 1e-6 * 26
    out["traffic"] = load_traffic("large")
    return out


def bench_fft2d(torch, estl_amd, stream, reps: int = 10):
    """2-D transform on the device (FFT2D's rows then columns, estl_plan2d_*): one
    4096 x 4096 complex64 array (128 MiB), uniform[-1,1) seed 5, forward, median of
    `reps` launches pairs timed with HIP events on the stream.  Accuracy: a unit 2-D
    tone transforms to exactly nx*ny at its bin (rel-L2 over the whole spectrum from
    Parseval) and forward + normalised backward returns the input."""
    nx = ny = 4096
    n = nx * ny
    x = torch.empty(ny, nx, dtype=torch.complex64, device="cuda:0")
    estl_amd.fill_uniform(x, 5, stream=stream)
    y = torch.empty_like(x)
    fwd = estl_amd.Plan2D(nx, ny)
    bwd = estl_amd.Plan2D(nx, ny, direction=estl_amd.BACKWARD, flags=estl_amd.FLAG_NORMALIZE)
    for _ in range(3):
        fwd.execute(x, y, stream)
    ev = []
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        fwd.execute(x, y, stream)
        e.record(stream)
        ev.append((s, e))
    stream.synchronize()
    ms = sorted(s.elapsed_time(e) for s, e in ev)[reps // 2]
    z = torch.empty_like(x)
    bwd.execute(y, z, stream)
    stream.synchronize()
    rt = (torch.linalg.vector_norm((z - x).to(torch.complex128)) /
          torch.linalg.vector_norm(x.to(torch.complex128))).item()
    fx, fy = 1234, 567
    j = torch.arange(ny, device="cuda:0", dtype=torch.float64).reshape(ny, 1)
    i = torch.arange(nx, device="cuda:0", dtype=torch.float64).reshape(1, nx)
    ph = 2 * math.pi * (torch.remainder(i * fx, nx) / nx + torch.remainder(j * fy, ny) / ny)
    t = torch.polar(torch.ones_like(ph), ph).to(torch.complex64)
    del ph
    torch.cuda.synchronize()  # t was made on torch's stream
    fwd.execute(t, y, stream)
    stream.synchronize()
    peak = complex(y[fy, fx].item())
    energy = (y.abs().double() ** 2).sum().item()
    tone = math.sqrt(max(energy - abs(peak) ** 2, 0.0) + abs(peak - n) ** 2) / n
    return {"workload": "2-D C2C fp32 forward, 4096 x 4096, uniform[-1,1) seed 5, HBM-resident",
            "ms": round(ms, 4), "launches": fwd.launches(),
            "gbs": round(16.0 * n / (ms * 1e-3) / 1e9, 1),
            "hbm_traffic_gbs": round(16.0 * n * fwd.launches() / (ms * 1e-3) / 1e9, 1),
            "gflops_5nlog2n": round(5.0 * n * math.log2(n) / (ms * 1e-3) / 1e9, 1),
            "tone_rel_l2": tone, "round_trip_rel_l2": rt, "bound": 1e-6 * math.log2(n)}


def _drive(srv, reqs, xp, hp, arrivals, filters=None, keep=(), fp16=None):
    """Submit request i at time arrivals[i] (s from start) -- with its own filter
    (hp[:k], spectrum computed inside the request) or, when `filters` maps (n, bank
    index) to a registered filter, with that -- and release each output as soon as the
    completion queue reports it; outputs of the indices in `keep` are held and returned."""
    import torch
    held, where = {}, {}
    keep = set(keep)
    i, count = 0, len(reqs.n)
    # the client holds its requests, its buffers' device addresses and its stream (the
    # pools are sized for the largest request)
    args = list(zip(reqs.n.tolist(), reqs.rows.tolist(), reqs.k.tolist(), reqs.flags.tolist()))
    banks = reqs.bank.tolist()
    px, ph = xp.data_ptr(), hp.data_ptr()
    p16 = (fp16[0].data_ptr(), fp16[1].data_ptr()) if fp16 is not None else None
    ps = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    while i < count or where:
        now = time.perf_counter() - t0
        while i < count and arrivals[i] <= now:
            n, rows, k, fl = args[i]
            if p16 is not None:        # (binary16 input pool, per-row scales)
                rid = srv.submit_conv_fp16(n, rows, p16[0], p16[1], ph, k, fl, stream=ps)
            elif filters is not None:
                rid = srv.submit_conv_filter(filters[(n, banks[i])], rows, px, fl, stream=ps)
            else:
                rid = srv.submit_conv(n, rows, px, ph, k, fl, stream=ps)
            where[rid] = i
            i += 1
        wait = max(0.0, arrivals[i] - (time.perf_counter() - t0)) if i < count else 0.05
        for rid in srv.next_done(1024, min(wait, 0.05)):
            j = where.pop(rid)
            if j in keep:
                held[j] = rid
            else:
                srv.release(rid)
    return time.perf_counter() - t0, held


def _capacity(srv, drive, runs: int = 3):
    """The stream submitted all at once, `runs` times: the median run's wall time and
    server statistics (a run of the 1000 requests takes ~30 ms, and runs differ by a
    few percent)."""
    res = []
    for _ in range(runs):
        srv.reset_stats()
        wall, _ = drive()
        res.append((wall, srv.stats()))
    res.sort(key=lambda r: r[0])
    return res[len(res) // 2]


def bench_serve(torch, estl_amd, stream, load: float = 0.7, budget: int = 96 << 30, streams: int = 8):
    """Served convolution stream: 1000 requests (seed 4), n = 2^U{10..22}, rows
    U{1..256} at every length (up to 2^30 values, 4 GiB, per request), each with its own
    filter of U{16..8192} values (uniform[-1,1)), 25 % correlations, fp32.  Every
    request reads its rows from one 4 GiB input pool.  All requests at once (capacity,
    the median of three runs), then
    Poisson arrivals at `load` x capacity: p50/p99 latency (submit -> completion
    observed), throughput, a bitwise alone-vs-mix check; beside two plain forms on one
    stream -- one blocking set_filter + execute per request, and one execute per
    transform (each real row on its own) -- the filter-bank form (32 filters
    registered once per n, resident in the arena) and the FP16-storage mode."""
    import estl_amd.workloads as wl
    reqs = wl.conv_requests()
    count = len(reqs.n)
    maxv = int(reqs.values().max())
    xp = torch.empty(maxv, dtype=torch.float32, device="cuda:0")
    hp = torch.empty(8192, dtype=torch.float32, device="cuda:0")
    estl_amd.fill_uniform(xp, 4, stream=stream)
    estl_amd.fill_uniform(hp, 4, offset=maxv, stream=stream)
    stream.synchronize()
    alg_bytes = float(8 * reqs.values().sum() + 4 * reqs.k.sum())
    srv = estl_amd.Server(budget, streams=streams)
    # warm the shapes once (one request per (n, rows, flags)), untimed
    shapes = {}
    for j in range(count):
        shapes.setdefault((int(reqs.n[j]), int(reqs.rows[j]), int(reqs.flags[j])), j)
    for j in shapes.values():
        rid = srv.submit_conv(int(reqs.n[j]), int(reqs.rows[j]), xp, hp, int(reqs.k[j]), int(reqs.flags[j]))
        srv.wait(rid, 60)
        srv.release(rid)
    srv.next_done(4096, 0.0)
    # capacity: everything submitted at t = 0 (median of three runs)
    with leg("serve.capacity"):
        wall_cap, cap = _capacity(srv, lambda: _drive(srv, reqs, xp, hp, [0.0] * count))
    capacity = count / wall_cap
    # open loop: Poisson arrivals at load x capacity
    rate = load * capacity
    arrivals = (reqs.gaps / rate).cumsum()
    sample = [0, 1, 2, 500, 999]
    srv.reset_stats()
    with leg("serve.load"):
        wall, held = _drive(srv, reqs, xp, hp, arrivals, keep=sample)
    st = srv.stats()
    same = True
    for j, rid in held.items():
        n, rows, k, fl = int(reqs.n[j]), int(reqs.rows[j]), int(reqs.k[j]), int(reqs.flags[j])
        y = srv.output(rid, (rows * n,), torch.float32)
        c = estl_amd.Convolution(n, rows, fl)
        c.set_filter(hp, k)
        z = torch.empty_like(y)
        c.execute(xp, z)
        torch.cuda.synchronize()
        same = same and torch.equal(y, z)
        srv.release(rid)
    # filter bank: 32 filters registered once per n (spectra resident), capacity only
    filters = {}
    for n, b in sorted(set(zip(reqs.n.tolist(), reqs.bank.tolist()))):
        filters[(n, b)] = srv.register_filter(n, hp, int(reqs.bank_k[b]))
    for f in filters.values():
        srv.wait(f, 60)
    with leg("serve.filter_bank"):
        wall_bank, bank = _capacity(srv, lambda: _drive(srv, reqs, xp, hp, [0.0] * count, filters))
    # FP16 storage: the same stream on binary16 inputs (unit row scales: the pool is
    # already in [-1, 1)), outputs binary16 with per-row scales
    x16 = xp.to(torch.float16)
    ones = torch.ones(256, dtype=torch.float32, device="cuda:0")
    for j in shapes.values():
        rid = srv.submit_conv_fp16(int(reqs.n[j]), int(reqs.rows[j]), x16, ones, hp, int(reqs.k[j]),
                                   int(reqs.flags[j]))
        srv.wait(rid, 60)
        srv.release(rid)
    srv.next_done(4096, 0.0)
    with leg("serve.fp16_capacity"):
        wall_cap16, cap16 = _capacity(srv, lambda: _drive(srv, reqs, xp, hp, [0.0] * count, fp16=(x16, ones)))
    rate16 = load * count / wall_cap16
    arrivals16 = (reqs.gaps / rate16).cumsum()
    srv.reset_stats()
    wall16, held16 = _drive(srv, reqs, xp, hp, arrivals16, keep=sample, fp16=(x16, ones))
    st16 = srv.stats()
    same16 = True
    for j, rid in held16.items():
        n, rows, k, fl = int(reqs.n[j]), int(reqs.rows[j]), int(reqs.k[j]), int(reqs.flags[j])
        y16, s16 = srv.output_fp16(rid, n, rows)
        c = estl_amd.Convolution(n, rows, fl, precision=estl_amd.FP16)
        c.set_filter(hp, k)
        z16, t16 = torch.empty_like(y16), torch.empty_like(s16)
        c.execute_fp16(x16, ones, z16, t16)
        torch.cuda.synchronize()
        same16 = same16 and torch.equal(y16, z16) and torch.equal(s16, t16)
        srv.release(rid)
    alg16 = float(4 * reqs.values().sum() + 4 * reqs.k.sum())
    fp16_mode = {"capacity_req_s": round(count / wall_cap16, 1), "capacity_gbs": round(alg16 / wall_cap16 / 1e9, 1),
                 "offered_req_s": round(rate16, 1), "p50_ms": st16["p50_ms"], "p99_ms": st16["p99_ms"],
                 "throughput_req_s": round(count / wall16, 1), "alone_equals_mix": same16,
                 "bytes_per_value": 4, "scaling": "per-row input scales; output scale = input scale x sum|h|"}
    failed = cap["failed"] + st["failed"] + bank["failed"] + cap16["failed"] + st16["failed"]
    srv.close()
    # plain form: one blocking call per request.  The objects (one per shape, each with
    # its own workspace) are built outside the timed calls and kept while their
    # workspace totals at most 32 GiB (least recently used dropped first); the plain
    # form's time is the sum of the blocking calls.
    y = torch.empty(maxv, dtype=torch.float32, device="cuda:0")
    sp = stream.cuda_stream
    objs, held_ws = {}, [0]

    def obj(key):
        c = objs.pop(key, None)
        if c is None:
            while objs and held_ws[0] > 32 << 30:
                old = objs.pop(next(iter(objs)))
                held_ws[0] -= old.workspace_bytes()
                old.close()
            c = estl_amd.Convolution(key[0], key[1], key[2])
            held_ws[0] += c.workspace_bytes()
        objs[key] = c
        return c

    def plain():
        spent = 0.0
        for j in range(count):
            n, rows, k, fl = int(reqs.n[j]), int(reqs.rows[j]), int(reqs.k[j]), int(reqs.flags[j])
            c = obj((n, rows, fl))
            stream.synchronize()
            t0 = time.perf_counter()
            c.set_filter(hp, k, sp)
            c.execute(xp, y, sp)
            stream.synchronize()
            spent += time.perf_counter() - t0
        return spent

    plain()
    wall_plain = plain()
    for c in objs.values():
        c.close()
    del objs
    singles = {}
    for j in range(count):
        key = (int(reqs.n[j]), int(reqs.flags[j]))
        if key not in singles:
            singles[key] = estl_amd.Convolution(key[0], 1, key[1])
    launches = 0

    def per_transform():
        nonlocal launches
        launches = 0
        for j in range(count):
            n, rows, k, fl = int(reqs.n[j]), int(reqs.rows[j]), int(reqs.k[j]), int(reqs.flags[j])
            c = singles[(n, fl)]
            c.set_filter(hp, k, sp)
            xb, yb = xp.data_ptr(), y.data_ptr()
            for r in range(rows):
                c.execute(xb + 4 * r * n, yb + 4 * r * n, sp)
            launches += rows
        stream.synchronize()

    per_transform()
    t0 = time.perf_counter()
    per_transform()
    wall_pt = time.perf_counter() - t0
    del singles
    return {"workload": "served convolution stream: 1000 requests (seed 4), n=2^U{10..22}, filter k=U{16..8192} "
                        "per request, rows U{1..256} at every n (up to 2^30 values per request), 25% correlation, "
                        "fp32, Poisson arrivals",
            "values": int(reqs.values().sum()), "max_request_values": maxv,
            "budget_bytes": budget, "streams": streams, "capacity_runs": 3,
            "capacity_req_s": round(capacity, 1), "capacity_gbs": round(alg_bytes / wall_cap / 1e9, 1),
            "capacity_p50_ms": cap["p50_ms"], "capacity_p99_ms": cap["p99_ms"],
            "capacity_groups": cap["groups"], "capacity_grouped_requests": cap["grouped_requests"],
            "offered_req_s": round(rate, 1), "load": load,
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "load_grouped_requests": st["grouped_requests"],
            "throughput_req_s": round(count / wall, 1), "throughput_gbs": round(alg_bytes / wall / 1e9, 1),
            "plain_per_request_req_s": round(count / wall_plain, 1),
            "served_over_plain_per_request": round(wall_plain / wall_cap, 2),
            "plain_per_transform_req_s": round(count / wall_pt, 1), "per_transform_executes": launches,
            "served_over_plain_per_transform": round(wall_pt / wall_cap, 2),
            "filter_bank_capacity_req_s": round(count / wall_bank, 1), "filter_bank_registered": len(filters),
            "arena_peak_bytes": st["arena_peak_bytes"], "peak_inflight": st["peak_inflight"],
            "rejected": st["rejected"], "failed": failed, "alone_equals_mix": same, "fp16_storage": fp16_mode,
            "traffic": load_traffic("served")}


def counters_leg(leg_name, torch, estl_amd, stream, fwd, x, y):
    """The launches one `rocprofv3 --pmc` run measures: set-up first (plans, tuning,
    inputs), then a k_fill_uniform launch that marks where the measured dispatches
    start, then the leg's launches.  Returns what the counter summary needs (the
    algorithmic bytes of the measured launches)."""
    marker = torch.empty(256, dtype=torch.float32, device="cuda:0")
    if leg_name == "rows":
        estl_amd.fill_uniform(marker, 0, stream=stream)
        for _ in range(3):
            fwd.execute(x, y, stream)
        stream.synchronize()
        return {"launches": 3, "algorithmic_bytes_per_launch": 16 * N * BATCH}
    if leg_name == "large":
        n = 1 << 26
        xl = torch.empty(n, dtype=torch.complex64, device="cuda:0")
        yl = torch.empty_like(xl)
        estl_amd.fill_uniform(xl, 2, stream=stream)
        p = estl_amd.Plan(n, 1, estl_amd.FORWARD, estl_amd.FP32, estl_amd.FLAG_AUTOTUNE)
        x16 = torch.empty(2 * n, dtype=torch.float16, device="cuda:0")
        y16 = torch.empty_like(x16)
        s_in = torch.empty(1, dtype=torch.float32, device="cuda:0")
        s_out = torch.empty_like(s_in)
        estl_amd.pack_fp16(xl, x16, s_in, n, 1, stream)
        p16 = estl_amd.Plan(n, 1, estl_amd.FORWARD, estl_amd.FP16, estl_amd.FLAG_AUTOTUNE)
        stream.synchronize()
        estl_amd.fill_uniform(marker, 0, stream=stream)
        for _ in range(3):
            p.execute(xl, yl, stream)
        for _ in range(3):
            p16.execute_fp16(x16, s_in, y16, s_out, stream)
        stream.synchronize()
        return {"fp32_passes": [[q["radix"], q["tile"]] for q in p.describe()["passes"]],
                "fp16_passes": [[q["radix"], q["tile"]] for q in p16.describe()["passes"]],
                "executes": 3, "fp32_algorithmic_bytes_per_execute": 16 * n,
                "fp16_algorithmic_bytes_per_execute": 8 * n}
    if leg_name == "uneven":
        import estl_amd.workloads as wl
        b = wl.uneven_batch()
        xu = torch.empty(b.total, dtype=torch.complex64, device="cuda:0")
        yu = torch.empty_like(xu)
        with torch.cuda.stream(stream):
            wl.fill_uneven(b, xu, torch)
        ub = estl_amd.UnevenBatch(b.sizes.tolist())
        ub.execute(xu, yu, stream)
        stream.synchronize()
        estl_amd.fill_uniform(marker, 0, stream=stream)
        for _ in range(3):
            ub.execute(xu, yu, stream)
        stream.synchronize()
        return {"launches": 3, "algorithmic_bytes_per_launch": 16 * int(b.total)}
    # served: one capacity run of the stream (every request with its own filter)
    import estl_amd.workloads as wl
    reqs = wl.conv_requests()
    count = len(reqs.n)
    maxv = int(reqs.values().max())
    xp = torch.empty(maxv, dtype=torch.float32, device="cuda:0")
    hp = torch.empty(8192, dtype=torch.float32, device="cuda:0")
    estl_amd.fill_uniform(xp, 4, stream=stream)
    estl_amd.fill_uniform(hp, 4, offset=maxv, stream=stream)
    stream.synchronize()
    srv = estl_amd.Server(96 << 30, streams=8)
    torch.cuda.synchronize()
    estl_amd.fill_uniform(marker, 0)
    torch.cuda.synchronize()
    wall, _ = _drive(srv, reqs, xp, hp, [0.0] * count)
    st = srv.stats()
    srv.close()
    return {"requests": count, "completed": st["completed"], "groups": st["groups"],
            "algorithmic_bytes": float(8 * reqs.values().sum() + 4 * reqs.k.sum())}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-size batch leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served convolution stream leg")
    ap.add_argument("--no-large", action="store_true", help="skip the N = 2^26 (FP32 and FP16 storage) leg")
    ap.add_argument("--no-2d", action="store_true", help="skip the 4096 x 4096 2-D transform leg")
    ap.add_argument("--percall-rows", type=int, default=BATCH)
    ap.add_argument("--graph", action="store_true",
                    help="replay the step's two launches as one HIP graph (measured: no gain over eager launches)")
    ap.add_argument("--counters", nargs="?", const="rows", default=None,
                    choices=["rows", "large", "uneven", "served"],
                    help="counter-collection mode for rocprofv3 --pmc: after a marker launch "
                         "(k_fill_uniform), the leg's launches -- rows: 3 forward launches of the batched "
                         "plan; large: 3 executes each of the tuned FP32 and FP16-storage 2^26 plans; "
                         "uneven: 3 eager work-queue launches of the 8192-size batch; served: one "
                         "capacity run of the served stream -- then 3 plain 512 MiB device copies "
                         "(the calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              f"(the batched workload would shard by batch row, each GPU transforming its own rows "
              f"with no exchange between GPUs, so its throughput is N times the one-GPU value)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import estl_amd
    import oracle

    estl_amd.init(0)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(dev)
    x = torch.empty(BATCH, N, dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    with torch.cuda.stream(stream):
        estl_amd.fill_uniform(x, SEED, stream=stream)
    stream.synchronize()

    # plans chosen by the autotuner (rows kernel vs pass splits and tiles, timed on the
    # device, cached per shape)
    fwd = estl_amd.Plan(N, BATCH, estl_amd.FORWARD, flags=estl_amd.FLAG_AUTOTUNE)
    bwd = estl_amd.Plan(N, BATCH, estl_amd.BACKWARD, flags=estl_amd.FLAG_AUTOTUNE)
    plan_desc = fwd.describe()

    if args.counters:
        info = counters_leg(args.counters, torch, estl_amd, stream, fwd, x, y)
        with torch.cuda.stream(stream):
            for _ in range(3):
                z.copy_(x)
        stream.synchronize()
        print(json.dumps({"counters_mode": args.counters, **info,
                          "copy_bytes_read": 8 * N * BATCH, "copy_bytes_written": 8 * N * BATCH}), flush=True)
        return 0

    def step():
        fwd.execute(x, y, stream)
        bwd.execute(y, z, stream)

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    run_step = step
    if args.graph:
        # the step's two launches captured once and replayed as one graph (no memset or
        # allocation is recorded: the plans' kernels need neither)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            step()
        graph.replay()
        stream.synchronize()
        run_step = graph.replay

    # timed region: exactly K steps, synchronised on both sides
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_step()
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    step_bytes = 2 * 16.0 * N * BATCH
    value = step_bytes / (wall / args.steps) / 1e9

    # per-launch durations with events on the launch stream (separate pass)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    for e in ev:
        e[0].record(stream)
        fwd.execute(x, y, stream)
        e[1].record(stream)
        bwd.execute(y, z, stream)
        e[2].record(stream)
    stream.synchronize()
    fwd_ms = [e[0].elapsed_time(e[1]) for e in ev]
    bwd_ms = [e[1].elapsed_time(e[2]) for e in ev]
    step_ms = [e[0].elapsed_time(e[2]) for e in ev]
    launch_bytes = 16.0 * N * BATCH
    fwd_mean = sum(fwd_ms) / len(fwd_ms)
    bwd_mean = sum(bwd_ms) / len(bwd_ms)
    achieved = launch_bytes / (fwd_mean * 1e-3) / 1e9

    # accuracy of the timed result: sampled rows against the FP64 oracle, round trip
    rows = [0, 1, 1234, 8191, BATCH - 1]
    xs = x[rows].cpu().numpy()
    err_fwd = max(oracle.rel_l2(y[r].cpu().numpy(), oracle.fft64(xs[i])) for i, r in enumerate(rows))
    rt = (torch.linalg.vector_norm((z / N - x).to(torch.complex128)) /
          torch.linalg.vector_norm(x.to(torch.complex128))).item()

    # the runtime's device-to-device copy of the same bytes (512 MiB read + 512 MiB
    # written, hipMemcpyAsync through torch), timed the same way: a plain read-once /
    # write-once reference beside the transform
    w = torch.empty_like(x)
    cev = []
    for _ in range(10):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            a.record(stream)
            w.copy_(x)
            b.record(stream)
        cev.append((a, b))
    stream.synchronize()
    copy_ms = sorted(a.elapsed_time(b) for a, b in cev)[5]
    copy_gbs = launch_bytes / (copy_ms * 1e-3) / 1e9
    # and the library's 16-byte streaming copy (plain and nontemporal loads / stores):
    # what a read-once / write-once kernel reaches, the achievable HBM rate
    stream_copy_gbs = {}
    for nt in (False, True):
        estl_amd.stream_copy(x, w, nt, stream)
        sev = []
        for _ in range(10):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            estl_amd.stream_copy(x, w, nt, stream)
            b.record(stream)
            sev.append((a, b))
        stream.synchronize()
        sc_ms = sorted(a.elapsed_time(b) for a, b in sev)[5]
        stream_copy_gbs["nontemporal" if nt else "plain"] = round(launch_bytes / (sc_ms * 1e-3) / 1e9, 1)
    assert torch.equal(w, x)
    copy16_gbs = max(stream_copy_gbs.values())
    del w

    # per-call form: one launch per transform, sequential, same stream, device-resident
    prow = min(args.percall_rows, BATCH)
    f1 = estl_amd.Plan(N, 1, estl_amd.FORWARD)
    b1 = estl_amd.Plan(N, 1, estl_amd.BACKWARD)
    xb, yb, zb = x.data_ptr(), y.data_ptr(), z.data_ptr()
    rb = N * 8

    def percall():
        for r in range(prow):
            f1.execute(xb + r * rb, yb + r * rb, stream)
        for r in range(prow):
            b1.execute(yb + r * rb, zb + r * rb, stream)

    percall()
    stream.synchronize()
    t0 = time.perf_counter()
    percall()
    stream.synchronize()
    pc_wall = time.perf_counter() - t0
    percall_gbs = 2 * 16.0 * N * prow / pc_wall / 1e9

    # single-transform latency, N = 65536 (batch 1), p50/p99 over 1000 launches
    n1 = 1 << 16
    a1 = torch.empty(1, n1, dtype=torch.complex64, device=dev)
    o1 = torch.empty_like(a1)
    estl_amd.fill_uniform(a1, 0, stream=stream)
    p1 = estl_amd.Plan(n1, 1, estl_amd.FORWARD)
    for _ in range(20):
        p1.execute(a1, o1, stream)
    lat = []
    for _ in range(1000):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        p1.execute(a1, o1, stream)
        e.record(stream)
        lat.append((s, e))
    stream.synchronize()
    lat_us = [1e3 * s.elapsed_time(e) for s, e in lat]
    err_c1 = oracle.rel_l2(o1.cpu().numpy()[0], oracle.fft64(a1.cpu().numpy()[0]))

    uneven = None if args.no_uneven else bench_uneven(torch, estl_amd, stream)
    serve = None if args.no_serve else bench_serve(torch, estl_amd, stream)
    large = None if args.no_large else bench_large(torch, estl_amd, stream)
    fft2d = None if args.no_2d else bench_fft2d(torch, estl_amd, stream)

    cpu = None if args.no_cpu else cpu_baseline(BATCH, 3)
    cpu_lib = None if args.no_cpu else cpu_library(BATCH)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched C2C fp32 forward+backward round trip, N=4096, batch 16384, "
                               "uniform[-1,1) seed 1, HBM-resident",
                   "n": N, "batch": BATCH, "seed": SEED, "plan": plan_desc, "hip_graph": bool(args.graph)},
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "gflops_5nlog2n": round(2 * 5.0 * N * math.log2(N) * BATCH / (wall / args.steps) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (load_traffic("rows") or {}).get("traffic_bytes_per_launch"),
                     "kernel": "k_fft_rows<12,-1,1> (forward, N=4096)",
                     "launch_ms_mean": round(fwd_mean, 5), "backward_launch_ms_mean": round(bwd_mean, 5),
                     "device_copy_gbs": round(copy_gbs, 1),
                     "over_device_copy": round(achieved / copy_gbs, 3),
                     "stream_copy16_gbs": stream_copy_gbs,
                     "over_stream_copy16": round(achieved / copy16_gbs, 3)},
        "per_call": {"value": round(percall_gbs, 2), "unit": "GB/s", "rows": prow,
                     "launches": 2 * prow, "batched_over_per_call": round(value / percall_gbs, 1)},
        "latency_us": {"step_p50": round(1e3 * percentile(step_ms, 50), 2),
                       "step_p99": round(1e3 * percentile(step_ms, 99), 2),
                       "n65536_batch1_p50": round(percentile(lat_us, 50), 2),
                       "n65536_batch1_p99": round(percentile(lat_us, 99), 2)},
        "rel_l2_vs_fp64": {"forward_sampled_rows": err_fwd, "round_trip": rt, "n65536": err_c1,
                           "bound": 1e-6 * math.log2(N)},
        "uneven_batch": uneven,
        "served_convolution": serve,
        "large_n": large,
        "fft2d": fft2d,
        "cpu_baseline": cpu,
        "cpu_library_all_cores": cpu_lib,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
