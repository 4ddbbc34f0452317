"""GPU tests of the serving loop (estl_server_*): each request's result, served in a
concurrent mix, is bit-identical to the same request run alone through the batched
entry points (estl_conv_*, estl_plan_*), which are themselves checked against the
oracle; plus admission by HBM footprint, hold-until-release and input-stream ordering.
"""
import math

import numpy as np
import pytest

import estl_amd
import oracle

pytestmark = pytest.mark.gpu

FULL = estl_amd.CONV_FULL_LENGTH

# (kind, n, rows, k, flags): convolution rows of n real values with a k-value filter,
# or a batched complex FFT of `rows` rows of n (k unused)
MIX = [
    ("conv", 2 ** 10, 256, 16, 0),
    ("conv", 2 ** 14, 8, 8192, 0),
    ("fft", 4096, 16, 0, estl_amd.FORWARD),
    ("conv", 2 ** 16, 4, 1000, estl_amd.CONV_CORRELATE),
    ("conv", 2 ** 12, 33, 77, estl_amd.CONV_CENTER),
    ("fft", 2 ** 18, 2, 0, estl_amd.BACKWARD),
    ("conv", 2 ** 20, 1, 4096, 0),
    ("conv", 2 ** 13, 64, 300, estl_amd.CONV_NORMALIZE_KERNEL),
    ("fft", 64, 1000, 0, estl_amd.FORWARD),
    ("conv", 2 ** 22, 1, 16, 0),
]


def make_inputs(torch, gpu, case, seed):
    kind, n, rows, k, flags = case
    if kind == "fft":
        x = torch.empty(rows * n, dtype=torch.complex64, device=gpu)
        estl_amd.fill_uniform(x, seed)
        return x, None
    x = torch.empty(rows * n, dtype=torch.float32, device=gpu)
    estl_amd.fill_uniform(x, seed)
    h = torch.empty(k, dtype=torch.float32, device=gpu)
    estl_amd.fill_uniform(h, seed + 1000)
    if flags & estl_amd.CONV_NORMALIZE_KERNEL:
        h += 1.0  # a filter with a nonzero sum
    return x, h


def alone(torch, gpu, case, x, h):
    kind, n, rows, k, flags = case
    if kind == "fft":
        y = torch.empty_like(x)
        estl_amd.Plan(n, rows, flags).execute(x, y)
    else:
        y = torch.empty_like(x)
        c = estl_amd.Convolution(n, rows, flags)
        c.set_filter(h)
        c.execute(x, y)
    torch.cuda.synchronize()
    return y


def submit(srv, case, x, h):
    kind, n, rows, k, flags = case
    if kind == "fft":
        return srv.submit_fft(n, rows, x, flags)
    return srv.submit_conv(n, rows, x, h, k, flags)


def result(srv, rid, case, torch):
    kind, n, rows, k, flags = case
    dtype = torch.complex64 if kind == "fft" else torch.float32
    t = srv.output(rid, (rows * n,), dtype)
    torch.cuda.synchronize()
    return t


def test_mix_equals_alone_bitwise(gpu):
    import torch
    inputs = [make_inputs(torch, gpu, c, 40 + i) for i, c in enumerate(MIX)]
    refs = [alone(torch, gpu, c, x, h) for c, (x, h) in zip(MIX, inputs)]
    srv = estl_amd.Server(2 << 30, streams=4)
    # every request twice, interleaved, so each runs beside different neighbours
    order = list(range(len(MIX))) * 2
    rng = np.random.default_rng(5)
    rng.shuffle(order)
    ids = [(i, submit(srv, MIX[i], *inputs[i])) for i in order]
    srv.drain(120)
    for i, rid in ids:
        srv.wait(rid, 60)
        y = result(srv, rid, MIX[i], torch)
        assert torch.equal(y, refs[i]), (MIX[i], rid)
        srv.release(rid)
    st = srv.stats()
    assert st["completed"] == len(ids) and st["failed"] == 0
    assert st["p50_ms"] > 0 and st["p99_ms"] >= st["p50_ms"]
    assert st["arena_used_bytes"] == 0
    assert st["peak_inflight"] >= 2
    srv.close()


def test_large_requests_beside_small_ones_bitwise(gpu):
    """Requests of the served stream's largest shapes -- 64 and 256 rows of 2^22 values
    (1 GiB and 4 GiB of input, up to 2^30 values) -- served beside small requests that
    arrive while they run: every result equals its request run alone, bit for bit; one
    sampled row of the 64-row request against the FP64 convolution theorem."""
    import torch
    cases = [("conv", 2 ** 22, 64, 777, 0), ("conv", 2 ** 10, 7, 16, 0), ("conv", 2 ** 16, 3, 5000, 0),
             ("fft", 4096, 8, 0, estl_amd.FORWARD), ("conv", 2 ** 22, 256, 8192, estl_amd.CONV_CORRELATE),
             ("conv", 2 ** 13, 256, 300, estl_amd.CONV_CORRELATE), ("conv", 2 ** 18, 1, 64, 0)]
    print("[large] inputs", flush=True)
    inputs = [make_inputs(torch, gpu, c, 90 + i) for i, c in enumerate(cases)]
    refs = []
    for c, (x, h) in zip(cases, inputs):
        refs.append(alone(torch, gpu, c, x, h))
    # FP64 check of one row of the 64-row request (convolution theorem, numpy)
    n, k = 2 ** 22, 777
    x0 = inputs[0][0][5 * n:6 * n].cpu().numpy().astype(np.float64)
    h0 = np.zeros(n)
    h0[:k] = inputs[0][1].cpu().numpy()
    want = np.fft.ifft(np.fft.fft(x0) * np.fft.fft(h0)).real
    got = refs[0][5 * n:6 * n].cpu().numpy()
    assert oracle.rel_l2(got, want) <= 2e-6 * math.log2(n)
    srv = estl_amd.Server(24 << 30, streams=4)
    print("[large] served", flush=True)
    ids = []
    for rep in range(2):
        for i, c in enumerate(cases):
            ids.append((i, submit(srv, c, *inputs[i])))
    srv.drain(300)
    for i, rid in ids:
        srv.wait(rid, 60)
        y = result(srv, rid, cases[i], torch)
        assert torch.equal(y, refs[i]), (cases[i], rid)
        del y
        srv.release(rid)
    st = srv.stats()
    assert st["completed"] == len(ids) and st["failed"] == 0 and st["rejected"] == 0
    assert st["arena_used_bytes"] == 0
    srv.close()


def test_served_results_match_oracle(gpu):
    import torch
    srv = estl_amd.Server(256 << 20, streams=2)
    n, rows, k = 4096, 3, 50
    x, h = make_inputs(torch, gpu, ("conv", n, rows, k, 0), 7)
    rid = srv.submit_conv(n, rows, x, h)
    srv.wait(rid, 60)
    y = result(srv, rid, ("conv", n, rows, k, 0), torch).cpu().numpy().astype(np.float64).reshape(rows, n)
    ref = oracle.conv_circular(x.cpu().numpy().astype(np.float64).reshape(rows, n), h.cpu().numpy().astype(np.float64))
    for r in range(rows):
        assert oracle.rel_l2(y[r], ref[r]) <= 2e-6 * math.log2(n)
    srv.release(rid)
    xf, _ = make_inputs(torch, gpu, ("fft", n, 2, 0, 0), 8)
    rid = srv.submit_fft(n, 2, xf, estl_amd.FORWARD)
    srv.wait(rid, 60)
    yf = result(srv, rid, ("fft", n, 2, 0, 0), torch).cpu().numpy().astype(np.complex128).reshape(2, n)
    ref = oracle.fft64(xf.cpu().numpy().astype(np.complex128).reshape(2, n), -1)
    for r in range(2):
        assert oracle.rel_l2(yf[r], ref[r]) <= 1e-6 * math.log2(n)
    srv.release(rid)
    srv.close()


def test_admission_over_budget_and_hold_until_release(gpu):
    import torch
    n, rows = 2 ** 16, 4                     # fused-pass convolution: output 1 MiB
    case = ("conv", n, rows, 64, FULL)       # whole-row transforms: two scratch buffers
    x, h = make_inputs(torch, gpu, case, 11)
    ref = alone(torch, gpu, case, x, h)
    # budget: an output plus one request's workspace, not two outputs plus workspace
    out_bytes = 4 * n * rows
    ws = 2 * 8 * n + 256 + 2 * 8 * n * (rows // 2)
    srv = estl_amd.Server(out_bytes + ws + 4096, streams=2)
    # a request that can never fit fails with NOMEM instead of waiting
    big = torch.empty(2 ** 22, dtype=torch.float32, device=gpu)
    rid_big = srv.submit_conv(2 ** 22, 1, big, h)
    with pytest.raises(estl_amd.EstlError) as e:
        srv.wait(rid_big, 30)
    assert e.value.code == estl_amd.ERR_NOMEM
    srv.release(rid_big)
    r1 = srv.submit_conv(n, rows, x, h, flags=FULL)
    r2 = srv.submit_conv(n, rows, x, h, flags=FULL)
    srv.wait(r1, 60)
    # r1's output still holds the arena: r2 cannot be admitted until it is released
    with pytest.raises(estl_amd.EstlError) as e:
        srv.wait(r2, 0.5)
    assert e.value.code == estl_amd.ERR_BUSY
    assert srv.poll(r2)[0] == estl_amd.REQ_QUEUED
    assert torch.equal(result(srv, r1, case, torch), ref)
    srv.release(r1)
    srv.wait(r2, 60)
    assert torch.equal(result(srv, r2, case, torch), ref)
    srv.release(r2)
    st = srv.stats()
    assert st["rejected"] == 1 and st["completed"] == 2
    assert st["arena_peak_bytes"] <= st["budget_bytes"]
    srv.close()


def test_input_stream_ordering(gpu):
    import torch
    n, rows, k = 2 ** 14, 16, 100
    x0, h = make_inputs(torch, gpu, ("conv", n, rows, k, 0), 21)
    ref = alone(torch, gpu, ("conv", n, rows, k, 0), x0, h)
    srv = estl_amd.Server(512 << 20, streams=2)
    side = torch.cuda.Stream()
    x = torch.empty_like(x0)
    with torch.cuda.stream(side):
        torch.cuda._sleep(2_000_000)   # the input arrives late on the producer's stream
        x.copy_(x0)
    rid = srv.submit_conv(n, rows, x, h, stream=side)
    srv.wait(rid, 60)
    assert torch.equal(result(srv, rid, ("conv", n, rows, k, 0), torch), ref)
    srv.release(rid)
    srv.close()


def test_invalid_requests_fail_cleanly(gpu):
    import torch
    srv = estl_amd.Server(64 << 20, streams=1)
    x = torch.zeros(1000, dtype=torch.float32, device=gpu)
    rid = srv.submit_conv(1000, 1, x, x, 10)   # not a power of two
    with pytest.raises(estl_amd.EstlError) as e:
        srv.wait(rid, 30)
    assert e.value.code == estl_amd.ERR_INVALID
    srv.release(rid)
    with pytest.raises(estl_amd.EstlError):
        srv.release(rid)                         # unknown after release
    assert srv.stats()["failed"] == 1
    srv.close()


def test_registered_filters_and_completion_queue(gpu):
    import torch
    srv = estl_amd.Server(1 << 30, streams=3)
    cases = [(2 ** 12, 40, 100, 0), (2 ** 16, 3, 5000, estl_amd.CONV_CORRELATE), (2 ** 10, 256, 16, 0),
             (2 ** 18, 2, 64, estl_amd.CONV_CORRELATE)]
    inputs, refs, fids = [], [], []
    for i, (n, rows, k, fl) in enumerate(cases):
        x, h = make_inputs(torch, gpu, ("conv", n, rows, k, fl), 60 + i)
        inputs.append((x, h))
        refs.append(alone(torch, gpu, ("conv", n, rows, k, fl), x, h))
        fids.append(srv.register_filter(n, h, k))  # not waited on: FIFO order covers it
    ids = {}
    # the requests wait on a busy input stream, so they are still in flight when the
    # filter's release is attempted below
    side = _hold(torch)
    for rep in range(3):
        for i, (n, rows, k, fl) in enumerate(cases):
            ids[srv.submit_conv_filter(fids[i], rows, inputs[i][0], fl, stream=side)] = i
    # a filter in use cannot be released
    with pytest.raises(estl_amd.EstlError) as e:
        srv.release(fids[0])
    assert e.value.code == estl_amd.ERR_BUSY
    seen = []
    while len(seen) < len(ids):
        got = srv.next_done(64, 30.0)
        assert got, "completion queue timed out"
        seen += got
    assert sorted(seen) == sorted(ids) and not set(fids) & set(seen)
    for rid, i in ids.items():
        n, rows, k, fl = cases[i]
        assert torch.equal(result(srv, rid, ("conv", n, rows, k, fl), torch), refs[i]), cases[i]
        srv.release(rid)
    for f in fids:
        srv.release(f)
    st = srv.stats()
    assert st["completed"] == len(ids) and st["arena_used_bytes"] == 0
    # registration options: centre / normalise apply to the registered spectrum
    n, rows, k = 2 ** 14, 4, 33
    x, h = make_inputs(torch, gpu, ("conv", n, rows, k, estl_amd.CONV_NORMALIZE_KERNEL), 70)
    for fl in (estl_amd.CONV_CENTER, estl_amd.CONV_NORMALIZE_KERNEL):
        ref = alone(torch, gpu, ("conv", n, rows, k, fl), x, h)
        f = srv.register_filter(n, h, k, fl)
        rid = srv.submit_conv_filter(f, rows, x)
        srv.wait(rid, 60)
        assert torch.equal(result(srv, rid, ("conv", n, rows, k, fl), torch), ref)
        srv.release(rid)
        srv.release(f)
    with pytest.raises(estl_amd.EstlError):
        srv.register_filter(n, h, k, estl_amd.CONV_CORRELATE)
    srv.close()


def test_fp16_requests_equal_fp16_object_bitwise(gpu):
    import torch
    srv = estl_amd.Server(1 << 30, streams=3)
    cases = [(2 ** 12, 7, 100, 0), (2 ** 16, 3, 3000, estl_amd.CONV_CORRELATE), (2 ** 20, 1, 64, 0)]
    refs, ids, keep = [], [], []
    for i, (n, rows, k, fl) in enumerate(cases):
        x = torch.empty(rows * n, dtype=torch.float32, device=gpu)
        estl_amd.fill_uniform(x, 80 + i)
        x16 = torch.empty(rows * n, dtype=torch.float16, device=gpu)
        s_in = torch.empty(rows, dtype=torch.float32, device=gpu)
        estl_amd.pack_fp16(x, x16, s_in, n // 2, rows)
        h = torch.empty(k, dtype=torch.float32, device=gpu)
        estl_amd.fill_uniform(h, 90 + i)
        c = estl_amd.Convolution(n, rows, fl, precision=estl_amd.FP16)
        c.set_filter(h)
        y16, s_out = torch.empty_like(x16), torch.empty_like(s_in)
        c.execute_fp16(x16, s_in, y16, s_out)
        torch.cuda.synchronize()
        refs.append((y16.clone(), s_out.clone()))
        keep.append((x, x16, s_in, h))   # inputs must outlive their requests
        # inline filter, and a registered FP16 filter, beside an FP32 request
        ids.append((i, "inline", srv.submit_conv_fp16(n, rows, x16, s_in, h, k, fl)))
        f = srv.register_filter(n, h, k, estl_amd.CONV_FP16)
        ids.append((i, "registered", srv.submit_conv_filter_fp16(f, rows, x16, s_in, fl)))
        ids.append((None, "fp32", srv.submit_conv(n, rows, x, h, k, fl)))
        with pytest.raises(estl_amd.EstlError):
            srv.submit_conv_filter(f, rows, x, fl)          # FP32 request on an FP16 filter
    srv.drain(120)
    for i, how, rid in ids:
        srv.wait(rid, 60)
        if i is not None:
            n, rows, k, fl = cases[i]
            y16, s_out = srv.output_fp16(rid, n, rows)
            torch.cuda.synchronize()
            assert torch.equal(y16, refs[i][0]), (cases[i], how)
            assert torch.equal(s_out, refs[i][1]), (cases[i], how, s_out.tolist(), refs[i][1].tolist())
        srv.release(rid)
    assert srv.stats()["failed"] == 0
    srv.close()


@pytest.mark.parametrize("fp16", [False, True])
def test_registered_filter_serves_both_window_lengths(gpu, fp16):
    """A filter registered for n = 2^16 with 4000 taps keeps the spectra of both
    overlap-save window lengths; requests on it of 1, 256 and 64 rows take the window
    length a call of their rows takes alone (2^13, 2^14 and 2^14 windows on a 256-CU
    part), and each result is bit-identical to the single Convolution call, in FP32 and
    in FP16 storage."""
    import torch
    n, k = 2 ** 16, 4000
    srv = estl_amd.Server(4 << 30, streams=3)
    h = torch.empty(k, dtype=torch.float32, device=gpu)
    estl_amd.fill_uniform(h, 95)
    f = srv.register_filter(n, h, k, estl_amd.CONV_FP16 if fp16 else 0)
    ids, refs, keep = [], [], []
    for i, rows in enumerate((1, 256, 64, 1)):
        fl = estl_amd.CONV_CORRELATE if i == 3 else 0
        x = torch.empty(rows * n, dtype=torch.float32, device=gpu)
        estl_amd.fill_uniform(x, 96 + i)
        c = estl_amd.Convolution(n, rows, fl, precision=estl_amd.FP16 if fp16 else estl_amd.FP32)
        c.set_filter(h)
        if fp16:
            x16 = torch.empty(rows * n, dtype=torch.float16, device=gpu)
            s_in = torch.empty(rows, dtype=torch.float32, device=gpu)
            estl_amd.pack_fp16(x, x16, s_in, n // 2, rows)
            y16, s_out = torch.empty_like(x16), torch.empty_like(s_in)
            c.execute_fp16(x16, s_in, y16, s_out)
            refs.append((y16, s_out))
            keep.append((x16, s_in))
            ids.append((rows, srv.submit_conv_filter_fp16(f, rows, x16, s_in, fl)))
        else:
            y = torch.empty_like(x)
            c.execute(x, y)
            refs.append((y,))
            keep.append((x,))
            ids.append((rows, srv.submit_conv_filter(f, rows, x, fl)))
        torch.cuda.synchronize()
        c.close()
    srv.drain(120)
    for (rows, rid), ref in zip(ids, refs):
        srv.wait(rid, 60)
        if fp16:
            y16, s_out = srv.output_fp16(rid, n, rows)
            torch.cuda.synchronize()
            assert torch.equal(y16, ref[0]) and torch.equal(s_out, ref[1]), rows
        else:
            y = srv.output(rid, (rows * n,), torch.float32)
            torch.cuda.synchronize()
            assert torch.equal(y, ref[0]), rows
        srv.release(rid)
    assert srv.stats()["failed"] == 0
    srv.close()


def _hold(torch):
    """A stream whose queue is busy for a while (a bounded device sleep): requests
    submitted with it as their input stream stay in flight until it ends, so the ones
    submitted after them queue up behind the server's in-flight limit."""
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        torch.cuda._sleep(20_000_000)
    return side


@pytest.mark.parametrize("n,flags", [(2 ** 15, FULL), (2 ** 15, estl_amd.CONV_CORRELATE | estl_amd.CONV_CENTER | FULL),
                                     (2 ** 12, 0), (2 ** 13, estl_amd.CONV_CORRELATE | estl_amd.CONV_CENTER),
                                     (2 ** 19, FULL)])
def test_request_groups_equal_alone_bitwise(gpu, n, flags):
    """Queued convolutions of one length run as one group (one launch per pass, or one
    fused launch, for all of them; rows above 2^13 with whole-row transforms,
    ESTL_CONV_FULL_LENGTH -- overlap-save requests run alone); every member's result is bit-identical to the
    request run alone, with odd and even row counts, distinct inputs and filters
    (inline, then registered), and a request of another length queued in between.  One
    stream (two requests in flight) whose first two requests wait on a busy input
    stream, so the rest queue up and group.  At 2^15 with no rotation, the filters of
    up to 256 taps skip the first spectrum pass alone but not in a group with longer
    ones; at 2^19 every filter skips it, alone and grouped."""
    import torch
    rows_list = [1, 2, 3, 7, 16, 5, 1, 9, 4, 2]
    ks = [16, 300, 4097, 64, 2048, 1, 999, 8192, 33, 100]
    cases = [("conv", n, r, k, flags) for r, k in zip(rows_list, ks)]
    other = ("conv", 2 ** 17, 3, 100, 0)
    inputs = [make_inputs(torch, gpu, c, 300 + i) for i, c in enumerate(cases)]
    xo, ho = make_inputs(torch, gpu, other, 399)
    refs = [alone(torch, gpu, c, x, h) for c, (x, h) in zip(cases, inputs)]
    ref_o = alone(torch, gpu, other, xo, ho)
    srv = estl_amd.Server(4 << 30, streams=1)
    side = _hold(torch)
    ids = []
    for i, c in enumerate(cases):
        kind, n_, rows, k, fl = c
        x, h = inputs[i]
        ids.append((i, srv.submit_conv(n_, rows, x, h, k, fl, stream=side if i < 2 else None)))
        if i == 5:
            ids.append((-1, submit(srv, other, xo, ho)))
    srv.drain(120)
    for i, rid in ids:
        srv.wait(rid, 60)
        y = result(srv, rid, other if i < 0 else cases[i], torch)
        assert torch.equal(y, ref_o if i < 0 else refs[i]), (i, rid)
        srv.release(rid)
    st = srv.stats()
    assert st["completed"] == len(ids) and st["failed"] == 0
    assert st["groups"] >= 1 and st["grouped_requests"] >= 4, st
    # registered filters: one per request, all ready before the requests queue up
    # (centring belongs to the registration, correlation to each request)
    fids = [srv.register_filter(n, h, int(h.numel()), flags & ~estl_amd.CONV_CORRELATE) for _, h in inputs]
    for f in fids:
        srv.wait(f, 60)
    srv.reset_stats()
    side = _hold(torch)
    ids = [(i, srv.submit_conv_filter(fids[i], cases[i][2], inputs[i][0], flags & estl_amd.CONV_CORRELATE,
                                      stream=side if i < 2 else None)) for i in range(len(cases))]
    srv.drain(120)
    for i, rid in ids:
        srv.wait(rid, 60)
        assert torch.equal(result(srv, rid, cases[i], torch), refs[i]), (i, rid)
        srv.release(rid)
    st = srv.stats()
    assert st["failed"] == 0 and st["groups"] >= 1 and st["grouped_requests"] >= 4, st
    for f in fids:
        srv.release(f)
    assert srv.stats()["arena_used_bytes"] == 0
    srv.close()


@pytest.mark.parametrize("n,flags,registered,short", [
    (2 ** 16, estl_amd.CONV_CORRELATE | FULL, False, False), (2 ** 11, estl_amd.CONV_CENTER, False, False),
    (2 ** 16, estl_amd.CONV_CORRELATE | estl_amd.CONV_CENTER | FULL, True, False),
    (2 ** 12, estl_amd.CONV_CENTER, True, False),
    (2 ** 16, FULL, False, True), (2 ** 19, estl_amd.CONV_CORRELATE | FULL, False, True),
])
def test_fp16_request_groups_equal_object_bitwise(gpu, n, flags, registered, short):
    """FP16-storage requests of one length, queued behind a busy input stream, run as a
    group: with inline filters the filters' L1 norms are kept in the group's descriptor
    and the spectra are stored divided by them, with registered filters (centring a
    filter option, correlation a request option) L follows each filter's spectrum.  With
    `short` filters every tap lies within the first spectrum pass's stride, so the
    group's spectra skip that pass (at 2^16 the one pass left both reads the taps and
    divides by L).  Each binary16 output and its row scales are bit-identical to the
    FP16 convolution object's."""
    import torch
    rows_list = [3, 1, 8, 5, 2, 7]
    ks = [100, 16, 200, 33, 256, 1] if short else [100, 16, 2000, 33, 8192, 1]
    refs, keep = [], []
    for i, (rows, k) in enumerate(zip(rows_list, ks)):
        x = torch.empty(rows * n, dtype=torch.float32, device=gpu)
        estl_amd.fill_uniform(x, 500 + i)
        x16 = torch.empty(rows * n, dtype=torch.float16, device=gpu)
        s_in = torch.empty(rows, dtype=torch.float32, device=gpu)
        estl_amd.pack_fp16(x, x16, s_in, n // 2, rows)
        h = torch.empty(k, dtype=torch.float32, device=gpu)
        estl_amd.fill_uniform(h, 600 + i)
        c = estl_amd.Convolution(n, rows, flags, precision=estl_amd.FP16)
        c.set_filter(h)
        y16, s_out = torch.empty_like(x16), torch.empty_like(s_in)
        c.execute_fp16(x16, s_in, y16, s_out)
        torch.cuda.synchronize()
        refs.append((y16.clone(), s_out.clone()))
        keep.append((x16, s_in, h))
    srv = estl_amd.Server(1 << 30, streams=1)
    fids = []
    if registered:
        fflags = (flags & (estl_amd.CONV_CENTER | FULL)) | estl_amd.CONV_FP16
        fids = [srv.register_filter(n, h, int(h.numel()), fflags) for (_, _, h) in keep]
        for f in fids:
            srv.wait(f, 60)
    side = _hold(torch)
    ids = []
    for i, (rows, (x16, s_in, h)) in enumerate(zip(rows_list, keep)):
        st = side if i < 2 else None
        if registered:
            ids.append(srv.submit_conv_filter_fp16(fids[i], rows, x16, s_in, flags & estl_amd.CONV_CORRELATE,
                                                   stream=st))
        else:
            ids.append(srv.submit_conv_fp16(n, rows, x16, s_in, h, int(h.numel()), flags, stream=st))
    srv.drain(120)
    for i, rid in enumerate(ids):
        srv.wait(rid, 60)
        y16, s_out = srv.output_fp16(rid, n, rows_list[i])
        torch.cuda.synchronize()
        assert torch.equal(y16, refs[i][0]), i
        assert torch.equal(s_out, refs[i][1]), (i, s_out.tolist(), refs[i][1].tolist())
        srv.release(rid)
    st = srv.stats()
    assert st["failed"] == 0 and st["groups"] >= 1 and st["grouped_requests"] >= 3, st
    for f in fids:
        srv.release(f)
    assert srv.stats()["arena_used_bytes"] == 0
    srv.close()


@pytest.mark.parametrize("n,flags,min_grouped", [(2 ** 16, FULL, 6),
                                                 (2 ** 20, estl_amd.CONV_CORRELATE | estl_amd.CONV_CENTER | FULL, 6),
                                                 (2 ** 14, estl_amd.CONV_CORRELATE | FULL, 2)])
def test_single_row_groups_equal_alone_bitwise(gpu, n, flags, min_grouped):
    """Single FP32 rows run through transforms of n/2 values; queued single rows of one
    length group with one another (each member's row read and written as one complex
    row of the group's passes, a spectrum merge per member), never with multi-row
    requests of the same length, and every result stays bit-identical to the request
    alone, with inline and then registered filters.  At n = 2^14 (transforms of 2^13
    in one launch each) single rows run alone and only the multi-row requests group."""
    import torch
    rows_list = [1, 1, 3, 1, 1, 2, 1, 1, 1, 4, 1]
    ks = [16, 300, 100, 4097, 64, 2048, 1, 999, 8192, 33, 700]
    cases = [("conv", n, r, k, flags) for r, k in zip(rows_list, ks)]
    inputs = [make_inputs(torch, gpu, c, 900 + i) for i, c in enumerate(cases)]
    refs = [alone(torch, gpu, c, x, h) for c, (x, h) in zip(cases, inputs)]
    srv = estl_amd.Server(4 << 30, streams=1)
    side = _hold(torch)
    ids = []
    for i, c in enumerate(cases):
        kind, n_, rows, k, fl = c
        x, h = inputs[i]
        ids.append((i, srv.submit_conv(n_, rows, x, h, k, fl, stream=side if i < 2 else None)))
    srv.drain(120)
    for i, rid in ids:
        srv.wait(rid, 60)
        assert torch.equal(result(srv, rid, cases[i], torch), refs[i]), (i, rid)
        srv.release(rid)
    st = srv.stats()
    assert st["failed"] == 0 and st["groups"] >= 1 and st["grouped_requests"] >= min_grouped, st
    fids = [srv.register_filter(n, h, int(h.numel()), flags & ~estl_amd.CONV_CORRELATE) for _, h in inputs]
    for f in fids:
        srv.wait(f, 60)
    srv.reset_stats()
    side = _hold(torch)
    ids = [(i, srv.submit_conv_filter(fids[i], cases[i][2], inputs[i][0], flags & estl_amd.CONV_CORRELATE,
                                      stream=side if i < 2 else None)) for i in range(len(cases))]
    srv.drain(120)
    for i, rid in ids:
        srv.wait(rid, 60)
        assert torch.equal(result(srv, rid, cases[i], torch), refs[i]), (i, rid)
        srv.release(rid)
    st = srv.stats()
    assert st["failed"] == 0 and st["groups"] >= 1 and st["grouped_requests"] >= min_grouped, st
    for f in fids:
        srv.release(f)
    assert srv.stats()["arena_used_bytes"] == 0
    srv.close()


def test_random_stream_equals_alone_bitwise(gpu):
    """Sixty seeded requests over the whole served space -- n = 2^10..2^20, 1..6 rows,
    filters of 1..3000 taps, every flag, FP32 and FP16 storage, inline and registered
    filters -- submitted at once (the first two held on a busy input stream, so the rest
    queue, group and run beside one another on two streams).  Every output (and FP16
    row scales) is bit-identical to the same convolution run alone through the object
    API, whatever group, path (pairs, single rows at half length, fused rows,
    overlap-save windows) or neighbours it had."""
    import torch
    rng = np.random.default_rng(99)
    flag_bits = (estl_amd.CONV_CORRELATE, estl_amd.CONV_CENTER, estl_amd.CONV_NORMALIZE_KERNEL, FULL)
    cases = []
    for i in range(60):
        n = 1 << int(rng.integers(10, 21))
        rows = int(rng.integers(1, 7))
        k = int(rng.integers(1, 3001))
        flags = sum(f for f in flag_bits if rng.random() < 0.3)
        fp16 = bool(rng.random() < 0.3)
        reg = bool(rng.random() < 0.4)
        cases.append((n, rows, k, flags, fp16, reg))
    srv = estl_amd.Server(4 << 30, streams=2)
    data, refs, fids = [], [], {}
    for i, (n, rows, k, flags, fp16, reg) in enumerate(cases):
        x = torch.empty(rows * n, dtype=torch.float32, device=gpu)
        estl_amd.fill_uniform(x, 1000 + i)
        h = torch.empty(k, dtype=torch.float32, device=gpu)
        estl_amd.fill_uniform(h, 2000 + i)
        if flags & estl_amd.CONV_NORMALIZE_KERNEL:
            h += 1.0
        c = estl_amd.Convolution(n, rows, flags, precision=estl_amd.FP16 if fp16 else estl_amd.FP32)
        c.set_filter(h)
        if fp16:
            x16 = torch.empty(rows * n, dtype=torch.float16, device=gpu)
            s_in = torch.empty(rows, dtype=torch.float32, device=gpu)
            estl_amd.pack_fp16(x, x16, s_in, n // 2, rows)
            y16, s_out = torch.empty_like(x16), torch.empty_like(s_in)
            c.execute_fp16(x16, s_in, y16, s_out)
            data.append((x16, s_in, h))
            refs.append((y16, s_out))
        else:
            y = torch.empty_like(x)
            c.execute(x, y)
            data.append((x, None, h))
            refs.append((y, None))
        if reg:
            ff = (flags & ~estl_amd.CONV_CORRELATE) | (estl_amd.CONV_FP16 if fp16 else 0)
            fids[i] = srv.register_filter(n, h, k, ff)
    torch.cuda.synchronize()
    for f in fids.values():
        srv.wait(f, 60)
    side = _hold(torch)
    ids = []
    for i, (n, rows, k, flags, fp16, reg) in enumerate(cases):
        x, s_in, h = data[i]
        st = side if i < 2 else None
        if reg and fp16:
            rid = srv.submit_conv_filter_fp16(fids[i], rows, x, s_in, flags & estl_amd.CONV_CORRELATE, stream=st)
        elif reg:
            rid = srv.submit_conv_filter(fids[i], rows, x, flags & estl_amd.CONV_CORRELATE, stream=st)
        elif fp16:
            rid = srv.submit_conv_fp16(n, rows, x, s_in, h, k, flags, stream=st)
        else:
            rid = srv.submit_conv(n, rows, x, h, k, flags, stream=st)
        ids.append(rid)
    srv.drain(120)
    for i, rid in enumerate(ids):
        n, rows, k, flags, fp16, reg = cases[i]
        srv.wait(rid, 60)
        if fp16:
            y16, s_out = srv.output_fp16(rid, n, rows)
            torch.cuda.synchronize()
            assert torch.equal(y16, refs[i][0]) and torch.equal(s_out, refs[i][1]), (i, cases[i])
        else:
            y = srv.output(rid, (rows * n,), torch.float32)
            torch.cuda.synchronize()
            assert torch.equal(y, refs[i][0]), (i, cases[i])
        srv.release(rid)
    st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == len(ids) and st["groups"] >= 1, st
    for f in fids.values():
        srv.release(f)
    assert srv.stats()["arena_used_bytes"] == 0
    srv.close()


def test_request_groups_at_member_limit(gpu):
    """Forty queued requests of one length form several groups of at most 16 members
    (the descriptor's limit); every result stays bit-identical to the request alone."""
    import torch
    n, count = 2 ** 14, 40
    rng = np.random.default_rng(77)
    rows_list = [int(r) for r in rng.integers(1, 9, size=count)]
    ks = [int(k) for k in rng.integers(1, 2000, size=count)]
    cases = [("conv", n, r, k, (estl_amd.CONV_CORRELATE if i % 3 == 0 else 0) | FULL)
             for i, (r, k) in enumerate(zip(rows_list, ks))]
    inputs = [make_inputs(torch, gpu, c, 700 + i) for i, c in enumerate(cases)]
    refs = [alone(torch, gpu, c, x, h) for c, (x, h) in zip(cases, inputs)]
    srv = estl_amd.Server(4 << 30, streams=1)
    side = _hold(torch)
    ids = []
    for i, c in enumerate(cases):
        kind, n_, rows, k, fl = c
        x, h = inputs[i]
        ids.append((i, srv.submit_conv(n_, rows, x, h, k, fl, stream=side if i < 2 else None)))
    srv.drain(120)
    for i, rid in ids:
        srv.wait(rid, 60)
        assert torch.equal(result(srv, rid, cases[i], torch), refs[i]), (i, rid)
        srv.release(rid)
    st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == count
    # two flag classes among 38 queued requests: at least three groups, none over 16
    assert st["groups"] >= 3 and st["grouped_requests"] >= 30, st
    assert st["grouped_requests"] <= 16 * st["groups"], st
    assert st["arena_used_bytes"] == 0
    srv.close()
