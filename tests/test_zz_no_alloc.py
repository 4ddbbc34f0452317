"""GPU test of estl_execute_device_noalloc (through the C ABI): the no-allocation variant
of estl_execute_device gives the same results bit for bit -- on two calls in a row, the
second with a smaller input, an in-place chain of three passes, the largest accepted
plan -- allocates no device memory on a stream it has not seen, and rejects plans whose
scratch exceeds its static array before any device work."""
import ctypes
import math

import numpy as np
import pytest

import estl_amd
import oracle

pytestmark = pytest.mark.gpu


def _noalloc(plan, x, y, stream=None):
    s = None if stream is None else stream.cuda_stream
    return estl_amd.lib().estl_execute_device_noalloc(ctypes.c_void_p(plan.handle), ctypes.c_void_p(x.data_ptr()),
                                                      ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(s))


def _regular(plan, x, y, stream=None):
    s = None if stream is None else stream.cuda_stream
    return estl_amd.lib().estl_execute_device(ctypes.c_void_p(plan.handle), ctypes.c_void_p(x.data_ptr()),
                                              ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(s))


def _input(torch, gpu, batch, n, seed):
    x = torch.empty(batch, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, seed)
    return x


def test_noalloc_equals_execute_device_two_calls(gpu):
    import torch
    # first call: 8 rows of 2^16 (two passes, 4 MiB of scratch); second: 2 rows of 2^14
    for (n, batch, seed) in ((1 << 16, 8, 101), (1 << 14, 2, 102)):
        p = estl_amd.Plan(n, batch, estl_amd.FORWARD)
        assert p.launches() >= 2
        x = _input(torch, gpu, batch, n, seed)
        ref = torch.empty_like(x)
        assert _regular(p, x, ref) == 0
        torch.cuda.synchronize()
        y = torch.full_like(x, float("nan"))
        # a stream the library's plans have not run on: the runtime's own first-launch
        # resources for it are set up by a kernel that allocates nothing (a fill)
        s = torch.cuda.Stream()
        estl_amd.fill_uniform(y, 1, stream=s)
        s.synchronize()
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        assert _noalloc(p, x, y, s) == 0, estl_amd.lib().estl_last_error()
        s.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 == free0, (free0, free1)  # nothing allocated
        assert torch.equal(y, ref), (n, batch)
        # against the FP64 oracle on two rows
        xs = x[[0, batch - 1]].cpu().numpy().astype(np.complex128)
        want = oracle.fft64(xs, -1)
        got = y[[0, batch - 1]].cpu().numpy()
        for r in range(2):
            assert oracle.rel_l2(got[r], want[r]) <= 1e-6 * math.log2(n)


def test_noalloc_in_place_odd_chain_and_limits(gpu):
    import torch
    # in place, three passes: the chain needs both halves of the array
    n, batch = 1 << 20, 2
    p = estl_amd.Plan(n, batch, estl_amd.BACKWARD, split=[7, 7, 6])
    x = _input(torch, gpu, batch, n, 103)
    ref = x.clone()
    assert _regular(p, ref, ref) == 0
    y = x.clone()
    assert _noalloc(p, y, y) == 0, estl_amd.lib().estl_last_error()
    torch.cuda.synchronize()
    assert torch.equal(y, ref)
    # the largest accepted plan: 2^23 values out of place (64 MiB of scratch)
    big = estl_amd.Plan(1 << 23, 1, estl_amd.FORWARD)
    xb = _input(torch, gpu, 1, 1 << 23, 104)
    rb, yb = torch.empty_like(xb), torch.empty_like(xb)
    assert _regular(big, xb, rb) == 0
    assert _noalloc(big, xb, yb) == 0, estl_amd.lib().estl_last_error()
    torch.cuda.synchronize()
    assert torch.equal(yb, rb)
    # rejected before any device work: 2^24 values; 2^23 in place with three passes
    too_big = estl_amd.Plan(1 << 24, 1, estl_amd.FORWARD)
    xt = torch.zeros(1 << 24, dtype=torch.complex64, device=gpu)
    assert _noalloc(too_big, xt, xt) == estl_amd.ERR_INVALID
    assert b"scratch" in estl_amd.lib().estl_last_error()
    odd = estl_amd.Plan(1 << 23, 1, estl_amd.FORWARD, split=[8, 8, 7])
    assert _noalloc(odd, xb, xb) == estl_amd.ERR_INVALID
    # FP16-storage plans are not run by it; a rows plan (no scratch) is
    h = estl_amd.Plan(4096, 4, estl_amd.FORWARD, estl_amd.FP16)
    x4 = _input(torch, gpu, 4, 4096, 105)
    assert _noalloc(h, x4, x4) == estl_amd.ERR_INVALID
    rows = estl_amd.Plan(4096, 4, estl_amd.FORWARD)
    r4, y4 = torch.empty_like(x4), torch.empty_like(x4)
    assert _regular(rows, x4, r4) == 0 and _noalloc(rows, x4, y4) == 0
    torch.cuda.synchronize()
    assert torch.equal(y4, r4)
