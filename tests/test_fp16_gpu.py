"""GPU tests of the FP16-storage / FP32-compute mode against the FP64 oracle.

Tolerance: every store to HBM rounds each component to binary16 (relative error at
most u = 2^-11); the transform is unitary in its stored scaling, so the relative L2
error of the result is bounded by about u * sqrt(stores) with stores = 1 (pack) + the
plan's passes (1 for n <= 2^13).  The tests use that bound, u * sqrt(stores), per row.
"""
import math

import numpy as np
import pytest

import estl_amd
import oracle

pytestmark = pytest.mark.gpu

U = 2.0 ** -11


def stores(n):
    # pack + one store per launch: the rows kernel for n <= 2^13, else the passes of the
    # default split (at most 9 bits each, at least two)
    if n <= 2 ** 13:
        return 2
    return 1 + max(2, (int(math.log2(n)) + 8) // 9)


def bound(n):
    return U * math.sqrt(stores(n))


def run_fp16(torch, gpu, x, n, batch, direction=estl_amd.FORWARD, flags=0):
    x16 = torch.empty(batch * n * 2, dtype=torch.float16, device=gpu)
    y16 = torch.empty_like(x16)
    s_in = torch.empty(batch, dtype=torch.float32, device=gpu)
    s_out = torch.empty_like(s_in)
    estl_amd.pack_fp16(x, x16, s_in, n, batch)
    plan = estl_amd.Plan(n, batch, direction, estl_amd.FP16, flags)
    plan.execute_fp16(x16, s_in, y16, s_out)
    y = torch.empty_like(x)
    estl_amd.unpack_fp16(y16, s_out, y, n, batch)
    torch.cuda.synchronize()
    return y, s_in


def test_pack_unpack_round_trip_and_scales(gpu):
    import torch
    n, batch = 1024, 6
    x = torch.empty(batch, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 3)
    mags = torch.tensor([1.0, 1e-6, 3e4, 0.0, 1.0, 7.5], device=gpu)
    x *= mags[:, None]
    x16 = torch.empty(batch * n * 2, dtype=torch.float16, device=gpu)
    s = torch.empty(batch, dtype=torch.float32, device=gpu)
    estl_amd.pack_fp16(x, x16, s, n, batch)
    y = torch.empty_like(x)
    estl_amd.unpack_fp16(x16, s, y, n, batch)
    torch.cuda.synchronize()
    xr = torch.view_as_real(x)
    ref_scale = xr.abs().amax(dim=(1, 2))
    ref_scale[ref_scale == 0] = 1.0
    assert torch.equal(s, ref_scale)                    # max-abs, exactly
    stored = x16.view(batch, n, 2).float()
    assert stored.abs().max().item() <= 1.0
    err = (y - x).abs().cpu().numpy()
    # per component: |x - y| <= (u/2 + 2^-22) * scale (round to nearest binary16 in
    # [-1, 1], plus the FP32 divide and multiply); complex modulus: sqrt(2) times that
    assert np.all(err <= math.sqrt(2) * (U / 2 + 2.0 ** -22) * ref_scale.cpu().numpy()[:, None])
    assert torch.count_nonzero(y[3]) == 0


@pytest.mark.parametrize("log2n", list(range(1, 14)) + [14, 16, 18, 20])
def test_fp16_forward_matches_fp64(gpu, log2n):
    import torch
    n = 1 << log2n
    batch = max(1, min(64, (1 << 20) // n))
    x = torch.empty(batch, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 100 + log2n)
    y, _ = run_fp16(torch, gpu, x, n, batch)
    xs, ys = x.cpu().numpy(), y.cpu().numpy()
    for r in sorted({0, batch // 2, batch - 1}):
        ref = oracle.fft64(xs[r].astype(np.complex128), -1)
        e = oracle.rel_l2(ys[r], ref)
        assert e <= bound(n), (log2n, r, e, bound(n))


@pytest.mark.parametrize("log2n", [6, 12, 15])
def test_fp16_round_trip_normalized(gpu, log2n):
    import torch
    n, batch = 1 << log2n, 4
    x = torch.empty(batch, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 7)
    y, _ = run_fp16(torch, gpu, x, n, batch)
    z, _ = run_fp16(torch, gpu, y, n, batch, estl_amd.BACKWARD, estl_amd.FLAG_NORMALIZE)
    rt = (torch.linalg.vector_norm((z - x).to(torch.complex128)) /
          torch.linalg.vector_norm(x.to(torch.complex128))).item()
    assert rt <= 2 * bound(n), rt


def test_fp16_rows_of_very_different_magnitude(gpu):
    # per-transform scaling: rows 1e-6 and 3e4 keep the relative accuracy of row 1.0
    import torch
    n, batch = 2 ** 16, 3
    x = torch.empty(batch, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 9)
    x *= torch.tensor([1.0, 1e-6, 3e4], device=gpu)[:, None]
    y, _ = run_fp16(torch, gpu, x, n, batch)
    xs, ys = x.cpu().numpy(), y.cpu().numpy()
    for r in range(batch):
        e = oracle.rel_l2(ys[r], oracle.fft64(xs[r].astype(np.complex128), -1))
        assert e <= bound(n), (r, e)


def test_fp16_tone_and_constant_no_overflow(gpu):
    # a unit tone concentrates N in one bin and a constant puts N |x| at bin 0: the
    # largest values a transform can produce; the 1/sqrt(B) pass scaling keeps them
    # inside binary16's range
    import torch
    n = 2 ** 20
    t = torch.arange(n, dtype=torch.float64, device=gpu)
    f = 12345
    tone = torch.polar(torch.ones_like(t), (2 * math.pi / n) * torch.remainder(t * f, n)).to(torch.complex64)
    const = torch.full((n,), 1 - 1j, dtype=torch.complex64, device=gpu)
    x = torch.stack([tone, const])
    y, _ = run_fp16(torch, gpu, x, n, 2)
    ys = y.cpu().numpy().astype(np.complex128)
    assert np.all(np.isfinite(ys))
    ref = np.zeros((2, n), np.complex128)
    ref[0, f] = n
    ref[1, 0] = n * (1 - 1j)
    for r in range(2):
        assert oracle.rel_l2(ys[r], ref[r]) <= bound(n), r


def test_fp16_plan_rejects_fp32_entry_point(gpu):
    import torch
    plan = estl_amd.Plan(4096, 1, estl_amd.FORWARD, estl_amd.FP16)
    x = torch.zeros(4096, dtype=torch.complex64, device=gpu)
    with pytest.raises(estl_amd.EstlError):
        plan.execute(x, x)
