"""GPU parity of the uneven-size batch (one work-queue launch for mixed sizes).

Tolerance: relative L2 <= 1e-6 * log2(N) per transform against the FP64 oracle
(sizes the oracle finishes quickly) or against the exact spectrum of the unit tones
(every size, full benchmark batch).
"""
import math

import numpy as np
import pytest

import estl_amd
import estl_amd.workloads as wl
import oracle

pytestmark = pytest.mark.gpu


def tol(n):
    return 1e-6 * max(1, int(math.log2(n)))


def run(torch, ub, x, graph=False):
    y = torch.empty_like(x)
    ub.execute(x, y, graph=graph)
    torch.cuda.synchronize()
    return y


def test_every_size_class_against_oracle(gpu):
    import torch
    # every size 2^1..2^20, a few of each, shuffled, packed
    rng = np.random.default_rng(5)
    logs = np.concatenate([np.arange(1, 21)] * 3)
    logs = logs[rng.permutation(len(logs))]
    sizes = [1 << int(k) for k in logs]
    total = sum(sizes)
    x = torch.empty(total, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 77)
    for direction in (estl_amd.FORWARD, estl_amd.BACKWARD):
        ub = estl_amd.UnevenBatch(sizes, direction=direction)
        assert ub.info["launches"] == 1 and ub.info["extent"] == total
        y = run(torch, ub, x).cpu().numpy()
        xh = x.cpu().numpy()
        off = 0
        for n in sizes:
            ref = oracle.fft64(xh[off:off + n], direction)
            assert oracle.rel_l2(y[off:off + n], ref) <= tol(n), (n, direction)
            off += n


def test_random_batches_against_oracle(gpu):
    """Eight seeded batches: 5..60 transforms of sizes 2^U{1..18} each, placed at random
    (sorted, non-overlapping) offsets with gaps, random direction, eager or graph
    replay -- every transform against the FP64 oracle, the gaps untouched."""
    import torch
    rng = np.random.default_rng(321)
    for case in range(8):
        count = int(rng.integers(5, 61))
        sizes = [1 << int(k) for k in rng.integers(1, 19, size=count)]
        offsets, pos = [], 0
        for n in sizes:
            pos += int(rng.integers(0, 100))
            offsets.append(pos)
            pos += n
        order = rng.permutation(count)          # offsets in any order of the size list
        sizes = [sizes[i] for i in order]
        offsets = [offsets[i] for i in order]
        extent = pos
        direction = estl_amd.FORWARD if rng.random() < 0.5 else estl_amd.BACKWARD
        graph = bool(rng.random() < 0.5)
        x = torch.empty(extent, dtype=torch.complex64, device=gpu)
        estl_amd.fill_uniform(x, 900 + case)
        y = torch.zeros_like(x)
        ub = estl_amd.UnevenBatch(sizes, offsets=offsets, direction=direction)
        ub.execute(x, y, graph=graph)
        torch.cuda.synchronize()
        xh, yh = x.cpu().numpy(), y.cpu().numpy()
        mask = np.ones(extent, bool)
        for o, n in zip(offsets, sizes):
            e = oracle.rel_l2(yh[o:o + n], oracle.fft64(xh[o:o + n], direction))
            assert e <= tol(n), (case, n, o, direction, graph, e)
            mask[o:o + n] = False
        assert not np.any(yh[mask]), case


def test_custom_offsets_and_in_place(gpu):
    import torch
    sizes = [256, 65536, 16, 8192, 4096]
    offsets = [70000, 0, 69900, 80000, 100000]  # gaps between transforms
    extent = max(o + n for o, n in zip(offsets, sizes))
    x = torch.zeros(extent, dtype=torch.complex64, device=gpu)
    src = torch.empty(extent, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(src, 4)
    for o, n in zip(offsets, sizes):
        x[o:o + n] = src[o:o + n]
    xh = x.cpu().numpy()
    ub = estl_amd.UnevenBatch(sizes, offsets=offsets)
    ub.execute(x, x)  # in place
    torch.cuda.synchronize()
    y = x.cpu().numpy()
    for o, n in zip(offsets, sizes):
        assert oracle.rel_l2(y[o:o + n], oracle.fft64(xh[o:o + n])) <= tol(n), n
    # values outside the transforms are untouched
    mask = np.ones(extent, bool)
    for o, n in zip(offsets, sizes):
        mask[o:o + n] = False
    assert np.array_equal(y[mask], xh[mask])


def test_repeat_and_graph_replay_identical(gpu):
    import torch
    b = wl.uneven_batch(count=300, seed=11, kmin=4, kmax=17)
    x = torch.empty(b.total, dtype=torch.complex64, device=gpu)
    wl.fill_uneven(b, x, torch)
    ub = estl_amd.UnevenBatch(b.sizes.tolist())
    eager = run(torch, ub, x)
    again = run(torch, ub, x)
    assert torch.equal(eager, again)  # the work queue resets itself between launches
    for rep in range(3):
        g = run(torch, ub, x, graph=True)
        assert torch.equal(g, eager), rep
        # unrelated eager work between graph replays
        torch.cuda.synchronize()
        _ = (x * 2).sum().item()


def test_uneven_batch_8192_tones_and_samples(gpu):
    """The full uneven workload: 8192 transforms, N = 2^8..2^18 (seed 3), 10 % tones."""
    import torch
    b = wl.uneven_batch()
    x = torch.empty(b.total, dtype=torch.complex64, device=gpu)
    wl.fill_uneven(b, x, torch)
    ub = estl_amd.UnevenBatch(b.sizes.tolist())
    y = run(torch, ub, x, graph=True)
    # every tone: spectrum N * delta_f exactly
    worst = 0.0
    for i in np.nonzero(b.tone)[0]:
        n, off, f = int(b.sizes[i]), int(b.offsets[i]), int(b.freq[i])
        yi = y[off:off + n]
        peak = yi[f].item()
        err2 = (yi.abs().double() ** 2).sum().item() - abs(peak) ** 2 + abs(peak - n) ** 2
        rel = math.sqrt(max(err2, 0.0)) / n
        worst = max(worst, rel / tol(n))
    assert worst <= 1.0, worst
    # a sample of the random transforms against the FP64 oracle
    rng = np.random.default_rng(0)
    cand = [i for i in np.nonzero(~b.tone)[0] if b.sizes[i] <= 1 << 15]
    for i in rng.choice(cand, size=24, replace=False):
        n, off = int(b.sizes[i]), int(b.offsets[i])
        xi = x[off:off + n].cpu().numpy()
        assert oracle.rel_l2(y[off:off + n].cpu().numpy(), oracle.fft64(xi)) <= tol(n), n
    # same results as one plan launch per transform, within tolerance
    for i in rng.choice(len(b.sizes), size=16, replace=False):
        n, off = int(b.sizes[i]), int(b.offsets[i])
        p = estl_amd.Plan(n, 1, estl_amd.FORWARD)
        yi = torch.empty(n, dtype=torch.complex64, device=gpu)
        p.execute(x[off:off + n], yi)
        torch.cuda.synchronize()
        assert oracle.rel_l2(yi.cpu().numpy(), y[off:off + n].cpu().numpy()) <= 2 * tol(n)


def test_rejects_bad_sizes(gpu):
    with pytest.raises(estl_amd.EstlError):
        estl_amd.UnevenBatch([256, 300])
    with pytest.raises(estl_amd.EstlError):
        estl_amd.UnevenBatch([1 << 21])
    with pytest.raises(estl_amd.EstlError):
        estl_amd.UnevenBatch([256, 512], flags=estl_amd.FLAG_NORMALIZE)
