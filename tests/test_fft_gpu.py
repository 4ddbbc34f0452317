"""GPU parity of the HIP FFT path against the oracle (run with -m gpu on an MI355X).

Tolerance (stated by the build target): relative L2 error <= 1e-6 * log2(N) for FP32,
against the FP64 oracle (oracle.fft64), per row.  The reference's FFTW3 path is absent
from this image, so its place is taken by the FP64 oracle and the FP32 restatement of
the reference's algorithm (oracle.fft32), both on the same seeded inputs.
At full benchmark sizes the checks are size-independent: round trip, Parseval,
linearity, exact known answers, and direct O(N) evaluation of sampled bins.
"""
import math

import numpy as np
import pytest

import estl_amd
import oracle

pytestmark = pytest.mark.gpu


def tol(n: int) -> float:
    return 1e-6 * max(1, int(math.log2(n)))


def dev_uniform(torch, batch, n, seed, device):
    x = torch.empty(batch, n, dtype=torch.complex64, device=device)
    estl_amd.fill_uniform(x, seed)
    return x


def run(torch, plan, x, out=None):
    y = torch.empty_like(x) if out is None else out
    plan.execute(x, y)
    torch.cuda.synchronize()
    return y


def max_row_err(got, want):
    return max(oracle.rel_l2(got[r], want[r]) for r in range(want.shape[0]))


def test_fill_uniform_matches_generator(gpu):
    import torch
    for count, seed, off in [(1, 0, 0), (1000, 1, 0), (123457, 3, 77)]:
        t = torch.empty(count, dtype=torch.float32, device=gpu)
        estl_amd.check(estl_amd.lib().estl_fill_uniform(t.data_ptr(), count, seed, off, None))
        torch.cuda.synchronize()
        assert np.array_equal(t.cpu().numpy(), oracle.fill_uniform(count, seed, off))


@pytest.mark.parametrize("n", [2 ** k for k in range(1, 13)])
def test_golden_vectors(gpu, golden, n):
    import torch
    for key in [k for k in golden if k.startswith(f"in_{n}_")]:
        s = key.split("_")[-1]
        x = torch.from_numpy(golden[key][None, :].copy()).to(gpu)
        for direction, ref in ((estl_amd.FORWARD, golden[f"fwd_{n}_{s}"]), (estl_amd.BACKWARD, golden[f"bwd_{n}_{s}"])):
            y = run(torch, estl_amd.Plan(n, 1, direction), x).cpu().numpy()[0]
            assert oracle.rel_l2(y, ref) <= tol(n), (n, s, direction)


@pytest.mark.parametrize("log2n", list(range(0, 14)))
@pytest.mark.parametrize("batch", [1, 3, 37])
def test_rows_parity(gpu, log2n, batch):
    import torch
    n = 1 << log2n
    x = dev_uniform(torch, batch, n, seed=100 + log2n, device=gpu)
    xh = x.cpu().numpy()
    assert np.array_equal(xh, oracle.uniform_complex(batch, n, 100 + log2n))
    for direction in (estl_amd.FORWARD, estl_amd.BACKWARD):
        p = estl_amd.Plan(n, batch, direction)
        y = run(torch, p, x).cpu().numpy()
        want = oracle.fft64(xh, direction)
        assert max_row_err(y, want) <= tol(n), (n, batch, direction)
        # the FP32 restatement of the reference's own algorithm agrees to the same bound
        assert max_row_err(y, oracle.fft32(xh, direction)) <= tol(n)


@pytest.mark.parametrize("n,split", [
    (2 ** 8, [4, 4]), (2 ** 12, [4, 4, 4]), (2 ** 12, [6, 6]), (2 ** 13, [9, 4]), (2 ** 13, [4, 9]),
    (2 ** 14, None), (2 ** 15, None), (2 ** 16, None), (2 ** 16, [4, 4, 4, 4]), (2 ** 17, None),
    (2 ** 18, None), (2 ** 18, [13, 5]), (2 ** 20, None), (2 ** 16, [12, 4]), (2 ** 16, [10, 6]),
])
@pytest.mark.parametrize("batch", [1, 2])
def test_pass_plans_parity(gpu, n, split, batch):
    import torch
    x = dev_uniform(torch, batch, n, seed=7, device=gpu)
    xh = x.cpu().numpy()
    for direction in (estl_amd.FORWARD, estl_amd.BACKWARD):
        p = estl_amd.Plan(n, batch, direction, split=split)
        if n > 2 ** 13 or split:
            assert p.describe()["kind"] == "passes"
        y = run(torch, p, x).cpu().numpy()
        assert max_row_err(y, oracle.fft64(xh, direction)) <= tol(n), (n, split, direction)


def test_random_plans_match_fp64(gpu):
    """Forty seeded plans over the plan space: n = 2^1..2^20, 1..5 rows, a random split
    into 1..4 passes of radix 2^1..2^13 (or the default plan), random tile widths
    within the limits, either direction, out of place or in place -- each row against
    the FP64 oracle."""
    import torch
    rng = np.random.default_rng(4242)
    for case in range(40):
        L = int(rng.integers(1, 21))
        n = 1 << L
        batch = int(rng.integers(1, 6)) if L <= 16 else 1
        split, tiles = None, None
        if L >= 2 and rng.random() < 0.7:
            P = int(rng.integers(2, min(4, L) + 1))
            while True:
                cuts = sorted(rng.choice(np.arange(1, L), size=P - 1, replace=False).tolist())
                split = [b - a for a, b in zip([0] + cuts, cuts + [L])]
                if max(split) <= 13:
                    break
            if rng.random() < 0.5:
                tiles = [int(rng.integers(0, max(0, min(6, 14 - b, L - b)) + 1)) for b in split]
        direction = estl_amd.FORWARD if rng.random() < 0.5 else estl_amd.BACKWARD
        in_place = bool(rng.random() < 0.3)
        x = dev_uniform(torch, batch, n, seed=500 + case, device=gpu)
        xh = x.cpu().numpy()
        p = estl_amd.Plan(n, batch, direction, split=split, tiles=tiles)
        y = run(torch, p, x, out=x if in_place else None)
        e = max_row_err(y.cpu().numpy(), oracle.fft64(xh, direction))
        assert e <= tol(n), (case, n, batch, split, tiles, direction, in_place, e)


@pytest.mark.parametrize("n,split", [(4096, None), (2 ** 16, None), (2 ** 16, [4, 4, 4, 4]), (2 ** 20, None)])
def test_in_place(gpu, n, split):
    import torch
    x = dev_uniform(torch, 2, n, seed=9, device=gpu)
    xh = x.cpu().numpy()
    p = estl_amd.Plan(n, 2, estl_amd.FORWARD, split=split)
    run(torch, p, x, out=x)
    assert max_row_err(x.cpu().numpy(), oracle.fft64(xh)) <= tol(n)


@pytest.mark.parametrize("n", [1, 64, 4096, 2 ** 16])
def test_normalized_round_trip(gpu, n):
    import torch
    x = dev_uniform(torch, 3, n, seed=12, device=gpu)
    f = estl_amd.Plan(n, 3, estl_amd.FORWARD)
    b = estl_amd.Plan(n, 3, estl_amd.BACKWARD, flags=estl_amd.FLAG_NORMALIZE)
    y = run(torch, f, x)
    z = run(torch, b, y)
    assert oracle.rel_l2(z.cpu().numpy(), x.cpu().numpy()) <= 2 * tol(max(n, 2))


@pytest.mark.parametrize("m", [1, 2, 4, 8, 12, 16, 20])
def test_reference_known_answer(gpu, m):
    # tests/Test.cpp:45-110: cos(2 pi effK q / M) -> M/2 at effK and M - effK, |diff|^2 < 1e-3 M
    import torch
    M = 1 << m
    B = 4 if m <= 16 else 2
    rows = []
    effs = []
    for bIx in range(B):
        effK = ((1 + bIx) % (M // 2 - 1)) + 1 if M > 2 else 1
        effs.append(effK)
        q = np.arange(M)
        rows.append(np.cos(2 * np.pi * ((1.0 * effK * q) / M)).astype(np.float32))
    x = torch.from_numpy(np.array(rows).astype(np.complex64)).to(gpu)
    y = run(torch, estl_amd.Plan(M, B, estl_amd.FORWARD), x).cpu().numpy()
    for bIx, effK in enumerate(effs):
        exp = np.zeros(M, dtype=complex)
        exp[effK] += M / 2.0
        exp[M - effK] += M / 2.0
        assert np.max(np.abs(y[bIx] - exp) ** 2) < 1e-3 * M


def test_host_entry_point_mirrors_backend(gpu):
    # FFT_HIP mirrors FFT<D>::setData/getData/execute (FFT.h:19-36)
    n, B = 1024, 3
    f = estl_amd.FFT_HIP(n, True, B)
    xh = oracle.uniform_complex(B, n, 21)
    for b in range(B):
        for q in range(n):
            f.set_data(q, xh[b, q], b)
    f.execute()
    got = np.array([[f.get_data(q, b) for q in range(n)] for b in range(B)])
    assert max_row_err(got, oracle.fft64(xh)) <= tol(n)
    # the backward direction is applied to every row (the reference conjugated row 0 only)
    g = estl_amd.FFT_HIP(n, False, B)
    g.data_in[:] = xh.ravel()
    g.execute()
    assert max_row_err(g.data_out.reshape(B, n), oracle.fft64(xh, oracle.BACKWARD)) <= tol(n)


def test_batched_4096_full_size(gpu):
    """N=4096, batch 16384 (512 MiB per direction), seed 1: forward + backward round trip."""
    import torch
    n, B = 4096, 16384
    x = dev_uniform(torch, B, n, seed=1, device=gpu)
    f = estl_amd.Plan(n, B, estl_amd.FORWARD)
    b = estl_amd.Plan(n, B, estl_amd.BACKWARD)
    y = run(torch, f, x)
    # sampled rows against the FP64 oracle
    rows = [0, 1, 4095, 8191, 16383]
    xs = x[rows].cpu().numpy()
    assert max_row_err(y[rows].cpu().numpy(), oracle.fft64(xs)) <= tol(n)
    # Parseval on every row, computed on the GPU in FP64
    ex = (x.abs().double() ** 2).sum(dim=1)
    ey = (y.abs().double() ** 2).sum(dim=1)
    assert torch.max(torch.abs(ey / (n * ex) - 1)).item() < 1e-5
    z = run(torch, b, y)
    num = torch.linalg.vector_norm((z / n - x).to(torch.complex128)).item()
    den = torch.linalg.vector_norm(x.to(torch.complex128)).item()
    assert num / den <= 2 * tol(n)


def test_large_single_2_26(gpu):
    """N = 2^26 single transform (seed 2): exact known answer, sampled bins, round trip."""
    import torch
    n = 1 << 26
    p = estl_amd.Plan(n, 1, estl_amd.FORWARD)
    b = estl_amd.Plan(n, 1, estl_amd.BACKWARD)
    # known answer: cos(2 pi k t / N) -> N/2 at k and N - k
    k = 123457
    t = torch.arange(n, device=gpu, dtype=torch.float64)
    c = torch.cos(2 * math.pi * ((k * t) % n) / n).to(torch.complex64).reshape(1, n)
    y = run(torch, p, c)
    yk = y[0].cpu().numpy()
    exp = np.zeros(n, dtype=np.complex128)
    exp[k] = exp[n - k] = n / 2
    assert oracle.rel_l2(yk, exp) <= tol(n)
    # random input: sampled bins evaluated directly in FP64
    x = dev_uniform(torch, 1, n, seed=2, device=gpu)
    y = run(torch, p, x)
    xh = x[0].cpu().numpy().astype(np.complex128)
    yh = y[0].cpu().numpy()
    tt = np.arange(n, dtype=np.int64)
    # a rel-L2 error e over the spectrum is e * ||x||_2 per bin on average (Parseval)
    xnorm = np.linalg.norm(xh)
    for kk in (1, n // 2, n - 1, 31337):
        w = np.exp(-2j * np.pi * ((tt * kk) % n) / n)
        want = np.dot(xh, w)
        assert abs(yh[kk] - want) <= 10 * tol(n) * xnorm, kk
    # Parseval and round trip over the whole transform
    ex = (x.abs().double() ** 2).sum().item()
    ey = (y.abs().double() ** 2).sum().item()
    assert abs(ey / (n * ex) - 1) < 1e-5
    z = run(torch, b, y)
    num = torch.linalg.vector_norm((z / n - x).to(torch.complex128)).item()
    den = torch.linalg.vector_norm(x.to(torch.complex128)).item()
    assert num / den <= 2 * tol(n)


def test_linearity(gpu):
    import torch
    n, B = 2 ** 16, 2
    x = dev_uniform(torch, B, n, seed=31, device=gpu)
    p = estl_amd.Plan(n, 1, estl_amd.FORWARD)
    a, c = 0.5 - 0.25j, -1.5 + 0.75j
    comb = (a * x[0] + c * x[1]).reshape(1, n).contiguous()
    y0 = run(torch, p, x[0:1].contiguous())
    y1 = run(torch, p, x[1:2].contiguous())
    yc = run(torch, p, comb)
    assert oracle.rel_l2(yc.cpu().numpy(), (a * y0 + c * y1).cpu().numpy()) <= tol(n)


@pytest.mark.parametrize("log2n", [28, 30])
def test_maximum_sizes_tone_and_round_trip(gpu, log2n):
    """The largest transforms (2^30 = 8 GiB per buffer; offsets inside a row are 32-bit):
    a unit tone transforms to exactly N at its bin (rel-L2 over the whole spectrum from
    Parseval: |y|^2 summed minus the peak), and backward/N returns the input."""
    import torch
    n = 1 << log2n
    f = 987654321 % n
    t = torch.arange(n, device=gpu, dtype=torch.int64)
    ph = (2 * math.pi / n) * torch.remainder(t * f, n).to(torch.float64)
    del t
    x = torch.polar(torch.ones_like(ph), ph).to(torch.complex64).reshape(1, n)
    del ph
    y = torch.empty_like(x)
    p = estl_amd.Plan(n, 1, estl_amd.FORWARD)
    assert p.launches() >= 3
    p.execute(x, y)
    torch.cuda.synchronize()
    peak = complex(y[0, f].item())
    energy = (y.abs().double() ** 2).sum().item()
    err = math.sqrt(max(energy - abs(peak) ** 2, 0.0) + abs(peak - n) ** 2) / n
    assert err <= tol(n), err
    b = estl_amd.Plan(n, 1, estl_amd.BACKWARD, flags=estl_amd.FLAG_NORMALIZE)
    b.execute(y, y)                     # in place
    torch.cuda.synchronize()
    num = torch.linalg.vector_norm((y - x).to(torch.complex128)).item()
    assert num / math.sqrt(n) <= 2 * tol(n)


def test_one_plan_on_concurrent_streams(gpu):
    """A multi-pass plan executed at once on several streams (each stream gets its own
    scratch): every result equals the same plan run alone, bit for bit."""
    import torch
    n, batch = 1 << 20, 4
    p = estl_amd.Plan(n, batch, estl_amd.FORWARD)
    assert p.launches() >= 2
    xs = [dev_uniform(torch, batch, n, seed=50 + i, device=gpu) for i in range(4)]
    refs = []
    for x in xs:
        y = torch.empty_like(x)
        p.execute(x, y)
        torch.cuda.synchronize()
        refs.append(y)
    streams = [torch.cuda.Stream() for _ in xs]
    outs = [torch.empty_like(x) for x in xs]
    torch.cuda.synchronize()
    for rep in range(3):
        for x, y, s in zip(xs, outs, streams):
            p.execute(x, y, s)
        torch.cuda.synchronize()
        for y, r in zip(outs, refs):
            assert torch.equal(y, r), rep
