"""GPU parity of the batched circular convolution (estl_conv_*) against the oracle's
direct O(n k) long-double sum (oracle.conv_circular) and the numpy golden fixtures.

Tolerance: relative L2 per row <= 2e-6 * log2(n) (a forward and an inverse FP32
transform plus a complex multiply between them).
"""
import math

import numpy as np
import pytest

import estl_amd
import oracle

pytestmark = pytest.mark.gpu


def tol(n):
    return 2e-6 * max(1, int(math.log2(n)))


def conv_gpu(torch, gpu, x, h, flags=0):
    rows, n = x.shape
    c = estl_amd.Convolution(n, rows, flags)
    dx = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(gpu)
    dh = torch.from_numpy(np.ascontiguousarray(h, dtype=np.float32)).to(gpu)
    dy = torch.empty_like(dx)
    c.set_filter(dh)
    c.execute(dx, dy)
    torch.cuda.synchronize()
    return dy.cpu().numpy().astype(np.float64)


def folded(h, n):
    hp = np.zeros(n)
    np.add.at(hp, np.arange(len(h)) % n, h)
    return hp


def check_rows(y, ref, n):
    for r in range(ref.shape[0]):
        assert oracle.rel_l2(y[r], ref[r]) <= tol(n), (r, oracle.rel_l2(y[r], ref[r]))


@pytest.mark.parametrize("n,rows,k", [
    (16, 1, 3), (64, 3, 16), (256, 4, 100), (1024, 5, 16), (4096, 2, 300), (8192, 3, 64),
    (2 ** 14, 3, 512), (2 ** 16, 2, 1000), (2 ** 17, 1, 64), (32, 2, 40), (2 ** 14, 2, 2 ** 14 + 77),
])
def test_conv_matches_direct_sum(gpu, n, rows, k):
    import torch
    x = oracle.fill_uniform(rows * n, 4).astype(np.float64).reshape(rows, n)
    h = oracle.fill_uniform(k, 4, offset=rows * n).astype(np.float64)
    y = conv_gpu(torch, gpu, x.astype(np.float32), h.astype(np.float32))
    ref = oracle.conv_circular(x.astype(np.float32).astype(np.float64), h.astype(np.float32).astype(np.float64))
    check_rows(y, ref, n)


def test_conv_golden_fixtures(gpu, golden):
    import torch
    for n, k in [(16, 3), (64, 16), (256, 100), (1024, 16), (32, 40)]:
        x, h, yref = golden[f"conv_x_{n}_{k}"], golden[f"conv_h_{n}_{k}"], golden[f"conv_y_{n}_{k}"]
        y = conv_gpu(torch, gpu, x.astype(np.float32), h.astype(np.float32))
        check_rows(y, yref, n)


@pytest.mark.parametrize("n", [256, 4096, 2 ** 15])
def test_correlation_center_and_normalize(gpu, n):
    import torch
    rows, k = 3, 33
    x = oracle.fill_uniform(rows * n, 8).astype(np.float32).reshape(rows, n)
    h = (oracle.fill_uniform(k, 9).astype(np.float32) + 1.5).astype(np.float32)  # nonzero sum
    x64, h64 = x.astype(np.float64), h.astype(np.float64)
    hp = folded(h64, n)
    # correlation: y[i] = sum_j hp[j] x[i + j] = convolution with hp reversed circularly
    g = np.roll(hp[::-1], 1)
    y = conv_gpu(torch, gpu, x, h, estl_amd.CONV_CORRELATE)
    check_rows(y, oracle.conv_circular(x64, g), n)
    # centre: the filter rotated so its index (k+1)/2 lands on 0 (Convolution.h:113-120)
    s = (k + 1) // 2
    y = conv_gpu(torch, gpu, x, h, estl_amd.CONV_CENTER)
    check_rows(y, oracle.conv_circular(x64, np.roll(hp, -s)), n)
    # normalise by the kernel sum (Convolution.h:141)
    y = conv_gpu(torch, gpu, x, h, estl_amd.CONV_NORMALIZE_KERNEL)
    check_rows(y, oracle.conv_circular(x64, hp / hp.sum()), n)


@pytest.mark.parametrize("n", [64, 2 ** 14])
def test_long_filter_centred_and_normalised(gpu, n):
    """A filter longer than n (folds modulo n) with the centre rotation and the kernel-sum
    normalisation together: the spectrum kernel builds the folded, rotated filter as it
    loads the taps (one workgroup for n <= 2^13, the first spectrum pass above)."""
    import torch
    rows, k = 3, n + 77
    x = oracle.fill_uniform(rows * n, 12).astype(np.float32).reshape(rows, n)
    h = (oracle.fill_uniform(k, 13).astype(np.float32) + 1.25).astype(np.float32)
    hp = folded(h.astype(np.float64), n)
    s = ((k + 1) // 2) % n
    y = conv_gpu(torch, gpu, x, h, estl_amd.CONV_CENTER | estl_amd.CONV_NORMALIZE_KERNEL)
    check_rows(y, oracle.conv_circular(x.astype(np.float64), np.roll(hp, -s) / hp.sum()), n)


@pytest.mark.parametrize("n,k,period", [
    (2 ** 15, 256, 256), (2 ** 15, 100, 256), (2 ** 19, 8192, 2 ** 13), (2 ** 19, 1000, 2 ** 13),
    (2 ** 20, 2 ** 14, 2 ** 14), (2 ** 22, 3000, 2 ** 15),
])
def test_spectrum_first_pass_skip_is_exact(gpu, n, k, period):
    """A filter whose taps all lie within the first spectrum pass's stride (`period`,
    n over the first radix of the default split) skips that pass: the second reads the
    taps.  The same filter padded with zero taps to one past the stride takes the full
    pass chain.  The two results are identical, and both match the convolution theorem
    in FP64 (numpy's FFT of the folded filter); at 2^15 also the oracle's direct sum."""
    import torch
    rows = 2 if n <= 2 ** 19 else 1
    x = oracle.fill_uniform(rows * n, 50).astype(np.float32).reshape(rows, n)
    h = oracle.fill_uniform(k, 51).astype(np.float32)
    hpad = np.concatenate([h, np.zeros(period + 1 - k, np.float32)])
    full = estl_amd.CONV_FULL_LENGTH  # whole-row transforms (their spectrum plan)
    y = conv_gpu(torch, gpu, x, h, full)
    ypad = conv_gpu(torch, gpu, x, hpad, full)
    assert np.array_equal(y, ypad)
    x64 = x.astype(np.float64)
    ref = np.real(np.fft.ifft(np.fft.fft(x64, axis=1) * np.fft.fft(folded(h.astype(np.float64), n))[None, :], axis=1))
    check_rows(y, ref, n)
    if n <= 2 ** 15:
        check_rows(y, oracle.conv_circular(x64, h.astype(np.float64)), n)


@pytest.mark.parametrize("n,k,flags", [
    (2 ** 14, 100, 0), (2 ** 14, 2 ** 14 + 5, estl_amd.CONV_CORRELATE), (2 ** 15, 300, estl_amd.CONV_CENTER),
    (2 ** 16, 1000, estl_amd.CONV_NORMALIZE_KERNEL), (2 ** 18, 64, estl_amd.CONV_CORRELATE | estl_amd.CONV_CENTER),
    (2 ** 20, 5000, 0), (2 ** 22, 8192, estl_amd.CONV_CORRELATE),
])
def test_single_row_half_length(gpu, n, k, flags):
    """One FP32 row of n >= 2^14 runs through complex transforms of n/2 values (its even
    and odd samples as one complex row) with the row's spectrum rebuilt between them.
    Checked against the convolution theorem in FP64 (every size; also the oracle's
    direct sum for a plain convolution at 2^14), for convolution, correlation, centring and normalisation; the
    same row at an odd float offset (not 8-byte aligned) takes the pair path and agrees
    within the same bound."""
    import torch
    x = oracle.fill_uniform(n, 60).astype(np.float32).reshape(1, n)
    h = (oracle.fill_uniform(k, 61) + (1.25 if flags & estl_amd.CONV_NORMALIZE_KERNEL else 0.0)).astype(np.float32)
    hp = folded(h.astype(np.float64), n)
    if flags & estl_amd.CONV_CENTER:
        hp = np.roll(hp, -(((k + 1) // 2) % n))
    if flags & estl_amd.CONV_NORMALIZE_KERNEL:
        hp = hp / hp.sum()
    H = np.fft.fft(hp)
    X = np.fft.fft(x.astype(np.float64), axis=1)
    ref = np.real(np.fft.ifft(X * (np.conj(H) if flags & estl_amd.CONV_CORRELATE else H)[None, :], axis=1))
    flags |= estl_amd.CONV_FULL_LENGTH  # whole-row transforms (not overlap-save windows)
    y = conv_gpu(torch, gpu, x, h, flags)
    check_rows(y, ref, n)
    if n <= 2 ** 16 and flags == estl_amd.CONV_FULL_LENGTH:
        check_rows(y, oracle.conv_circular(x.astype(np.float64), h.astype(np.float64)), n)
    # misaligned row: the pair path
    c = estl_amd.Convolution(n, 1, flags)
    c.set_filter(torch.from_numpy(h).to(gpu))
    buf = torch.zeros(2 * n + 4, dtype=torch.float32, device=gpu)
    xin = buf[1:1 + n]
    xin.copy_(torch.from_numpy(x[0]).to(gpu))
    yout = buf[n + 3:2 * n + 3]
    c.execute(xin, yout)
    torch.cuda.synchronize()
    check_rows(yout.cpu().numpy().astype(np.float64).reshape(1, n), ref, n)


def test_random_configurations_match_fp64(gpu):
    """Forty seeded configurations over the whole shape space -- n = 2^4..2^22, 1..9 rows
    (capped at 2^22 values), filters of 1..2n taps, every flag combination -- each
    checked row by row against the convolution theorem in FP64 (numpy's FFT of the
    folded, rotated, normalised filter).  Covers every pipeline: the fused rows kernel,
    overlap-save windows, and (whole-row transforms) the pass chain with its middle
    pass, single rows through half-length transforms, and the spectrum plan with and
    without its first pass."""
    import torch
    rng = np.random.default_rng(2024)
    for case in range(40):
        L = int(rng.integers(4, 23))
        n = 1 << L
        rows = int(min(rng.integers(1, 10), max(1, (1 << 22) // n)))
        k = int(rng.integers(1, 2 * n + 1)) if rng.random() < 0.3 else int(rng.integers(1, min(n, 8192) + 1))
        bits = int(rng.integers(0, 16))
        flags = ((estl_amd.CONV_CORRELATE if bits & 1 else 0) | (estl_amd.CONV_NORMALIZE_KERNEL if bits & 2 else 0) |
                 (estl_amd.CONV_CENTER if bits & 4 else 0) | (estl_amd.CONV_FULL_LENGTH if bits & 8 else 0))
        x = oracle.fill_uniform(rows * n, 100 + case).astype(np.float32).reshape(rows, n)
        h = oracle.fill_uniform(k, 200 + case).astype(np.float32)
        if flags & estl_amd.CONV_NORMALIZE_KERNEL:
            h = (h + 1.5).astype(np.float32)
        hp = folded(h.astype(np.float64), n)
        if flags & estl_amd.CONV_CENTER:
            hp = np.roll(hp, -(((k + 1) // 2) % n))
        if flags & estl_amd.CONV_NORMALIZE_KERNEL:
            hp = hp / hp.sum()
        H = np.fft.fft(hp)
        if flags & estl_amd.CONV_CORRELATE:
            H = np.conj(H)
        ref = np.real(np.fft.ifft(np.fft.fft(x.astype(np.float64), axis=1) * H[None, :], axis=1))
        y = conv_gpu(torch, gpu, x, h, flags)
        for r in range(rows):
            e = oracle.rel_l2(y[r], ref[r])
            assert e <= tol(n), (case, n, rows, k, flags, r, e)


def test_identity_and_delay_closed_forms(gpu):
    import torch
    n, rows = 2 ** 14, 5
    x = oracle.fill_uniform(rows * n, 10).astype(np.float32).reshape(rows, n)
    y = conv_gpu(torch, gpu, x, np.array([1.0], np.float32))
    check_rows(y, x.astype(np.float64), n)
    d = np.zeros(8, np.float32)
    d[7] = 1.0
    y = conv_gpu(torch, gpu, x, d)
    check_rows(y, np.roll(x.astype(np.float64), 7, axis=1), n)


def test_filter_reuse_and_odd_rows(gpu):
    import torch
    n, rows, k = 4096, 7, 50
    c = estl_amd.Convolution(n, rows)
    h = torch.from_numpy(oracle.fill_uniform(k, 12)).to(gpu)
    c.set_filter(h)
    ys = []
    for seed in (1, 2):
        x = torch.from_numpy(oracle.fill_uniform(rows * n, seed)).to(gpu)
        y = torch.empty_like(x)
        c.execute(x, y)
        torch.cuda.synchronize()
        ref = oracle.conv_circular(x.cpu().numpy().astype(np.float64).reshape(rows, n),
                                   h.cpu().numpy().astype(np.float64))
        check_rows(y.cpu().numpy().reshape(rows, n).astype(np.float64), ref, n)
        ys.append(y)
    # a second execute with the same inputs is bit-identical
    y2 = torch.empty_like(ys[-1])
    c.execute(x, y2)
    torch.cuda.synchronize()
    assert torch.equal(y2, ys[-1])


# ---------------------------------------------------------------- FP16 storage
U16 = 2.0 ** -11


def conv_fp16(torch, gpu, x, h, flags=0):
    """x: (rows, n) float32 (values of any magnitude); returns (y float64, x as
    dequantised from its binary16 form, launches)."""
    rows, n = x.shape
    c = estl_amd.Convolution(n, rows, flags, precision=estl_amd.FP16)
    dx = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(gpu)
    x16 = torch.empty(rows * n, dtype=torch.float16, device=gpu)
    s_in = torch.empty(rows, dtype=torch.float32, device=gpu)
    estl_amd.pack_fp16(dx, x16, s_in, n // 2, rows)        # real rows as n/2 complex values
    dh = torch.from_numpy(np.ascontiguousarray(h, dtype=np.float32)).to(gpu)
    c.set_filter(dh)
    y16 = torch.empty_like(x16)
    s_out = torch.empty_like(s_in)
    c.execute_fp16(x16, s_in, y16, s_out)
    torch.cuda.synchronize()
    y = (y16.float().view(rows, n) * s_out[:, None]).cpu().numpy().astype(np.float64)
    xq = (x16.float().view(rows, n) * s_in[:, None]).cpu().numpy().astype(np.float64)
    assert float(y16.float().abs().max()) <= 1.0
    return y, xq


def fp16_tol(n):
    # one binary16 store per launch: 1 (n <= 2^13) or 2P - 1 for a P-pass split
    P = 0 if n <= 2 ** 13 else max(2, (int(math.log2(n)) + 8) // 9)
    stores = 1 if P == 0 else 2 * P - 1
    return U16 * math.sqrt(stores + 1)


@pytest.mark.parametrize("n,rows,k,flags", [
    (16, 1, 3, 0), (1024, 5, 16, 0), (4096, 3, 300, estl_amd.CONV_CORRELATE), (8192, 4, 8192, 0),
    (2 ** 14, 3, 512, 0), (2 ** 16, 2, 1000, estl_amd.CONV_CORRELATE), (2 ** 18, 1, 64, 0),
    (2 ** 15, 5, 33, estl_amd.CONV_CENTER), (4096, 2, 33, estl_amd.CONV_NORMALIZE_KERNEL),
    (2 ** 14, 2, 33, estl_amd.CONV_NORMALIZE_KERNEL),
])
def test_conv_fp16_matches_direct_sum(gpu, n, rows, k, flags):
    import torch
    x = oracle.fill_uniform(rows * n, 30).astype(np.float64).reshape(rows, n)
    x[0] *= 1e-4                     # per-row scaling: rows of very different magnitude
    if rows > 1:
        x[-1] *= 3e3
    h = oracle.fill_uniform(k, 31).astype(np.float64)
    if flags & estl_amd.CONV_NORMALIZE_KERNEL:
        h += 1.5
    y, xq = conv_fp16(torch, gpu, x.astype(np.float32), h.astype(np.float32), flags)
    h32 = h.astype(np.float32).astype(np.float64)
    # references through the oracle's direct sum over the k taps (O(n k)):
    #   correlation y[i] = sum_j h[j] x[i + j] = (x_r (*) h)[-i] with x_r[m] = x[-m];
    #   centring rotates the output: y[i] = (x (*) h)[i + (k + 1)/2];
    #   normalising divides the filter by its sum
    if flags & estl_amd.CONV_CORRELATE:
        xr = np.roll(xq[:, ::-1], 1, axis=1)
        ref = np.roll(oracle.conv_circular(xr, h32)[:, ::-1], 1, axis=1)
    elif flags & estl_amd.CONV_CENTER:
        ref = np.roll(oracle.conv_circular(xq, h32), -((k + 1) // 2), axis=1)
    elif flags & estl_amd.CONV_NORMALIZE_KERNEL:
        ref = oracle.conv_circular(xq, h32 / h32.sum())
    else:
        ref = oracle.conv_circular(xq, h32)
    for r in range(rows):
        e = oracle.rel_l2(y[r], ref[r])
        assert e <= fp16_tol(n), (n, r, e, fp16_tol(n))


def test_conv_fp16_output_scale_is_input_scale_times_l1(gpu):
    import torch
    n, rows, k = 4096, 2, 100
    x = oracle.fill_uniform(rows * n, 40).astype(np.float32).reshape(rows, n)
    h = oracle.fill_uniform(k, 41).astype(np.float32)
    c = estl_amd.Convolution(n, rows, 0, precision=estl_amd.FP16)
    dx = torch.from_numpy(x).to(gpu)
    x16 = torch.empty(rows * n, dtype=torch.float16, device=gpu)
    s_in = torch.empty(rows, dtype=torch.float32, device=gpu)
    estl_amd.pack_fp16(dx, x16, s_in, n // 2, rows)
    c.set_filter(torch.from_numpy(h).to(gpu))
    y16, s_out = torch.empty_like(x16), torch.empty_like(s_in)
    c.execute_fp16(x16, s_in, y16, s_out)
    torch.cuda.synchronize()
    l1 = float(np.abs(h).astype(np.float32).sum(dtype=np.float32))
    np.testing.assert_allclose(s_out.cpu().numpy(), s_in.cpu().numpy() * l1, rtol=1e-6)
    with pytest.raises(estl_amd.EstlError):
        c.execute(dx, dx)            # an FP16 object has no FP32 entry point


@pytest.mark.parametrize("n", [4096, 2 ** 14])
def test_conv_fp16_normalised_output_scale(gpu, n):
    """With kernel normalisation the filter is h / sum(h), whose L1 bound is
    sum|h| / |sum h|: the output scale is the input scale times that (the spectrum of
    n <= 2^13 is normalised inside its one launch, which also records the kernel sum;
    larger n normalise after the spectrum passes)."""
    import torch
    rows, k = 3, 500
    x = oracle.fill_uniform(rows * n, 42).astype(np.float32).reshape(rows, n)
    h = (oracle.fill_uniform(k, 43) + 0.5).astype(np.float32)     # mixed signs, nonzero sum
    c = estl_amd.Convolution(n, rows, estl_amd.CONV_NORMALIZE_KERNEL, precision=estl_amd.FP16)
    dx = torch.from_numpy(x).to(gpu)
    x16 = torch.empty(rows * n, dtype=torch.float16, device=gpu)
    s_in = torch.empty(rows, dtype=torch.float32, device=gpu)
    estl_amd.pack_fp16(dx, x16, s_in, n // 2, rows)
    c.set_filter(torch.from_numpy(h).to(gpu))
    y16, s_out = torch.empty_like(x16), torch.empty_like(s_in)
    c.execute_fp16(x16, s_in, y16, s_out)
    torch.cuda.synchronize()
    h64 = h.astype(np.float64)
    L = np.abs(h64).sum() / abs(h64.sum())
    np.testing.assert_allclose(s_out.cpu().numpy(), s_in.cpu().numpy() * L, rtol=1e-5)
    y = (y16.float().view(rows, n) * s_out[:, None]).cpu().numpy().astype(np.float64)
    xq = (x16.float().view(rows, n) * s_in[:, None]).cpu().numpy().astype(np.float64)
    ref = oracle.conv_circular(xq, h64 / h64.sum())
    for r in range(rows):
        assert oracle.rel_l2(y[r], ref[r]) <= fp16_tol(n), r


def fp64_reference(x, h, n, flags):
    """The circular convolution (or correlation) of each row of x with the folded,
    optionally rotated and normalised filter, by the convolution theorem in FP64."""
    hp = folded(h.astype(np.float64), n)
    k = len(h)
    if flags & estl_amd.CONV_CENTER:
        hp = np.roll(hp, -(((k + 1) // 2) % n))
    if flags & estl_amd.CONV_NORMALIZE_KERNEL:
        hp = hp / hp.sum()
    H = np.fft.fft(hp)
    if flags & estl_amd.CONV_CORRELATE:
        H = np.conj(H)
    return np.real(np.fft.ifft(np.fft.fft(x.astype(np.float64), axis=1) * H[None, :], axis=1))


OLS_CASES = [(2 ** 14, 3, 1, 0), (2 ** 14, 2, 17, estl_amd.CONV_CORRELATE), (2 ** 15, 1, 4097, 0),
             (2 ** 16, 3, 1000, estl_amd.CONV_CENTER), (2 ** 17, 2, 4096, estl_amd.CONV_CORRELATE | estl_amd.CONV_CENTER),
             (2 ** 18, 1, 333, estl_amd.CONV_NORMALIZE_KERNEL), (2 ** 20, 2, 4000, 0),
             (2 ** 22, 1, 2048, estl_amd.CONV_CORRELATE), (2 ** 22, 2, 8, estl_amd.CONV_CENTER),
             # windows of 2^14 values (n >= 2^15, 2049..8193 taps) and the bounds between them
             (2 ** 14, 2, 4097, 0), (2 ** 15, 2, 8193, 0), (2 ** 15, 1, 2049, estl_amd.CONV_CORRELATE),
             (2 ** 16, 3, 6000, estl_amd.CONV_CORRELATE | estl_amd.CONV_CENTER),
             (2 ** 20, 2, 8192, estl_amd.CONV_NORMALIZE_KERNEL), (2 ** 22, 1, 5001, estl_amd.CONV_CENTER)]


@pytest.mark.parametrize("n,rows,k,flags", OLS_CASES)
def test_overlap_save_matches_fp64(gpu, n, rows, k, flags):
    """Rows longer than a window with filters of at most 8193 taps run overlap-save:
    windows of 2^13 values (at most 4097 taps; the only length for rows of 2^14) or 2^14
    values (rows of 2^15 and more; the only length above 4097 taps), two windows per
    complex transform, from any rows, one launch, the outputs of the last window cut at
    n and the circular wrap read across the row's ends.  Checked against
    the convolution theorem in FP64 for convolution, correlation, centring and
    normalisation, at the window-size bound (log2 2^14 terms), against the oracle's
    direct sum at 2^14, and against the whole-row pipeline (ESTL_CONV_FULL_LENGTH)
    within the two paths' bounds."""
    import torch
    x = oracle.fill_uniform(rows * n, 70).astype(np.float32).reshape(rows, n)
    h = (oracle.fill_uniform(k, 71) + (1.25 if flags & estl_amd.CONV_NORMALIZE_KERNEL else 0.0)).astype(np.float32)
    ref = fp64_reference(x, h, n, flags)
    y = conv_gpu(torch, gpu, x, h, flags)
    for r in range(rows):
        assert oracle.rel_l2(y[r], ref[r]) <= 2e-6 * 14, (r, oracle.rel_l2(y[r], ref[r]))
    if n == 2 ** 14 and not flags & (estl_amd.CONV_CORRELATE | estl_amd.CONV_CENTER):
        check_rows(y, oracle.conv_circular(x.astype(np.float64), h.astype(np.float64)), n)
    yf = conv_gpu(torch, gpu, x, h, flags | estl_amd.CONV_FULL_LENGTH)
    for r in range(rows):
        assert oracle.rel_l2(y[r], yf[r]) <= 2 * tol(n)


def test_overlap_save_closed_forms(gpu):
    """Exact answers: a unit impulse filter returns the rows unchanged, a delayed impulse
    (k taps, the last one 1) rotates them by k - 1 (a correlation by -(k - 1)), and a
    constant row convolved with any filter is the filter sum everywhere -- across the
    window boundaries and the circular wrap."""
    import torch
    n, rows = 2 ** 16, 3
    x = oracle.fill_uniform(rows * n, 72).astype(np.float32).reshape(rows, n)
    one = np.array([1.0], np.float32)
    assert np.allclose(conv_gpu(torch, gpu, x, one), x, rtol=0, atol=2e-6)
    for k in (2000, 3001, 8193):  # windows of 2^13, 2^14 and 2^14 at the tap bound
        d = np.zeros(k, np.float32)
        d[-1] = 1.0
        assert np.allclose(conv_gpu(torch, gpu, x, d), np.roll(x, k - 1, axis=1), rtol=0, atol=2e-6), k
        assert np.allclose(conv_gpu(torch, gpu, x, d, estl_amd.CONV_CORRELATE), np.roll(x, -(k - 1), axis=1), rtol=0,
                           atol=2e-6), k
    c = np.full((1, n), 0.5, np.float32)
    h = oracle.fill_uniform(777, 73).astype(np.float32)
    y = conv_gpu(torch, gpu, c, h)
    assert np.allclose(y, 0.5 * h.astype(np.float64).sum(), rtol=0, atol=1e-4)


@pytest.mark.parametrize("n,rows,k,flags", [(2 ** 15, 3, 500, 0), (2 ** 20, 2, 4097, estl_amd.CONV_CORRELATE),
                                            (2 ** 17, 1, 64, estl_amd.CONV_CENTER | estl_amd.CONV_NORMALIZE_KERNEL),
                                            (2 ** 16, 2, 8000, estl_amd.CONV_CENTER)])
def test_overlap_save_fp16_storage(gpu, n, rows, k, flags):
    """FP16 storage through overlap-save windows: binary16 rows and per-row scales in,
    binary16 rows in [-1, 1] and scales (input scale x L) out, within the FP16 bound of
    the pack and the one store against the FP64 convolution theorem."""
    import torch
    x = oracle.fill_uniform(rows * n, 74).astype(np.float32).reshape(rows, n)
    h = (oracle.fill_uniform(k, 75) + (1.25 if flags & estl_amd.CONV_NORMALIZE_KERNEL else 0.0)).astype(np.float32)
    dx = torch.from_numpy(x.reshape(-1)).to(gpu)
    x16 = torch.empty(rows * n, dtype=torch.float16, device=gpu)
    s_in = torch.empty(rows, dtype=torch.float32, device=gpu)
    estl_amd.pack_fp16(dx, x16, s_in, n // 2, rows)
    c = estl_amd.Convolution(n, rows, flags, precision=estl_amd.FP16)
    c.set_filter(torch.from_numpy(h).to(gpu))
    y16, s_out = torch.empty_like(x16), torch.empty_like(s_in)
    c.execute_fp16(x16, s_in, y16, s_out)
    torch.cuda.synchronize()
    assert y16.abs().max().item() <= 1.0
    y = (y16.float().reshape(rows, n) * s_out[:, None]).cpu().numpy().astype(np.float64)
    # the stored input (binary16 x scale), convolved exactly
    xs = (x16.float().reshape(rows, n) * s_in[:, None]).cpu().numpy()
    ref = fp64_reference(xs, h, n, flags)
    for r in range(rows):
        # one binary16 store of the output relative to its L1 bound, the inputs exact
        assert oracle.rel_l2(y[r], ref[r]) <= 2.0 ** -11 * 4, (r, oracle.rel_l2(y[r], ref[r]))
