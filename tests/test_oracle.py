"""Pins the oracle (oracle/oracle.c) before anything is checked against it.

Each check is chosen so a plausible mistake fails it: closed forms and the reference's
own known-answer test (tests/Test.cpp:45-110), numpy FP64 golden vectors, brute force,
invariants, the per-pass Stockham state, and planted mistakes that must be caught.
"""
import numpy as np
import pytest

import oracle


def test_plan_radices_match_reference_rule():
    # FFT_OpenCL_Contiguous.h:55-60: floor(log2N/4) radix-16 passes, 2^(log2N mod 4) last
    assert oracle.plan_radices(2 ** 16) == [16, 16, 16, 16]
    assert oracle.plan_radices(4096) == [16, 16, 16]
    assert oracle.plan_radices(2 ** 26) == [16] * 6 + [4]
    assert oracle.plan_radices(2) == [2]
    assert oracle.plan_radices(32) == [16, 2]
    assert oracle.plan_radices(2 ** 11) == [16, 16, 8]
    A, N = oracle.prefix([16, 16, 8])
    assert A == [1, 16, 256] and N == [16, 256, 2048]


@pytest.mark.parametrize("n", [2 ** k for k in range(1, 13)])
def test_fft64_matches_numpy_golden(golden, n):
    for key in [k for k in golden if k.startswith(f"in_{n}_")]:
        s = key.split("_")[-1]
        x = golden[key]
        assert oracle.rel_l2(oracle.fft64(x), golden[f"fwd_{n}_{s}"]) < 1e-14
        assert oracle.rel_l2(oracle.fft64(x, oracle.BACKWARD), golden[f"bwd_{n}_{s}"]) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512])
def test_fft64_matches_bruteforce(n):
    x = oracle.uniform_complex(3, n, seed=11)
    for sign in (oracle.FORWARD, oracle.BACKWARD):
        ref = oracle.dft(x, sign)
        got = oracle.fft64(x, sign)
        assert oracle.rel_l2(got, ref) < 1e-14


def test_bruteforce_matches_golden(golden):
    for n in (8, 64, 512):
        assert oracle.rel_l2(oracle.dft(golden[f"in_{n}_0"]), golden[f"fwd_{n}_0"]) < 1e-15


def test_generator_matches_golden(golden):
    for s in range(5):
        assert np.array_equal(oracle.fill_uniform(64, s), golden[f"gen_{s}"])
    assert np.array_equal(oracle.fill_uniform(16, 1, 1000), golden["gen_offset_1_1000"])
    u = oracle.fill_uniform(1 << 16, 3)
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(float(u.mean())) < 0.01  # uniform on [-1, 1)


@pytest.mark.parametrize("m", [2, 3, 4, 5, 6, 8, 10, 12])
def test_reference_known_answer(m):
    # tests/Test.cpp:60-77: row bIx holds cos(2 pi effK q / M); the forward transform
    # is M/2 at bins effK and M - effK and 0 elsewhere
    M = 1 << m
    for k in (1, 2, M // 2 - 1):
        rows = []
        for bIx in range(3):
            effK = ((k + bIx) % (M // 2 - 1)) + 1 if M > 2 else 1
            q = np.arange(M)
            rows.append(np.cos(2 * np.pi * effK * q / M))
        y = oracle.fft64(np.array(rows))
        for bIx in range(3):
            effK = ((k + bIx) % (M // 2 - 1)) + 1 if M > 2 else 1
            exp = np.zeros(M, dtype=complex)
            exp[effK] += M / 2
            exp[M - effK] += M / 2
            assert np.max(np.abs(y[bIx] - exp) ** 2) < 1e-3 * M


@pytest.mark.parametrize("n", [2, 16, 64, 1024, 4096])
def test_closed_forms(n):
    d = np.zeros(n, complex); d[0] = 1
    assert np.allclose(oracle.fft64(d), np.ones(n), atol=1e-13)            # delta -> ones
    assert np.allclose(oracle.fft64(np.ones(n)), n * d, atol=1e-10)          # const -> N delta
    k = 3 % n
    e = np.exp(2j * np.pi * k * np.arange(n) / n)
    exp = np.zeros(n, complex); exp[k] = n
    assert np.allclose(oracle.fft64(e), exp, atol=1e-9)                      # e^{2 pi i k t/N} -> N delta_k


@pytest.mark.parametrize("n", [16, 256, 4096, 2 ** 14])
def test_invariants(n):
    x = oracle.uniform_complex(2, n, seed=5).astype(np.complex128)
    y = oracle.fft64(x)
    # Parseval
    assert np.allclose((np.abs(y) ** 2).sum(axis=1), n * (np.abs(x) ** 2).sum(axis=1), rtol=1e-12)
    # round trip
    assert oracle.rel_l2(oracle.fft64(y, oracle.BACKWARD) / n, x) < 1e-14
    # linearity
    a, b = 0.3 - 0.7j, -1.1 + 0.2j
    assert oracle.rel_l2(oracle.fft64(a * x[0] + b * x[1]), a * y[0] + b * y[1]) < 1e-14
    # shift theorem: x[t - s] -> X[k] W^{k s}
    s = 5
    ks = np.arange(n)
    assert oracle.rel_l2(oracle.fft64(np.roll(x[0], s)), y[0] * np.exp(-2j * np.pi * ks * s / n)) < 1e-13
    # real input -> Hermitian spectrum
    r = oracle.fft64(x[0].real)
    assert np.allclose(r[1:], np.conj(r[1:][::-1]), atol=1e-10 * n)


@pytest.mark.parametrize("n,log2radix", [(256, 4), (2048, 4), (4096, 4), (512, 3), (1024, 5)])
def test_pass_states_are_decimated_dfts(n, log2radix):
    # after pass q: d_q[(L/N_q) f + z] = DFT_{N_q}(x[(L/N_q) t + z])[f]
    # (docs/ImplementationStrategy.pdf eq. 5/7; FFT_OpenCL_Contiguous.h:167-180)
    x = oracle.uniform_complex(1, n, seed=7)[0]
    st = oracle.pass_states(x, oracle.FORWARD, log2radix)
    rad = oracle.plan_radices(n, log2radix)
    _, Nq = oracle.prefix(rad)
    for q, nq in enumerate(Nq):
        stride = n // nq
        for z in (0, stride - 1):
            want = oracle.dft(x[z::stride].astype(np.complex128))
            got = st[q][z::stride]
            assert oracle.rel_l2(got, want) < 1e-13, (q, z)


@pytest.mark.parametrize("inject", [1, 2, 3])
def test_planted_mistakes_are_caught(golden, inject):
    # the checks above must fail on a dropped pre-twiddle (1), a wrong local twiddle
    # sign (2) and a wrong output stride (3)
    x = golden["in_4096_1"]
    bad = oracle.fft64(x, oracle.FORWARD, 4, inject)
    assert oracle.rel_l2(bad, golden["fwd_4096_1"]) > 1e-2


def test_fft32_restatement_accuracy(golden):
    for n in (64, 1024, 4096):
        x = golden[f"in_{n}_1"]
        err = oracle.rel_l2(oracle.fft32(x), golden[f"fwd_{n}_1"])
        assert err < 1e-6 * np.log2(n)


def test_conv_direct_matches_golden(golden):
    for n, k in [(16, 3), (64, 16), (256, 100), (1024, 16), (32, 40)]:
        x, h, y = golden[f"conv_x_{n}_{k}"], golden[f"conv_h_{n}_{k}"], golden[f"conv_y_{n}_{k}"]
        got = oracle.conv_circular(x, h)
        assert np.max(np.abs(got - y)) < 1e-12 * max(1.0, np.max(np.abs(y)))


def test_conv_direct_closed_forms():
    x = oracle.fill_uniform(64, 9).astype(np.float64)
    # identity filter
    assert np.allclose(oracle.conv_circular(x, np.array([1.0])), x)
    # delay by 3
    h = np.zeros(4); h[3] = 1
    assert np.allclose(oracle.conv_circular(x, h), np.roll(x, 3))
    # filter longer than n folds modulo n
    h = np.zeros(64 + 2); h[64 + 1] = 2.0
    assert np.allclose(oracle.conv_circular(x, h), 2 * np.roll(x, 1))


def test_planar_layout_formula():
    # Complex.h:42-44: ((i >> 5) << 6) | (e << 5) | (i & 31)
    assert oracle.planar_pos(0, 0) == 0 and oracle.planar_pos(0, 1) == 32
    assert oracle.planar_pos(31, 1) == 63 and oracle.planar_pos(32, 0) == 64
    assert oracle.planar_pos(100, 1) == 3 * 64 + 32 + 4
