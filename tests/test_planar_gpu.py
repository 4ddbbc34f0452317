"""GPU tests of the tile-planarized layout (ESTL_FLAG_PLANAR): the layout of the
reference's PlannarizedComplexArray<5> (Complex.h:34-48,97-130) -- per 32 values, 32
reals then their 32 imaginaries -- read and written by the kernels directly.  A planar
plan must give bit-identical values to the interleaved plan of the same shape (the
arithmetic is the same; only the addressing differs), and leave every float outside
the values' positions untouched.
"""
import numpy as np
import pytest

import estl_amd
import oracle

pytestmark = pytest.mark.gpu


def planar_floats(count):
    last = count - 1
    return (((last >> 5) << 6) | (last & 31)) + 33


def positions(count, element):
    """Float positions of values 0..count-1 (element 0 = real, 1 = imaginary); the
    vectorised map is checked against the oracle's (pinned to the reference's own
    ArrayPlannarizer) at every index up to 2^10 and a sample beyond."""
    i = np.arange(count, dtype=np.int64)
    pos = ((i >> 5) << 6) | (element << 5) | (i & 31)
    for k in list(range(min(count, 1024))) + list(range(1024, count, 4093)):
        assert pos[k] == oracle.planar_pos(int(k), element)
    return pos


def to_planar(z):
    """complex64 (count,) -> float32 planar array."""
    count = z.size
    out = np.full(planar_floats(count), np.nan, np.float32)
    out[positions(count, 0)] = z.real
    out[positions(count, 1)] = z.imag
    return out


def from_planar(f, count):
    return (f[positions(count, 0)] + 1j * f[positions(count, 1)]).astype(np.complex64)


CASES = [(1, 1), (2, 1), (8, 3), (16, 1), (32, 1), (64, 5), (4096, 7), (8192, 2), (2 ** 14, 3), (2 ** 16, 1),
         (2 ** 20, 1), (2, 16)]


@pytest.mark.parametrize("n,batch", CASES)
@pytest.mark.parametrize("direction", [estl_amd.FORWARD, estl_amd.BACKWARD])
def test_planar_equals_interleaved(gpu, n, batch, direction):
    import torch
    count = n * batch
    z = oracle.uniform_complex(batch, n, 9).astype(np.complex64).reshape(-1)
    ref_in = torch.from_numpy(z).to(gpu)
    ref_out = torch.empty_like(ref_in)
    estl_amd.Plan(n, batch, direction).execute(ref_in, ref_out)
    pin = torch.from_numpy(to_planar(z)).to(gpu)
    pout = torch.full_like(pin, float("nan"))
    # pad slots between the values must stay untouched: fill the output with a marker
    marker = np.float32(-12345.0)
    pout.fill_(float(marker))
    p = estl_amd.Plan(n, batch, direction, flags=estl_amd.FLAG_PLANAR)
    p.execute(pin, pout)
    torch.cuda.synchronize()
    got = pout.cpu().numpy()
    assert np.array_equal(from_planar(got, count), ref_out.cpu().numpy())
    if n > 1:  # (n = 1 is a copy of the whole span, pad slots included)
        used = np.zeros(got.size, bool)
        used[positions(count, 0)] = True
        used[positions(count, 1)] = True
        assert np.all(got[~used] == marker)


def test_planar_random_pass_plans(gpu):
    """Twelve seeded pass plans (random split and tile widths, so the planar first-pass
    loads and last-pass stores take both their direct and their staged paths) give
    bit-identical values to the interleaved plan of the same split and tiles."""
    import torch
    rng = np.random.default_rng(17)
    for case in range(12):
        L = int(rng.integers(6, 19))
        n = 1 << L
        batch = int(rng.integers(1, 4))
        P = int(rng.integers(2, 4))
        while True:
            cuts = sorted(rng.choice(np.arange(1, L), size=P - 1, replace=False).tolist())
            split = [b - a for a, b in zip([0] + cuts, cuts + [L])]
            if max(split) <= 13:
                break
        tiles = [int(rng.integers(0, max(0, min(6, 14 - b, L - b)) + 1)) for b in split]
        direction = estl_amd.FORWARD if rng.random() < 0.5 else estl_amd.BACKWARD
        count = n * batch
        z = oracle.uniform_complex(batch, n, 40 + case).astype(np.complex64).reshape(-1)
        ref_in = torch.from_numpy(z).to(gpu)
        ref_out = torch.empty_like(ref_in)
        estl_amd.Plan(n, batch, direction, split=split, tiles=tiles).execute(ref_in, ref_out)
        pin = torch.from_numpy(to_planar(z)).to(gpu)
        pout = torch.zeros_like(pin)
        estl_amd.Plan(n, batch, direction, flags=estl_amd.FLAG_PLANAR, split=split, tiles=tiles).execute(pin, pout)
        torch.cuda.synchronize()
        got = from_planar(pout.cpu().numpy(), count)
        assert np.array_equal(got, ref_out.cpu().numpy()), (case, n, batch, split, tiles, direction)


def test_planar_in_place_and_host(gpu):
    import torch
    n, batch = 2 ** 15, 2
    z = oracle.uniform_complex(batch, n, 3).astype(np.complex64)
    p = estl_amd.Plan(n, batch, flags=estl_amd.FLAG_PLANAR)
    buf = torch.from_numpy(to_planar(z.reshape(-1))).to(gpu)
    p.execute(buf, buf)
    torch.cuda.synchronize()
    y = from_planar(buf.cpu().numpy(), n * batch).reshape(batch, n)
    for r in range(batch):
        assert oracle.rel_l2(y[r], oracle.fft64(z[r].astype(np.complex128), -1)) <= 1e-6 * 15
    # host entry point on planar host arrays
    h_in = to_planar(z.reshape(-1))
    h_out = np.zeros_like(h_in)
    p.execute_host(h_in, h_out)
    assert np.array_equal(from_planar(h_out, n * batch), y.reshape(-1))
    with pytest.raises(estl_amd.EstlError):
        estl_amd.Plan(64, 1, precision=estl_amd.FP16, flags=estl_amd.FLAG_PLANAR)
