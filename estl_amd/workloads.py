"""Seeded synthetic workloads of the benchmark configurations.

Every random draw comes from the same splitmix64 stream as ``estl_fill_uniform``
(value i of seed s: u = splitmix64(splitmix64(s) + i)), so host and device agree
bit for bit and nothing depends on a library's RNG.

* ``uneven_batch(count, seed)`` -- transform sizes 2^k with k uniform on {8..18},
  packed back to back; 10 % of the transforms are unit-amplitude complex tones
  exp(2 pi i f t / N) with f uniform on [0, N) (their spectrum is exactly N at bin f).
* ``conv_requests(count, seed, ...)`` -- the served convolution stream (see there).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def stream(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """count raw 64-bit draws of stream `seed` starting at `offset`."""
    key = splitmix64(np.array([seed], dtype=np.uint64))[0]
    idx = np.arange(count, dtype=np.uint64) + np.uint64(offset)
    with np.errstate(over="ignore"):
        return splitmix64(key + idx)


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """uniform [0, 1) doubles with 53-bit resolution."""
    return (stream(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class UnevenBatch:
    sizes: np.ndarray      # int64, transform lengths
    offsets: np.ndarray    # int64, packed offsets
    tone: np.ndarray       # bool, transform is a unit tone
    freq: np.ndarray       # int64, tone frequency (0 where not a tone)
    seed: int

    @property
    def total(self) -> int:
        return int(self.sizes.sum())


def uneven_batch(count: int = 8192, seed: int = 3, kmin: int = 8, kmax: int = 18,
                 tone_fraction: float = 0.1) -> UnevenBatch:
    k = kmin + (stream(seed, count) % np.uint64(kmax - kmin + 1)).astype(np.int64)
    sizes = np.left_shift(np.int64(1), k)
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    tone = uniform01(seed ^ 0x5EED, count) < tone_fraction
    f = (stream(seed ^ 0xF4E9, count) % sizes.astype(np.uint64)).astype(np.int64)
    freq = np.where(tone, f, 0)
    return UnevenBatch(sizes=sizes, offsets=offsets, tone=tone, freq=freq, seed=seed)


def fill_uneven(batch: UnevenBatch, buf, torch) -> None:
    """Fill a complex64 device buffer: uniform[-1,1) re/im from the batch seed (the
    estl_fill_uniform stream over the packed buffer), tones overwritten."""
    import estl_amd
    estl_amd.fill_uniform(buf, batch.seed)
    for i in np.nonzero(batch.tone)[0]:
        n, off, f = int(batch.sizes[i]), int(batch.offsets[i]), int(batch.freq[i])
        t = torch.arange(n, device=buf.device, dtype=torch.float64)
        ph = (2.0 * math.pi / n) * torch.remainder(t * f, n)
        buf[off:off + n] = torch.polar(torch.ones_like(ph), ph).to(torch.complex64)


@dataclass
class ConvRequests:
    n: np.ndarray          # int64, row length 2^U{log2n_min..log2n_max}
    rows: np.ndarray       # int64, rows per request, U{1..rows_max} (capped at max_values / n)
    k: np.ndarray          # int64, filter length U{k_min..k_max}
    flags: np.ndarray      # int64, 0 or correlation (correlate_fraction of requests)
    bank: np.ndarray       # int64, index into a bank of `bank_size` filters (bank mode)
    bank_k: np.ndarray     # int64, lengths of the bank's filters, U{k_min..k_max}
    gaps: np.ndarray       # float64, unit-rate exponential inter-arrival gaps
    seed: int

    def values(self) -> np.ndarray:
        return self.n * self.rows


def conv_requests(count: int = 1000, seed: int = 4, log2n=(10, 22), k_range=(16, 8192), rows_max: int = 256,
                  max_values: int = 1 << 30, correlate_fraction: float = 0.25, bank_size: int = 32) -> ConvRequests:
    """The served convolution stream: request i convolves rows[i] real rows of n[i]
    values with its own k[i]-value filter (or, in bank mode, with filter bank[i] of a
    fixed bank); arrivals are a Poisson process (exponential gaps of unit mean, scaled
    by 1 / rate by the driver).  Rows are U{1..rows_max} at every length: the default
    cap of 2^30 values (4 GiB of FP32 input) is what 256 rows of 2^22 reach, so it only
    binds for smaller `max_values`."""
    import estl_amd
    ln = log2n[0] + (stream(seed, count) % np.uint64(log2n[1] - log2n[0] + 1)).astype(np.int64)
    span = np.uint64(k_range[1] - k_range[0] + 1)
    k = k_range[0] + (stream(seed ^ 0x4B, count) % span).astype(np.int64)
    n = np.left_shift(np.int64(1), ln)
    rows = 1 + (stream(seed ^ 0x52, count) % np.uint64(rows_max)).astype(np.int64)
    rows = np.minimum(rows, np.maximum(1, max_values // n))
    corr = uniform01(seed ^ 0xC0, count) < correlate_fraction
    flags = np.where(corr, estl_amd.CONV_CORRELATE, 0).astype(np.int64)
    bank = (stream(seed ^ 0xBA, count) % np.uint64(bank_size)).astype(np.int64)
    bank_k = k_range[0] + (stream(seed ^ 0xBB, bank_size) % span).astype(np.int64)
    gaps = -np.log1p(-uniform01(seed ^ 0x9A, count))
    return ConvRequests(n=n, rows=rows, k=k, flags=flags, bank=bank, bank_k=bank_k, gaps=gaps, seed=seed)
