"""Write the committed golden fixtures under tests/golden/.

The fixtures come from numpy's FP64 pocketfft (``numpy.fft``) and an independent
numpy implementation of the splitmix64 input generator -- not from the oracle and
not from the reference (the reference ships no vectors and its parity path, FFTW3,
is absent here).  They pin:

* ``gen_*``      : the seeded uniform[-1,1) generator (first values of seeds 0..4);
* ``fwd_<n>_<s>``: the unnormalised forward DFT (sign -1, FFTW_FORWARD convention of
                   FFT_FFTW3.h:26-28) of the generator's complex64 rows;
* ``bwd_<n>_<s>``: the unnormalised backward DFT (sign +1) of the same rows;
* ``conv_*``     : circular convolutions of real rows with real filters (FP64).

Run:  python tests/golden/gen_golden.py   (rewrites tests/golden/golden.npz)
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(count: int, seed: int, offset: int = 0) -> np.ndarray:
    key = splitmix64(np.array([seed], dtype=np.uint64))[0]
    idx = np.arange(count, dtype=np.uint64) + np.uint64(offset)
    with np.errstate(over="ignore"):
        u = splitmix64(key + idx)
    return ((u >> np.uint64(40)).astype(np.float64) * (1.0 / 8388608.0) - 1.0).astype(np.float32)


def main() -> None:
    out: dict[str, np.ndarray] = {}
    for seed in range(5):
        out[f"gen_{seed}"] = uniform(64, seed)
    out["gen_offset_1_1000"] = uniform(16, 1, offset=1000)

    for n in [2 ** k for k in range(1, 13)]:
        seeds = [0, 1] if n <= 1024 else [1]
        for s in seeds:
            x = uniform(2 * n, s).view(np.complex64).astype(np.complex128)
            out[f"in_{n}_{s}"] = x.astype(np.complex64)
            out[f"fwd_{n}_{s}"] = np.fft.fft(x)
            out[f"bwd_{n}_{s}"] = np.fft.ifft(x) * n

    # circular convolution fixtures (real rows, real filters), FP64 via FFT
    cases = [(16, 3, 4), (64, 16, 4), (256, 100, 4), (1024, 16, 4), (32, 40, 4)]
    for n, k, seed in cases:
        x = uniform(2 * n, seed).astype(np.float64).reshape(2, n)
        h = uniform(k, seed, offset=2 * n).astype(np.float64)
        hp = np.zeros(n)
        np.add.at(hp, np.arange(k) % n, h)
        y = np.real(np.fft.ifft(np.fft.fft(x, axis=-1) * np.fft.fft(hp)[None, :], axis=-1))
        out[f"conv_x_{n}_{k}"] = x
        out[f"conv_h_{n}_{k}"] = h
        out[f"conv_y_{n}_{k}"] = y

    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
