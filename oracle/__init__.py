"""Oracle: a CPU restatement of ESTL's FFT hot path (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import this package, and only as the checker (or the timed CPU baseline).  The
product library (``estl_amd``) never loads it.

The numerics live in ``oracle.c`` (compiled to ``liboracle.so`` by ``build.py``);
this module is a thin ctypes/numpy wrapper.  See the header of ``oracle.c`` for the
reference file:line each function restates.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FORWARD = -1
BACKWARD = +1


def lib():
    """Load liboracle.so (built by oracle/build.py)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} missing: run oracle/build.py (or __graft_entry__.build())")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        lp = ctypes.POINTER(ctypes.c_long)
        cl, ci = ctypes.c_long, ctypes.c_int
        L.oracle_plan_radices.argtypes = [cl, ci, ip]
        L.oracle_plan_radices.restype = ci
        L.oracle_prefix.argtypes = [ci, ip, lp, lp]
        L.oracle_fft64_ex.argtypes = [cl, cl, dp, dp, ci, ci, ci]
        L.oracle_fft64_ex.restype = ci
        L.oracle_fft32_ex.argtypes = [cl, cl, fp, fp, ci, ci, ci]
        L.oracle_fft32_ex.restype = ci
        L.oracle_pass_states64.argtypes = [cl, dp, dp, ci, ci]
        L.oracle_pass_states64.restype = ci
        L.oracle_dft_ld.argtypes = [cl, cl, dp, dp, ci]
        L.oracle_dft_ld.restype = ci
        L.oracle_conv_circular.argtypes = [cl, cl, dp, cl, dp, dp]
        L.oracle_conv_circular.restype = ci
        L.oracle_planar_pos.argtypes = [cl, ci]
        L.oracle_planar_pos.restype = cl
        L.oracle_fill_uniform.argtypes = [fp, cl, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_fill_uniform.restype = None
        _lib = L
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def plan_radices(n: int, log2radix: int = 4) -> list[int]:
    """Global radix list of FFT_OpenCL_Contiguous.h:55-60."""
    buf = (ctypes.c_int * 64)()
    p = lib().oracle_plan_radices(n, log2radix, buf)
    return [buf[i] for i in range(p)]


def prefix(radices: list[int]) -> tuple[list[int], list[int]]:
    """computeParameters of OpenCLFFTAlgorithm.h:9-12 -> (A_q, N_q)."""
    P = len(radices)
    r = (ctypes.c_int * max(P, 1))(*radices)
    A = (ctypes.c_long * max(P, 1))()
    N = (ctypes.c_long * max(P, 1))()
    lib().oracle_prefix(P, r, A, N)
    return [A[i] for i in range(P)], [N[i] for i in range(P)]


def _as_c128(x):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    return x


def fft64(x, sign: int = FORWARD, log2radix: int = 4, inject: int = 0) -> np.ndarray:
    """FP64 Stockham FFT along the last axis (restates FFT_OpenCL_Contiguous)."""
    x = _as_c128(x)
    n = x.shape[-1]
    batch = x.size // n if n else 0
    out = np.empty_like(x)
    xin = x.view(np.float64)
    rc = lib().oracle_fft64_ex(n, batch, _ptr(xin, ctypes.c_double),
                               _ptr(out.view(np.float64), ctypes.c_double), sign, log2radix, inject)
    if rc:
        raise ValueError(f"oracle_fft64 failed ({rc}) for n={n}")
    return out


def fft32(x, sign: int = FORWARD, log2radix: int = 4) -> np.ndarray:
    """FP32 restatement of the same algorithm (correctly rounded FP32 twiddles)."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex64))
    n = x.shape[-1]
    batch = x.size // n
    out = np.empty_like(x)
    rc = lib().oracle_fft32_ex(n, batch, _ptr(x.view(np.float32), ctypes.c_float),
                               _ptr(out.view(np.float32), ctypes.c_float), sign, log2radix, 0)
    if rc:
        raise ValueError(f"oracle_fft32 failed ({rc}) for n={n}")
    return out


def pass_states(x, sign: int = FORWARD, log2radix: int = 4) -> np.ndarray:
    """State d_q after every global pass of one row: array [P, n]."""
    x = _as_c128(x).reshape(-1)
    n = x.size
    P = len(plan_radices(n, log2radix))
    st = np.empty((P, n), dtype=np.complex128)
    rc = lib().oracle_pass_states64(n, _ptr(x.view(np.float64), ctypes.c_double),
                                    _ptr(st.view(np.float64), ctypes.c_double), sign, log2radix)
    if rc != P:
        raise ValueError(f"oracle_pass_states64 failed ({rc})")
    return st


def dft(x, sign: int = FORWARD) -> np.ndarray:
    """Brute-force O(N^2) DFT in long double along the last axis."""
    x = _as_c128(x)
    n = x.shape[-1]
    batch = x.size // n
    out = np.empty_like(x)
    rc = lib().oracle_dft_ld(n, batch, _ptr(x.view(np.float64), ctypes.c_double),
                             _ptr(out.view(np.float64), ctypes.c_double), sign)
    if rc:
        raise ValueError("oracle_dft_ld failed")
    return out


def conv_circular(x, h) -> np.ndarray:
    """Direct circular convolution of real rows x[..., n] with real filter h[K]."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    h = np.ascontiguousarray(np.asarray(h, dtype=np.float64))
    n = x.shape[-1]
    batch = x.size // n
    y = np.empty_like(x)
    lib().oracle_conv_circular(n, batch, _ptr(x, ctypes.c_double), h.size,
                               _ptr(h, ctypes.c_double), _ptr(y, ctypes.c_double))
    return y


def planar_pos(index: int, element: int) -> int:
    """ArrayPlannarizer::getPosition (Complex.h:42-44), plannarLevel 5."""
    return lib().oracle_planar_pos(index, element)


def fill_uniform(count: int, seed: int, offset: int = 0) -> np.ndarray:
    """Seeded uniform[-1,1) float32 stream (splitmix64 over the flat index)."""
    out = np.empty(count, dtype=np.float32)
    lib().oracle_fill_uniform(_ptr(out, ctypes.c_float), count, seed, offset)
    return out


def uniform_complex(batch: int, n: int, seed: int) -> np.ndarray:
    """batch x n complex64 with re, im ~ uniform[-1,1) from `seed` (interleaved stream)."""
    return fill_uniform(2 * batch * n, seed).view(np.complex64).reshape(batch, n)


def rel_l2(a, b) -> float:
    """||a - b||_2 / ||b||_2 computed in FP64."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)
