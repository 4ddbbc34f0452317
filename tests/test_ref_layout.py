"""Checks the oracle's restatements against the reference's own host headers.

oracle/_ref/ref_layout is compiled by oracle/build.py from Complex.h, DataInterface.h
and SimpleMath.h of the reference tree (a driver of our own around them); it prints
what those headers compute.  Skipped when the checker has not been built.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_layout")


@pytest.fixture(scope="module")
def ref_lines():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref/ref_layout not built (reference tree absent)")
    out = subprocess.run([REF], capture_output=True, text=True, check=True, timeout=60).stdout
    return [l.split() for l in out.splitlines() if l.strip()]


def test_ceilog_and_plan(ref_lines):
    for l in ref_lines:
        if l[0] == "ceilog":
            n, r = int(l[1]), int(l[2])
            if n >= 2:
                # the plan's passes consume exactly ceilog(n, 2) bits
                rad = oracle.plan_radices(n)
                assert int(np.log2(np.prod(rad))) == r


def test_planarized_layout(ref_lines):
    for l in ref_lines:
        if l[0] in ("planar", "pdi"):
            size = int(l[1])
            raw = [float(v) for v in l[2:]]
            for i in range(size):
                re_pos, im_pos = oracle.planar_pos(i, 0), oracle.planar_pos(i, 1)
                if l[0] == "planar":
                    assert raw[re_pos] == i and raw[im_pos] == -i - 0.5
                else:
                    assert raw[re_pos] == 2.0 * i and raw[im_pos] == 0.25 * i


def test_interleaved_and_split_layout(ref_lines):
    for l in ref_lines:
        if l[0] == "interleaved":
            size = int(l[1])
            raw = np.array([float(v) for v in l[2:]])
            assert np.array_equal(raw[0::2], np.arange(size)) and np.array_equal(raw[1::2], -np.arange(size) - 0.5)
        if l[0] == "split":
            size = int(l[1])
            raw = np.array([float(v) for v in l[2:]])
            assert np.array_equal(raw[:size], np.arange(size)) and np.array_equal(raw[size:], -np.arange(size) - 0.5)


def test_complex_multiply(ref_lines):
    # Complex<float> operator* (Complex.h:91-93) against numpy complex64 products
    n = 0
    for l in ref_lines:
        if l[0] == "cmul":
            v = [float.fromhex(t) for t in l[1:]]
            a = complex(np.float32(v[0]), np.float32(v[1]))
            b = complex(np.float32(v[2]), np.float32(v[3]))
            exact = a * b
            got = complex(v[4], v[5])
            assert abs(got - exact) <= 4 * np.finfo(np.float32).eps * (abs(a) * abs(b) + 1e-30)
            assert v[6] == -v[1]
            n += 1
    assert n == 64
