"""GPU parity of the device 2-D transform (estl_plan2d_*) and of the profiling surface.

The 2-D reference is the FP64 oracle applied along x then along y (a 2-D DFT is
separable; the 1-D oracle is pinned by its own tests), and a brute-force 2-D DFT for
tiny sizes.  Tolerance: relative L2 <= 1e-6 * log2(nx * ny) (two FP32 transforms).
"""
import math

import numpy as np
import pytest

import estl_amd
import oracle

pytestmark = pytest.mark.gpu


def ref2d(x, sign=-1):
    """x: (batch, ny, nx) complex128 -> 2-D DFT along the last two axes via the oracle."""
    b, ny, nx = x.shape
    y = oracle.fft64(x.reshape(b * ny, nx), sign).reshape(b, ny, nx)
    yt = np.ascontiguousarray(y.transpose(0, 2, 1)).reshape(b * nx, ny)
    return oracle.fft64(yt, sign).reshape(b, nx, ny).transpose(0, 2, 1)


def tol(nx, ny):
    return 1e-6 * max(1, int(math.log2(nx * ny)))


SIZES = [(1, 16), (16, 1), (2, 2), (4, 8), (8, 2), (64, 64), (128, 32), (32, 4096), (4096, 8), (8, 8192),
         (1024, 1024), (16384, 4), (4, 16384), (256, 32768), (1, 1)]


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("batch", [1, 3])
def test_plan2d_matches_oracle(gpu, nx, ny, batch):
    import torch
    if nx * ny * batch > (1 << 23):
        batch = 1
    x = torch.empty(batch, ny, nx, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 50 + nx + ny)
    y = torch.empty_like(x)
    p = estl_amd.Plan2D(nx, ny, batch)
    p.execute(x, y)
    torch.cuda.synchronize()
    ref = ref2d(x.cpu().numpy().astype(np.complex128))
    got = y.cpu().numpy()
    for b in range(batch):
        e = oracle.rel_l2(got[b], ref[b])
        assert e <= tol(nx, ny), (nx, ny, b, e)
    if ny > 1 and ny <= 8192 and nx > 1 and nx <= 8192:
        # one row launch; one column pass up to ny = 2^9 (or nx < 4), two above
        assert p.launches() == (2 if ny <= 512 or nx < 4 else 3), p.launches()


def test_plan2d_random_shapes(gpu):
    """Twenty seeded 2-D shapes (nx, ny = 2^U{0..14}, at most 2^22 values, 1..3 arrays),
    either direction, normalised or not, in place or not, against the oracle's
    separable FP64 transform."""
    import torch
    rng = np.random.default_rng(77)
    for case in range(20):
        while True:
            lx, ly = int(rng.integers(0, 15)), int(rng.integers(0, 15))
            if lx + ly <= 22:
                break
        nx, ny = 1 << lx, 1 << ly
        batch = int(rng.integers(1, 4)) if lx + ly <= 20 else 1
        direction = estl_amd.FORWARD if rng.random() < 0.5 else estl_amd.BACKWARD
        norm = bool(rng.random() < 0.3)
        in_place = bool(rng.random() < 0.3)
        x = torch.empty(batch, ny, nx, dtype=torch.complex64, device=gpu)
        estl_amd.fill_uniform(x, 300 + case)
        xh = x.cpu().numpy().astype(np.complex128)
        p = estl_amd.Plan2D(nx, ny, batch, direction, estl_amd.FLAG_NORMALIZE if norm else 0)
        y = x if in_place else torch.empty_like(x)
        p.execute(x, y)
        torch.cuda.synchronize()
        ref = ref2d(xh, direction)
        if norm:
            ref = ref / (nx * ny)
        got = y.cpu().numpy()
        for b in range(batch):
            e = oracle.rel_l2(got[b], ref[b])
            assert e <= tol(nx, ny), (case, nx, ny, batch, direction, norm, in_place, e)


def test_plan2d_bruteforce_tiny(gpu):
    import torch
    nx, ny = 8, 4
    x = torch.empty(ny, nx, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 77)
    y = torch.empty_like(x)
    estl_amd.Plan2D(nx, ny).execute(x, y)
    torch.cuda.synchronize()
    xs = x.cpu().numpy().astype(np.complex128)
    ky, kx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ref = np.zeros((ny, nx), np.complex128)
    for v in range(ny):
        for u in range(nx):
            ref += xs[v, u] * np.exp(-2j * np.pi * (kx * u / nx + ky * v / ny))
    assert oracle.rel_l2(y.cpu().numpy(), ref) <= tol(nx, ny)


@pytest.mark.parametrize("nx,ny", [(64, 64), (2048, 16), (16, 16384)])
def test_plan2d_in_place_round_trip(gpu, nx, ny):
    import torch
    x = torch.empty(2, ny, nx, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 5)
    y = x.clone()
    estl_amd.Plan2D(nx, ny, 2).execute(y, y)
    estl_amd.Plan2D(nx, ny, 2, estl_amd.BACKWARD, estl_amd.FLAG_NORMALIZE).execute(y, y)
    torch.cuda.synchronize()
    rt = (torch.linalg.vector_norm((y - x).to(torch.complex128)) /
          torch.linalg.vector_norm(x.to(torch.complex128))).item()
    assert rt <= 2 * tol(nx, ny), rt


def test_plan2d_host_entry_and_rejects(gpu):
    nx, ny = 32, 16
    x = (oracle.fill_uniform(2 * nx * ny, 3).astype(np.float32)).view(np.complex64).reshape(ny, nx)
    y = np.empty_like(x)
    estl_amd.Plan2D(nx, ny).execute_host(x, y)
    assert oracle.rel_l2(y, ref2d(x[None].astype(np.complex128))[0]) <= tol(nx, ny)
    for bad in [(3, 4), (4, 0), (0, 4)]:
        with pytest.raises(estl_amd.EstlError):
            estl_amd.Plan2D(*bad)


@pytest.mark.parametrize("n,batch", [(4096, 4096), (1 << 20, 4)])
def test_profile_surface(gpu, n, batch):
    import torch
    p = estl_amd.Plan(n, batch)
    rep = p.profile(5)
    ks = rep["kernels"]
    assert len(ks) == p.launches() >= 1
    assert all(k["ms"] > 0 and k["gflops"] > 0 for k in ks)
    # flops per launch: 5 n log2(radix) per transform; they add up to 5 n log2 n
    assert sum(k["flops"] for k in ks) == pytest.approx(5.0 * n * math.log2(n) * batch)
    assert abs(sum(k["share"] for k in ks) - 1.0) < 1e-6
    assert ("k_fft_rows" in ks[0]["info"]) == (n <= 8192)
    # profiling on caller buffers leaves a correct transform behind
    x = torch.empty(batch, n, dtype=torch.complex64, device=gpu)
    estl_amd.fill_uniform(x, 1)
    y = torch.empty_like(x)
    p.profile(2, x, y)
    torch.cuda.synchronize()
    e = oracle.rel_l2(y[0].cpu().numpy(), oracle.fft64(x[0].cpu().numpy().astype(np.complex128), -1))
    assert e <= 1e-6 * math.log2(n)
