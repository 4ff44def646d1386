"""Shared fixtures for the MoFFT test suite.

`gpu`-marked tests need an MI355X (gfx950) and the built libmofft.so; the
rest run on any host (oracle, generator, planner, recipe parser, ABI).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
COMPILER = os.path.join(ROOT, "mofft_amd", "compiler")
if COMPILER not in sys.path:
    sys.path.insert(0, COMPILER)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and libmofft.so")


def seeded_complex(shape, seed, dist="uniform", dtype=np.complex128):
    """iid components from a seed: uniform[-1,1) or N(0,1)."""
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        re = rng.uniform(-1.0, 1.0, shape)
        im = rng.uniform(-1.0, 1.0, shape)
    else:
        re = rng.standard_normal(shape)
        im = rng.standard_normal(shape)
    return (re + 1j * im).astype(dtype)


def _torch_first():
    """PyTorch-ROCm ships its own HIP runtime.  When a process uses both, torch
    must load first so libmofft binds to the already-loaded runtime (same
    soname); loading libmofft first leaves torch without a device."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


@pytest.fixture(scope="session")
def mofft():
    _torch_first()
    import mofft_amd
    mofft_amd.lib()
    return mofft_amd


@pytest.fixture(scope="session")
def gpu(mofft):
    if not mofft.device_available():
        pytest.fail("gpu test ran without a usable gfx950 device")
    return mofft


_SM_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def seeded_device_uniform(seed, count, bits, start=0):
    """Host restatement of the library's device input generator
    (mofft_hip_fill_uniform / mofft_server_generate): component c of complex
    element i is 2u - 1, u = top 53 bits of
    splitmix64(seed + (2i + c + 1) * 0x9E3779B97F4A7C15) * 2^-53.  Returns
    elements start .. start + count - 1."""
    with np.errstate(over="ignore"):
        k = np.arange(2 * start + 1, 2 * (start + count) + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _SM_GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    v = 2.0 * u - 1.0
    if bits == 32:
        v = v.astype(np.float32)
        return v.view(np.complex64)
    return v.view(np.complex128)


# powers of the small radices with kernels of their own beyond 1024 points
SMALL_POWERS = [1296, 2187, 2401, 3125]


def mixed_kernel_plans(mofft, bits, rows):
    """(kernel index, N, outer-to-inner radices, pass mask) for every
    compiled kernel of a length that is not a power of two: a whole-transform
    kernel (rows = 1) runs as the single pass of N = L, a strided kernel as
    the first pass of N = 16 L (radix 16 outermost, the boundary right after
    it)."""
    out = []
    for i, k in enumerate(mofft.kernels()):
        if (k.precision_bits != bits or k.rows != rows
                or k.length & (k.length - 1) == 0):
            continue
        inner = list(reversed(k.stages[: k.stage_count]))
        if k.rows:
            out.append((i, k.length, inner, 0))
        else:
            out.append((i, 16 * k.length, [16] + inner, 1))
    return out


def random_smooth_lengths(seed, count, limit):
    """`count` distinct lengths in [2, limit] whose prime factors are radices
    (2, 3, 5, 7, 11, 13), drawn log-uniformly from a seed, sorted."""
    rng = np.random.default_rng(seed)
    out = set()
    while len(out) < count:
        n = 1
        target = np.exp(rng.uniform(np.log(2), np.log(limit)))
        while n * 2 <= target:
            q = int(rng.choice([2, 2, 2, 3, 3, 5, 5, 7, 11, 13]))
            if n * q > limit:
                break
            n *= q
        if n >= 2:
            out.add(n)
    return sorted(out)
