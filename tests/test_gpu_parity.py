"""GPU parity: libmofft.so on gfx950 against the FP64 oracle.

Bars: plain FP32 relative L2 error <= 1e-6*log2(N), FP64
<= 1e-15*log2(N), and forward->backward round trips back to N*x within the
same bounds.  The oracle is FP64 direct summation (N <= 1024) or the
compensated double-double FFT (larger N, and large batches), both pinned
against each other and against closed forms in tests/test_oracle.py.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from conftest import seeded_complex, mixed_kernel_plans, random_smooth_lengths

pytestmark = pytest.mark.gpu


def tol(n, precision_bits):
    return (1e-6 if precision_bits == 32 else 1e-15) * max(1.0, math.log2(n))


def reference(x, sign):
    n = x.shape[-1]
    if n <= 1024 and x.size <= 4 * 1024:
        return oracle.dft(x, sign)
    return oracle.fft_dd(x, sign)


def run_host(mf, n, batch, bits, direction, in_place, x, radices=None, mask=None):
    dtype = np.complex64 if bits == 32 else np.complex128
    xin = np.ascontiguousarray(x.astype(dtype))
    plan = mf.Plan(n, batch, bits, direction, in_place, radices=radices,
                   pass_mask=mask)
    try:
        if in_place:
            buf = xin.copy()
            plan.execute(buf, buf)
            return buf, plan.description()
        out = np.empty_like(xin)
        plan.execute(xin, out)
        # out-of-place must leave the input untouched
        assert np.array_equal(xin, x.astype(dtype))
        return out, plan.description()
    finally:
        plan.destroy()


SINGLE = [2, 3, 4, 5, 7, 8, 11, 13, 16, 32, 45, 64, 128, 256, 512, 1024, 2048, 4096]
MULTI = [8192, 16384, 65536, 1 << 18, 1 << 20]
MIXED = [30, 121, 144, 169, 196, 225, 1728, 2197, 2744, 3375, 20736]


def batch_for(n):
    return max(1, min(64, (1 << 16) // n))


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n", SINGLE + MIXED + MULTI)
def test_forward_out_of_place(gpu, n, bits):
    batch = batch_for(n)
    x = seeded_complex((batch, n), seed=1000 + n)
    y, desc = run_host(gpu, n, batch, bits, gpu.FORWARD, False, x)
    ref = reference(x.astype(np.complex64 if bits == 32 else np.complex128), -1)
    err = oracle.rel_l2(y, ref)
    print(f"N={n} batch={batch} f{bits} passes={desc.passes} err={err:.3e}")
    assert err <= tol(n, bits), (n, bits, err, desc.passes)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n", [16, 1024, 4096, 16384, 65536, 1728, 20736])
@pytest.mark.parametrize("in_place", [False, True])
def test_backward(gpu, n, bits, in_place):
    batch = batch_for(n)
    x = seeded_complex((batch, n), seed=2000 + n, dist="normal")
    y, desc = run_host(gpu, n, batch, bits, gpu.BACKWARD, in_place, x)
    ref = reference(x.astype(np.complex64 if bits == 32 else np.complex128), 1)
    err = oracle.rel_l2(y, ref)
    assert err <= tol(n, bits), (n, bits, in_place, err, desc.passes)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n", [64, 1024, 8192, 65536, 1 << 18])
def test_in_place_forward(gpu, n, bits):
    batch = batch_for(n)
    x = seeded_complex((batch, n), seed=3000 + n)
    y, desc = run_host(gpu, n, batch, bits, gpu.FORWARD, True, x)
    ref = reference(x.astype(np.complex64 if bits == 32 else np.complex128), -1)
    err = oracle.rel_l2(y, ref)
    assert err <= tol(n, bits), (n, bits, err, desc.passes)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n", [256, 4096, 1 << 16, 1 << 20])
def test_roundtrip(gpu, n, bits):
    batch = max(1, min(8, (1 << 18) // n))
    dtype = np.complex64 if bits == 32 else np.complex128
    x = seeded_complex((batch, n), seed=4000 + n).astype(dtype)
    y, _ = run_host(gpu, n, batch, bits, gpu.FORWARD, False, x)
    z, _ = run_host(gpu, n, batch, bits, gpu.BACKWARD, False, y)
    err = oracle.rel_l2(z / n, x)
    assert err <= tol(n, bits), (n, bits, err)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("chain", [[32, 8], [8, 32], [16, 16, 16], [64, 2],
                                   [13, 16, 2], [4, 4, 4, 4, 4, 4], [2] * 12])
def test_explicit_chains(gpu, chain, bits):
    n = int(np.prod(chain))
    batch = 3
    x = seeded_complex((batch, n), seed=5000 + n)
    y, desc = run_host(gpu, n, batch, bits, gpu.FORWARD, False, x, radices=chain)
    assert desc.chain == chain
    ref = reference(x.astype(np.complex64 if bits == 32 else np.complex128), -1)
    err = oracle.rel_l2(y, ref)
    assert err <= tol(n, bits), (chain, bits, err, desc.passes)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n", [2, 3, 5, 8, 12, 16])
@pytest.mark.parametrize("batch", [1, 127, 128, 129, 300])
def test_small_transform_tiles(gpu, n, bits, batch):
    """Lengths whose whole transform is one thread's work: the tile of 128
    transforms is read and written coalesced and handed over through LDS,
    so full, partial and single-transform tiles, both directions and both
    placements all have to agree with the oracle."""
    x = seeded_complex((batch, n), seed=7000 + 13 * n + batch)
    xr = x.astype(np.complex64 if bits == 32 else np.complex128)
    for direction, sign in ((gpu.FORWARD, -1), (gpu.BACKWARD, 1)):
        ref = oracle.dft(xr, sign)
        for in_place in (False, True):
            y, desc = run_host(gpu, n, batch, bits, direction, in_place, x)
            err = oracle.rel_l2(y, ref)
            assert err <= tol(n, bits), (n, batch, bits, direction, in_place, err)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("rows", [1, 0])
def test_mixed_kernels(gpu, bits, rows):
    """Each kernel of a length that is not a power of two (every smooth
    length up to 1024 points, the reference's mixed lengths, the powers of 3,
    5, 6, 7 up to 3125) against the oracle, forward out of place and
    backward in place: whole-transform kernels on N = L, strided kernels as
    the first pass of N = 16 L."""
    plans = mixed_kernel_plans(gpu, bits, rows)
    assert plans
    for i, n, radices, mask in plans:
        batch = max(1, min(8, 8192 // n))
        x = seeded_complex((batch, n), seed=9100 + n + bits)
        xr = x.astype(np.complex64 if bits == 32 else np.complex128)
        for direction, sign, in_place in ((gpu.FORWARD, -1, False),
                                          (gpu.BACKWARD, 1, True)):
            y, desc = run_host(gpu, n, batch, bits, direction, in_place, x,
                               radices=radices, mask=mask)
            assert desc.pass_kernel[0] == i
            err = oracle.rel_l2(y, reference(xr, sign))
            assert err <= tol(n, bits), (n, radices, mask, direction, err)


# plans whose strided tiles are not whole output groups (C columns per tile
# not a multiple of the previous passes' product ns): 2-, 3- and 4-pass model
# plans of smooth lengths, the tuned 12^5 plan, and both pass orders of 16 x 36
GROUP_CASES = [(18, None, None), (20, None, None), (24, None, None),
               (35, None, None), (280, None, None), (315, None, None),
               (420, None, None), (33600, None, None), (67200, None, None),
               (134400, None, None), (198450, None, None), (248832, None, None),
               (576, [16, 6, 6], 1), (576, [6, 6, 16], 2), (1296, [6, 6, 6, 6], 2)]


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n,radices,mask", GROUP_CASES)
def test_partial_output_groups(gpu, n, radices, mask, bits):
    batch = max(2, min(8, (1 << 16) // n))
    x = seeded_complex((batch, n), seed=9300 + n)
    xr = x.astype(np.complex64 if bits == 32 else np.complex128)
    for direction, sign, in_place in ((gpu.FORWARD, -1, False),
                                      (gpu.BACKWARD, 1, True)):
        y, desc = run_host(gpu, n, batch, bits, direction, in_place, x,
                           radices=radices, mask=mask)
        err = oracle.rel_l2(y, reference(xr, sign))
        assert err <= tol(n, bits), (n, desc.passes, direction, err)


def _wisdom_cases():
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "mofft_amd", "wisdom", "gfx950.json")
    with open(path) as f:
        entries = json.load(f)["entries"]
    return sorted({(e["length"], 32 if e["precision"] == "fp32" else 64)
                   for e in entries})


@pytest.mark.parametrize("n,bits", _wisdom_cases())
def test_every_wisdom_plan(gpu, n, bits):
    """Each tuned PRETUNED plan (the tuner times plans, it does not check
    them) against the oracle, one transform forward."""
    x = seeded_complex((1, n), seed=9500 + n)
    xr = x.astype(np.complex64 if bits == 32 else np.complex128)
    plan = gpu.Plan(n, 1, bits, gpu.FORWARD)
    try:
        assert plan.diagnostics.used_wisdom == 1
        y = np.empty_like(xr)
        plan.execute(xr, y)
    finally:
        plan.destroy()
    err = oracle.rel_l2(y, reference(xr, -1))
    assert err <= tol(n, bits), (n, bits, err)


@pytest.mark.parametrize("bits", [32, 64])
def test_random_smooth_lengths(gpu, bits):
    """40 seeded 13-smooth lengths up to 2^20 on their PRETUNED plans (one to
    three passes of any mix of radices) against the oracle, forward out of
    place and backward in place, batches of 1 to 3."""
    for n in random_smooth_lengths(500 + bits, 40, 1 << 20):
        batch = 1 + n % 3
        x = seeded_complex((batch, n), seed=9700 + n)
        xr = x.astype(np.complex64 if bits == 32 else np.complex128)
        for direction, sign, in_place in ((gpu.FORWARD, -1, False),
                                          (gpu.BACKWARD, 1, True)):
            y, desc = run_host(gpu, n, batch, bits, direction, in_place, x)
            err = oracle.rel_l2(y, reference(xr, sign))
            assert err <= tol(n, bits), (n, desc.passes, direction, err)
