"""Every compiled pass kernel against the FP64 oracle.

The reference compiles each generated kernel candidate and checks it against
a direct radix-r DFT before it may be used
(compiler/mofft_compiler/validate.py:78-226).  Here every registry entry -
each (precision, length, whole-transform or strided, radix chain, tile-shape
bank) instantiation - runs once inside a plan built to use exactly it (the
plan's description names the kernel):

* a whole-transform kernel as the single pass of N = L, on one full tile of
  transforms and one partial tile, forward out of place and backward in
  place;
* a strided kernel as the first pass of N = 16 L (no pass twiddles, forward
  out of place) and as the second pass of N = L x 16 (pass twiddles of span
  16, backward in place).

Bars: relative L2 error <= 1e-6 log2 N (FP32) / 1e-15 log2 N (FP64), the
matrix-core kernels (bank 3, FP32 split precision) included.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import seeded_complex

pytestmark = pytest.mark.gpu

WIDE_BANK = 4    # kernel bank of the four-times-wider strided tiles


def tol(n, bits):
    return (1e-6 if bits == 32 else 1e-15) * max(1.0, math.log2(n))


def reference(x, sign):
    n = x.shape[-1]
    if n <= 256 and x.size <= 1 << 14:
        return oracle.dft(x, sign)
    return oracle.fft_dd(x, sign)


def _set_bank(gpu, plan, bits, bank):
    st = getattr(gpu.lib(), f"mofft_plan_set_kernel_bank_f{bits}")(plan.handle, bank)
    assert st == gpu.SUCCESS


def _run(gpu, n, batch, bits, direction, in_place, x, radices, mask, bank, kernel, slot):
    dtype = np.complex64 if bits == 32 else np.complex128
    xin = np.ascontiguousarray(x.astype(dtype))
    with gpu.Plan(n, batch, bits, direction, in_place, radices=radices, pass_mask=mask) as p:
        if bank:
            _set_bank(gpu, p, bits, bank)
            # regroup under the bank: a chain whose whole-transform kernel
            # exists only there (the radix-32 matrix-core ones) is one pass
            p.set_blocked_transpose_mask(mask)
        d = p.description()
        assert d.pass_kernel[slot] == kernel, (n, radices, mask, bank, list(d.pass_kernel[:d.pass_count]))
        if in_place:
            y = xin.copy()
            p.execute(y, y)
        else:
            y = np.empty_like(xin)
            p.execute(xin, y)
    return y


def _cases(gpu, bits, rows):
    out = []
    for i, k in enumerate(gpu.kernels()):
        if k.precision_bits == bits and k.rows == rows:
            out.append((i, k.length, list(reversed(k.stages[:k.stage_count])),
                        k.columns_per_block, k.variant))
    return out


@pytest.mark.parametrize("bits", [32, 64])
def test_every_row_kernel(gpu, bits):
    cases = _cases(gpu, bits, 1)
    assert cases
    for i, L, chain, cols, bank in cases:
        batch = max(1, min(cols + 1, (1 << 17) // L))
        x = seeded_complex((batch, L), seed=20000 + i)
        xr = x.astype(np.complex64 if bits == 32 else np.complex128)
        for direction, sign, in_place in ((gpu.FORWARD, -1, False), (gpu.BACKWARD, 1, True)):
            y = _run(gpu, L, batch, bits, direction, in_place, x, chain, 0, bank, i, 0)
            err = oracle.rel_l2(y, reference(xr, sign))
            assert err <= tol(L, bits), ("rows", i, L, chain, bank, direction, err)


@pytest.mark.parametrize("bits", [32, 64])
def test_every_strided_kernel(gpu, bits):
    cases = _cases(gpu, bits, 0)
    assert cases
    for i, L, chain, cols, bank in cases:
        n = 16 * L
        batch = 2
        x = seeded_complex((batch, n), seed=30000 + i)
        xr = x.astype(np.complex64 if bits == 32 else np.complex128)
        # first pass (stride 16, no pass twiddles), forward out of place;
        # the wide bank (4) widens only the passes after the first, so its
        # kernels run in the second-pass case alone
        if bank != WIDE_BANK:
            y = _run(gpu, n, batch, bits, gpu.FORWARD, False, x, [16] + chain, 1, bank, i, 0)
            err = oracle.rel_l2(y, reference(xr, -1))
            assert err <= tol(n, bits), ("first pass", i, L, chain, bank, err)
        # second pass (span 16, pass twiddles), backward in place
        mask = 1 << (len(chain) - 1)
        y = _run(gpu, n, batch, bits, gpu.BACKWARD, True, x, chain + [16], mask, bank, i, 1)
        err = oracle.rel_l2(y, reference(xr, 1))
        assert err <= tol(n, bits), ("second pass", i, L, chain, bank, err)
