"""Write the golden DFT fixtures used by tests/test_oracle.py.

The fixtures are computed with mpmath at 60 significant digits (an
independent arbitrary-precision evaluation of the DFT definition of
tests/test_fft.c:20-26 / include/mofft.h:26), then rounded to double.
Inputs are seeded uniform[-1,1) components rounded to float32, so the same
inputs serve FP32 and FP64 checks.  No reference code is run.

Usage: python tests/golden/make_golden.py   (rewrites dft_mpmath.npz)
"""
from pathlib import Path

import mpmath
import numpy as np

SIZES = (1, 2, 3, 4, 5, 7, 8, 12, 13, 16, 30, 45, 64)
OUT = Path(__file__).with_name("dft_mpmath.npz")


def mp_dft(x, sign):
    n = len(x)
    with mpmath.workdps(60):
        xs = [mpmath.mpc(float(v.real), float(v.imag)) for v in x]
        out = []
        for k in range(n):
            acc = mpmath.mpc(0)
            for j in range(n):
                acc += xs[j] * mpmath.expjpi(sign * mpmath.mpf(2 * ((j * k) % n)) / n)
            out.append(complex(float(acc.real), float(acc.imag)))
    return np.array(out, dtype=np.complex128)


def main():
    rng = np.random.default_rng(20261203)
    arrays = {}
    for n in SIZES:
        x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
        x = x.astype(np.complex128)
        arrays[f"x_{n}"] = x
        arrays[f"fwd_{n}"] = mp_dft(x, -1)
        arrays[f"bwd_{n}"] = mp_dft(x, +1)
    np.savez(OUT, **arrays)
    print(f"wrote {OUT} ({len(SIZES)} sizes)")


if __name__ == "__main__":
    main()
