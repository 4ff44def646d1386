"""Pin the CPU oracle before trusting it (CPU only).

Every oracle function is checked against something it does not share code
with: committed mpmath golden vectors, numpy/pocketfft, closed-form spectra
(tests/test_binary_banks.c:17-36 tones, examples/consumer/main.c:169-176
impulse), invariants (linearity, Parseval, round trip), the reference's
known-answer byte counts (tests/test_fft.c:531-541), and the reference's
own error bounds for its executor (tests/test_fft.c:45, 91).  Mutation tests
plant plausible mistakes in an oracle copy and check the KATs catch them.
"""
import math
from pathlib import Path

import mpmath
import numpy as np
import pytest

import oracle
from conftest import seeded_complex

GOLDEN = np.load(Path(__file__).parent / "golden" / "dft_mpmath.npz")
GOLDEN_SIZES = sorted(int(k.split("_")[1]) for k in GOLDEN.files if k.startswith("x_"))


@pytest.mark.parametrize("n", GOLDEN_SIZES)
def test_direct_matches_mpmath_golden(n):
    x = GOLDEN[f"x_{n}"]
    for sign, key in ((-1, "fwd"), (1, "bwd")):
        want = GOLDEN[f"{key}_{n}"]
        got = oracle.dft_direct(x, sign)
        # double-double accumulation rounds once: equal to the correctly
        # rounded mpmath value up to the last bit
        assert np.max(np.abs(got - want)) <= 2 * np.finfo(float).eps * max(1, np.max(np.abs(want)))


@pytest.mark.parametrize("n", GOLDEN_SIZES)
def test_fft_dd_matches_mpmath_golden(n):
    x = GOLDEN[f"x_{n}"]
    got = oracle.fft_dd(x, -1)
    want = GOLDEN[f"fwd_{n}"]
    assert np.max(np.abs(got - want)) <= 2 * np.finfo(float).eps * max(1, np.max(np.abs(want)))


@pytest.mark.parametrize("n", [6, 17, 59, 60, 128, 243, 1000, 1024, 2187, 4096])
def test_direct_and_compensated_fft_agree(n):
    x = seeded_complex((2, n), seed=n)
    for sign in (-1, 1):
        a = oracle.dft_direct(x, sign)
        b = oracle.fft_dd(x, sign)
        # both round a value accurate to ~1e-19 * |X|max once: components
        # differ by at most one ulp of the largest coefficient
        big = np.max(np.abs(np.concatenate([a.real.ravel(), a.imag.ravel()])))
        assert np.max(np.abs(a - b)) <= 2 * np.spacing(big)
        assert oracle.rel_l2(a, b) < 2e-16


@pytest.mark.parametrize("n", [16, 1000, 4096, 1 << 15, 3 ** 9, 1 << 18])
def test_oracle_vs_pocketfft(n):
    x = seeded_complex((1, n), seed=7 * n)
    got = oracle.dft(x, -1)
    # numpy's own FP64 error is ~1e-16 * sqrt(log N)
    assert oracle.rel_l2(got, np.fft.fft(x)) < 1e-15
    assert oracle.rel_l2(oracle.dft(x, 1), np.fft.ifft(x) * n) < 1e-15


def tones(n, sign):
    j = np.arange(n)
    x = 0.25 + np.exp(2j * np.pi * 3 * j / n) + 0.5 * np.exp(-2j * np.pi * 5 * j / n)
    want = np.zeros(n, dtype=np.complex128)
    want[0] += 0.25
    want[(3 % n) if sign < 0 else (n - 3 % n) % n] += 1.0
    want[(n - 5 % n) % n if sign < 0 else 5 % n] += 0.5
    return x, want


@pytest.mark.parametrize("n", [121, 169, 196, 225, 256, 1024, 1728, 2197, 20736, 65536])
@pytest.mark.parametrize("sign", [-1, 1])
def test_closed_form_tones(n, sign):
    x, want = tones(n, sign)
    got = oracle.dft(x, sign) / n
    # the tone inputs carry numpy's exp() rounding (~1e-16 per sample); the
    # reference's own bound here is 2e-11 (tests/test_binary_banks.c:49-50)
    assert np.max(np.abs(got - want)) < 1e-14


@pytest.mark.parametrize("n", [8, 1000, 4096, 65536])
def test_impulse(n):
    rng = np.random.default_rng(n)
    n0 = int(rng.integers(0, n))
    x = np.zeros(n, dtype=np.complex128)
    x[n0] = 1.0
    got = oracle.dft(x, -1)
    # X[k] = exp(-2*pi*i*k*n0/N), checked at 64 bins against mpmath
    for k in rng.integers(0, n, 64):
        with mpmath.workdps(40):
            ang = -2 * mpmath.pi * ((int(k) * n0) % n) / n
            want = complex(float(mpmath.cos(ang)), float(mpmath.sin(ang)))
        assert abs(got[k] - want) <= 2.3e-16


def test_linearity_parseval_roundtrip():
    n = 1 << 14
    x = seeded_complex((1, n), seed=1)
    y = seeded_complex((1, n), seed=2)
    a, b = 0.75 - 0.5j, -1.25 + 2j
    lhs = oracle.dft(a * x + b * y, -1)
    rhs = a * oracle.dft(x, -1) + b * oracle.dft(y, -1)
    assert oracle.rel_l2(lhs, rhs) < 2e-16
    X = oracle.dft(x, -1)
    assert abs(np.vdot(X, X).real / (n * np.vdot(x, x).real) - 1) < 1e-14
    assert oracle.rel_l2(oracle.dft(X, 1) / n, x) < 2e-16


def test_omega_correctly_rounded():
    cases = [(1, 3), (1, 7), (5, 12), (7, 16), (3, 1024), (1000, 4096), (12345, 65536),
             (99, 1 << 20), (123457, 1 << 24), (7, 13 ** 5), (31, 3375)]
    for m, big_m in cases:
        with mpmath.workdps(50):
            ang = -2 * mpmath.pi * m / big_m
            c, s = mpmath.cos(ang), mpmath.sin(ang)
            with mpmath.workprec(24):
                c32, s32 = float(+c), float(+s)
            c64, s64 = float(c), float(s)
        w32 = oracle.omega(m, big_m, -1, 32)
        w64 = oracle.omega(m, big_m, -1, 64)
        assert (w32.real, w32.imag) == (c32, s32), (m, big_m)
        assert (w64.real, w64.imag) == (c64, s64), (m, big_m)


# ---------------------------------------------------------------------
# reference executor restatement (src/mofft.c:1932-2045)


@pytest.mark.parametrize("chain", [[16], [32, 8], [8, 32], [16, 16, 16], [13, 16, 2],
                                   [4, 4, 4, 4, 4, 4], [15, 15], [64, 2], [5, 3, 2]])
def test_reference_executor_restatement(chain):
    n = int(np.prod(chain))
    x = seeded_complex((n,), seed=n)
    want = oracle.dft(x, -1)
    y64 = oracle.ref_execute(x, chain, -1, 64)
    y32 = oracle.ref_execute(x.astype(np.complex64), chain, -1, 32)
    # the reference's own bounds: tests/test_fft.c:45 and :91
    assert np.max(np.abs(y64 - want)) <= 96 * np.finfo(np.float64).eps * n
    want32 = oracle.dft(x.astype(np.complex64), -1)
    assert np.max(np.abs(y32 - want32)) <= 32 * np.finfo(np.float32).eps * n
    back = oracle.ref_execute(y64, chain, 1, 64) / n
    assert oracle.rel_l2(back, x) < 512 * np.finfo(np.float64).eps * math.log2(n)


def test_reference_plan_byte_kats():
    # tests/test_fft.c:531-541 (N=4096, FP64, {8,8,8,8})
    sec = oracle.plan_bytes(4096, 64, [8, 8, 8, 8], 0, 0, True)
    lin = oracle.plan_bytes(4096, 64, [8, 8, 8, 8], 0, 0, False)
    blk = oracle.plan_bytes(4096, 64, [8, 8, 8, 8], 7, 0, False)
    assert sec == (336896, 262144, 336896, 262144, 196608)
    assert lin[:4] == (533504, 458752, 992256, 458752)
    assert blk[2] == 533504
    assert oracle.plan_bytes(4096, 64, [8, 8, 8], 0, 1) is None


# ---------------------------------------------------------------------
# mutation checks: a plausible oracle mistake must fail the KATs


def _kats_pass(dft_fn):
    for n in (12, 64, 225):
        for sign in (-1, 1):
            x, want = tones(n, sign)
            if np.max(np.abs(dft_fn(x, sign) / n - want)) > 1e-12:
                return False
    for n in GOLDEN_SIZES:
        if n < 3:
            continue
        if np.max(np.abs(dft_fn(GOLDEN[f"x_{n}"], -1) - GOLDEN[f"fwd_{n}"])) > 1e-12:
            return False
    return True


def test_kats_accept_the_oracle():
    assert _kats_pass(oracle.dft)


@pytest.mark.parametrize("mutation", ["sign", "conj_input", "index_shift", "transpose",
                                      "dropped_term", "normalized"])
def test_kats_reject_planted_mistakes(mutation):
    def bad(x, sign):
        x = np.asarray(x, dtype=np.complex128)
        if mutation == "sign":
            return oracle.dft(x, -sign)
        if mutation == "conj_input":
            return oracle.dft(np.conj(x), sign)
        if mutation == "index_shift":
            return np.roll(oracle.dft(x, sign), 1)
        if mutation == "transpose":
            n = x.shape[-1]
            p = next((d for d in range(2, n) if n % d == 0), None)
            y = oracle.dft(x, sign)
            return y.reshape(p, n // p).T.reshape(-1) if p else y[::-1]
        if mutation == "dropped_term":
            x = x.copy()
            x[..., -1] = 0
            return oracle.dft(x, sign)
        if mutation == "normalized":
            return oracle.dft(x, sign) / x.shape[-1]
        raise AssertionError(mutation)
    assert not _kats_pass(bad)
