"""GPU tests of the matrix-core (MFMA) split-precision form: FP32 plans whose
radix-16 sub-stages run as FP16 hi/lo products accumulated in FP32
(MOFFT_PLAN_MFMA, kernel bank 3).

The bar is the plain FP32 path's: relative L2 error against the FP64 oracle
<= 1e-6*log2N for every N = 2^4..2^20 (uniform[-1,1) inputs and a unit
impulse at a seeded index), and the round trip within the same bound.  The
measured errors of both forms are printed per N (pytest -s).
"""
import math

import numpy as np
import pytest

import oracle
from conftest import seeded_complex

pytestmark = pytest.mark.gpu


def tol(n):
    return 1e-6 * max(1.0, math.log2(n))


@pytest.mark.parametrize("k", list(range(4, 21)))
def test_mfma_error_per_length(gpu, k):
    n = 1 << k
    batch = max(1, (1 << 15) // n)
    x = seeded_complex((batch, n), seed=300 + k, dtype=np.complex64)
    ref = oracle.dft(x.astype(np.complex128), -1)
    errs = {}
    for form in ("plain", "mfma"):
        y = np.empty_like(x)
        z = np.empty_like(x)
        with gpu.Plan(n, batch, 32, gpu.FORWARD, mfma=form == "mfma") as f, \
                gpu.Plan(n, batch, 32, gpu.BACKWARD, mfma=form == "mfma") as b:
            if form == "mfma":
                assert f.diagnostics.kernel_bank == gpu.MFMA_BANK
            f.execute(x, y)
            b.execute(y, z)
        errs[form] = (oracle.rel_l2(y, ref), oracle.rel_l2(z / n, x))
    # unit impulse at a seeded index: X[q] = w^(q*n0)
    n0 = int(np.random.default_rng(k).integers(0, n))
    imp = np.zeros((1, n), np.complex64)
    imp[0, n0] = 1
    yi = np.empty_like(imp)
    with gpu.Plan(n, 1, 32, gpu.FORWARD, mfma=True) as f:
        f.execute(imp, yi)
    want = oracle.omega_table(n, -1, 64, indices=(np.arange(n) * n0) % n)
    e_imp = oracle.rel_l2(yi[0], want)
    print(f"N=2^{k}: forward rel-L2 plain {errs['plain'][0]:.2e} mfma {errs['mfma'][0]:.2e}; "
          f"round trip plain {errs['plain'][1]:.2e} mfma {errs['mfma'][1]:.2e}; "
          f"impulse mfma {e_imp:.2e}; bar {tol(n):.1e}")
    for form in ("plain", "mfma"):
        assert errs[form][0] <= tol(n), (form, errs[form])
        assert errs[form][1] <= tol(n), (form, errs[form])
    assert e_imp <= tol(n)


def test_mfma_flag_rules(gpu):
    # FP64 has no matrix-core form; MEASURE/STRICT cannot be combined with it
    with pytest.raises(gpu.MofftError) as e:
        gpu.Plan(1024, 1, 32, mode=gpu.MEASURE, flags=0, mfma=True)
    assert e.value.status == gpu.INVALID_ARGUMENT
    with pytest.raises(gpu.MofftError):
        gpu.Plan(1024, 1, 32, flags=gpu.STRICT_WISDOM, mfma=True)


def test_mfma_bank_on_reference_setter(gpu):
    """The reference's kernel-bank setter selects the same form."""
    import ctypes
    n = 4096
    x = seeded_complex((8, n), seed=1, dtype=np.complex64)
    with gpu.Plan(n, 8, 32, radices=[16, 16, 16]) as p:
        assert gpu.lib().mofft_plan_set_kernel_bank_f32(p.handle, gpu.MFMA_BANK) == 0
        got = ctypes.c_uint()
        gpu.lib().mofft_plan_get_kernel_bank_f32(p.handle, ctypes.byref(got))
        assert got.value == gpu.MFMA_BANK
        y = np.empty_like(x)
        p.execute(x, y)
    with gpu.Plan(n, 8, 32, mfma=True) as q:
        assert q.description().chain == [16, 16, 16]
        y2 = np.empty_like(x)
        q.execute(x, y2)
    assert np.array_equal(y, y2)
    assert oracle.rel_l2(y, oracle.dft(x.astype(np.complex128), -1)) <= tol(n)
