"""GPU tests of the matrix-core (MFMA) split-precision form: FP32 plans whose
radix-16 (and, in the radix-32 chains, radix-32) sub-stages run as FP16
hi/lo products accumulated in FP32 (MOFFT_PLAN_MFMA, kernel bank 3).

The bar is the plain FP32 path's: relative L2 error against the FP64 oracle
<= 1e-6*log2N for every N = 2^4..2^20 at batch 2^24/N (uniform[-1,1) inputs
and a unit impulse at a seeded index), and the round trip within the same
bound.  The measured errors of both forms are printed per N (pytest -s) and,
with MOFFT_ERROR_TABLE set, written as one JSON line per N.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from conftest import seeded_complex

pytestmark = pytest.mark.gpu


def tol(n):
    return 1e-6 * max(1.0, math.log2(n))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    return torch


def _sample_rows(batch, points):
    """Indices of the transforms checked against the oracle: about `points`
    points' worth spread evenly over the batch, always the first and last."""
    n_rows = max(1, min(batch, points))
    return sorted(set(np.linspace(0, batch - 1, n_rows).round().astype(int).tolist()))


def _record(row):
    """Append one per-N error row to the table named by MOFFT_ERROR_TABLE
    (a JSON-lines file; unset: nothing is written)."""
    path = os.environ.get("MOFFT_ERROR_TABLE")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps(row) + "\n")


@pytest.mark.parametrize("k", list(range(4, 21)))
def test_mfma_error_per_length(gpu, torch_dev, k):
    """C3's precision sweep at its full size: N = 2^k, batch 2^24 / N,
    components iid uniform[-1, 1) from a seed, forward out of place with the
    plain FP32 form and the matrix-core form.  Every transform runs on the
    device; the forward error is measured against the FP64 oracle on
    transforms sampled evenly over the batch (2^20 points' worth, first and
    last included); the round trip (backward in place, / N) over the whole
    batch on the device; and a unit impulse at a seeded index against its
    closed form X[q] = w_N^(q n0), for both forms."""
    torch = torch_dev
    n = 1 << k
    batch = (1 << 24) // n
    s = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(300 + k)
    x = torch.view_as_complex(torch.rand(batch, n, 2, device="cuda", generator=g) * 2 - 1)
    rows = _sample_rows(batch, max(1, (1 << 20) // n))
    xs = x[rows].cpu().numpy()
    ref = oracle.fft_dd(xs.astype(np.complex128), -1)
    n0 = int(np.random.default_rng(k).integers(0, n))
    imp = torch.zeros(1, n, dtype=torch.complex64, device="cuda")
    imp[0, n0] = 1
    want = oracle.omega_table(n, -1, 64, indices=(np.arange(n, dtype=np.uint64) * n0) % n)
    errs = {}
    for form in ("plain", "mfma"):
        y = torch.empty_like(x)
        yi = torch.empty_like(imp)
        with gpu.Plan(n, batch, 32, gpu.FORWARD, mfma=form == "mfma") as f, \
                gpu.Plan(n, batch, 32, gpu.BACKWARD, in_place=True, mfma=form == "mfma") as b, \
                gpu.Plan(n, 1, 32, gpu.FORWARD, mfma=form == "mfma") as fi:
            if form == "mfma":
                assert f.diagnostics.kernel_bank == gpu.MFMA_BANK
            f.execute_async(x, y, s)
            fi.execute_async(imp, yi, s)
            torch.cuda.synchronize()
            fwd = oracle.rel_l2(y[rows].cpu().numpy(), ref)
            b.execute_async(y, y, s)
            torch.cuda.synchronize()
        rt = (torch.linalg.vector_norm(y / n - x) / torch.linalg.vector_norm(x)).item()
        e_imp = oracle.rel_l2(yi[0].cpu().numpy(), want)
        errs[form] = (fwd, rt, e_imp)
    row = {"n": n, "batch": batch, "sampled_transforms": len(rows), "bar": tol(n)}
    for form in ("plain", "mfma"):
        row[f"{form}_forward"], row[f"{form}_round_trip"], row[f"{form}_impulse"] = errs[form]
    _record(row)
    print(f"N=2^{k} x {batch}: forward rel-L2 plain {errs['plain'][0]:.2e} mfma "
          f"{errs['mfma'][0]:.2e}; round trip plain {errs['plain'][1]:.2e} mfma "
          f"{errs['mfma'][1]:.2e}; impulse plain {errs['plain'][2]:.2e} mfma "
          f"{errs['mfma'][2]:.2e}; bar {tol(n):.1e}")
    for form in ("plain", "mfma"):
        for what, e in zip(("forward", "round trip", "impulse"), errs[form]):
            assert e <= tol(n), (form, what, e)


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


# radix-32 matrix-core chains (one whole-transform pass each), plan order
# (outermost radix first)
MFMA32 = [(1024, [32, 32]), (2048, [4, 16, 32]), (4096, [4, 32, 32]),
          (8192, [16, 16, 32]), (16384, [16, 32, 32])]


def _mfma_chain_plan(gpu, n, batch, radices, direction=None, in_place=False):
    p = gpu.Plan(n, batch, 32, gpu.FORWARD if direction is None else direction,
                 in_place=in_place, radices=radices)
    assert gpu.lib().mofft_plan_set_kernel_bank_f32(p.handle, gpu.MFMA_BANK) == 0
    # one pass: the chain's plain kernels group it into several (there is no
    # plain whole-transform kernel with a radix-32 sub-stage)
    p.set_blocked_transpose_mask(0)
    return p


@pytest.mark.parametrize("n,radices", MFMA32)
def test_mfma_radix32_chains(gpu, n, radices):
    """Radix-32 sub-stages on the matrix cores (64 x 64 real split-FP16
    products): the plan is one pass of the radix-32 kernel; forward out of
    place against the FP64 oracle, backward in place back to the input,
    both within the plain FP32 bar; a unit impulse against its closed form."""
    batch = max(2, (1 << 18) // n)
    x = seeded_complex((batch, n), seed=n + 32, dtype=np.complex64)
    ref = oracle.fft_dd(x.astype(np.complex128), -1)
    with _mfma_chain_plan(gpu, n, batch, radices) as f, \
            _mfma_chain_plan(gpu, n, batch, radices, gpu.BACKWARD, True) as b:
        d = f.description()
        assert d.passes == [n] and d.chain == radices
        info = gpu.kernels()[d.pass_kernel[0]]
        assert info.variant == gpu.MFMA_BANK
        assert list(info.stages[:info.stage_count]) == radices[::-1]
        y = np.empty_like(x)
        f.execute(x, y)
        fwd = oracle.rel_l2(y, ref)
        z = y.copy()
        b.execute(z, z)
        rt = oracle.rel_l2(z / n, x)
    n0 = int(np.random.default_rng(n).integers(0, n))
    imp = np.zeros((1, n), np.complex64)
    imp[0, n0] = 1
    want = oracle.omega_table(n, -1, 64, indices=(np.arange(n, dtype=np.uint64) * n0) % n)
    with _mfma_chain_plan(gpu, n, 1, radices) as fi:
        yi = np.empty_like(imp)
        fi.execute(imp, yi)
    e_imp = oracle.rel_l2(yi[0], want)
    print(f"N={n} radix-32 matrix-core {radices}: forward {fwd:.2e} round trip {rt:.2e} "
          f"impulse {e_imp:.2e} bar {tol(n):.1e}")
    assert fwd <= tol(n) and rt <= tol(n) and e_imp <= tol(n), (fwd, rt, e_imp)


@pytest.mark.parametrize("n,radices", [(4096, [16, 16, 16]), (16384, [16, 32, 32]),
                                       (1024, [32, 32]), (256, [16, 16]),
                                       (1 << 16, None), (1 << 18, None)])
def test_mfma_column_scale_per_transform(gpu, n, radices):
    """Chains of several matrix-core sub-stages scale each column (each
    transform of a whole-transform tile, each strided column of a multi-pass
    one) once by its own power of two: a batch whose transforms differ in
    magnitude by up to 2^100 either way meets the bar transform by
    transform, not only over the batch.  radices None: the MOFFT_PLAN_MFMA
    plan (two strided passes of radix-16 pairs at 2^16 and 2^18)."""
    batch = 64 if n <= 16384 else 8
    x = seeded_complex((batch, n), seed=n + 7, dtype=np.complex128)
    ex = np.random.default_rng(n).integers(-100, 101, batch)
    ex[0], ex[1] = -100, 100
    x = (x * np.ldexp(1.0, ex)[:, None]).astype(np.complex64)
    ref = oracle.fft_dd(x.astype(np.complex128), -1)
    plan = (_mfma_chain_plan(gpu, n, batch, radices) if radices
            else gpu.Plan(n, batch, 32, gpu.FORWARD, mfma=True))
    with plan as f:
        if radices is None:
            assert len(f.description().passes) == 2
            assert f.diagnostics.kernel_bank == gpu.MFMA_BANK
        y = np.empty_like(x)
        f.execute(x, y)
    worst = max(oracle.rel_l2(y[i], ref[i]) for i in range(batch))
    assert worst <= tol(n), worst
