"""GPU tests of the API surface around the transform: device pointers and
streams (mofft_hip.h), caller workspace, pass regrouping and kernel banks,
recipes, MEASURE and wisdom-strict planning, twiddle tables.

Every numeric check compares against the FP64 oracle with the path's bars
(FP32 1e-6*log2N, FP64 1e-15*log2N relative L2) or asserts bit identity where
only data movement differs (reference tests/test_fft.c:253-306).
"""
import ctypes
import math

import numpy as np
import pytest

import oracle
from conftest import seeded_complex

pytestmark = pytest.mark.gpu


def tol(n, bits):
    return (1e-6 if bits == 32 else 1e-15) * max(1.0, math.log2(n))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    return torch


def to_dev(torch, x, bits):
    t = torch.from_numpy(np.ascontiguousarray(x.astype(np.complex64 if bits == 32 else np.complex128)))
    return t.to("cuda")


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n,batch", [(1024, 64), (1 << 16, 8), (4096, 3)])
def test_async_device_pointers(gpu, torch_dev, n, batch, bits):
    torch = torch_dev
    x = seeded_complex((batch, n), seed=n + batch)
    d_in = to_dev(torch, x, bits)
    d_out = torch.empty_like(d_in)
    stream = torch.cuda.Stream()
    with gpu.Plan(n, batch, bits, gpu.FORWARD) as p:
        with torch.cuda.stream(stream):
            p.execute_async(d_in, d_out, stream.cuda_stream)
        stream.synchronize()
        y = d_out.cpu().numpy()
    ref = oracle.fft_dd(x.astype(np.complex64 if bits == 32 else np.complex128), -1)
    assert oracle.rel_l2(y, ref) <= tol(n, bits)


def test_async_rejects_host_pointers(gpu):
    x = np.zeros(64, np.complex64)
    with gpu.Plan(64) as p:
        with pytest.raises(gpu.MofftError) as e:
            p.execute_async(x, x.copy())
        assert e.value.status == gpu.INVALID_ARGUMENT


@pytest.mark.parametrize("bits", [32, 64])
def test_caller_workspace_concurrent_streams(gpu, torch_dev, bits):
    """One in-place multi-pass plan, two streams, two workspaces."""
    torch = torch_dev
    n, batch = 1 << 16, 4
    xs = [seeded_complex((batch, n), seed=s) for s in (1, 2)]
    bufs = [to_dev(torch, x, bits) for x in xs]
    with gpu.Plan(n, batch, bits, gpu.FORWARD, in_place=True) as p:
        need = p.workspace_size()
        assert need > 0
        works = [torch.empty(need, dtype=torch.uint8, device="cuda") for _ in bufs]
        streams = [torch.cuda.Stream() for _ in bufs]
        for b, w, s in zip(bufs, works, streams):
            p.execute_async_workspace(b, b, w, need, s.cuda_stream)
        for s in streams:
            s.synchronize()
        with pytest.raises(gpu.MofftError):
            p.execute_async_workspace(bufs[0], bufs[0], works[0], need - 1, 0)
    for x, b in zip(xs, bufs):
        ref = oracle.fft_dd(x.astype(np.complex64 if bits == 32 else np.complex128), -1)
        assert oracle.rel_l2(b.cpu().numpy(), ref) <= tol(n, bits)


@pytest.mark.parametrize("bits", [32, 64])
def test_pass_grouping_agrees(gpu, bits):
    """Same chain, different HBM pass boundaries (a different four-step
    factorisation of the same stages): equal to rounding and each within the
    oracle bar; the all-fused and all-split extremes are both checked."""
    n = 1 << 12
    x = seeded_complex((2, n), seed=9).astype(np.complex64 if bits == 32 else np.complex128)
    outs = []
    for mask in (0, 1, 2, 3):
        with gpu.Plan(n, 2, bits, radices=[16, 16, 16], pass_mask=mask) as p:
            y = np.empty_like(x)
            p.execute(x, y)
            outs.append((p.description().passes, y))
    assert [o[0] for o in outs] == [[4096], [256, 16], [16, 256], [16, 16, 16]]
    ref = oracle.dft(x, -1)
    for _, y in outs:
        assert oracle.rel_l2(y, ref) <= tol(n, bits)
        assert oracle.rel_l2(y, outs[0][1]) <= tol(n, bits)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n", [1024, 1 << 16])
def test_every_kernel_bank(gpu, n, bits):
    """Every kernel bank on the same chain and pass grouping.  Banks 0, 1, 2
    and 4 change only the tile shape (columns per workgroup): the same
    arithmetic on the same values, so their outputs are bit-identical (the
    GPU form of the reference's linear-vs-blocked transpose identity,
    tests/test_fft.c:253-306).  Bank 3 (matrix-core split precision, FP32)
    rounds differently and is checked against the oracle bar only."""
    x = seeded_complex((4, n), seed=n).astype(np.complex64 if bits == 32 else np.complex128)
    ref = oracle.fft_dd(x, -1)
    with gpu.Plan(n, 4, bits) as p0:
        count = p0.diagnostics.kernel_bank_count
    assert count >= 2
    L = gpu.lib()
    outs = {}
    for bank in range(count):
        with gpu.Plan(n, 4, bits) as p:
            assert getattr(L, f"mofft_plan_set_kernel_bank_f{bits}")(p.handle, bank) == 0
            got = ctypes.c_uint()
            getattr(L, f"mofft_plan_get_kernel_bank_f{bits}")(p.handle, ctypes.byref(got))
            assert got.value == bank
            y = np.empty_like(x)
            p.execute(x, y)
        assert oracle.rel_l2(y, ref) <= tol(n, bits)
        outs[bank] = y
    for bank, y in outs.items():
        if bank != gpu.MFMA_BANK:
            assert np.array_equal(y, outs[0]), ("bank", bank, "differs from bank 0")
    with gpu.Plan(n, 4, bits) as p:
        assert getattr(L, f"mofft_plan_set_kernel_bank_f{bits}")(p.handle, count) == gpu.INVALID_ARGUMENT


def _chain_plan(gpu, n, batch, bits, passes, bank):
    """Plan whose passes run the given sub-stage chains (execution order,
    first pass first) under kernel bank `bank`."""
    chain, mask, acc = [], 0, 0
    for pc in reversed(passes):
        chain += list(reversed(pc))
    for pc in reversed(passes[1:]):
        acc += len(pc)
        mask |= 1 << (acc - 1)
    p = gpu.Plan(n, batch, bits, radices=chain, pass_mask=mask)
    assert getattr(gpu.lib(), f"mofft_plan_set_kernel_bank_f{bits}")(p.handle, bank) == 0
    return p


@pytest.mark.parametrize("bits,n,batch,passes,bank", [
    (32, 1 << 20, 2, [(8, 8, 4, 4), (8, 8, 4, 4)], 5),     # 64-byte runs
    (32, 1 << 20, 2, [(16, 16, 4), (16, 16, 4)], 2),
    (32, 1 << 22, 1, [(8, 8, 8, 4), (8, 8, 8, 4)], 5),     # 32-byte runs
    (64, 1 << 20, 1, [(4, 4, 4, 4, 4), (4, 4, 4, 4, 4)], 5),
    # tile counts that leave a partial group of 8 S workgroups
    (32, 12 * 1024, 22, [(8, 8, 4, 4), (12,)], 5),
    (32, 12 * 2048, 30, [(8, 8, 8, 4), (12,)], 5),
])
def test_tile_order_bit_identical(gpu, bits, n, batch, passes, bank):
    """Strided tiles narrower than a 128-byte line are dealt to workgroups
    so that the tiles sharing a line run on one XCD (MOFFT_PASS_XCD_ORDER).
    The order moves whole tiles between workgroups, never the arithmetic:
    the output is bit-identical to the linear order's and within the bar;
    a tile dropped or run twice by the remap would show as a mismatch."""
    L = gpu.lib()
    dt = np.complex64 if bits == 32 else np.complex128
    x = seeded_complex((batch, n), seed=n + bank).astype(dt)
    outs = []
    try:
        with _chain_plan(gpu, n, batch, bits, passes, bank) as p:
            d = p.description()
            runs = [gpu.kernels()[d.pass_kernel[i]] for i in range(d.pass_count)]
            assert any(not k.rows and k.columns_per_block * (bits // 4) < 128
                       for k in runs), "no narrow tile in this plan"
            for order in (1, 0):
                L.mofft_internal_set_tile_order(order)
                y = np.empty_like(x)
                p.execute(x, y)
                outs.append(y)
    finally:
        L.mofft_internal_set_tile_order(-1)
    assert np.array_equal(outs[0], outs[1])
    assert oracle.rel_l2(outs[0], oracle.fft_dd(x, -1)) <= tol(n, bits)


@pytest.mark.parametrize("bits", [32, 64])
def test_recipe_roundtrip_replays_bits(gpu, bits):
    n = 20736
    x = seeded_complex((1, n), seed=3).astype(np.complex64 if bits == 32 else np.complex128)
    with gpu.Plan(n, 1, bits, gpu.BACKWARD) as p:
        recipe = p.export()
        y0 = np.empty_like(x)
        p.execute(x, y0)
    assert recipe.startswith("MOFFT_PLAN_V1 " + gpu.build_id() + " 00000950 f%d 20736 1 0 " % bits)
    L = gpu.lib()
    h = ctypes.c_void_p()
    d = gpu.PlanDiagnostics()
    raw = recipe.encode()
    st = getattr(L, f"mofft_plan_import_f{bits}")(ctypes.byref(h), raw, len(raw) + 1, ctypes.byref(d))
    assert st == gpu.SUCCESS and d.length == n and d.direction == gpu.BACKWARD
    y1 = np.empty_like(x)
    assert getattr(L, f"mofft_execute_f{bits}")(h, x.ctypes.data, y1.ctypes.data) == 0
    getattr(L, f"mofft_plan_destroy_f{bits}")(h)
    assert np.array_equal(y0, y1)


@pytest.mark.parametrize("bits,n", [(32, 1 << 20), (64, 1 << 16), (32, 1 << 18), (64, 3 << 14)])
def test_stage_layout_selects_factored_pass_twiddles(gpu, bits, n):
    """mofft_plan_set_stage_layout: SECTION makes every later pass read the
    two-level (factored) pass-twiddle table instead of its full table, so
    the plan's twiddle memory shrinks; the transform stays within the bar
    against the oracle, and the layout survives export and import (the
    imported plan replays the same bits)."""
    batch = 2
    dt = np.complex64 if bits == 32 else np.complex128
    x = seeded_complex((batch, n), seed=n + bits).astype(dt)
    L = gpu.lib()
    setter = getattr(L, f"mofft_plan_set_stage_layout_f{bits}")
    with gpu.Plan(n, batch, bits, gpu.FORWARD) as p:
        # The compiled wisdom may already pick SECTION for this length.
        assert setter(p.handle, 0) == gpu.SUCCESS          # MOFFT_STAGE_LAYOUT_TRANSPOSE
        d0 = p.description()
        assert d0.pass_count >= 2
        y0 = np.empty_like(x)
        p.execute(x, y0)
        assert setter(p.handle, 1) == gpu.SUCCESS          # MOFFT_STAGE_LAYOUT_SECTION
        d1 = p.description()
        assert list(d1.passes) == list(d0.passes) and d1.twiddle_bytes < d0.twiddle_bytes
        y1 = np.empty_like(x)
        p.execute(x, y1)
        recipe = p.export()
        assert setter(p.handle, 0) == gpu.SUCCESS          # back to TRANSPOSE
        assert p.description().twiddle_bytes == d0.twiddle_bytes
        y2 = np.empty_like(x)
        p.execute(x, y2)
    ref = oracle.fft_dd(x[:1].astype(np.complex128), -1)
    assert oracle.rel_l2(y1[:1], ref) <= tol(n, bits)
    assert oracle.rel_l2(y0[:1], ref) <= tol(n, bits)
    assert np.array_equal(y0, y2)
    h = ctypes.c_void_p()
    d = gpu.PlanDiagnostics()
    raw = recipe.encode()
    st = getattr(L, f"mofft_plan_import_f{bits}")(ctypes.byref(h), raw, len(raw) + 1, ctypes.byref(d))
    assert st == gpu.SUCCESS and d.stage_layout == 1
    y3 = np.empty_like(x[:1])
    assert getattr(L, f"mofft_execute_f{bits}")(h, x[:1].ctypes.data, y3.ctypes.data) == 0
    getattr(L, f"mofft_plan_destroy_f{bits}")(h)
    assert np.array_equal(y3, y1[:1])


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n,batch,in_place", [(30, 1, False), (4096, 16, True), (1 << 16, 4, False)])
def test_measure_planner(gpu, n, batch, in_place, bits):
    p = gpu.Plan(n, batch, bits, gpu.FORWARD, in_place=in_place, mode=gpu.MEASURE,
                 candidate_limit=6, timing_samples=3)
    try:
        d = p.diagnostics
        assert d.measured == 1 and 1 <= d.candidates_measured <= 6
        assert d.selected_median_seconds > 0 and d.planning_seconds > 0
        assert math.prod(d.chain) == n
        x = seeded_complex((batch, n), seed=n).astype(np.complex64 if bits == 32 else np.complex128)
        if in_place:
            y = x.copy()
            p.execute(y, y)
        else:
            y = np.empty_like(x)
            p.execute(x, y)
    finally:
        p.destroy()
    assert oracle.rel_l2(y, oracle.fft_dd(x, -1)) <= tol(n, bits)


def test_strict_wisdom_and_model_fallback(gpu):
    # 1000 has no tuned entry: STRICT fails, the default falls back to the model
    with pytest.raises(gpu.MofftError) as e:
        gpu.Plan(1000, 1, 32, flags=gpu.STRICT_WISDOM)
    assert e.value.status == gpu.WISDOM_NOT_FOUND
    with gpu.Plan(1000) as p:
        assert p.diagnostics.used_model_fallback == 1 and p.diagnostics.used_wisdom == 0
    # 1024 has one: STRICT succeeds and reports the wisdom plan
    with gpu.Plan(1024, 1, 32, flags=gpu.STRICT_WISDOM) as p:
        assert p.diagnostics.used_wisdom == 1 and p.diagnostics.used_model_fallback == 0


def _device_table(gpu, kind, bits, a, b, count):
    out = np.empty(count, np.complex64 if bits == 32 else np.complex128)
    st = gpu.lib().mofft_internal_twiddle_table(
        ctypes.c_int(kind), ctypes.c_int(bits), ctypes.c_int64(a), ctypes.c_int64(b),
        out.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(count))
    assert st == gpu.SUCCESS
    return out


def _ulps(got, want):
    """Largest distance in units of the last place of `want`, per part."""
    got = np.asarray(got)
    want = np.asarray(want)
    ft = np.float32 if got.dtype == np.complex64 else np.float64
    worst = 0.0
    for a, b in ((got.real, want.real), (got.imag, want.imag)):
        sp = np.spacing(np.abs(b).astype(ft))
        worst = max(worst, float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / sp)))
    return worst


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("kind,a,b", [(0, 1 << 20, 1), (0, 1 << 16, 1 << 6), (0, 3 * 5 * 7 * 11 * 13 * 16, 7),
                                      (1, 256, 16), (1, 729, 3), (1, 1000, 12),
                                      (2, 128, 512), (2, 12, 1728), (2, 1024, 1024)])
def test_twiddle_tables_bitwise(gpu, bits, kind, a, b):
    """Every twiddle table the plans use is generated on the device in FP64
    and rounded once (src/mofft.c:1097-1105 rounds cos/sin of a double
    angle once).  Each table kind - the two-level powers w_M^((i*step) mod M),
    the sub-stage table w_(ns*r)^(a*k) and the full pass table
    w_(ns*L)^(a*k) - is compared with the oracle's long-double values
    rounded once: FP32 tables are equal bit for bit, FP64 ones within one
    ulp (the device sincospi is not correctly rounded)."""
    if kind == 0:
        count = min(a, 1 << 20)
        m = (np.arange(count, dtype=np.uint64) * np.uint64(b)) % np.uint64(a)
        big_m = a
    elif kind == 1:
        count = a * (b - 1)
        i = np.arange(count, dtype=np.uint64)
        m = (i // np.uint64(b - 1)) * (i % np.uint64(b - 1) + np.uint64(1))
        big_m = a * b
    else:
        count = a * b
        i = np.arange(count, dtype=np.uint64)
        m = (i % np.uint64(a)) * (i // np.uint64(a))
        big_m = a * b
    got = _device_table(gpu, kind, bits, a, b, count)
    want = oracle.omega_table(big_m, -1, bits, indices=m % np.uint64(big_m))
    if bits == 32:
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), _ulps(got, want)
    else:
        assert _ulps(got, want) <= 1.0


@pytest.mark.parametrize("bits", [32, 64])
def test_twiddle_tables_rounded_once(gpu, torch_dev, bits):
    """The GPU-generated tables equal the correctly rounded values (FP32
    exactly; FP64 within one ulp - the device sincospi is not correctly
    rounded).  Read back through a two-pass plan's impulse response."""
    n = 1 << 16
    # X[k] = w_N^(k*n0) for an impulse at n0: exercises stage and pass tables
    x = np.zeros((1, n), dtype=np.complex64 if bits == 32 else np.complex128)
    x[0, 1] = 1.0
    with gpu.Plan(n, 1, bits) as p:
        y = np.empty_like(x)
        p.execute(x, y)
    k = np.arange(0, n, 997)
    want = oracle.omega_table(n, -1, 64, indices=k)
    err = np.max(np.abs(y[0, k] - want))
    assert err <= (8 * np.finfo(np.float32).eps if bits == 32 else 8 * np.finfo(np.float64).eps)


@pytest.mark.parametrize("bits,n,batch,in_place", [(32, 1024, 64, False), (64, 1 << 16, 4, True),
                                                   (32, 1 << 21, 2, False)])
def test_captured_graph_replays(gpu, torch_dev, bits, n, batch, in_place):
    """A plan captured into a HIP graph: every replay (with eager work and
    synchronisation between replays) equals the direct execution."""
    torch = torch_dev
    x = seeded_complex((batch, n), seed=n, dtype=np.complex64 if bits == 32 else np.complex128)
    src = to_dev(torch, x, bits)
    buf = src.clone() if in_place else src
    out = buf if in_place else torch.empty_like(src)
    with gpu.Plan(n, batch, bits, gpu.FORWARD, in_place=in_place) as p:
        p.execute_async(buf, out, 0)
        torch.cuda.synchronize()
        want = out.clone()
        with p.capture(buf, out) as g:
            for rep in range(3):
                if in_place:
                    buf.copy_(src)
                else:
                    out.zero_()
                junk = torch.randn(1 << 18, device="cuda").sum()
                torch.cuda.synchronize()
                g.launch(0)
                torch.cuda.synchronize()
                assert torch.equal(out, want), rep
                del junk
    ref = oracle.fft_dd(x, -1)
    assert oracle.rel_l2(want.cpu().numpy(), ref) <= tol(n, bits)


def test_capture_before_first_launch(gpu, torch_dev):
    """Capturing a plan whose kernel has never been launched in the process
    (a 139 KiB-LDS matrix-core kernel: its launch attribute is set before the
    capture) and replaying it equals a direct execution."""
    torch = torch_dev
    n, batch = 1 << 15, 4
    x = seeded_complex((batch, n), seed=15, dtype=np.complex64)
    src = to_dev(torch, x, 32)
    out = torch.empty_like(src)
    with gpu.Plan(n, batch, 32, gpu.FORWARD, mfma=True) as p:
        with p.capture(src, out) as g:
            g.launch(0)
            torch.cuda.synchronize()
            got = out.clone()
        out.zero_()
        p.execute_async(src, out, 0)
        torch.cuda.synchronize()
    assert torch.equal(got, out)
    assert oracle.rel_l2(got.cpu().numpy(), oracle.fft_dd(x, -1)) <= tol(n, 32)


@pytest.mark.parametrize("n,bits", [(1 << 30, 32), (1 << 28, 32), (1 << 26, 64),
                                    (3 ** 17, 32)])
def test_largest_lengths(gpu, torch_dev, n, bits):
    """Lengths of 2^26 points up to the largest one (2^30, 8 GiB per FP32
    transform; three or four passes) through size-independent checks on the
    device:
    an impulse at n0 gives X[k] = w_N^(k n0) (closed form, FP64 on the
    device), and forward out of place then backward in place returns N x."""
    torch = torch_dev
    cdt = torch.complex64 if bits == 32 else torch.complex128
    s = torch.cuda.current_stream().cuda_stream
    n0 = 12345
    x = torch.zeros(n, dtype=cdt, device="cuda")
    x[n0] = 1.0
    y = torch.empty_like(x)
    with gpu.Plan(n, 1, bits, gpu.FORWARD) as p:
        assert p.description().pass_count in (3, 4)
        p.execute_async(x, y, s)
        torch.cuda.synchronize()
    num = 0.0
    for lo in range(0, n, 1 << 26):
        hi = min(n, lo + (1 << 26))
        k = torch.arange(lo, hi, dtype=torch.int64, device="cuda")
        ang = (k * n0 % n).to(torch.float64) * (-2.0 * math.pi / n)
        ys = y[lo:hi]
        num += ((ys.real.double() - torch.cos(ang)) ** 2
                + (ys.imag.double() - torch.sin(ang)) ** 2).sum().item()
    err = math.sqrt(num / n)
    assert err <= tol(n, bits), ("impulse", n, bits, err)
    g = torch.Generator(device="cuda")
    g.manual_seed(n)
    x = torch.randn(n, dtype=cdt, device="cuda", generator=g)
    with gpu.Plan(n, 1, bits, gpu.FORWARD) as f, \
            gpu.Plan(n, 1, bits, gpu.BACKWARD, in_place=True) as b:
        f.execute_async(x, y, s)
        b.execute_async(y, y, s)
        torch.cuda.synchronize()
    rt = (torch.linalg.vector_norm(y / n - x) / torch.linalg.vector_norm(x)).item()
    assert rt <= tol(n, bits), ("round trip", n, bits, rt)


@pytest.mark.parametrize("n,batch,bits", [(20736, 1000, 64), (1 << 21, 20, 32),
                                          (1 << 16, 300, 64), (98415, 200, 32)])
def test_multi_chunk_batches(gpu, torch_dev, n, batch, bits):
    """Batches larger than one 128 MiB chunk (2-3 chunks, the last one
    partial; two- and three-pass plans): forward in place, then a sample
    of transforms (first, last, the first of the last chunk) against the
    oracle, then backward in place back to N x for the whole batch."""
    torch = torch_dev
    cdt = torch.complex64 if bits == 32 else torch.complex128
    s = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(n + batch)
    x = torch.randn(batch, n, dtype=cdt, device="cuda", generator=g)
    y = x.clone()
    with gpu.Plan(n, batch, bits, gpu.FORWARD, in_place=True) as f, \
            gpu.Plan(n, batch, bits, gpu.BACKWARD, in_place=True) as b:
        d = f.description()
        chunks = -(-batch // d.chunk_transforms)
        assert chunks >= 2 and batch % d.chunk_transforms != 0
        f.execute_async(y, y, s)
        torch.cuda.synchronize()
        last_chunk = (chunks - 1) * d.chunk_transforms
        for t in (0, last_chunk - 1, last_chunk, batch - 1):
            ref = oracle.fft_dd(x[t:t + 1].cpu().numpy().astype(np.complex128), -1)
            err = oracle.rel_l2(y[t:t + 1].cpu().numpy(), ref)
            assert err <= tol(n, bits), ("forward", t, err)
        b.execute_async(y, y, s)
        torch.cuda.synchronize()
    rt = (torch.linalg.vector_norm(y / n - x) / torch.linalg.vector_norm(x)).item()
    assert rt <= tol(n, bits), ("round trip", rt)


@pytest.mark.parametrize("bits", [32, 64])
def test_graph_outlives_regroup_and_plan(gpu, torch_dev, bits):
    """A captured graph keeps the twiddle tables and workspace it was
    captured with: after a setter regroups the plan's passes (new tables for
    the plan) and after the plan is destroyed, replays still equal the
    direct execution of the original grouping."""
    torch = torch_dev
    n, batch = 1 << 16, 4
    x = seeded_complex((batch, n), seed=77, dtype=np.complex64 if bits == 32 else np.complex128)
    src = to_dev(torch, x, bits)
    buf = src.clone()
    L = gpu.lib()
    p = gpu.Plan(n, batch, bits, gpu.FORWARD, in_place=True, radices=[16, 16, 16, 16], pass_mask=2)
    try:
        p.execute_async(buf, buf, 0)
        torch.cuda.synchronize()
        want = buf.clone()
        g = p.capture(buf, buf)
        # one pass per radix: a different schedule, new tables for the plan
        p.set_blocked_transpose_mask(7)
        assert p.description().pass_count == 4
        buf.copy_(src)
        p.execute_async(buf, buf, 0)
        torch.cuda.synchronize()
        assert oracle.rel_l2(buf.cpu().numpy(), oracle.fft_dd(x, -1)) <= tol(n, bits)
        buf.copy_(src)
        torch.cuda.synchronize()
        g.launch(0)
        torch.cuda.synchronize()
        assert torch.equal(buf, want)
    finally:
        p.destroy()
    buf.copy_(src)
    scratch = torch.randn(1 << 22, device="cuda")  # reuse freed memory
    torch.cuda.synchronize()
    g.launch(0)
    torch.cuda.synchronize()
    assert torch.equal(buf, want)
    g.destroy()
    del scratch
