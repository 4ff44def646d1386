"""Host-side logic of libmofft.so that needs no GPU (CPU only).

Covers: the C ABI exports every entry point declared in include/*.h, the
reference's argument/status conventions (src/mofft.c:739-762, 1232-1331,
1384-1432), the planner and pass schedule, the cost-model byte accounting
known answers (tests/test_fft.c:493-560) and the recipe parser's rejection
rules (src/mofft.c:1675-1758, tests/test_plan_recipe.c).  No compute call is
made: on a machine without a gfx950 device plan creation must report
MOFFT_UNSUPPORTED_HARDWARE.
"""
import ctypes
import math
import re
from pathlib import Path

import pytest

import mofft_amd as mf
from conftest import SMALL_POWERS, mixed_kernel_plans, random_smooth_lengths

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mofft_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(mf.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 42
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the reference's 41 drop-in entry points (include/mofft.h:107-231)
    ref_api = []
    for s in ("f32", "f64"):
        ref_api += [f"mofft_plan_create_{s}", f"mofft_plan_create_with_options_{s}",
                    f"mofft_plan_dft_1d_{s}", f"mofft_execute_plan_{s}",
                    f"mofft_plan_create_with_radices_{s}",
                    f"mofft_plan_set_transpose_strategy_{s}",
                    f"mofft_plan_set_blocked_transpose_mask_{s}",
                    f"mofft_plan_set_direct_input_{s}", f"mofft_plan_set_stage_layout_{s}",
                    f"mofft_plan_set_kernel_bank_{s}", f"mofft_plan_get_kernel_bank_{s}",
                    f"mofft_plan_export_{s}", f"mofft_plan_import_{s}",
                    f"mofft_execute_{s}", f"mofft_plan_destroy_{s}"]
    ref_api += ["mofft_plan_options_init", "mofft_build_id", "mofft_estimated_radix_cost",
                "mofft_estimated_stage_cost", "mofft_estimated_plan_cost",
                "mofft_estimated_plan_cost_with_layout",
                "mofft_estimated_plan_cost_breakdown_with_layout",
                "mofft_wisdom_is_formal", "mofft_wisdom_source", "mofft_status_string",
                "mofft_version"]
    assert len(ref_api) == 41
    assert set(ref_api) <= set(names)
    assert all(hasattr(lib, n) for n in ref_api)


def test_binding_table_matches_headers():
    assert set(mf.exported_symbols()) <= set(declared_functions())


def test_strings_and_ids():
    assert mf.status_string(mf.SUCCESS) == "success"
    assert mf.status_string(mf.BUFFER_TOO_SMALL) == "output buffer too small"
    assert mf.status_string(99) == "unknown status"
    assert mf.lib().mofft_version().decode() == "0.4.0"
    bid = mf.build_id()
    assert re.fullmatch(r"[0-9a-f]{64}", bid)
    assert (ROOT / "mofft_amd" / "lib" / "BUILD_ID").read_text().strip() == bid
    assert mf.lib().mofft_wisdom_source().decode() in ("none", "embedded-plan-wisdom") or True


def test_no_device_means_unsupported_hardware():
    if mf.device_available():
        pytest.skip("a gfx950 device is present")
    L = mf.lib()
    p = ctypes.c_void_p(1)
    assert L.mofft_plan_create_f32(ctypes.byref(p), 1024, mf.FORWARD, mf.OUT_OF_PLACE) == mf.UNSUPPORTED_HARDWARE
    assert p.value is None
    # argument errors are reported before the hardware probe
    assert L.mofft_plan_create_f32(None, 1024, mf.FORWARD, mf.OUT_OF_PLACE) == mf.INVALID_ARGUMENT
    assert L.mofft_plan_create_f32(ctypes.byref(p), 1024, 0, mf.OUT_OF_PLACE) == mf.INVALID_ARGUMENT
    opts = mf.PlanOptions(mf.MEASURE, 1, 0, 0)   # MEASURE with flags
    assert L.mofft_plan_create_with_options_f32(ctypes.byref(p), 1024, mf.FORWARD, 0,
                                                ctypes.byref(opts), None) == mf.INVALID_ARGUMENT
    opts = mf.PlanOptions(mf.PRETUNED, 0, 65, 0)  # candidate limit > 64
    assert L.mofft_plan_create_with_options_f64(ctypes.byref(p), 1024, mf.FORWARD, 0,
                                                ctypes.byref(opts), None) == mf.INVALID_ARGUMENT
    opts = mf.PlanOptions(mf.PRETUNED, 0, 0, 32)  # samples > 31
    assert L.mofft_plan_create_with_options_f64(ctypes.byref(p), 1024, mf.FORWARD, 0,
                                                ctypes.byref(opts), None) == mf.INVALID_ARGUMENT
    # bound-array constructor returns NULL on failure / bad flags
    buf = (ctypes.c_float * 64)()
    assert L.mofft_plan_dft_1d_f32(32, buf, buf, mf.FORWARD, 0) is None
    assert L.mofft_plan_dft_1d_f32(32, buf, buf, mf.FORWARD, 3) is None
    # destroy(NULL) is a no-op; executing NULL plans is an argument error
    L.mofft_plan_destroy_f32(None)
    assert L.mofft_execute_f32(None, buf, buf) == mf.INVALID_ARGUMENT
    assert L.mofft_execute_plan_f64(None) == mf.INVALID_ARGUMENT


@pytest.mark.parametrize("n,bits,passes", [
    (1 << 25, 32, [512, 256, 256]), (1 << 25, 64, [512, 256, 256]),
    (45, 32, [45]), (24, 64, [24]), (1000, 64, [1000]),
    (59049, 32, [243, 243]), (117649, 64, [343, 343]),
    (100000, 32, [400, 250]),
    # MI355X, 2^24 points: [192, 175] 130-134 us, [105, 320] 134-135 us
    (33600, 32, [[192, 175], [105, 320]]),
    (161051, 64, [121, 121, 11]), (8748, 64, [108, 81]),
    # 1000-point FP32 strided kernels have 64-byte runs: three wide passes
    # (MI355X, 2^24 points: [125, 64, 125] 182-188 us, [125, 125, 64] 185,
    # [64, 125, 125] 198; the nominal weights' [100, 100, 100] 199): either
    # of the two within the measurement's noise; the compiled wisdom entry
    # (tuned: [125, 250, 32], 172 us) takes precedence where it exists
    (10 ** 6, 32, [[125, 64, 125], [125, 125, 64], [125, 250, 32]]),
    (10 ** 6, 64, [1000, 1000]),
    # 2048-point passes cost three short ones
    (3 << 21, 32, [384, 128, 128]),
])
def test_model_schedule(n, bits, passes):
    """Lengths without a wisdom entry: the analytical planner's schedule
    (lowest weighted HBM cost, then fewest passes, then the most balanced
    split; the split the measured pass costs prefer where they predict it
    at least 15 % faster)."""
    d = mf.describe(n, 1, bits)
    if isinstance(passes[0], list):
        assert d.passes in passes
        passes = d.passes
    else:
        assert d.passes == passes
    assert math.prod(d.chain) == n
    # the mask marks exactly the pass boundaries
    assert bin(d.pass_mask).count("1") == len(passes) - 1


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("rows", [1, 0])
def test_mixed_kernels_selected(bits, rows):
    """Every kernel of a length that is not a power of two is the one its
    plan runs (whole-transform kernels as the single pass, strided kernels as
    the first pass), so the GPU parity test over these plans covers each of
    them: every 2,3,5,7,11,13-smooth length up to 1024 and the larger powers."""
    plans = mixed_kernel_plans(mf, bits, rows)
    lengths = {mf.kernels()[i].length for i, *_ in plans}
    smooth = [n for n in range(3, 1025) if n & (n - 1)
              and max(p for p in range(2, n + 1) if n % p == 0
                      and all(p % q for q in range(2, p))) <= 13]
    assert set(smooth) <= lengths and set(SMALL_POWERS) <= lengths
    for i, n, radices, mask in plans:
        d = mf.describe(n, 1, bits, radices, mask)
        assert d.pass_kernel[0] == i, (n, radices, mask)
        assert d.passes == ([n] if mask == 0 else [n // 16, 16])


@pytest.mark.parametrize("bits", [32, 64])
def test_random_smooth_length_schedules(bits):
    """300 seeded 13-smooth lengths up to 2^26: every one has a schedule
    whose passes and radix chain multiply to N, whose pass boundaries are
    the mask bits, with every pass a compiled kernel of that length; up to
    4096 points it is one pass (every smooth length there has a row
    kernel)."""
    ks = mf.kernels()
    for n in random_smooth_lengths(77 + bits, 300, 1 << 26):
        d = mf.describe(n, 1, bits)
        assert math.prod(d.passes) == n and math.prod(d.chain) == n, n
        assert bin(d.pass_mask).count("1") == d.pass_count - 1, n
        for i in range(d.pass_count):
            k = ks[d.pass_kernel[i]]
            assert k.precision_bits == bits and k.length == d.passes[i], (n, i)
        if n <= 4096:
            assert d.pass_count == 1, n


@pytest.mark.parametrize("bits", [32, 64])
def test_pretuned_pass_counts(bits):
    """PRETUNED plans (compiled wisdom or the model): one pass while a
    transform is small, at least two once it cannot fit one workgroup's LDS
    (FP32 > 32768, FP64 > 16384 points), at most three up to 2^24; the
    passes multiply to N.  Which count wins in between is the tuner's
    measurement (FP32 32768 is tuned to one pass, 20480 to two, for instance)."""
    lds_max = 32768 if bits == 32 else 16384
    for k in range(3, 25):
        n = 1 << k
        d = mf.describe(n, 1, bits)
        assert math.prod(d.passes) == n and math.prod(d.chain) == n
        if n <= 1024:
            assert d.pass_count == 1, n
        if n > lds_max:
            assert d.pass_count >= 2, n
        assert d.pass_count <= 3, n


def test_unsupported_sizes():
    for n in (1, 17, 34, 2 * 19, 1 << 40 | 1):
        with pytest.raises(mf.MofftError) as e:
            mf.describe(n, 1, 32)
        assert e.value.status == mf.UNSUPPORTED_SIZE


def test_explicit_chain_grouping():
    # greedy longest compiled run from the innermost stage
    assert mf.describe(4096, 1, 32, [16, 16, 16]).passes == [4096]
    # a set mask bit forces an HBM pass boundary before that level
    d = mf.describe(4096, 1, 32, [16, 16, 16], 1)
    assert d.passes == [256, 16] and d.pass_mask == 1
    d = mf.describe(4096, 1, 32, [16, 16, 16], 3)
    assert d.passes == [16, 16, 16]
    # radix 32 and 64 run as their own L = 32 / 64 passes
    assert mf.describe(256, 1, 32, [32, 8]).passes == [8, 32]
    assert mf.describe(4096, 1, 64, [2] * 12).pass_count >= 1
    for bad in ([17, 16], [4, 16, 8], [], [16, 16, 16, 1]):
        with pytest.raises(mf.MofftError):
            mf.describe(4096 if bad else 1, 1, 32, bad)
    with pytest.raises(mf.MofftError):
        mf.describe(4096, 1, 32, [16, 16, 16], 4)   # mask beyond depth-1 levels


def test_workspace_rules():
    # out-of-place two-pass plans need no workspace; in-place ones do
    d = mf.describe(1 << 20, 16, 32)
    assert d.pass_count == 2 and d.workspace_bytes > 0      # describe() reports in-place
    assert d.chunk_transforms * (1 << 20) * 8 <= 128 << 20   # Infinity-Cache-sized chunks
    assert d.workspace_bytes == d.chunk_transforms * (1 << 20) * 8


def test_cost_model_byte_kats():
    # tests/test_fft.c:531-541
    sec = mf.estimated_plan_breakdown(4096, 64, [8, 8, 8, 8], 0, 0, True)
    lin = mf.estimated_plan_breakdown(4096, 64, [8, 8, 8, 8], 0, 0, False)
    blk = mf.estimated_plan_breakdown(4096, 64, [8, 8, 8, 8], 7, 0, False)
    assert (sec.logical_read_bytes, sec.logical_write_bytes,
            sec.transferred_read_bytes, sec.transferred_write_bytes,
            sec.peak_working_set_bytes) == (336896, 262144, 336896, 262144, 196608)
    assert sec.layout_cost == 0.0
    assert (lin.logical_read_bytes, lin.logical_write_bytes,
            lin.transferred_read_bytes, lin.transferred_write_bytes) == (533504, 458752, 992256, 458752)
    assert blk.transferred_read_bytes == 533504
    assert lin.layout_cost > 0.0
    L = mf.lib()
    arr = (ctypes.c_int * 4)(8, 8, 8, 8)
    assert abs(L.mofft_estimated_plan_cost(4096, 64, arr, 4, 0, 0) - lin.total_cost) < 1e-12
    assert abs(L.mofft_estimated_plan_cost_with_layout(4096, 64, arr, 4, 0, 0, 1) - sec.total_cost) < 1e-12
    assert L.mofft_estimated_plan_cost(4096, 64, arr, 3, 0, 1) > 1e300
    assert L.mofft_estimated_plan_cost_with_layout(4096, 64, arr, 4, 0, 0, 99) > 1e300
    assert L.mofft_estimated_plan_cost_breakdown_with_layout(4096, 64, arr, 4, 0, 0, 1, None) == mf.INVALID_ARGUMENT
    small = L.mofft_estimated_stage_cost(8, 64, 0, 4096, 512)
    large = L.mofft_estimated_stage_cost(8, 64, 0, 262144, 32768)
    assert math.isfinite(small) and math.isfinite(large) and large > small
    assert L.mofft_estimated_stage_cost(8, 16, 0, 4096, 512) > 1e300
    assert L.mofft_estimated_radix_cost(17, 32, 0) >= 1e300


def test_kernel_registry_introspection():
    ks = mf.kernels()
    assert len(ks) == mf.lib().mofft_kernel_count()
    for k in ks:
        assert math.prod(k.stages[: k.stage_count]) == k.length
    bad = mf.KernelInfo()
    assert mf.lib().mofft_kernel_get_info(len(ks), ctypes.byref(bad)) == mf.INVALID_ARGUMENT


def _import(recipe: bytes, size=None, bits=32):
    p = ctypes.c_void_p(1)
    fn = getattr(mf.lib(), f"mofft_plan_import_f{bits}")
    st = fn(ctypes.byref(p), recipe, len(recipe) + 1 if size is None else size, None)
    return st, p.value


def test_recipe_parser_rejections():
    bid = mf.build_id().encode()
    good = b"MOFFT_PLAN_V1 " + bid + b" 00000950 f32 1024 -1 0 0 0000000000000000 0 0 3 4 16 16\n"
    cases = [
        (b"", mf.INVALID_ARGUMENT),
        (good[:-1], mf.INVALID_ARGUMENT),                       # no trailing newline
        (good.replace(b"MOFFT_PLAN_V1", b"MOFFT_PLAN_V2"), mf.INCOMPATIBLE_RECIPE),
        (good.replace(bid, b"0" * 64), mf.INCOMPATIBLE_RECIPE),
        (good.replace(b" -1 ", b" 2 "), mf.INVALID_ARGUMENT),
        (good.replace(b" 3 4 16 16", b" 3 4 16 17"), mf.INVALID_ARGUMENT),
        (good.replace(b" 3 4 16 16", b" 2 4 16 16"), mf.INVALID_ARGUMENT),
        (good.replace(b"0000000000000000", b"0000000000000004"), mf.INVALID_ARGUMENT),
        (good.replace(b"00000950", b"0950"), mf.INVALID_ARGUMENT),
        (good.replace(b"f32", b"f64"), mf.INCOMPATIBLE_RECIPE),
        (good.replace(b" 1024 ", b"  1024 "), mf.INVALID_ARGUMENT),
        (good.replace(b" 0 0 0000", b" 0 6 0000"), mf.INVALID_ARGUMENT),   # bank
    ]
    for recipe, want in cases:
        st, p = _import(recipe)
        assert (st, p) == (want, None), (recipe, st)
    # a recipe that parses still needs the gfx950 family to be confirmed
    if not mf.device_available():
        assert _import(good) == (mf.INCOMPATIBLE_RECIPE, None)
    # embedded NUL / wrong size
    assert _import(good, size=len(good))[0] == mf.INVALID_ARGUMENT


def test_compiled_wisdom_matches_json():
    """PRETUNED reads the autotuner's wisdom (mofft_amd/wisdom/gfx950.json)
    compiled into the library: every entry's chain and pass boundaries."""
    import json
    data = json.loads((ROOT / "mofft_amd" / "wisdom" / "gfx950.json").read_text())
    assert mf.lib().mofft_wisdom_source().decode() == data.get("source", "none")
    for e in data["entries"]:
        bits = 32 if e["precision"] == "fp32" else 64
        d = mf.describe(e["length"], 1, bits)
        assert d.chain == e["radices"], e
        assert d.pass_mask == e["blocked_transpose_mask"], e


# ---- persistent uneven batch: host-side planning ----

@pytest.mark.parametrize("bits", [32, 64])
def test_batch_schedules(bits):
    """Every power of two up to 4096 is one gathered row tile; up to 2^18
    two strided passes of at most 512 points; longer ones three or more.
    Every kind is a 256-thread tile of the batch registry."""
    for k in range(1, 23):
        n = 1 << k
        d = mf.batch_describe_length(n, bits)
        assert math.prod(d.chain) == n and math.prod(d.passes) == n
        want = 1 if n <= 4096 else 2 if n <= 1 << 18 else 3
        assert d.pass_count == want, (n, d.passes)
        if d.pass_count > 1:
            assert max(d.passes) <= 512
        # workspace of an in-place transform: one transform when multi-pass
        assert d.workspace_bytes == (0 if n <= 4096 else n * (8 if bits == 32 else 16))


def test_batch_rejections_without_device():
    L = mf.lib()
    d = mf.PlanDescription()
    for n in (0, 1, 3, 12, 1000):
        assert L.mofft_batch_describe_length(32, n, ctypes.byref(d)) == mf.UNSUPPORTED_SIZE
    assert L.mofft_batch_describe_length(16, 64, ctypes.byref(d)) == mf.INVALID_ARGUMENT
    p = ctypes.c_void_p(1)
    lengths = (ctypes.c_size_t * 2)(64, 128)
    dirs = (ctypes.c_int * 2)(-1, 1)
    ptrs = (ctypes.c_void_p * 2)(16, 32)
    create = L.mofft_batch_plan_create
    assert create(None, 32, 2, lengths, dirs, ptrs, ptrs, None) == mf.INVALID_ARGUMENT
    assert create(ctypes.byref(p), 32, 0, lengths, dirs, ptrs, ptrs, None) == mf.INVALID_ARGUMENT
    assert p.value is None
    assert create(ctypes.byref(p), 48, 2, lengths, dirs, ptrs, ptrs, None) == mf.INVALID_ARGUMENT
    bad_dirs = (ctypes.c_int * 2)(-1, 0)
    assert create(ctypes.byref(p), 32, 2, lengths, bad_dirs, ptrs, ptrs, None) == mf.INVALID_ARGUMENT
    odd = (ctypes.c_size_t * 2)(64, 96)
    assert create(ctypes.byref(p), 32, 2, odd, dirs, ptrs, ptrs, None) == mf.UNSUPPORTED_SIZE
    if not mf.device_available():
        assert create(ctypes.byref(p), 32, 2, lengths, dirs, ptrs, ptrs, None) == mf.UNSUPPORTED_HARDWARE
    assert L.mofft_batch_execute(None) == mf.INVALID_ARGUMENT
    assert L.mofft_batch_check(None) == mf.INVALID_ARGUMENT
    L.mofft_batch_plan_destroy(None)


def test_batch_kinds_are_256_thread_tiles():
    import json
    man = json.loads((ROOT / "mofft_amd" / "csrc" / "generated" / "manifest.json").read_text())
    kinds = man["batch_kinds"]
    assert len(kinds) == 42
    for k in kinds:
        assert k["threads_per_column"] * k["columns_per_block"] == 256
        assert math.prod(k["stages"]) == k["length"]
        assert k["lds_bytes"] <= 64 * 1024


def test_server_without_device_and_options():
    L = mf.lib()
    s = ctypes.c_void_p(1)
    o = mf.ServerOptions()
    L.mofft_server_options_init(ctypes.byref(o))
    assert o.budget_bytes == 1 << 30 and o.flags == 0
    assert L.mofft_server_create(None, ctypes.byref(o)) == mf.INVALID_ARGUMENT
    bad = mf.ServerOptions(16, 0, 0, 0, 0)
    assert L.mofft_server_create(ctypes.byref(s), ctypes.byref(bad)) == mf.INVALID_ARGUMENT
    assert s.value is None
    if not mf.device_available():
        assert L.mofft_server_create(ctypes.byref(s), ctypes.byref(o)) == mf.UNSUPPORTED_HARDWARE
    L.mofft_server_destroy(None)
    assert L.mofft_server_poll(None, None, 0, None, 0) == mf.INVALID_ARGUMENT


def test_device_generator_restatement_properties():
    """The host restatement of the seeded device generator: values in
    [-1, 1), deterministic per seed, independent across seeds, FP32 the
    rounding of FP64 (the GPU tests pin it bit-exactly to the device)."""
    import numpy as np
    from conftest import seeded_device_uniform
    a = seeded_device_uniform(1, 100000, 64)
    assert np.array_equal(a, seeded_device_uniform(1, 100000, 64))
    v = a.view(np.float64)
    assert v.min() >= -1.0 and v.max() < 1.0
    assert abs(v.mean()) < 0.01 and abs(v.var() - 1 / 3) < 0.01
    b = seeded_device_uniform(2, 100000, 64).view(np.float64)
    assert abs(np.corrcoef(v, b)[0, 1]) < 0.01
    f = seeded_device_uniform(1, 1000, 32).view(np.float32)
    assert np.array_equal(f, v[:2000].astype(np.float32))
    # first value by hand: splitmix64(1 + gamma)
    m = (1 << 64) - 1
    z = (1 + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    assert v[0] == 2.0 * ((z >> 11) * 2.0 ** -53) - 1.0


DIV_CHECK = r"""
#include <stdio.h>
#include <stdint.h>
#include "mofft_abi.h"
/* the device's div_by, compiled for the host */
static uint32_t div_by(uint32_t x, uint32_t m, uint32_t s) {
  if (s == 0) return x;
  uint32_t t = (uint32_t)(((uint64_t)x * m) >> 32);
  return (t + ((x - t) >> 1)) >> (s - 1);
}
static uint64_t rng = 0x9e3779b97f4a7c15ull;
static uint32_t next(void) { rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17; return (uint32_t)rng; }
int main(void) {
  uint32_t ds[6000]; int nd = 0;
  for (uint32_t d = 1; d <= 4200; ++d) ds[nd++] = d;
  for (int k = 12; k <= 30; ++k) { ds[nd++] = 1u << k; ds[nd++] = (1u << k) - 1; ds[nd++] = (1u << k) + 1; }
  const uint32_t odd[] = {371293u, 759375u, 1594323u, 2147483647u, 2147483629u, 1073741827u, 3u * 5u * 7u * 11u * 13u * 13u};
  for (unsigned i = 0; i < sizeof(odd) / sizeof(odd[0]); ++i) ds[nd++] = odd[i];
  while (nd < 6000) ds[nd++] = 1u + (next() >> (next() % 31));
  long bad = 0;
  for (int i = 0; i < nd; ++i) {
    const uint32_t d = ds[i];
    uint32_t m, s;
    mofft_div_magic(d, &m, &s);
    uint32_t xs[64]; int nx = 0;
    const uint32_t edge[] = {0u, 1u, d - 1u, d, d + 1u, 2u * d - 1u, 2u * d, 0x7fffffffu, 0x80000000u, 0xffffffffu, 0xfffffffeu};
    for (unsigned k = 0; k < sizeof(edge) / sizeof(edge[0]); ++k) xs[nx++] = edge[k];
    const uint32_t top = 0xffffffffu / d;
    xs[nx++] = top * d; xs[nx++] = top * d - 1u; xs[nx++] = top * d + (d - 1u);
    while (nx < 64) xs[nx++] = next();
    for (int k = 0; k < nx; ++k)
      if (div_by(xs[k], m, s) != xs[k] / d) { if (bad < 5) printf("d=%u x=%u\n", d, xs[k]); ++bad; }
    for (uint32_t x = 0; x < 20000u; ++x)
      if (div_by(x, m, s) != x / d) { if (bad < 5) printf("d=%u x=%u\n", d, x); ++bad; }
  }
  printf("bad %ld\n", bad);
  return bad != 0;
}
"""


def test_invariant_divisor_multipliers(tmp_path):
    """The multiply-high division of the pass kernels (mofft_div_magic and
    the device's div_by) equals integer division for every divisor class a
    launch uses (1, small, powers of two and their neighbours, the odd
    mixed-radix lengths, primes near 2^31) at edge and random dividends
    across the whole 32-bit range."""
    import subprocess
    src = tmp_path / "div.c"
    src.write_text(DIV_CHECK)
    exe = tmp_path / "div"
    csrc = Path(mf.__file__).parent / "csrc"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-I", str(csrc), str(src), "-o", str(exe)],
                   check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout
    assert res.stdout.strip() == "bad 0"


def test_launch_element_bound():
    """Element offsets inside one launch are 32-bit: lengths above 2^30 are
    rejected, and a batch is split so that one launch covers fewer than 2^30
    elements (single-pass plans included)."""
    with pytest.raises(mf.MofftError) as e:
        mf.describe((1 << 30) * 2, 1, 32)
    assert e.value.status == mf.UNSUPPORTED_SIZE
    d = mf.describe(64, 1 << 25, 32)              # 2^31 elements in all
    assert d.pass_count == 1 and d.chunk_transforms == (1 << 30) // 64
    d = mf.describe(1 << 12, 1 << 20, 64)
    assert d.chunk_transforms * (1 << 12) <= 1 << 30


def _check_output(out, ref, bits, bound):
    fn = ctypes.CDLL(mf.LIB_PATH).mofft_internal_check_output
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                   ctypes.c_double]
    return fn(out.ctypes.data, None if ref is None else ref.ctypes.data, out.size, bits, bound)


@pytest.mark.parametrize("bits", [32, 64])
def test_measure_result_check(bits):
    """MEASURE's result check (the reference rejects a candidate whose
    output is not finite, src/mofft.c:725-737, 990-993; here a candidate
    that disagrees with the analytical plan is also dropped), driven on host
    buffers: finite and equal passes, a NaN or an infinity anywhere is
    'not finite', a difference above the bound is 'disagree'."""
    import numpy as np
    dt = np.complex64 if bits == 32 else np.complex128
    rng = np.random.default_rng(1)
    ref = (rng.standard_normal(4096) + 1j * rng.standard_normal(4096)).astype(dt)
    bound = 4e-6 if bits == 32 else 4e-15
    assert _check_output(ref.copy(), ref, bits, bound) == 0
    assert _check_output(ref.copy(), None, bits, bound) == 0
    for bad in (np.nan, np.inf, -np.inf):
        for where in (0, 1234, 4095):
            y = ref.copy()
            y[where] = complex(bad, 0) if where % 2 else complex(0, bad)
            assert _check_output(y, ref, bits, bound) == 1
            assert _check_output(y, None, bits, bound) == 1
    y = ref.copy()
    y[7] += 1e-2                      # rel-L2 ~1e-4: far above the bound
    assert _check_output(y, ref, bits, bound) == 2
    y = ref.copy()
    y[7] *= 1 + (1e-7 if bits == 32 else 1e-16)
    assert _check_output(y, ref, bits, bound) == 0


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n,chain", [(1024, [4, 16, 16]), (1 << 16, [8, 8, 8, 8, 16]),
                                     (30, [2, 15]), (20736, [12, 12, 12, 12]),
                                     (64, [64]), (96, [32, 3])])
def test_cpu_port_against_oracle(bits, n, chain):
    """The timed CPU baseline (the reference executor with plan-time
    tables, oracle/mofft_cpu_port.c) computes the DFT: forward against the
    oracle and the round trip within the path's bars."""
    import numpy as np
    import oracle
    dt = np.complex64 if bits == 32 else np.complex128
    rng = np.random.default_rng(n)
    x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(dt)
    y = np.empty_like(x)
    z = np.empty_like(x)
    oracle.PortPlan(n, chain, -1, bits).execute(x, y)
    oracle.PortPlan(n, chain, 1, bits).execute(y, z)
    bar = (1e-6 if bits == 32 else 1e-15) * math.log2(n)
    assert oracle.rel_l2(y, oracle.dft(x.astype(np.complex128), -1)) <= bar
    assert oracle.rel_l2(z / n, x) <= bar
    with pytest.raises(ValueError):
        oracle.PortPlan(n * 17, chain + [17], -1, bits)


def test_cost_model_is_measured_and_ranks_tuned_plans():
    """The planner's cost model (mofft_estimated_schedule_seconds) is built
    from the compiled pass-cost table (every kernel timed by mofft_tune
    --pass-costs) and generated codelet op counts.  Against the tuned
    wisdom - each entry the fastest of up to 1000 timed plans - it ranks the
    tuned plan first or second among the one-, two- and three-pass
    candidates of its length for most entries, and predicts nine in ten
    tuned plans within 10 % of the best prediction and every one within
    30 % (the model is of single passes and misses their interplay, so
    exact ranks differ)."""
    import json
    L = mf.lib()
    assert L.mofft_estimated_pass_costs_measured() == 1
    # generated codelet counts: radix-16 = 144 adds + 24 multiplies
    assert L.mofft_estimated_radix_cost(16, 32, 1) < L.mofft_estimated_radix_cost(16, 32, 0)
    entries = json.loads((ROOT / "mofft_amd" / "wisdom" / "gfx950.json").read_text())["entries"]
    ks = mf.kernels()

    def kernels_of(bits, rows):
        out = {}
        for k in ks:
            if k.precision_bits == bits and k.rows == rows and k.variant == 0:
                out.setdefault(k.length, []).append(tuple(k.stages[:k.stage_count]))
        return out

    top2 = 0
    ratios = []
    for e in entries:
        n, bits = e["length"], 32 if e["precision"] == "fp32" else 64
        batch = max(1, ((1 << 24) if bits == 32 else (1 << 23)) // n)
        rows, cols = kernels_of(bits, 1), kernels_of(bits, 0)
        cands = {(tuple(reversed(st)), 0) for st in rows.get(n, [])}
        for a in cols:
            if n % a:
                continue
            if n // a in cols:
                for sa in cols[a]:
                    for sb in cols[n // a]:
                        cands.add((tuple(reversed(sb)) + tuple(reversed(sa)), 1 << (len(sb) - 1)))
            if n > 1 << 20:
                for b in cols:
                    c = n // a // b if (n // a) % b == 0 else 0
                    if min(a, b, c) >= 64 and c in cols:
                        for sa in cols[a][:2]:
                            for sb in cols[b][:2]:
                                for sc in cols[c][:2]:
                                    ch = tuple(reversed(sc)) + tuple(reversed(sb)) + tuple(reversed(sa))
                                    cands.add((ch, (1 << (len(sc) - 1)) | (1 << (len(sc) + len(sb) - 1))))
        costs = {mf.estimated_schedule_seconds(n, batch, bits, list(ch), m, bank)
                 for ch, m in cands for bank in (0, 1, 2, 4, 5)}
        costs.discard(float("inf"))
        tuned = mf.estimated_schedule_seconds(n, batch, bits, e["radices"],
                                              e["blocked_transpose_mask"], e["kernel_bank"])
        assert tuned < float("inf") and costs
        rank = 1 + sum(1 for c in costs if c < tuned * (1 - 1e-9))
        top2 += rank <= 2
        ratios.append(tuned / min(costs))
    assert top2 > len(entries) // 2, (top2, len(entries))
    ratios.sort()
    assert ratios[int(0.9 * len(ratios))] <= 1.1, ratios[-12:]
    assert ratios[-1] <= 1.3, ratios[-12:]
