"""The gfx950 pass-kernel registry: which LDS-resident kernels get compiled.

An entry is one instantiation of ``fft_pass_kernel`` (csrc/hip/fft_pass.h):

* ``length`` L: points per column kept in LDS;
* ``stages``: the radix chain of the L-point DFT, execution order (stage 0
  first, i.e. innermost first - the reverse of the reference's outer-to-inner
  radix lists, include/mofft.h:138-139);
* ``rows``: 1 when a column is a whole contiguous transform (single-pass
  plans), 0 for strided columns of a multi-pass plan;
* ``tpt`` threads per column and ``cols`` columns per workgroup;
* ``variant``: the tile-shape family (the plan's kernel bank).  Bank 0 is
  the default shape; bank 1 doubles and bank 2 halves the columns per
  workgroup, for power-of-two lengths >= 64 where the alternative fits;
  bank 4 quadruples the columns of short strided tiles and bank 5 gives
  long strided tiles more threads per column and fewer columns; bank 3 is
  the matrix-core form.

The reference's radix set {64, 32, 16, 15, ..., 2} (src/mofft.c:27-28) is
covered: every radix has a strided single-radix entry (radix 32 and 64 run as
two-stage L = 32 / 64 kernels), so any plan chain the reference accepts has a
schedule here.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass

from .codelets import RADICES

LDS_CAP = 150 * 1024         # bytes of dynamic LDS a registry entry may use
                             # (gfx950: 160 KiB per CU)
STAGE_RADICES = (16, 15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2)

# lengths the reference's own benchmark and tests use (tools/benchmark.cpp:28-41,
# tests/test_fft.c:490-491, tests/test_binary_banks.c:60-62) and their
# four-step factors
MIXED_ROWS = (30, 45, 121, 144, 169, 196, 225, 1728, 2197, 2744, 3375)
MIXED_COLS = (121, 144, 169, 196, 225, 1728, 2197, 2744, 3375)
# powers of the small odd and composite radices up to 4096 points, so that a
# length such as 3^10 or 10^5 gets a two-pass schedule (729 x 81, 1000 x 100)
# instead of one pass per radix stage
SMALL_POWERS = (25, 27, 36, 49, 81, 100, 125, 216, 243, 343, 625, 729,
                1000, 1296, 2187, 2401, 3125)
# 3 x 2^k and 5 x 2^k lengths above 4096 (the autotuner's wisdom covers
# them) as one HBM pass: FP32 up to 24576 points, FP64 up to 6144; radix-16
# led and radix-8 led chains, the tuner picks
LONG_SMOOTH_ROWS = {32: (5120, 6144, 10240, 12288, 20480, 24576),
                    64: (5120, 6144)}
SMOOTH_MAX = 1024          # single-pass and strided kernels
SMOOTH_ROWS_MAX = 4096     # single-pass kernels only (a pass of a longer
                           # transform is at most 1024 points here)


def smooth_lengths(limit: int = SMOOTH_MAX) -> tuple[int, ...]:
    """Every length up to `limit` whose prime factors are radices (2, 3, 5,
    7, 11, 13).  Those up to 1024 get a single-pass and a strided kernel,
    those up to 4096 a single-pass one: a smooth transform of up to 4096
    points is one HBM pass and one of up to about a million points two."""
    out = []
    for n in range(2, limit + 1):
        m = n
        for q in (2, 3, 5, 7, 11, 13):
            while m % q == 0:
                m //= q
        if m == 1:
            out.append(n)
    return tuple(out)
# Besides the radix-16 chains, every power-of-two length from 64 has kernels
# with radix-8 and with radix-4 sub-stages: half (a quarter)
# of the registers per butterfly - an FP64 radix-16 butterfly holds 64 VGPRs
# of data - for one more LDS exchange.  The planner times both; measured on
# MI355X the radix-8 chains win for most multi-pass and for FP32 N=2^10..2^12
# single-pass plans.
ALT_RADIX4 = ((64, (4, 4, 4)), (256, (4, 4, 4, 4)), (512, (8, 4, 4, 4)),
              (1024, (4, 4, 4, 4, 4)), (2048, (8, 4, 4, 4, 4)),
              (4096, (4, 4, 4, 4, 4, 4)),
              # four sub-stages without a radix-2 one for the long single
              # passes (the radix-16 chains end in 2 and 4)
              (8192, (16, 8, 8, 8)), (16384, (16, 16, 8, 8)))


@dataclass(frozen=True)
class Entry:
    precision: int
    length: int
    rows: int
    tpt: int
    cols: int
    stages: tuple[int, ...]
    variant: int = 0

    @property
    def complex_bytes(self) -> int:
        return 8 if self.precision == 32 else 16

    @property
    def stage_tw_entries(self) -> int:
        """Entries of the sub-stage twiddle table (sub-stages with ns > 1)."""
        ns, count = 1, 0
        for r in self.stages:
            if ns > 1:
                count += ns * (r - 1)
            ns *= r
        return count

    @property
    def lds_bytes(self) -> int:
        """Tile exchange area, plus the sub-stage twiddle table when it is at
        most a quarter of that and the sum stays within LDS_CAP (the rule of
        StageTw in csrc/hip/fft_pass.h)."""
        real = self.complex_bytes // 2
        # one pad element every 32 (FP32) / 16 (FP64): csrc/hip/fft_pass.h Tile
        lp = self.length + (self.length >> (5 if self.precision == 32 else 4)) + 1
        row_words = self.cols * lp
        col_words = self.length * self.cols + ((self.length * self.cols) >> 5) + 1
        words = row_words if self.rows else max(row_words, col_words)
        tile = words * real
        table = self.stage_tw_entries * self.complex_bytes
        offset = (tile + 15) // 16 * 16
        if table > 0 and 4 * table <= tile and offset + table <= LDS_CAP:
            return offset + table
        return tile

    @property
    def threads(self) -> int:
        return self.tpt * self.cols

    @property
    def symbol(self) -> str:
        kind = "rows" if self.rows else "cols"
        chain = "x".join(str(r) for r in self.stages)
        return (f"mofft_pass_f{self.precision}_{kind}_L{self.length}_{chain}"
                f"_t{self.tpt}_c{self.cols}_v{self.variant}")


def chain_for(length: int, cap: int = 16) -> tuple[int, ...]:
    """Fewest-stage ordered factorisation of `length` into radices <= cap.

    Ties prefer the most balanced chain (largest smallest radix), then
    larger radices first."""
    if length == 1:
        return ()
    best = None
    allowed = [r for r in STAGE_RADICES if r <= cap]

    def rec(n, acc):
        nonlocal best
        if n == 1:
            key = (len(acc), -min(acc), tuple(-r for r in acc))
            if best is None or key < best[0]:
                best = (key, tuple(acc))
            return
        if best is not None and len(acc) >= best[0][0]:
            return
        for r in allowed:
            if n % r == 0 and (not acc or r <= acc[-1]):
                rec(n // r, acc + [r])

    rec(length, [])
    if best is None:
        raise ValueError(f"{length} does not factor into radices <= {cap}")
    return best[1]


def _pow2_chain(length: int) -> tuple[int, ...]:
    if length <= 16:
        return (length,)
    if length == 32:
        return (8, 4)
    if length == 64:
        return (8, 8)
    chain = []
    n = length
    while n % 16 == 0 and n > 16:
        chain.append(16)
        n //= 16
    if n > 1:
        chain.append(n)
    return tuple(chain)


def _pow2_chain8(length: int) -> tuple[int, ...]:
    """Radix-8 chain of a power of two >= 64: 8s, then the remainder, with a
    trailing (8, 2) rebalanced to (4, 4)."""
    chain = []
    n = length
    while n % 8 == 0 and n > 8:
        chain.append(8)
        n //= 8
    chain.append(n)
    if chain[-1] == 2 and len(chain) > 1:
        chain[-2:] = [4, 4]
    return tuple(chain)


def _tpt(stages: tuple[int, ...], length: int) -> int:
    # one butterfly of the largest radix per thread, at most 1024 threads
    # per column (larger lengths give each thread several butterflies)
    return max(1, min(1024, length // max(stages)))


def _make(precision: int, length: int, rows: int,
          stages: tuple[int, ...]) -> Entry | None:
    tpt = _tpt(stages, length)
    real = 4 if precision == 32 else 8
    if rows:
        # 128-thread workgroups measured best for whole-transform kernels
        cols = max(1, 128 // tpt)
    else:
        # at least 128 B of adjacent columns per access, at least 256 threads;
        # long strided columns give each thread several butterflies rather
        # than narrowing the runs
        cmin = 16 if precision == 32 else 8
        while (tpt * cmin > 1024 and tpt > 1 and length & (length - 1) == 0):
            tpt //= 2
        cols = max(cmin, 256 // tpt)
        while cols * tpt > 1024 and cols > 1:
            cols //= 2
    e = Entry(precision, length, rows, tpt, cols, stages)
    while e.lds_bytes > LDS_CAP and e.cols > 1:
        e = Entry(precision, length, rows, tpt, e.cols // 2, stages)
    if e.lds_bytes > LDS_CAP or e.threads > 1024:
        return None
    del real
    return e


MFMA_VARIANT = 3
WIDE_VARIANT = 4
NARROW_VARIANT = 5


# Matrix-core chains with radix-32 sub-stages (bank 3 only: a VALU plan
# runs a radix-32 level as a two-sub-stage L = 32 kernel), whole-transform
# tiles: (length, chain, threads per column).  A radix-32 sub-stage is a
# 64 x 64 real product per 32 butterflies: 2^10 .. 2^14 points need one LDS
# exchange fewer than with radix-16 sub-stages.  Each chain has at least 32
# butterflies per column in every sub-stage, so its tile loads and stores
# HBM in the matrix layout.  (N = 2^15 as three radix-32 sub-stages holds
# 64 values per lane in 1024-thread tiles and spills; not built.)
MFMA32_ROWS = ((1024, (32, 32), 64), (2048, (32, 16, 4), 128),
               (4096, (32, 32, 4), 256), (8192, (32, 16, 16), 512),
               (16384, (32, 32, 16), 1024))


def mfma_prefix(stages: tuple[int, ...]) -> int:
    """Number of leading radix-16 / radix-32 sub-stages (the ones a bank-3
    kernel runs on the matrix cores: they lead its chain)."""
    n = 0
    while n < len(stages) and stages[n] in (16, 32):
        n += 1
    return n


def mfma_fits(e: Entry) -> bool:
    """Tile shape constraints of the matrix-core sub-stages: FP32, radix-16
    / radix-32 sub-stages only at the head of the chain, whole waves, whole
    32-butterfly groups, the same number of groups per wave."""
    k = mfma_prefix(e.stages)
    if e.precision != 32 or k == 0:
        return False
    if any(r in (16, 32) for r in e.stages[k:]):
        return False
    if e.threads % 64:
        return False
    waves = e.threads // 64
    for r in e.stages[:k]:
        nb = e.length // r
        groups = (e.cols * nb) // 32
        if (e.cols * nb) % 32 or groups % waves:
            return False
    return True


def mfma32_entries() -> list[Entry]:
    out = []
    for length, chain, tpt in MFMA32_ROWS:
        e = Entry(32, length, 1, tpt, max(1, 128 // tpt), chain, MFMA_VARIANT)
        assert mfma_fits(e) and e.lds_bytes <= LDS_CAP, e
        out.append(e)
    return out


def _variants(e: Entry) -> list[Entry]:
    """Bank 1 (twice the columns) and bank 2 (half the columns) of a
    power-of-two entry, where they fit the thread and LDS limits; bank 5
    (narrow strided tiles, below); bank 3:
    the same tile with its radix-16 sub-stages on the matrix cores (FP32
    split-precision form), where the tile shape allows it."""
    out = []
    if not e.rows and e.length >= 512 and e.length & (e.length - 1) == 0:
        # bank 5: long strided columns with one butterfly of the largest
        # radix per thread (at most 256 threads per column, at most 8
        # points per thread) in 1024-thread workgroups - runs of 4-8
        # adjacent columns instead of 8-16, so a thread holds half (a
        # quarter) of the points and two workgroups can share a CU; the
        # kernel deals the tiles of one cache line to one XCD
        # (tile_of_block in csrc/hip/fft_pass.h)
        tpt = min(e.length // max(e.stages), 256)
        n = Entry(e.precision, e.length, 0, tpt, 1024 // tpt, e.stages, NARROW_VARIANT)
        if (tpt > e.tpt and e.length // tpt <= 8 and n.cols >= 4
                and n.lds_bytes <= LDS_CAP):
            out.append(n)
    if not e.rows and 64 <= e.length <= 256 and e.length & (e.length - 1) == 0:
        # bank 4: four times the columns (512-byte+ runs per access)
        w = Entry(e.precision, e.length, e.rows, e.tpt, e.cols * 4, e.stages, WIDE_VARIANT)
        if w.threads <= 1024 and w.lds_bytes <= LDS_CAP:
            out.append(w)
    m = Entry(e.precision, e.length, e.rows, e.tpt, e.cols, e.stages, MFMA_VARIANT)
    if e.length & (e.length - 1) == 0 and mfma_fits(m):
        out.append(m)
    if e.length < 64 or e.length & (e.length - 1):
        return out
    for variant, cols in ((1, e.cols * 2), (2, e.cols // 2)):
        if cols < 1 or (not e.rows and cols < 4):
            continue
        v = Entry(e.precision, e.length, e.rows, e.tpt, cols, e.stages, variant)
        if v.threads <= 1024 and v.lds_bytes <= LDS_CAP:
            out.append(v)
    return out


def entries() -> list[Entry]:
    out: list[Entry] = []
    seen = set()

    def add(e):
        if e is not None and (e.precision, e.length, e.rows, e.stages) not in seen:
            seen.add((e.precision, e.length, e.rows, e.stages))
            out.append(e)
            out.extend(_variants(e))

    for precision, rows_max, cols_max in ((32, 32768, 4096), (64, 16384, 4096)):
        # single-pass kernels over whole transforms
        for k in range(1, rows_max.bit_length()):
            length = 1 << k
            add(_make(precision, length, 1, _pow2_chain(length)))
        for r in STAGE_RADICES:
            add(_make(precision, r, 1, (r,)))
        for length in MIXED_ROWS + SMALL_POWERS + smooth_lengths(SMOOTH_ROWS_MAX):
            add(_make(precision, length, 1, chain_for(length)))
        for length in LONG_SMOOTH_ROWS[precision]:
            add(_make(precision, length, 1, chain_for(length)))
            add(_make(precision, length, 1, chain_for(length, cap=8)))
        for k in range(7, rows_max.bit_length()):
            add(_make(precision, 1 << k, 1, _pow2_chain8(1 << k)))
        for length, chain in ALT_RADIX4:
            add(_make(precision, length, 1, chain))
        # 32 points as 4 x 4 x 2 with eight threads per transform and 32
        # transforms per workgroup (FP32 2^24 points: 38.7 us against 41.0
        # for the 8 x 4 tile, four threads per transform)
        add(Entry(precision, 32, 1, 8, 32, (4, 4, 2)))
        # 128 points as 4 x 4 x 4 x 2, 32 threads per transform (38.8 us
        # against 39.3 for the 8 x 4 x 4 tile)
        add(Entry(precision, 128, 1, 32, 8, (4, 4, 4, 2)))
        # the persistent batch's chains, so a single-transform plan can
        # reproduce a batch transform exactly
        for k in range(1, BATCH_ROWS_MAX.bit_length()):
            add(_make(precision, 1 << k, 1, BATCH_CHAIN(precision, 1 << k)))
        # strided kernels for multi-pass schedules
        for k in range(1, cols_max.bit_length()):
            length = 1 << k
            add(_make(precision, length, 0, _pow2_chain(length)))
        for r in STAGE_RADICES:
            add(_make(precision, r, 0, (r,)))
        for length in MIXED_COLS + SMALL_POWERS + smooth_lengths():
            add(_make(precision, length, 0, chain_for(length)))
        for k in range(7, cols_max.bit_length()):
            add(_make(precision, 1 << k, 0, _pow2_chain8(1 << k)))
        for length, chain in ALT_RADIX4:
            # strided passes stop at cols_max: a longer strided tile gives a
            # thread 16+ butterflies of registers (the 16384-point ones took
            # over ten minutes each to compile) and no plan uses one
            if length <= cols_max:
                add(_make(precision, length, 0, chain))
        for k in range(1, BATCH_COLS_MAX.bit_length()):
            add(_make(precision, 1 << k, 0, BATCH_CHAIN(precision, 1 << k)))
    out.extend(mfma32_entries())
    return out


# Persistent batch kernel (csrc/hip/batch.h): every tile kind runs with the
# same 256-thread workgroup.  Whole-transform (row) kinds cover power-of-two
# lengths up to 4096, strided kinds up to 512 (so every power of two up to
# 2^18 is one or two passes; longer ones take three).
BATCH_THREADS = 256
BATCH_ROWS_MAX = 4096
BATCH_COLS_MAX = 512


def _batch_chain(precision: int, length: int) -> tuple[int, ...]:
    """FP32: the radix-16 chains (measured on 4096 transforms of N=2^6..2^18:
    2.46 ms, against 2.57 ms with radix-8 chains - the kernel's registers
    stay bound by its strided kinds, so the extra exchange is not repaid).
    FP64: radix-8 chains from 64 points, (8, 4) and (4, 4) below, which take
    the kernel from 187 to 141 VGPRs (radix-16 FP64 butterflies hold 64)."""
    if precision == 32 or length < 16:
        return _pow2_chain(length)
    if length >= 64:
        return _pow2_chain8(length)
    return {32: (8, 4), 16: (4, 4)}[length]


BATCH_CHAIN = _batch_chain


def _pow2_floor(x: int) -> int:
    return 1 << (max(1, x).bit_length() - 1)


def batch_entries() -> list[Entry]:
    """Tile kinds of the persistent batch kernel: the bank-0 chains of the
    power-of-two entries, with threads per column a power of two and
    columns * threads = 256.  Strided kinds keep at least 16 (FP32) / 8
    (FP64) adjacent columns, i.e. 128-byte runs per access, unless that
    leaves a thread more than 16 (FP32) / 8 (FP64) points of a sub-stage
    (register pressure of the one kernel is that of its largest kind)."""
    out: list[Entry] = []

    def points_per_thread(length, stages, tpt):
        return max(-(-(length // r) // tpt) * r for r in stages)

    for precision in (32, 64):
        cmin = 16 if precision == 32 else 8
        ppt_max = 16 if precision == 32 else 8
        for rows, lmax in ((1, BATCH_ROWS_MAX), (0, BATCH_COLS_MAX)):
            for k in range(1, lmax.bit_length()):
                length = 1 << k
                stages = BATCH_CHAIN(precision, length)
                tpt = min(BATCH_THREADS, _pow2_floor(length // max(stages)))
                if not rows:
                    tpt = min(tpt, BATCH_THREADS // cmin)
                    while (points_per_thread(length, stages, tpt) > ppt_max and
                           2 * tpt <= BATCH_THREADS and
                           (BATCH_THREADS // (2 * tpt)) * length % BATCH_THREADS == 0):
                        tpt *= 2
                e = Entry(precision, length, rows, tpt, BATCH_THREADS // tpt, stages)
                assert e.threads == BATCH_THREADS and e.lds_bytes <= LDS_CAP, e
                out.append(e)
    return out


def check(es: list[Entry]) -> None:
    for e in es:
        prod = 1
        for r in e.stages:
            assert r in RADICES and (r <= 16 or (r == 32 and e.variant == MFMA_VARIANT)), e
            prod *= r
        if e.variant == MFMA_VARIANT:
            assert mfma_fits(e), e
        assert prod == e.length, e
        assert e.threads <= 1024 and e.lds_bytes <= LDS_CAP, e
        # the cooperative tile store gives every thread the same number of
        # tile elements
        assert e.rows or (e.cols * e.length) % e.threads == 0, e
    # every reference radix has a strided single-radix (or L = radix) entry
    for precision in (32, 64):
        for r in RADICES:
            assert any(e.precision == precision and e.length == r and not e.rows
                       for e in es), (precision, r)


def describe(es: list[Entry]) -> str:
    lines = []
    for i, e in enumerate(es):
        lines.append(f"{i:3d} f{e.precision} {'rows' if e.rows else 'cols'} "
                     f"L={e.length:5d} stages={list(e.stages)} tpt={e.tpt} "
                     f"cols={e.cols} threads={e.threads} lds={e.lds_bytes} "
                     f"bank={e.variant}")
    return "\n".join(lines)


__all__ = ["Entry", "entries", "batch_entries", "check", "describe", "chain_for"]
_ = itertools  # kept for callers that enumerate chains
