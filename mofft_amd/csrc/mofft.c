/*
 * mofft.c - host runtime of the MI355X MoFFT library (plain C).
 *
 * Implements every entry point of include/mofft.h (the reference's drop-in
 * ABI, include/mofft.h:107-231) and the additive ones of mofft_hip.h.  All
 * device work goes through the thin C ABI of mofft_abi.h.
 *
 * The reference executes a plan as a sequence of radix stages with a
 * materialising transpose between stages (src/mofft.c:1932-2160).  Here the
 * same radix chain is grouped into HBM passes: each pass is one launch of an
 * LDS-resident kernel that runs several Stockham radix stages on-chip, so an
 * N <= 8192 (FP32) transform costs one read and one write of HBM and larger
 * ones a few (a "four-step" scheme generalised to any number of passes).
 *
 *   chain (outer -> inner)      e.g. {4, 16, 16} for N = 1024
 *   execution order             innermost radix first (src/mofft.c:1936)
 *   pass boundary mask          bit l: data round-trips HBM before level l
 *                               (the GPU meaning of the reference's
 *                               blocked-transpose mask, include/mofft.h:152)
 */
#include "mofft_internal.h"
#include "mofft_wisdom.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifndef MOFFT_BINARY_ID
#define MOFFT_BINARY_ID "0000000000000000000000000000000000000000000000000000000000000000"
#endif

#define MOFFT_DEFAULT_SEARCH_CANDIDATES 16
#define MOFFT_DEFAULT_TIMING_SAMPLES 5
#define MOFFT_MAX_SEARCH_CANDIDATES 64
#define MOFFT_MAX_TIMING_SAMPLES 31
#define MOFFT_RECIPE_MAX_BYTES 1024
/* Device family token of recipes: the gfx target number in hex digits. */
#define MOFFT_GPU_FAMILY 0x00000950u
/* Multi-pass schedules stream the batch in chunks of at most this many bytes
 * (the workspace of an in-place plan is one chunk).  128 MiB keeps a chunk
 * and the intermediate its first pass writes inside the 256 MB Infinity
 * Cache, so the second pass reads it from there; smaller chunks lose more
 * to launch ramp-up than they gain.  Measured on MI355X, two-pass plans:
 * FP64 N=2^16 x 256 fwd+bwd 5.57 TFLOP/s at 128 MiB vs 5.22 at 256, 5.29 at
 * 64, 5.06 at 160; FP32 N=2^16..2^20 over 256 MiB 9-11% faster than at
 * 256 MiB.  MOFFT_CHUNK_BYTES in the environment overrides it (a tuning
 * knob; read once per process). */
#define MOFFT_DEFAULT_CHUNK_BYTES ((size_t)128 << 20)
/* Kernel bank of the matrix-core (split FP16) radix-16 kernels, and of
 * four-times-wider strided tiles. */
#define MOFFT_MFMA_BANK 3u
#define MOFFT_WIDE_BANK 4u
#define MOFFT_NARROW_BANK 5u

static size_t chunk_bytes(void) {
  static size_t cached = 0;
  if (cached == 0) {
    size_t v = MOFFT_DEFAULT_CHUNK_BYTES;
    const char *env = getenv("MOFFT_CHUNK_BYTES");
    if (env && *env) {
      char *end = NULL;
      const unsigned long long x = strtoull(env, &end, 10);
      if (end && *end == '\0' && x >= ((size_t)1 << 20)) v = (size_t)x;
    }
    cached = v;
  }
  return cached;
}

/* Largest full pass-twiddle table (bytes) of a precision;
 * MOFFT_PASS_TABLE_BYTES in the environment overrides both (a tuning knob;
 * 0 = always the two-level table; read once per process). */
static size_t pass_table_bytes(int precision_bits) {
  static int cached = 0;
  static int overridden = 0;
  static size_t v = 0;
  if (!cached) {
    const char *env = getenv("MOFFT_PASS_TABLE_BYTES");
    if (env && *env) {
      char *end = NULL;
      const unsigned long long x = strtoull(env, &end, 10);
      if (end && *end == '\0') {
        v = (size_t)x;
        overridden = 1;
      }
    }
    cached = 1;
  }
  if (overridden) return v;
  return precision_bits == 64 ? MOFFT_PASS_TABLE_BYTES_F64 : MOFFT_PASS_TABLE_BYTES;
}

/* XCD-aware tile order of narrow strided tiles (MOFFT_PASS_XCD_ORDER);
 * MOFFT_TILE_ORDER=linear in the environment turns it off (a tuning knob
 * for comparing the two orders; read once per process).  The order never
 * changes results. */
static int tile_order_override = -1;

static int xcd_tile_order(void) {
  static int cached = -1;
  if (tile_order_override >= 0) return tile_order_override;
  if (cached < 0) {
    const char *env = getenv("MOFFT_TILE_ORDER");
    cached = !(env && strcmp(env, "linear") == 0);
  }
  return cached;
}

void mofft_internal_set_tile_order(int xcd) {
  tile_order_override = xcd < 0 ? -1 : xcd != 0;
}

/* src/mofft.c:27-28 - the radix set the API accepts. */
static const int k_radices[] = {64, 32, 16, 15, 14, 13, 12, 11, 10,
                                9, 8, 7, 6, 5, 4, 3, 2};

/* Device tables of a plan: twiddles and the in-place workspace.  Shared,
 * reference-counted, with the HIP graphs captured from the plan: a setter
 * that regroups the passes gives the plan new tables, and a graph captured
 * before keeps replaying on the old ones until it is destroyed. */
typedef struct {
  void *twiddles;
  void *workspace;
  unsigned int refs;
} mofft_device_tables;

static void tables_drop(mofft_device_tables *t) {
  if (!t || --t->refs > 0) return;
  mofft_hip_free(t->twiddles);
  mofft_hip_free(t->workspace);
  free(t);
}

typedef struct mofft_plan_common {
  mofft_schedule s;
  int direction;
  mofft_placement placement;
  int direct_input;
  mofft_stage_layout stage_layout;
  unsigned int kernel_bank;
  mofft_device_tables *tables;  /* owner of the next two */
  void *twiddles;
  void *workspace;
  void *staging_in;
  void *staging_out;
  const void *bound_input;
  void *bound_output;
  int has_bound_arrays;
} mofft_plan_common;

struct mofft_plan_f32 { mofft_plan_common c; };
struct mofft_plan_f64 { mofft_plan_common c; };

/* ------------------------------------------------------------------ */
/* small helpers                                                        */
/* ------------------------------------------------------------------ */

static size_t sat_mul(size_t a, size_t b) {
  return a != 0 && b > SIZE_MAX / a ? SIZE_MAX : a * b;
}

static size_t sat_add(size_t a, size_t b) {
  return b > SIZE_MAX - a ? SIZE_MAX : a + b;
}

static double monotonic_seconds(void) {
  struct timespec now;
  if (clock_gettime(CLOCK_MONOTONIC, &now) != 0) return -1.0;
  return (double)now.tv_sec + (double)now.tv_nsec * 1e-9;
}

static int is_supported_radix(int radix) {
  for (size_t i = 0; i < sizeof(k_radices) / sizeof(k_radices[0]); ++i)
    if (k_radices[i] == radix) return 1;
  return 0;
}

static size_t complex_bytes(int precision_bits) {
  return precision_bits == 32 ? 8 : 16;
}

/* Kernel banks are the tile-shape families of the registry (variant field):
 * bank b runs every pass with its variant-b kernel where one is compiled,
 * else with variant 0. */
static unsigned int kernel_bank_count(void) {
  static unsigned int cached = 0;
  if (cached == 0) {
    int v = 0;
    for (int i = 0; i < mofft_registry_count; ++i)
      if (mofft_registry[i].variant > v) v = mofft_registry[i].variant;
    cached = (unsigned int)v + 1;
  }
  return cached;
}

int mofft_device_available(void) { return mofft_hip_device_available(); }

/* ------------------------------------------------------------------ */
/* kernel registry lookups                                              */
/* ------------------------------------------------------------------ */

const mofft_kernel_set mofft_pass_kernel_set = {mofft_registry,
                                                &mofft_registry_count};
const mofft_kernel_set mofft_batch_kernel_set = {mofft_batch_registry,
                                                 &mofft_batch_registry_count};
#define PASS_KERNELS (&mofft_pass_kernel_set)

static int find_kernel_variant(const mofft_kernel_set *ks, int precision_bits,
                               int rows, const int *stages, int count,
                               int variant) {
  for (int i = 0; i < *ks->count; ++i) {
    const mofft_registry_entry *e = &ks->entries[i];
    if (e->precision_bits != precision_bits || e->rows != rows ||
        e->variant != variant)
      continue;
    if (e->stage_count == count &&
        memcmp(e->stages, stages, (size_t)count * sizeof(int)) == 0)
      return i;
  }
  /* radix 32 / 64 levels run as the two-stage L = 32 / 64 kernels */
  if (count == 1 && (stages[0] == 32 || stages[0] == 64)) {
    for (int i = 0; i < *ks->count; ++i) {
      const mofft_registry_entry *e = &ks->entries[i];
      if (e->precision_bits == precision_bits && e->rows == rows &&
          e->variant == variant && e->length == stages[0])
        return i;
    }
  }
  return -1;
}

/* Registry entry whose sub-stage chain equals stages[0..count) (execution
 * order), or, for a single radix, any entry of that length; the bank's
 * variant when compiled, else variant 0. */
static int find_kernel(const mofft_kernel_set *ks, int precision_bits,
                       int rows, const int *stages, int count,
                       unsigned int bank) {
  int k = -1;
  if (bank != 0)
    k = find_kernel_variant(ks, precision_bits, rows, stages, count, (int)bank);
  if (k < 0) k = find_kernel_variant(ks, precision_bits, rows, stages, count, 0);
  return k;
}

static int find_kernel_by_length(const mofft_kernel_set *ks, int precision_bits,
                                 int rows, size_t length) {
  for (int i = 0; i < *ks->count; ++i) {
    const mofft_registry_entry *e = &ks->entries[i];
    if (e->precision_bits == precision_bits && e->rows == rows &&
        e->variant == 0 && (size_t)e->length == length)
      return i;
  }
  return -1;
}

int mofft_kernel_count(void) { return mofft_registry_count; }

mofft_status mofft_kernel_get_info(int index, mofft_kernel_info *info) {
  if (!info || index < 0 || index >= mofft_registry_count)
    return MOFFT_INVALID_ARGUMENT;
  const mofft_registry_entry *e = &mofft_registry[index];
  memset(info, 0, sizeof(*info));
  info->precision_bits = e->precision_bits;
  info->length = e->length;
  info->rows = e->rows;
  info->threads_per_column = e->threads_per_column;
  info->columns_per_block = e->columns_per_block;
  info->stage_count = e->stage_count;
  for (int i = 0; i < e->stage_count && i < 8; ++i) info->stages[i] = e->stages[i];
  info->variant = e->variant;
  return MOFFT_SUCCESS;
}

/* ------------------------------------------------------------------ */
/* schedule: chain + pass mask -> passes                                */
/* ------------------------------------------------------------------ */

static mofft_status validate_chain(size_t length, const int *radices,
                                   size_t count) {
  if (!radices || count == 0 || count > MOFFT_MAX_PLAN_DEPTH || length < 2)
    return MOFFT_INVALID_ARGUMENT;
  size_t remaining = length;
  for (size_t i = 0; i < count; ++i) {
    if (!is_supported_radix(radices[i]) || remaining % (size_t)radices[i])
      return MOFFT_INVALID_ARGUMENT;
    remaining /= (size_t)radices[i];
  }
  return remaining == 1 ? MOFFT_SUCCESS : MOFFT_INVALID_ARGUMENT;
}

/* Group the chain into passes.  Level boundaries named in `mask` always
 * split; elsewhere each pass takes the longest run of stages that a compiled
 * kernel implements (greedy from the innermost stage). */
static mofft_status build_schedule_layout(const mofft_kernel_set *ks,
                                          size_t length, size_t batch,
                                          int precision_bits,
                                          const int *radices, size_t depth,
                                          uint64_t mask, unsigned int bank,
                                          int factored, mofft_schedule *s) {
  mofft_status st = validate_chain(length, radices, depth);
  if (st != MOFFT_SUCCESS) return st;
  if (batch == 0 || (precision_bits != 32 && precision_bits != 64))
    return MOFFT_INVALID_ARGUMENT;
  if (depth < 64 && depth > 0 && (mask >> (depth - 1)) != 0)
    return MOFFT_INVALID_ARGUMENT;
  if (bank >= kernel_bank_count()) return MOFFT_INVALID_ARGUMENT;
  /* element offsets inside a launch are 32-bit */
  if (length > (size_t)MOFFT_PASS_MAX_ELEMENTS / 2) return MOFFT_UNSUPPORTED_SIZE;
  memset(s, 0, sizeof(*s));
  s->ks = ks;
  s->length = length;
  s->batch = batch;
  s->precision_bits = precision_bits;
  s->bank = bank;
  s->factored_twiddles = factored;
  s->depth = (int)depth;
  memcpy(s->factors, radices, depth * sizeof(int));
  int exec[MOFFT_MAX_PLAN_DEPTH];
  for (size_t i = 0; i < depth; ++i) exec[i] = radices[depth - 1 - i];
  size_t start = 0, span = 1;
  uint64_t used_mask = 0;
  while (start < depth) {
    /* forced boundary: execution index i = depth-1-l starts level l */
    size_t limit = depth - start;
    for (size_t len = 1; len < depth - start; ++len) {
      const size_t level = depth - 1 - (start + len);
      if ((mask >> level) & 1u) { limit = len; break; }
    }
    int kernel = -1;
    size_t take = 0;
    for (size_t len = limit; len >= 1; --len) {
      const int rows = (start == 0 && len == depth) ? 1 : 0;
      if (len > 8) continue;
      /* the wide bank widens only the passes after the first: their
       * stores are column runs, the first pass leaves through LDS as
       * whole blocks and prefers the occupancy of the narrow tile
       * (measured, FP64 N = 2^16: 130 vs 171 us for pass 0, 139 vs
       * 161 us for pass 1) */
      const unsigned int b =
          (bank == MOFFT_WIDE_BANK && s->pass_count == 0) ? 0u : bank;
      kernel = find_kernel(ks, precision_bits, rows, exec + start, (int)len, b);
      if (kernel >= 0) { take = len; break; }
    }
    if (kernel < 0) return MOFFT_UNSUPPORTED_SIZE;
    mofft_pass *p = &s->passes[s->pass_count++];
    size_t pl = 1;
    for (size_t i = 0; i < take; ++i) pl *= (size_t)exec[start + i];
    p->kernel = kernel;
    p->length = pl;
    p->stride = length / pl;
    p->span = span;
    span *= pl;
    start += take;
    if (start < depth) used_mask |= (uint64_t)1 << (depth - 1 - start);
  }
  s->pass_mask = used_mask;
  /* twiddle layout: pass twiddles (two-level table of w_N), then the
   * sub-stage tables of every pass */
  size_t count = 0;
  if (s->pass_count > 1) {
    uint32_t bits = 0;
    while (((size_t)1 << bits) < length) ++bits;
    s->split_shift = (bits + 1) / 2;
    s->tw_lo_count = (size_t)1 << s->split_shift;
    if (s->tw_lo_count > length) s->tw_lo_count = length;
    s->tw_hi_count = (length + ((size_t)1 << s->split_shift) - 1) >> s->split_shift;
    s->tw_hi_offset = 0;
    s->tw_lo_offset = s->tw_hi_count;
    count = s->tw_hi_count + s->tw_lo_count;
  }
  const size_t cb0 = complex_bytes(precision_bits);
  for (int i = 1; i < s->pass_count; ++i) {
    mofft_pass *p = &s->passes[i];
    const size_t entries = p->span * p->length;
    if (!factored && sat_mul(entries, cb0) <= pass_table_bytes(precision_bits)) {
      p->pass_table = 1;
      p->pass_offset = count;
      count += entries;
    }
  }
  for (int i = 0; i < s->pass_count; ++i) {
    mofft_pass *p = &s->passes[i];
    const mofft_registry_entry *e = &ks->entries[p->kernel];
    p->stage_offset = count;
    size_t ns = 1, n_entries = 0;
    for (int k = 0; k < e->stage_count; ++k) {
      if (ns > 1) n_entries += ns * (size_t)(e->stages[k] - 1);
      ns *= (size_t)e->stages[k];
    }
    p->stage_count = n_entries;
    count += n_entries;
  }
  s->twiddle_count = count;
  const size_t cb = complex_bytes(precision_bits);
  const size_t transform_bytes = sat_mul(length, cb);
  const size_t cbytes = chunk_bytes();
  size_t chunk = transform_bytes >= cbytes ? 1 : cbytes / transform_bytes;
  if (chunk > batch) chunk = batch;
  if (chunk == 0) chunk = 1;
  s->chunk = s->pass_count > 1 ? chunk : batch;
  /* ... so one launch covers fewer than MOFFT_PASS_MAX_ELEMENTS / 2 */
  const size_t max_chunk = (size_t)(MOFFT_PASS_MAX_ELEMENTS / 2) / length;
  if (s->chunk > max_chunk) s->chunk = max_chunk;
  s->workspace_bytes = s->pass_count > 1 ? sat_mul(s->chunk, transform_bytes) : 0;
  return MOFFT_SUCCESS;
}

/* The default twiddle layout: full pass tables up to pass_table_bytes. */
static mofft_status build_schedule(const mofft_kernel_set *ks, size_t length,
                                   size_t batch, int precision_bits,
                                   const int *radices, size_t depth,
                                   uint64_t mask, unsigned int bank,
                                   mofft_schedule *s) {
  return build_schedule_layout(ks, length, batch, precision_bits, radices,
                               depth, mask, bank, 0, s);
}

/* Workspace is needed by in-place multi-pass plans and by out-of-place
 * plans with three or more passes. */
static size_t schedule_workspace(const mofft_schedule *s,
                                 mofft_placement placement) {
  if (s->pass_count <= 1) return 0;
  if (placement == MOFFT_OUT_OF_PLACE && s->pass_count == 2) return 0;
  return s->workspace_bytes;
}

/* ------------------------------------------------------------------ */
/* analytical planner (PRETUNED without wisdom)                          */
/* ------------------------------------------------------------------ */

typedef struct {
  int count;
  long long weight; /* sum of pass weights (pass_weight tenths, or measured
                       ns per byte in millionths) */
  size_t lengths[MOFFT_MAX_PLAN_DEPTH];
  int kernels[MOFFT_MAX_PLAN_DEPTH];
} pass_split;

/* Cost of a strided pass in units of one read + write of the batch at the
 * speed of the short-column kernels, from MI355X pass timings: L <= 512
 * runs at 5.0-5.8 TB/s (1); L = 1024 at 3.6-4.0 TB/s over 128 MiB (one
 * 1024-thread workgroup per CU) but close to the others on a single
 * multi-GiB transform (2^30: three 1024-point passes 4 % faster than four
 * shorter ones), hence 1.3; L = 2048 three times slower than L = 256 (3);
 * runs narrower than 128 bytes (C adjacent columns) cost at least 2.5
 * (FP32 3 x 2^20 as 4096 (64-byte runs) x 768: 305 us, as three passes of
 * wide runs 132 us). */
static int pass_weight(const mofft_registry_entry *e) {
  const int cb = e->precision_bits == 32 ? 8 : 16;
  int w = e->length <= 512 ? 10 : e->length <= 1024 ? 13 : 30;
  if (e->columns_per_block * cb < 128 && w < 25) w = 25;
  return w; /* tenths: integer sums compare exactly */
}

static double codelet_flops(int radix);

/* Measured pass cost of registry entry `kernel` in a role (MOFFT_ROLE_*)
 * at a stride (first pass) or span (later pass) `span`, in ns per byte
 * moved, from the pass-cost table `mofft_tune --pass-costs` writes (each
 * kernel timed as the pass it is in plans, 2^24 FP32 / 2^23 FP64 points,
 * strided ones at spans 16, 64, 256, 1024): linear in log2 span between
 * the measured spans, the nearest one outside them; 0 when the table has
 * no entry for it. */
static double measured_pass_ns(int kernel, int role, size_t span) {
  if (kernel < 0 || kernel >= mofft_registry_count) return 0.0;
  const double *row = mofft_pass_cost_ns[kernel];
  if (role == MOFFT_ROLE_ROWS) return row[0];
  const double *v = row + (role == MOFFT_ROLE_FIRST ? 1 : 5);
  const double x = log2((double)(span < 1 ? 1 : span));
  static const double at[4] = {4.0, 6.0, 8.0, 10.0};   /* log2 of 16 .. 1024 */
  double lo_x = 0, lo_v = 0, hi_x = 0, hi_v = 0;
  int have_lo = 0, have_hi = 0;
  for (int q = 0; q < 4; ++q) {
    if (v[q] <= 0) continue;
    if (at[q] <= x && (!have_lo || at[q] > lo_x)) { lo_x = at[q]; lo_v = v[q]; have_lo = 1; }
    if (at[q] >= x && (!have_hi || at[q] < hi_x)) { hi_x = at[q]; hi_v = v[q]; have_hi = 1; }
  }
  if (have_lo && have_hi)
    return hi_x > lo_x ? lo_v + (hi_v - lo_v) * (x - lo_x) / (hi_x - lo_x) : lo_v;
  if (have_lo) return lo_v;
  if (have_hi) return hi_v;
  return 0.0;
}

/* Nominal cost of a pass the table does not cover: the measured pass
 * weights below at 5.5 TB/s (ns per byte). */
static double nominal_pass_ns(const mofft_registry_entry *e) {
  return (e->rows ? 10.0 : (double)pass_weight(e)) / 10.0 / 5500.0;
}

static double pass_ns_per_byte(int kernel, int role, size_t span) {
  const double m = measured_pass_ns(kernel, role, span);
  return m > 0 ? m : nominal_pass_ns(&mofft_registry[kernel]);
}

/* Role of pass i of a schedule over the pass kernels, and its stride
 * (first pass) or span (later passes). */
static int pass_role(const mofft_schedule *s, int i, size_t *span) {
  *span = i == 0 ? s->passes[i].stride : s->passes[i].span;
  if (mofft_registry[s->passes[i].kernel].rows) return MOFFT_ROLE_ROWS;
  return i == 0 ? MOFFT_ROLE_FIRST : MOFFT_ROLE_LATER;
}

/* Predicted device seconds of a schedule over the pass kernels: every pass
 * moves the batch once (read + write) at its kernel's measured cost in its
 * role, plus the codelet arithmetic of its stages at the VALU peak (a small
 * term; it separates plans whose passes cost the same). */
static double schedule_seconds(const mofft_schedule *s) {
  const double bytes = 2.0 * (double)s->length * (double)s->batch *
                       (double)complex_bytes(s->precision_bits);
  double ns = 0.0;
  for (int i = 0; i < s->pass_count; ++i) {
    size_t span = 1;
    const int role = pass_role(s, i, &span);
    ns += bytes * pass_ns_per_byte(s->passes[i].kernel, role, span);
  }
  double flops = 0.0;
  for (int l = 0; l < s->depth; ++l) {
    const double f = codelet_flops(s->factors[l]);
    if (f > 0) flops += f * (double)(s->length / (size_t)s->factors[l]);
  }
  const double peak = s->precision_bits == 32 ? 157.3e3 : 78.6e3; /* flop/ns */
  ns += flops * (double)s->batch / peak;
  return ns * 1e-9;
}

/* Integer weight of a pass for the split search: measured ns per byte in
 * millionths (sums compare exactly). */
static long long pass_weight_measured(int kernel, int role, size_t span) {
  return llround(pass_ns_per_byte(kernel, role, span) * 1e6);
}

/* better = lower HBM cost, then fewer passes, then larger smallest pass,
 * then larger first pass */
static int split_better(const pass_split *a, const pass_split *b) {
  if (b->count == 0) return 1;
  if (a->weight != b->weight) return a->weight < b->weight;
  if (a->count != b->count) return a->count < b->count;
  size_t amin = SIZE_MAX, bmin = SIZE_MAX;
  for (int i = 0; i < a->count; ++i) if (a->lengths[i] < amin) amin = a->lengths[i];
  for (int i = 0; i < b->count; ++i) if (b->lengths[i] < bmin) bmin = b->lengths[i];
  if (amin != bmin) return amin > bmin;
  for (int i = 0; i < a->count; ++i)
    if (a->lengths[i] != b->lengths[i]) return a->lengths[i] > b->lengths[i];
  return 0;
}

/* 1 when the compiled pass-cost table holds measurements and
 * MOFFT_PASS_COSTS=0 in the environment does not switch it off (a tuning
 * knob for comparing against the nominal weights; read once per process) */
static int pass_costs_measured(void) {
  static int cached = -1;
  if (cached < 0) {
    const char *env = getenv("MOFFT_PASS_COSTS");
    cached = strcmp(mofft_pass_cost_source, "none") != 0 &&
             !(env && strcmp(env, "0") == 0);
  }
  return cached;
}

/* Weight of pass-kernel entry i at position `count` of a split, at
 * stride (first pass) or span (later) `span`: the measured cost when the
 * table is present, else the nominal tenths. */
static long long split_weight(const mofft_kernel_set *ks, int i, int count,
                              size_t span, int measured) {
  if (ks != PASS_KERNELS) return 10; /* batch kinds: passes of equal cost */
  if (measured)
    return pass_weight_measured(i, count == 0 ? MOFFT_ROLE_FIRST : MOFFT_ROLE_LATER,
                                span);
  return pass_weight(&ks->entries[i]);
}

/* the cheapest any pass can be (prunes the split search) */
static long long min_split_weight(const mofft_kernel_set *ks, int precision_bits,
                                  int measured) {
  long long m = -1;
  static const size_t spans[] = {16, 64, 256, 1024};
  for (int i = 0; i < *ks->count; ++i) {
    const mofft_registry_entry *e = &ks->entries[i];
    if (e->precision_bits != precision_bits || e->rows || e->variant != 0) continue;
    for (int count = 0; count < 2; ++count)
      for (int q = 0; q < 4; ++q) {
        const long long w = split_weight(ks, i, count, spans[q], measured);
        if (m < 0 || w < m) m = w;
      }
  }
  return m < 0 ? 10 : m;
}

/* Distinct strided pass lengths of a kernel set and precision, in
 * registry order of their first bank-0 kernel (built once per set). */
typedef struct {
  int count;
  size_t length[512];
  int first[512];       /* first bank-0 strided kernel of that length */
  int nk[512];          /* bank-0 strided kernels of that length (stage chains) */
  int kernels[512][8];
} split_lengths;

static const split_lengths *strided_lengths(const mofft_kernel_set *ks,
                                            int precision_bits) {
  static split_lengths table[2][2];      /* [pass / batch set][FP32 / FP64] */
  static int built[2][2];
  static pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
  const int a = ks == PASS_KERNELS ? 0 : 1, b = precision_bits == 32 ? 0 : 1;
  pthread_mutex_lock(&lock);
  if (!built[a][b]) {
    split_lengths *t = &table[a][b];
    t->count = 0;
    for (int i = 0; i < *ks->count && t->count < 512; ++i) {
      const mofft_registry_entry *e = &ks->entries[i];
      if (e->precision_bits != precision_bits || e->rows || e->variant != 0 ||
          e->length < 2)
        continue;
      int seen = -1;
      for (int j = 0; j < t->count && seen < 0; ++j)
        if (t->length[j] == (size_t)e->length) seen = j;
      if (seen >= 0) {
        if (t->nk[seen] < 8) t->kernels[seen][t->nk[seen]++] = i;
        continue;
      }
      t->length[t->count] = (size_t)e->length;
      t->first[t->count] = i;
      t->nk[t->count] = 1;
      t->kernels[t->count][0] = i;
      t->count++;
    }
    built[a][b] = 1;
  }
  pthread_mutex_unlock(&lock);
  return &table[a][b];
}

static void search_splits(const mofft_kernel_set *ks, int precision_bits,
                          size_t remaining, pass_split *cur, pass_split *best,
                          int max_passes, long long min_w, int measured) {
  /* product of the passes chosen so far: the next pass's span */
  size_t done = 1;
  for (int i = 0; i < cur->count; ++i) done *= cur->lengths[i];
  if (remaining == 1) {
    if (cur->count > 0 && split_better(cur, best)) *best = *cur;
    return;
  }
  if (cur->count >= max_passes) return;
  /* every further pass costs at least min_w */
  if (best->count != 0 && cur->weight + min_w > best->weight) return;
  const split_lengths *sl = strided_lengths(ks, precision_bits);
  for (int q = 0; q < sl->count; ++q) {
    const size_t l = sl->length[q];
    if (remaining % l) continue;
    /* one candidate per pass length: its chosen kernel; a first pass runs
     * at stride N / l, a later one at span `done` */
    const size_t span = cur->count == 0 ? remaining / l : done;
    int k = sl->first[q];
    if (measured && ks == PASS_KERNELS) {
      long long bw = split_weight(ks, k, cur->count, span, measured);
      for (int j = 1; j < sl->nk[q]; ++j) {
        const long long w = split_weight(ks, sl->kernels[q][j], cur->count, span, measured);
        if (w < bw) { bw = w; k = sl->kernels[q][j]; }
      }
    }
    const long long w = split_weight(ks, k, cur->count, span, measured);
    cur->lengths[cur->count] = l;
    cur->kernels[cur->count] = k;
    cur->count++;
    cur->weight += w;
    search_splits(ks, precision_bits, remaining / l, cur, best, max_passes, min_w,
                  measured);
    cur->weight -= w;
    cur->count--;
  }
}

/* The single-pass kernel of a length: with measured costs the cheapest
 * whole-transform kernel of any stage chain (the plain tile banks), else
 * the first bank-0 one compiled. */
static int single_pass_kernel(const mofft_kernel_set *ks, int precision_bits,
                              size_t length, int measured) {
  if (ks != PASS_KERNELS || !measured)
    return find_kernel_by_length(ks, precision_bits, 1, length);
  int best = -1;
  double bn = 0.0;
  for (int i = 0; i < *ks->count; ++i) {
    const mofft_registry_entry *e = &ks->entries[i];
    if (e->precision_bits != precision_bits || !e->rows ||
        (size_t)e->length != length || e->variant == (int)MOFFT_MFMA_BANK)
      continue;
    const double n = pass_ns_per_byte(i, MOFFT_ROLE_ROWS, 1);
    if (best < 0 || n < bn) { best = i; bn = n; }
  }
  return best;
}

/* The multi-pass split of the lowest weight (measured or nominal), as a
 * chain (outer -> inner) and pass mask. */
static mofft_status best_split(const mofft_kernel_set *ks, size_t length,
                               int precision_bits, int measured, int *radices,
                               size_t *depth, uint64_t *mask) {
  pass_split cur, best;
  memset(&cur, 0, sizeof(cur));
  memset(&best, 0, sizeof(best));
  const long long min_w = min_split_weight(ks, precision_bits, measured);
  /* a plan of more than max_passes passes costs at least
   * min_w * (max_passes + 1), so the search stops once the best plan found
   * costs no more than that */
  for (int max_passes = 2; max_passes <= 12; max_passes += 2) {
    search_splits(ks, precision_bits, length, &cur, &best, max_passes, min_w,
                  measured);
    if (best.count != 0 && best.weight <= min_w * (max_passes + 1)) break;
  }
  if (best.count == 0) return MOFFT_UNSUPPORTED_SIZE;
  /* concatenate: the last pass is outermost */
  size_t d = 0;
  uint64_t m = 0;
  for (int p = best.count - 1; p >= 0; --p) {
    const mofft_registry_entry *e = &ks->entries[best.kernels[p]];
    if (d + (size_t)e->stage_count > MOFFT_MAX_PLAN_DEPTH)
      return MOFFT_UNSUPPORTED_SIZE;
    for (int i = e->stage_count - 1; i >= 0; --i) radices[d++] = e->stages[i];
    if (p > 0) m |= (uint64_t)1 << (d - 1);
  }
  *depth = d;
  *mask = m;
  return MOFFT_SUCCESS;
}

static double chain_seconds(size_t length, int precision_bits, const int *radices,
                            size_t depth, uint64_t mask) {
  mofft_schedule s;
  if (build_schedule(PASS_KERNELS, length, 1, precision_bits, radices, depth,
                     mask, 0, &s) != MOFFT_SUCCESS)
    return DBL_MAX;
  return schedule_seconds(&s);
}

/* A multi-pass split the measured model prefers is taken over the split of
 * the nominal pass weights only when it is predicted this much faster: the
 * per-pass measurements (each kernel alone at four strides / spans) miss
 * effects of the actual neighbours, and on 58 untuned lengths (2^24
 * points, MI355X) taking every measured-model split ran 3 % faster in the
 * geometric mean but 17 lengths more than 5 % slower (up to 24 %); with
 * this margin 5 % faster, 2 lengths more than 5 % slower (at most 9 %). */
#define MOFFT_MODEL_SPLIT_MARGIN 0.85

/* Model plan: a single-pass kernel when one covers the length (with
 * measured costs the cheapest of its stage chains), otherwise the pass
 * decomposition of the lowest predicted HBM cost.  `use_measured` = 0
 * gives the nominal plan: the first compiled (radix-16) chain of a single
 * pass and the split of the nominal weights - the chains the matrix-core
 * bank has kernels for. */
static mofft_status model_chain_ex(const mofft_kernel_set *ks, size_t length,
                                   int precision_bits, int *radices,
                                   size_t *depth, uint64_t *mask,
                                   int use_measured) {
  if (length < 2) return MOFFT_UNSUPPORTED_SIZE;
  const int measured = use_measured && pass_costs_measured();
  const int single = single_pass_kernel(ks, precision_bits, length, measured);
  if (single >= 0) {
    const mofft_registry_entry *e = &ks->entries[single];
    for (int i = 0; i < e->stage_count; ++i)
      radices[i] = e->stages[e->stage_count - 1 - i];
    *depth = (size_t)e->stage_count;
    *mask = 0;
    return MOFFT_SUCCESS;
  }
  mofft_status st = best_split(ks, length, precision_bits, 0, radices, depth, mask);
  if (st != MOFFT_SUCCESS || ks != PASS_KERNELS || !measured)
    return st;
  int r2[MOFFT_MAX_PLAN_DEPTH];
  size_t d2 = 0;
  uint64_t m2 = 0;
  if (best_split(ks, length, precision_bits, 1, r2, &d2, &m2) != MOFFT_SUCCESS)
    return st;
  const double t_nominal = chain_seconds(length, precision_bits, radices, *depth, *mask);
  const double t_measured = chain_seconds(length, precision_bits, r2, d2, m2);
  if (t_measured < MOFFT_MODEL_SPLIT_MARGIN * t_nominal) {
    memcpy(radices, r2, d2 * sizeof(int));
    *depth = d2;
    *mask = m2;
  }
  return MOFFT_SUCCESS;
}

static mofft_status model_chain(const mofft_kernel_set *ks, size_t length,
                                int precision_bits, int *radices,
                                size_t *depth, uint64_t *mask) {
  return model_chain_ex(ks, length, precision_bits, radices, depth, mask, 1);
}

/* The tile bank of the model plan: with measured costs, the plain bank
 * (0, 1, 2, 4, 5) whose schedule of the chain is predicted fastest; else 0. */
static unsigned int model_bank(size_t length, int precision_bits,
                               const int *radices, size_t depth,
                               uint64_t mask) {
  if (!pass_costs_measured()) return 0;
  static const unsigned int banks[] = {0u, 1u, 2u, MOFFT_WIDE_BANK,
                                       MOFFT_NARROW_BANK};
  unsigned int best = 0;
  double bt = 0.0;
  for (size_t i = 0; i < sizeof(banks) / sizeof(banks[0]); ++i) {
    mofft_schedule s;
    if (build_schedule(PASS_KERNELS, length, 1, precision_bits, radices, depth,
                       mask, banks[i], &s) != MOFFT_SUCCESS)
      continue;
    const double t = schedule_seconds(&s);
    if (i == 0 || t < bt) { best = banks[i]; bt = t; }
  }
  return best;
}

/* ------------------------------------------------------------------ */
/* option handling (src/mofft.c:739-762, 1384-1396)                      */
/* ------------------------------------------------------------------ */

static mofft_status validate_options(const mofft_plan_options *given,
                                     mofft_plan_options *resolved) {
  if (given) *resolved = *given;
  else mofft_plan_options_init(resolved);
  if ((resolved->mode != MOFFT_PLANNER_PRETUNED &&
       resolved->mode != MOFFT_PLANNER_MEASURE) ||
      (resolved->flags & ~((uint32_t)MOFFT_PLAN_STRICT_WISDOM)) != 0 ||
      (resolved->mode == MOFFT_PLANNER_MEASURE && resolved->flags != 0))
    return MOFFT_INVALID_ARGUMENT;
  if (resolved->candidate_limit == 0)
    resolved->candidate_limit = MOFFT_DEFAULT_SEARCH_CANDIDATES;
  if (resolved->timing_samples == 0)
    resolved->timing_samples = MOFFT_DEFAULT_TIMING_SAMPLES;
  if (resolved->candidate_limit > MOFFT_MAX_SEARCH_CANDIDATES ||
      resolved->timing_samples > MOFFT_MAX_TIMING_SAMPLES)
    return MOFFT_INVALID_ARGUMENT;
  return MOFFT_SUCCESS;
}

void mofft_plan_options_init(mofft_plan_options *options) {
  if (!options) return;
  options->mode = MOFFT_PLANNER_PRETUNED;
  options->flags = 0;
  options->candidate_limit = 0;
  options->timing_samples = 0;
}

static int options_from_bound_flags(uint32_t flags,
                                    mofft_plan_options *options) {
  if ((flags & ~((uint32_t)MOFFT_FLAG_MEASURE |
                 (uint32_t)MOFFT_FLAG_WISDOM_ONLY)) != 0 ||
      ((flags & MOFFT_FLAG_MEASURE) && (flags & MOFFT_FLAG_WISDOM_ONLY)))
    return 0;
  mofft_plan_options_init(options);
  options->mode = (flags & MOFFT_FLAG_MEASURE) ? MOFFT_PLANNER_MEASURE
                                               : MOFFT_PLANNER_PRETUNED;
  if (flags & MOFFT_FLAG_WISDOM_ONLY)
    options->flags |= MOFFT_PLAN_STRICT_WISDOM;
  return 1;
}

/* Wisdom: exact (length, precision) entry.  1 hit, 0 miss, -1 stale. */
static int wisdom_lookup(size_t length, int precision_bits, int *radices,
                         size_t *depth, uint64_t *mask, int *direct,
                         int *layout, unsigned int *bank) {
  uint64_t m = 0;
  int di = 0, sl = 0;
  unsigned int b = 0;
  const size_t count = mofft_generated_wisdom_lookup(
      length, precision_bits, radices, MOFFT_MAX_PLAN_DEPTH, &m, &di, &sl, &b);
  if (count == 0) return 0;
  if (b >= kernel_bank_count()) return -1;
  if (validate_chain(length, radices, count) != MOFFT_SUCCESS) return -1;
  *depth = count;
  *mask = m;
  if (direct) *direct = di;
  if (layout) *layout = sl;
  if (bank) *bank = b;
  return 1;
}

int mofft_wisdom_is_formal(void) { return mofft_generated_wisdom_is_formal(); }
const char *mofft_wisdom_source(void) { return mofft_generated_wisdom_source(); }

/* ------------------------------------------------------------------ */
/* plan lifetime                                                        */
/* ------------------------------------------------------------------ */

static void plan_release(mofft_plan_common *p) {
  if (!p) return;
  tables_drop(p->tables);
  mofft_hip_free(p->staging_in);
  mofft_hip_free(p->staging_out);
  free(p);
}

static mofft_status fill_schedule_twiddles(const mofft_schedule *s,
                                           void **twiddles) {
  const size_t cb = complex_bytes(s->precision_bits);
  *twiddles = NULL;
  if (s->twiddle_count == 0) return MOFFT_SUCCESS;
  if (mofft_hip_malloc(twiddles, sat_mul(s->twiddle_count, cb)) != 0)
    return MOFFT_ALLOCATION_FAILURE;
  char *base = (char *)*twiddles;
  int rc = 0;
  if (s->pass_count > 1) {
    rc |= mofft_hip_fill_powers(base + s->tw_hi_offset * cb, s->precision_bits,
                                (int64_t)s->tw_hi_count, (int64_t)s->length,
                                (int64_t)1 << s->split_shift, NULL);
    rc |= mofft_hip_fill_powers(base + s->tw_lo_offset * cb, s->precision_bits,
                                (int64_t)s->tw_lo_count, (int64_t)s->length,
                                1, NULL);
  }
  for (int i = 0; i < s->pass_count && rc == 0; ++i) {
    const mofft_pass *ps = &s->passes[i];
    if (ps->pass_table)
      rc |= mofft_hip_fill_pass(base + ps->pass_offset * cb, s->precision_bits,
                                (int64_t)ps->span, (int64_t)ps->length, NULL);
  }
  for (int i = 0; i < s->pass_count && rc == 0; ++i) {
    const mofft_pass *ps = &s->passes[i];
    const mofft_registry_entry *e = &s->ks->entries[ps->kernel];
    size_t ns = 1, off = ps->stage_offset;
    for (int k = 0; k < e->stage_count; ++k) {
      const int r = e->stages[k];
      if (ns > 1) {
        rc |= mofft_hip_fill_stage(base + off * cb, s->precision_bits,
                                   (int64_t)ns, r, NULL);
        off += ns * (size_t)(r - 1);
      }
      ns *= (size_t)r;
    }
  }
  if (rc == 0) rc = mofft_hip_synchronize(NULL);
  if (rc != 0) {
    mofft_hip_free(*twiddles);
    *twiddles = NULL;
    return MOFFT_INTERNAL_ERROR;
  }
  return MOFFT_SUCCESS;
}

/* New device tables (twiddles and workspace) for p's schedule; the plan
 * keeps its current ones on failure. */
static mofft_status make_tables(mofft_plan_common *p, const mofft_schedule *s) {
  mofft_device_tables *t =
      (mofft_device_tables *)calloc(1, sizeof(mofft_device_tables));
  if (!t) return MOFFT_ALLOCATION_FAILURE;
  t->refs = 1;
  mofft_status st = fill_schedule_twiddles(s, &t->twiddles);
  if (st == MOFFT_SUCCESS) {
    const size_t ws = schedule_workspace(s, p->placement);
    if (ws && mofft_hip_malloc(&t->workspace, ws) != 0)
      st = MOFFT_ALLOCATION_FAILURE;
  }
  if (st != MOFFT_SUCCESS) {
    tables_drop(t);
    return st;
  }
  tables_drop(p->tables);
  p->tables = t;
  p->twiddles = t->twiddles;
  p->workspace = t->workspace;
  return MOFFT_SUCCESS;
}

/* Build a plan from an explicit chain + mask (device resources included). */
static mofft_status plan_from_chain_layout(mofft_plan_common **out,
                                           size_t length, size_t batch,
                                           int precision_bits, int direction,
                                           mofft_placement placement,
                                           const int *radices, size_t depth,
                                           uint64_t mask, unsigned int bank,
                                           mofft_stage_layout layout) {
  *out = NULL;
  if (!mofft_hip_device_available()) return MOFFT_UNSUPPORTED_HARDWARE;
  mofft_plan_common *p = (mofft_plan_common *)calloc(1, sizeof(*p));
  if (!p) return MOFFT_ALLOCATION_FAILURE;
  mofft_status st = build_schedule_layout(
      PASS_KERNELS, length, batch, precision_bits, radices, depth, mask, bank,
      layout == MOFFT_STAGE_LAYOUT_SECTION, &p->s);
  if (st != MOFFT_SUCCESS) { free(p); return st; }
  p->direction = direction;
  p->placement = placement;
  p->direct_input = 0;
  p->stage_layout = layout;
  p->kernel_bank = bank;
  st = make_tables(p, &p->s);
  if (st != MOFFT_SUCCESS) { plan_release(p); return st; }
  *out = p;
  return MOFFT_SUCCESS;
}

static mofft_status plan_from_chain(mofft_plan_common **out, size_t length,
                                    size_t batch, int precision_bits,
                                    int direction, mofft_placement placement,
                                    const int *radices, size_t depth,
                                    uint64_t mask, unsigned int bank) {
  return plan_from_chain_layout(out, length, batch, precision_bits, direction,
                                placement, radices, depth, mask, bank,
                                MOFFT_STAGE_LAYOUT_TRANSPOSE);
}

/* ------------------------------------------------------------------ */
/* execution                                                            */
/* ------------------------------------------------------------------ */

/* Optional per-launch timing: events[k] is recorded before launch k and
 * events[k + 1] after it (k counts launches over chunks and passes). */
typedef struct {
  void **events;
  size_t capacity;
  size_t used;
} launch_timing;

static int launch_schedule_timed(const mofft_plan_common *p, const void *in,
                                 void *out, void *work, void *stream,
                                 size_t count, launch_timing *timing);

static int launch_schedule_count(const mofft_plan_common *p, const void *in,
                                 void *out, void *work, void *stream,
                                 size_t count) {
  return launch_schedule_timed(p, in, out, work, stream, count, NULL);
}

/* Arguments of pass i over `cnt` transforms from src to dst. */
static void fill_pass_args(const mofft_plan_common *p, int i, const void *src,
                           void *dst, size_t cnt, mofft_pass_args *a) {
  const mofft_schedule *s = &p->s;
  const size_t cb = complex_bytes(s->precision_bits);
  const char *tw = (const char *)p->twiddles;
  const int backward = p->direction == MOFFT_BACKWARD;
  const int P = s->pass_count;
  const mofft_pass *ps = &s->passes[i];
  memset(a, 0, sizeof(*a));
  a->in = src;
  a->out = dst;
  a->stage_tw = tw ? tw + ps->stage_offset * cb : NULL;
  a->tw_hi = tw ? tw + s->tw_hi_offset * cb : NULL;
  a->tw_lo = tw ? tw + s->tw_lo_offset * cb : NULL;
  a->tw_pass = (tw && ps->pass_table) ? tw + ps->pass_offset * cb : NULL;
  a->n = (int64_t)s->length;
  a->stride = (int64_t)ps->stride;
  a->ns = (int64_t)ps->span;
  a->tw_scale = (int64_t)(s->length / (ps->span * ps->length));
  a->columns = (int64_t)(cnt * ps->stride);
  a->split_shift = s->split_shift;
  mofft_pass_args_set_divisors(a, (int64_t)ps->length);
  a->flags = 0;
  if (backward && i == 0) a->flags |= MOFFT_PASS_SWAP_IN;
  if (backward && i == P - 1) a->flags |= MOFFT_PASS_SWAP_OUT;
  if (ps->span > 1) a->flags |= MOFFT_PASS_TWIDDLE;
  if (ps->span > 1 && ps->pass_table) a->flags |= MOFFT_PASS_TWIDDLE_TABLE;
  if (xcd_tile_order()) a->flags |= MOFFT_PASS_XCD_ORDER;
}

/* The passes of one chunk of `cnt` transforms starting at transform c0. */
static int launch_chunk(const mofft_plan_common *p, const void *in, void *out,
                        void *work, void *stream, size_t c0, size_t cnt,
                        launch_timing *timing) {
  const mofft_schedule *s = &p->s;
  const size_t cb = complex_bytes(s->precision_bits);
  const int P = s->pass_count;
  const int in_place = in == out;
  const char *src_chunk = (const char *)in + c0 * s->length * cb;
  char *dst_chunk = (char *)out + c0 * s->length * cb;
  const void *src = src_chunk;
  for (int i = 0; i < P; ++i) {
    void *dst;
    if (i == P - 1) {
      dst = dst_chunk;
    } else if (in_place) {
      dst = (i % 2 == 0) ? work : dst_chunk;
    } else {
      dst = ((P - 2 - i) % 2 == 0) ? (void *)dst_chunk : work;
    }
    const mofft_pass *ps = &s->passes[i];
    mofft_pass_args a;
    fill_pass_args(p, i, src, dst, cnt, &a);
    if (timing && timing->used == 0) {
      if (timing->capacity < 1) return 1;
      if (mofft_hip_event_record(timing->events[0], stream) != 0) return 1;
    }
    const int rc = mofft_hip_launch_pass(ps->kernel, &a, stream);
    if (rc != 0) return rc;
    if (timing) {
      if (timing->used + 2 > timing->capacity) return 1;
      if (mofft_hip_event_record(timing->events[++timing->used], stream) != 0)
        return 1;
    }
    src = dst;
  }
  return 0;
}

/* The first `count` transforms in chunks of at most `chunk` (the workspace
 * `work` holds one chunk's intermediate). */
static int launch_schedule_chunks(const mofft_plan_common *p, const void *in,
                                  void *out, void *work, void *stream,
                                  size_t count, size_t chunk,
                                  launch_timing *timing) {
  for (size_t c0 = 0; c0 < count; c0 += chunk) {
    const size_t cnt = count - c0 < chunk ? count - c0 : chunk;
    const int rc = launch_chunk(p, in, out, work, stream, c0, cnt, timing);
    if (rc != 0) return rc;
  }
  return 0;
}

static int launch_schedule_timed(const mofft_plan_common *p, const void *in,
                                 void *out, void *work, void *stream,
                                 size_t count, launch_timing *timing) {
  return launch_schedule_chunks(p, in, out, work, stream, count, p->s.chunk,
                                timing);
}

static int launch_schedule(const mofft_plan_common *p, const void *in,
                           void *out, void *work, void *stream) {
  return launch_schedule_count(p, in, out, work, stream, p->s.batch);
}

static mofft_status execute_device(const mofft_plan_common *p, const void *in,
                                   void *out, void *work, size_t work_bytes,
                                   void *stream) {
  const size_t need = schedule_workspace(&p->s, p->placement);
  if (need && (!work || work_bytes < need)) return MOFFT_INVALID_ARGUMENT;
  return launch_schedule(p, in, out, work, stream) == 0 ? MOFFT_SUCCESS
                                                        : MOFFT_INTERNAL_ERROR;
}

/* Synchronous execute of mofft.h: host buffers are staged through device
 * memory the plan keeps for later calls. */
static mofft_status execute_sync(const mofft_plan_common *cp, const void *in,
                                 void *out) {
  if (!cp || !in || !out) return MOFFT_INVALID_ARGUMENT;
  if ((cp->placement == MOFFT_IN_PLACE) != (in == out))
    return MOFFT_INVALID_ARGUMENT;
  if (!mofft_hip_device_available()) return MOFFT_UNSUPPORTED_HARDWARE;
  mofft_plan_common *p = (mofft_plan_common *)cp;
  const size_t bytes = sat_mul(sat_mul(p->s.length, p->s.batch),
                               complex_bytes(p->s.precision_bits));
  const int din = mofft_hip_is_device_pointer(in);
  const int dout = mofft_hip_is_device_pointer(out);
  const void *dev_in = in;
  void *dev_out = out;
  if (!din) {
    if (!p->staging_in && mofft_hip_malloc(&p->staging_in, bytes) != 0)
      return MOFFT_ALLOCATION_FAILURE;
    if (mofft_hip_copy(p->staging_in, in, bytes) != 0)
      return MOFFT_INTERNAL_ERROR;
    dev_in = p->staging_in;
  }
  if (!dout) {
    if (in == out) {
      dev_out = (void *)dev_in;
    } else {
      if (!p->staging_out && mofft_hip_malloc(&p->staging_out, bytes) != 0)
        return MOFFT_ALLOCATION_FAILURE;
      dev_out = p->staging_out;
    }
  }
  mofft_status st = execute_device(p, dev_in, dev_out, p->workspace,
                                   schedule_workspace(&p->s, p->placement), NULL);
  if (st != MOFFT_SUCCESS) return st;
  if (!dout && mofft_hip_copy(out, dev_out, bytes) != 0)
    return MOFFT_INTERNAL_ERROR;
  return mofft_hip_synchronize(NULL) == 0 ? MOFFT_SUCCESS
                                          : MOFFT_INTERNAL_ERROR;
}

/* ------------------------------------------------------------------ */
/* diagnostics                                                          */
/* ------------------------------------------------------------------ */

static void fill_diagnostics(mofft_plan_diagnostics *d,
                             const mofft_plan_common *p) {
  if (!d) return;
  d->radix_count = (size_t)p->s.depth;
  memcpy(d->radices, p->s.factors, (size_t)p->s.depth * sizeof(int));
  d->blocked_transpose_mask = p->s.pass_mask;
  d->direct_input = p->direct_input;
  d->stage_layout = p->stage_layout;
  d->kernel_bank = p->kernel_bank;
  d->kernel_bank_count = kernel_bank_count();
  d->length = p->s.length;
  d->precision_bits = p->s.precision_bits;
  d->direction = (mofft_direction)p->direction;
  d->placement = p->placement;
}

/* ------------------------------------------------------------------ */
/* MEASURE planner (src/mofft.c:798-1052 restated for GPU schedules)    */
/* ------------------------------------------------------------------ */

typedef struct {
  int radices[MOFFT_MAX_PLAN_DEPTH];
  size_t depth;
  uint64_t mask;
  unsigned int bank;
  double model;
  double samples[MOFFT_MAX_TIMING_SAMPLES];
  double median;
} measured_candidate;

static double scalar_median(double *v, size_t n) {
  if (n == 0) return DBL_MAX;
  for (size_t i = 1; i < n; ++i) {
    const double x = v[i];
    size_t j = i;
    while (j > 0 && x < v[j - 1]) { v[j] = v[j - 1]; --j; }
    v[j] = x;
  }
  return (n & 1) ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

static double schedule_model_cost(const mofft_schedule *s) {
  /* the measured pass costs when the table is present */
  if (pass_costs_measured()) return schedule_seconds(s);
  /* HBM round trips dominate: one read + one write per pass, each strided
   * pass weighted by its measured cost (pass_weight), plus a small
   * per-stage compute term */
  const double bytes = 2.0 * (double)s->length * (double)s->batch *
                       (double)complex_bytes(s->precision_bits);
  double passes = 0.0;
  for (int i = 0; i < s->pass_count; ++i) {
    const mofft_registry_entry *e = &mofft_registry[s->passes[i].kernel];
    passes += e->rows ? 1.0 : 0.1 * (double)pass_weight(e);
  }
  return passes * bytes / 8.0e12 + (double)s->depth * 1e-7;
}

static int same_candidate(const measured_candidate *a,
                          const measured_candidate *b) {
  return a->depth == b->depth && a->mask == b->mask && a->bank == b->bank &&
         memcmp(a->radices, b->radices, a->depth * sizeof(int)) == 0;
}

/* Adds the chain once per kernel bank whose schedule differs from bank 0's
 * (a bank with no variant kernel for any pass is the same plan). */
static void add_candidate(measured_candidate *pool, size_t *count, size_t cap,
                          size_t length, size_t batch, int precision_bits,
                          const int *radices, size_t depth, uint64_t mask) {
  mofft_schedule base;
  memset(&base, 0, sizeof(base));
  /* the matrix-core bank computes in split FP16 (other rounding): MEASURE
   * only times numerically equivalent plain plans */
  for (unsigned int bank = 0; bank < kernel_bank_count(); ++bank) {
    if (bank == MOFFT_MFMA_BANK) continue;
    if (*count >= cap) return;
    mofft_schedule s;
    if (build_schedule(PASS_KERNELS, length, batch, precision_bits, radices,
                       depth, mask, bank, &s) != MOFFT_SUCCESS)
      return;
    if (bank == 0) {
      base = s;
    } else {
      int same = s.pass_count == base.pass_count;
      for (int i = 0; same && i < s.pass_count; ++i)
        same = s.passes[i].kernel == base.passes[i].kernel;
      if (same) continue;
    }
    measured_candidate c;
    memset(&c, 0, sizeof(c));
    memcpy(c.radices, radices, depth * sizeof(int));
    c.depth = depth;
    c.mask = s.pass_mask; /* effective grouping */
    c.bank = bank;
    c.model = pass_costs_measured() ? schedule_model_cost(&s)
                                    : schedule_model_cost(&s) * (1.0 + 0.01 * bank);
    int dup = 0;
    for (size_t i = 0; i < *count; ++i)
      if (same_candidate(&pool[i], &c)) { dup = 1; break; }
    if (!dup) pool[(*count)++] = c;
  }
}

static void enumerate_splits(int precision_bits, size_t remaining,
                             pass_split *cur, int max_passes,
                             measured_candidate *pool, size_t *count,
                             size_t cap, size_t length, size_t batch) {
  if (*count >= cap) return;
  if (remaining == 1) {
    int radices[MOFFT_MAX_PLAN_DEPTH];
    size_t d = 0;
    uint64_t mask = 0;
    for (int p = cur->count - 1; p >= 0; --p) {
      const mofft_registry_entry *e = &mofft_registry[cur->kernels[p]];
      if (d + (size_t)e->stage_count > MOFFT_MAX_PLAN_DEPTH) return;
      for (int i = e->stage_count - 1; i >= 0; --i) radices[d++] = e->stages[i];
      if (p > 0) mask |= (uint64_t)1 << (d - 1);
    }
    add_candidate(pool, count, cap, length, batch, precision_bits, radices, d,
                  mask);
    return;
  }
  if (cur->count >= max_passes) return;
  for (int i = 0; i < mofft_registry_count; ++i) {
    const mofft_registry_entry *e = &mofft_registry[i];
    if (e->precision_bits != precision_bits || e->variant != 0) continue;
    /* row kernels are whole single-pass transforms; column kernels are the
     * passes of a split (and single-pass chains where a row kernel with
     * the same stages exists) */
    if (e->rows && !(cur->count == 0 && (size_t)e->length == length)) continue;
    const size_t l = (size_t)e->length;
    if (l < 2 || remaining % l) continue;
    cur->lengths[cur->count] = l;
    cur->kernels[cur->count] = i;
    cur->count++;
    enumerate_splits(precision_bits, remaining / l, cur, max_passes, pool,
                     count, cap, length, batch);
    cur->count--;
  }
}

static int candidate_model_compare(const void *l, const void *r) {
  const measured_candidate *a = (const measured_candidate *)l;
  const measured_candidate *b = (const measured_candidate *)r;
  if (a->model < b->model) return -1;
  if (a->model > b->model) return 1;
  if (a->depth != b->depth) return a->depth < b->depth ? -1 : 1;
  for (size_t i = 0; i < a->depth; ++i)
    if (a->radices[i] != b->radices[i]) return a->radices[i] < b->radices[i] ? -1 : 1;
  if (a->mask != b->mask) return a->mask < b->mask ? -1 : 1;
  if (a->bank != b->bank) return a->bank < b->bank ? -1 : 1;
  return 0;
}

static int candidate_time_compare(const void *l, const void *r) {
  const measured_candidate *a = (const measured_candidate *)l;
  const measured_candidate *b = (const measured_candidate *)r;
  if (a->median < b->median) return -1;
  if (a->median > b->median) return 1;
  return candidate_model_compare(l, r);
}

/* The MEASURE result check (the reference rejects a non-finite candidate
 * output, src/mofft.c:725-737, 990-993): compares `count` complex values of
 * a candidate's output with the same values of the analytical plan's.
 * Returns 0 when every value is finite and the relative L2 difference is
 * at most `bound`, 1 when a value is not finite, 2 when they disagree. */
int mofft_internal_check_output(const void *out, const void *ref,
                                size_t count, int precision_bits,
                                double bound) {
  double diff = 0.0, norm = 0.0;
  for (size_t i = 0; i < 2 * count; ++i) {
    const double a = precision_bits == 32 ? (double)((const float *)out)[i]
                                          : ((const double *)out)[i];
    if (!isfinite(a)) return 1;
    if (!ref) continue;
    const double b = precision_bits == 32 ? (double)((const float *)ref)[i]
                                          : ((const double *)ref)[i];
    diff += (a - b) * (a - b);
    norm += b * b;
  }
  if (!ref) return 0;
  return sqrt(diff) <= bound * sqrt(norm) ? 0 : 2;
}

/* Relative L2 error bound of one precision at length n: 1e-6 log2 n (FP32)
 * or 1e-15 log2 n (FP64); two plans of one transform agree within twice
 * that, and the search allows four times it. */
static double precision_bound(size_t length, int precision_bits) {
  const double l = log2((double)(length < 2 ? 2 : length));
  return (precision_bits == 32 ? 1e-6 : 1e-15) * l;
}

static mofft_status search_best_plan(mofft_plan_common **out, size_t length,
                                     size_t batch, int precision_bits,
                                     int direction, mofft_placement placement,
                                     const mofft_plan_options *options,
                                     mofft_plan_diagnostics *diagnostics) {
  const size_t cap = 4 * MOFFT_MAX_SEARCH_CANDIDATES;
  measured_candidate *pool =
      (measured_candidate *)calloc(cap, sizeof(measured_candidate));
  if (!pool) return MOFFT_ALLOCATION_FAILURE;
  size_t count = 0;
  int radices[MOFFT_MAX_PLAN_DEPTH];
  size_t depth = 0;
  uint64_t mask = 0;
  if (model_chain(PASS_KERNELS, length, precision_bits, radices, &depth, &mask) ==
      MOFFT_SUCCESS)
    add_candidate(pool, &count, cap, length, batch, precision_bits, radices,
                  depth, mask);
  pass_split cur;
  memset(&cur, 0, sizeof(cur));
  for (int max_passes = 1; max_passes <= 4 && count < cap; ++max_passes)
    enumerate_splits(precision_bits, length, &cur, max_passes, pool, &count,
                     cap, length, batch);
  /* the all-blocked variant of the model chain: one pass per radix */
  if (depth > 0)
    add_candidate(pool, &count, cap, length, batch, precision_bits, radices,
                  depth, depth > 1 ? (((uint64_t)1 << (depth - 1)) - 1) : 0);
  if (count == 0) { free(pool); return MOFFT_UNSUPPORTED_SIZE; }
  qsort(pool, count, sizeof(pool[0]), candidate_model_compare);
  size_t selected = count < options->candidate_limit ? count
                                                     : options->candidate_limit;
  /* deterministic shuffle of the measurement order (src/mofft.c:929-942) */
  uint32_t state = (uint32_t)length ^ ((uint32_t)precision_bits << 16) ^
                   UINT32_C(0x9e3779b9);
  for (size_t i = selected; i > 1; --i) {
    state ^= state << 13; state ^= state >> 17; state ^= state << 5;
    const size_t j = (size_t)state % i;
    measured_candidate t = pool[i - 1]; pool[i - 1] = pool[j]; pool[j] = t;
  }
  const size_t cb = complex_bytes(precision_bits);
  const size_t bytes = sat_mul(sat_mul(length, batch), cb);
  void *din = NULL, *dout = NULL, *dref = NULL;
  void *hcheck = NULL, *href = NULL;
  size_t check_count = 0, check_bytes = 0;
  int have_ref = 0;
  mofft_status st = MOFFT_SUCCESS;
  if (mofft_hip_malloc(&din, bytes) || mofft_hip_malloc(&dout, bytes) ||
      (placement == MOFFT_IN_PLACE && mofft_hip_malloc(&dref, bytes))) {
    st = MOFFT_ALLOCATION_FAILURE;
    goto done;
  }
  {
    /* deterministic input (src/mofft.c:942-957) generated on the host */
    const size_t host_elems = length * batch * 2;
    void *host = malloc(bytes);
    if (!host) { st = MOFFT_ALLOCATION_FAILURE; goto done; }
    uint32_t x = UINT32_C(0x7f4a7c15);
    for (size_t i = 0; i < host_elems; ++i) {
      x ^= x << 13; x ^= x >> 17; x ^= x << 5;
      const double v = ((double)(x & UINT32_C(0xffff)) / 32768.0) - 1.0;
      if (precision_bits == 32) ((float *)host)[i] = (float)v;
      else ((double *)host)[i] = v;
    }
    const int rc = mofft_hip_copy(din, host, bytes);
    free(host);
    if (rc) { st = MOFFT_INTERNAL_ERROR; goto done; }
    if (dref && mofft_hip_copy(dref, din, bytes)) { st = MOFFT_INTERNAL_ERROR; goto done; }
  }
  /* the result check covers the first transforms of the batch (at most
   * 8 MiB of output, at least one transform) */
  check_count = (size_t)8 << 20;
  check_count = check_count / (length * cb) < batch ? check_count / (length * cb)
                                                     : batch;
  if (check_count == 0) check_count = 1;
  check_bytes = check_count * length * cb;
  hcheck = malloc(check_bytes);
  href = malloc(check_bytes);
  if (!hcheck || !href) { st = MOFFT_ALLOCATION_FAILURE; goto done; }
  {
    /* reference output: the analytical plan on the same input */
    mofft_plan_common *p = NULL;
    if (depth > 0 &&
        plan_from_chain(&p, length, batch, precision_bits, direction, placement,
                        radices, depth, mask, 0) == MOFFT_SUCCESS) {
      const void *xin = placement == MOFFT_IN_PLACE ? dout : din;
      st = MOFFT_SUCCESS;
      if (placement == MOFFT_IN_PLACE && mofft_hip_copy(dout, dref, bytes))
        st = MOFFT_INTERNAL_ERROR;
      if (st == MOFFT_SUCCESS)
        st = execute_device(p, xin, dout, p->workspace,
                            schedule_workspace(&p->s, placement), NULL);
      if (st == MOFFT_SUCCESS && (mofft_hip_synchronize(NULL) ||
                                  mofft_hip_copy(href, dout, check_bytes)))
        st = MOFFT_INTERNAL_ERROR;
      plan_release(p);
      if (st != MOFFT_SUCCESS) goto done;
      if (mofft_internal_check_output(href, NULL, check_count * length,
                                      precision_bits, 0.0) != 0) {
        st = MOFFT_INTERNAL_ERROR;
        goto done;
      }
      have_ref = 1;
    }
  }
  {
    const size_t iterations = length * batch < 65536 ? 8 : (length * batch < (1u << 22) ? 4 : 1);
    const double bound = 4.0 * precision_bound(length, precision_bits);
    size_t accepted = 0;
    for (size_t i = 0; i < selected; ++i) {
      measured_candidate *c = &pool[i];
      mofft_plan_common *p = NULL;
      st = plan_from_chain(&p, length, batch, precision_bits, direction,
                           placement, c->radices, c->depth, c->mask, c->bank);
      if (st != MOFFT_SUCCESS) goto done;
      const void *xin = placement == MOFFT_IN_PLACE ? dout : din;
      if (placement == MOFFT_IN_PLACE && mofft_hip_copy(dout, dref, bytes))
        st = MOFFT_INTERNAL_ERROR;
      if (st == MOFFT_SUCCESS)
        st = execute_device(p, xin, dout, p->workspace,
                            schedule_workspace(&p->s, placement), NULL);
      if (st == MOFFT_SUCCESS && (mofft_hip_synchronize(NULL) ||
                                  mofft_hip_copy(hcheck, dout, check_bytes)))
        st = MOFFT_INTERNAL_ERROR;
      if (st == MOFFT_SUCCESS) {
        /* a non-finite output fails the search (as the reference does); an
         * output that disagrees with the analytical plan's drops the
         * candidate: a timing is never taken for a result */
        const int chk = mofft_internal_check_output(
            hcheck, have_ref ? href : NULL, check_count * length,
            precision_bits, bound);
        if (chk == 1) st = MOFFT_INTERNAL_ERROR;
        if (chk == 2) {
          plan_release(p);
          c->median = DBL_MAX;
          continue;
        }
      }
      for (uint32_t smp = 0; smp < options->timing_samples && st == MOFFT_SUCCESS; ++smp) {
        double times[8];
        for (size_t it = 0; it < iterations && st == MOFFT_SUCCESS; ++it) {
          if (placement == MOFFT_IN_PLACE) {
            if (mofft_hip_copy(dout, dref, bytes) || mofft_hip_synchronize(NULL)) {
              st = MOFFT_INTERNAL_ERROR;
              break;
            }
          }
          /* device time of the launches (HIP events on the null stream) */
          double secs = 0.0;
          if (mofft_hip_timer_begin(NULL)) st = MOFFT_INTERNAL_ERROR;
          if (st == MOFFT_SUCCESS)
            st = execute_device(p, xin, dout, p->workspace,
                                schedule_workspace(&p->s, placement), NULL);
          if (st == MOFFT_SUCCESS && mofft_hip_timer_end(NULL, &secs))
            st = MOFFT_INTERNAL_ERROR;
          times[it] = secs;
        }
        c->samples[smp] = scalar_median(times, iterations);
      }
      plan_release(p);
      if (st != MOFFT_SUCCESS) goto done;
      c->median = scalar_median(c->samples, options->timing_samples);
      accepted++;
    }
    if (accepted == 0) { st = MOFFT_INTERNAL_ERROR; goto done; }
  }
  qsort(pool, selected, sizeof(pool[0]), candidate_time_compare);
  st = plan_from_chain(out, length, batch, precision_bits, direction, placement,
                       pool[0].radices, pool[0].depth, pool[0].mask,
                       pool[0].bank);
  if (st == MOFFT_SUCCESS && diagnostics) {
    diagnostics->measured = 1;
    diagnostics->candidates_measured = selected;
    diagnostics->selected_median_seconds = pool[0].median;
    fill_diagnostics(diagnostics, *out);
  }
done:
  mofft_hip_free(din);
  mofft_hip_free(dout);
  mofft_hip_free(dref);
  free(hcheck);
  free(href);
  free(pool);
  return st;
}

/* ------------------------------------------------------------------ */
/* plan creation entry points                                           */
/* ------------------------------------------------------------------ */

static mofft_status create_with_options(mofft_plan_common **out,
                                        size_t length, size_t batch,
                                        int precision_bits, int direction,
                                        mofft_placement placement,
                                        const mofft_plan_options *given,
                                        mofft_plan_diagnostics *diagnostics) {
  if (!out) return MOFFT_INVALID_ARGUMENT;
  *out = NULL;
  if (diagnostics) memset(diagnostics, 0, sizeof(*diagnostics));
  if ((direction != MOFFT_FORWARD && direction != MOFFT_BACKWARD) ||
      (placement != MOFFT_IN_PLACE && placement != MOFFT_OUT_OF_PLACE) ||
      batch == 0)
    return MOFFT_INVALID_ARGUMENT;
  mofft_plan_options options;
  mofft_status st = validate_options(given, &options);
  if (st != MOFFT_SUCCESS) return st;
  if (!mofft_hip_device_available()) return MOFFT_UNSUPPORTED_HARDWARE;
  const double start = monotonic_seconds();
  if (length < 2) return MOFFT_UNSUPPORTED_SIZE;
  if (options.mode == MOFFT_PLANNER_MEASURE) {
    st = search_best_plan(out, length, batch, precision_bits, direction,
                          placement, &options, diagnostics);
    if (st == MOFFT_SUCCESS && diagnostics)
      diagnostics->planning_seconds = monotonic_seconds() - start;
    return st;
  }
  int radices[MOFFT_MAX_PLAN_DEPTH];
  size_t depth = 0;
  uint64_t mask = 0;
  int direct = 0, layout = 0;
  unsigned int bank = 0;
  const int w = wisdom_lookup(length, precision_bits, radices, &depth, &mask,
                              &direct, &layout, &bank);
  if (w < 0) return MOFFT_WISDOM_NOT_FOUND;
  if (w == 0) {
    if (options.flags & MOFFT_PLAN_STRICT_WISDOM) return MOFFT_WISDOM_NOT_FOUND;
    st = model_chain(PASS_KERNELS, length, precision_bits, radices, &depth, &mask);
    if (st != MOFFT_SUCCESS) return st;
    bank = model_bank(length, precision_bits, radices, depth, mask);
  }
  st = plan_from_chain_layout(out, length, batch, precision_bits, direction,
                              placement, radices, depth, mask, bank,
                              w > 0 ? (mofft_stage_layout)layout
                                    : MOFFT_STAGE_LAYOUT_TRANSPOSE);
  if (st != MOFFT_SUCCESS) return st;
  if (w > 0) (*out)->direct_input = direct;
  if (diagnostics) {
    diagnostics->used_wisdom = w > 0;
    diagnostics->used_model_fallback = w == 0;
    fill_diagnostics(diagnostics, *out);
    diagnostics->planning_seconds = monotonic_seconds() - start;
  }
  return MOFFT_SUCCESS;
}

static mofft_status create_with_radices(mofft_plan_common **out,
                                        size_t length, size_t batch,
                                        int precision_bits, int direction,
                                        mofft_placement placement,
                                        const int *radices, size_t count) {
  if (!out || (direction != MOFFT_FORWARD && direction != MOFFT_BACKWARD) ||
      (placement != MOFFT_IN_PLACE && placement != MOFFT_OUT_OF_PLACE) ||
      batch == 0)
    return MOFFT_INVALID_ARGUMENT;
  *out = NULL;
  mofft_status st = validate_chain(length, radices, count);
  if (st != MOFFT_SUCCESS) return st;
  return plan_from_chain(out, length, batch, precision_bits, direction,
                         placement, radices, count, 0, 0);
}

mofft_status mofft_plan_create_f32(mofft_plan_f32 **plan, size_t length,
                                   mofft_direction direction,
                                   mofft_placement placement) {
  if (!plan) return MOFFT_INVALID_ARGUMENT;
  return create_with_options((mofft_plan_common **)(void *)plan, length, 1, 32,
                             direction, placement, NULL, NULL);
}

mofft_status mofft_plan_create_f64(mofft_plan_f64 **plan, size_t length,
                                   mofft_direction direction,
                                   mofft_placement placement) {
  if (!plan) return MOFFT_INVALID_ARGUMENT;
  return create_with_options((mofft_plan_common **)(void *)plan, length, 1, 64,
                             direction, placement, NULL, NULL);
}

mofft_status mofft_plan_create_with_options_f32(
    mofft_plan_f32 **plan, size_t length, mofft_direction direction,
    mofft_placement placement, const mofft_plan_options *options,
    mofft_plan_diagnostics *diagnostics) {
  return create_with_options((mofft_plan_common **)(void *)plan, length, 1, 32,
                             direction, placement, options, diagnostics);
}

mofft_status mofft_plan_create_with_options_f64(
    mofft_plan_f64 **plan, size_t length, mofft_direction direction,
    mofft_placement placement, const mofft_plan_options *options,
    mofft_plan_diagnostics *diagnostics) {
  return create_with_options((mofft_plan_common **)(void *)plan, length, 1, 64,
                             direction, placement, options, diagnostics);
}

/* MOFFT_PLAN_MFMA (FP32): the analytical chain (its passes are the
 * radix-16-led registry kernels) on the matrix-core kernel bank. */
static mofft_status create_mfma(mofft_plan_common **out, size_t length,
                                size_t batch, mofft_direction direction,
                                mofft_placement placement,
                                mofft_plan_diagnostics *diagnostics) {
  if (!out) return MOFFT_INVALID_ARGUMENT;
  *out = NULL;
  if (diagnostics) memset(diagnostics, 0, sizeof(*diagnostics));
  if ((direction != MOFFT_FORWARD && direction != MOFFT_BACKWARD) ||
      (placement != MOFFT_IN_PLACE && placement != MOFFT_OUT_OF_PLACE) ||
      batch == 0)
    return MOFFT_INVALID_ARGUMENT;
  if (!mofft_hip_device_available()) return MOFFT_UNSUPPORTED_HARDWARE;
  int radices[MOFFT_MAX_PLAN_DEPTH];
  size_t depth = 0;
  uint64_t mask = 0;
  const double start = monotonic_seconds();
  /* the nominal chain: radix-16 sub-stages, whose kernels have
   * matrix-core variants (the radix-32 whole-transform kernels are slower
   * at 1024-4096 and 16384 points and 6 % faster only at 8192; a plan
   * selects them with their chain and kernel bank 3) */
  mofft_status st = model_chain_ex(PASS_KERNELS, length, 32, radices, &depth,
                                   &mask, 0);
  if (st != MOFFT_SUCCESS) return st;
  st = plan_from_chain(out, length, batch, 32, direction, placement, radices,
                       depth, mask, MOFFT_MFMA_BANK);
  if (st != MOFFT_SUCCESS) return st;
  if (diagnostics) {
    diagnostics->used_model_fallback = 1;
    fill_diagnostics(diagnostics, *out);
    diagnostics->planning_seconds = monotonic_seconds() - start;
  }
  return MOFFT_SUCCESS;
}

mofft_status mofft_plan_create_batched_f32(
    mofft_plan_f32 **plan, size_t length, size_t batch,
    mofft_direction direction, mofft_placement placement,
    const mofft_plan_options *options, mofft_plan_diagnostics *diagnostics) {
  if (options && (options->flags & MOFFT_PLAN_MFMA)) {
    if (options->mode != MOFFT_PLANNER_PRETUNED ||
        (options->flags & ~(uint32_t)MOFFT_PLAN_MFMA) != 0) {
      if (plan) *plan = NULL;
      return MOFFT_INVALID_ARGUMENT;
    }
    return create_mfma((mofft_plan_common **)(void *)plan, length, batch,
                       direction, placement, diagnostics);
  }
  return create_with_options((mofft_plan_common **)(void *)plan, length, batch,
                             32, direction, placement, options, diagnostics);
}

mofft_status mofft_plan_create_batched_f64(
    mofft_plan_f64 **plan, size_t length, size_t batch,
    mofft_direction direction, mofft_placement placement,
    const mofft_plan_options *options, mofft_plan_diagnostics *diagnostics) {
  return create_with_options((mofft_plan_common **)(void *)plan, length, batch,
                             64, direction, placement, options, diagnostics);
}

mofft_status mofft_plan_create_with_radices_f32(
    mofft_plan_f32 **plan, size_t length, mofft_direction direction,
    mofft_placement placement, const int *radices, size_t radix_count) {
  return create_with_radices((mofft_plan_common **)(void *)plan, length, 1, 32,
                             direction, placement, radices, radix_count);
}

mofft_status mofft_plan_create_with_radices_f64(
    mofft_plan_f64 **plan, size_t length, mofft_direction direction,
    mofft_placement placement, const int *radices, size_t radix_count) {
  return create_with_radices((mofft_plan_common **)(void *)plan, length, 1, 64,
                             direction, placement, radices, radix_count);
}

mofft_status mofft_plan_create_batched_with_radices_f32(
    mofft_plan_f32 **plan, size_t length, size_t batch,
    mofft_direction direction, mofft_placement placement, const int *radices,
    size_t radix_count) {
  return create_with_radices((mofft_plan_common **)(void *)plan, length, batch,
                             32, direction, placement, radices, radix_count);
}

mofft_status mofft_plan_create_batched_with_radices_f64(
    mofft_plan_f64 **plan, size_t length, size_t batch,
    mofft_direction direction, mofft_placement placement, const int *radices,
    size_t radix_count) {
  return create_with_radices((mofft_plan_common **)(void *)plan, length, batch,
                             64, direction, placement, radices, radix_count);
}

static mofft_plan_common *bound_plan(size_t length, const void *input,
                                     void *output, mofft_direction direction,
                                     uint32_t flags, int precision_bits) {
  if (!input || !output) return NULL;
  mofft_plan_options options;
  if (!options_from_bound_flags(flags, &options)) return NULL;
  mofft_plan_common *p = NULL;
  const mofft_placement placement =
      input == output ? MOFFT_IN_PLACE : MOFFT_OUT_OF_PLACE;
  if (create_with_options(&p, length, 1, precision_bits, direction, placement,
                          &options, NULL) != MOFFT_SUCCESS)
    return NULL;
  p->bound_input = input;
  p->bound_output = output;
  p->has_bound_arrays = 1;
  return p;
}

mofft_plan_f32 *mofft_plan_dft_1d_f32(size_t length,
                                      const mofft_complex_f32 *input,
                                      mofft_complex_f32 *output,
                                      mofft_direction direction,
                                      uint32_t flags) {
  return (mofft_plan_f32 *)bound_plan(length, input, output, direction, flags, 32);
}

mofft_plan_f64 *mofft_plan_dft_1d_f64(size_t length,
                                      const mofft_complex_f64 *input,
                                      mofft_complex_f64 *output,
                                      mofft_direction direction,
                                      uint32_t flags) {
  return (mofft_plan_f64 *)bound_plan(length, input, output, direction, flags, 64);
}

mofft_status mofft_execute_plan_f32(const mofft_plan_f32 *plan) {
  if (!plan || !plan->c.has_bound_arrays) return MOFFT_INVALID_ARGUMENT;
  return execute_sync(&plan->c, plan->c.bound_input, plan->c.bound_output);
}

mofft_status mofft_execute_plan_f64(const mofft_plan_f64 *plan) {
  if (!plan || !plan->c.has_bound_arrays) return MOFFT_INVALID_ARGUMENT;
  return execute_sync(&plan->c, plan->c.bound_input, plan->c.bound_output);
}

mofft_status mofft_execute_f32(const mofft_plan_f32 *plan,
                               const mofft_complex_f32 *input,
                               mofft_complex_f32 *output) {
  if (!plan) return MOFFT_INVALID_ARGUMENT;
  return execute_sync(&plan->c, input, output);
}

mofft_status mofft_execute_f64(const mofft_plan_f64 *plan,
                               const mofft_complex_f64 *input,
                               mofft_complex_f64 *output) {
  if (!plan) return MOFFT_INVALID_ARGUMENT;
  return execute_sync(&plan->c, input, output);
}

static mofft_status execute_async(const mofft_plan_common *p, const void *in,
                                  void *out, void *work, size_t work_bytes,
                                  int own_workspace, void *stream) {
  if (!p || !in || !out) return MOFFT_INVALID_ARGUMENT;
  if ((p->placement == MOFFT_IN_PLACE) != (in == out))
    return MOFFT_INVALID_ARGUMENT;
  if (!mofft_hip_is_device_pointer(in) || !mofft_hip_is_device_pointer(out))
    return MOFFT_INVALID_ARGUMENT;
  if (own_workspace) {
    work = p->workspace;
    work_bytes = schedule_workspace(&p->s, p->placement);
  }
  return execute_device(p, in, out, work, work_bytes, stream);
}

mofft_status mofft_execute_async_f32(const mofft_plan_f32 *plan,
                                     const mofft_complex_f32 *input,
                                     mofft_complex_f32 *output, void *stream) {
  return execute_async(plan ? &plan->c : NULL, input, output, NULL, 0, 1, stream);
}

mofft_status mofft_execute_async_f64(const mofft_plan_f64 *plan,
                                     const mofft_complex_f64 *input,
                                     mofft_complex_f64 *output, void *stream) {
  return execute_async(plan ? &plan->c : NULL, input, output, NULL, 0, 1, stream);
}

mofft_status mofft_plan_workspace_size_f32(const mofft_plan_f32 *plan,
                                           size_t *bytes) {
  if (!plan || !bytes) return MOFFT_INVALID_ARGUMENT;
  *bytes = schedule_workspace(&plan->c.s, plan->c.placement);
  return MOFFT_SUCCESS;
}

mofft_status mofft_plan_workspace_size_f64(const mofft_plan_f64 *plan,
                                           size_t *bytes) {
  if (!plan || !bytes) return MOFFT_INVALID_ARGUMENT;
  *bytes = schedule_workspace(&plan->c.s, plan->c.placement);
  return MOFFT_SUCCESS;
}

mofft_status mofft_execute_async_workspace_f32(
    const mofft_plan_f32 *plan, const mofft_complex_f32 *input,
    mofft_complex_f32 *output, void *workspace, size_t workspace_bytes,
    void *stream) {
  return execute_async(plan ? &plan->c : NULL, input, output, workspace,
                       workspace_bytes, 0, stream);
}

mofft_status mofft_execute_async_workspace_f64(
    const mofft_plan_f64 *plan, const mofft_complex_f64 *input,
    mofft_complex_f64 *output, void *workspace, size_t workspace_bytes,
    void *stream) {
  return execute_async(plan ? &plan->c : NULL, input, output, workspace,
                       workspace_bytes, 0, stream);
}

/* The allocation-free execute: the plan's workspace replaced by the
 * library's per-device scratch array (MOFFT_SCRATCH_BYTES), chunks shrunk
 * to what it holds.  No device allocation and no synchronisation. */
_Static_assert(MOFFT_SCRATCH_BYTES == MOFFT_SCRATCH_MAX_BYTES,
               "one scratch size in the ABI and the public header");

static mofft_status execute_scratch(const mofft_plan_common *p, const void *in,
                                    void *out, void *stream) {
  if (!p || !in || !out) return MOFFT_INVALID_ARGUMENT;
  if ((p->placement == MOFFT_IN_PLACE) != (in == out))
    return MOFFT_INVALID_ARGUMENT;
  if (!mofft_hip_is_device_pointer(in) || !mofft_hip_is_device_pointer(out))
    return MOFFT_INVALID_ARGUMENT;
  const size_t need = schedule_workspace(&p->s, p->placement);
  if (need == 0)
    return launch_schedule(p, in, out, NULL, stream) == 0 ? MOFFT_SUCCESS
                                                          : MOFFT_INTERNAL_ERROR;
  /* workspace per transform: one transform's intermediate */
  const size_t per = need / p->s.chunk;
  void *scratch = NULL;
  size_t scratch_bytes = 0;
  if (mofft_hip_scratch(&scratch, &scratch_bytes) != 0 || !scratch)
    return MOFFT_INTERNAL_ERROR;
  if (per == 0 || per > scratch_bytes) return MOFFT_INVALID_ARGUMENT;
  size_t chunk = scratch_bytes / per;
  if (chunk > p->s.chunk) chunk = p->s.chunk;
  return launch_schedule_chunks(p, in, out, scratch, stream, p->s.batch, chunk,
                                NULL) == 0
             ? MOFFT_SUCCESS
             : MOFFT_INTERNAL_ERROR;
}

mofft_status mofft_execute_async_scratch_f32(const mofft_plan_f32 *plan,
                                             const mofft_complex_f32 *input,
                                             mofft_complex_f32 *output,
                                             void *stream) {
  return execute_scratch(plan ? &plan->c : NULL, input, output, stream);
}

mofft_status mofft_execute_async_scratch_f64(const mofft_plan_f64 *plan,
                                             const mofft_complex_f64 *input,
                                             mofft_complex_f64 *output,
                                             void *stream) {
  return execute_scratch(plan ? &plan->c : NULL, input, output, stream);
}

mofft_status mofft_plan_get_batch_f32(const mofft_plan_f32 *plan,
                                      size_t *batch) {
  if (!plan || !batch) return MOFFT_INVALID_ARGUMENT;
  *batch = plan->c.s.batch;
  return MOFFT_SUCCESS;
}

mofft_status mofft_plan_get_batch_f64(const mofft_plan_f64 *plan,
                                      size_t *batch) {
  if (!plan || !batch) return MOFFT_INVALID_ARGUMENT;
  *batch = plan->c.s.batch;
  return MOFFT_SUCCESS;
}

void mofft_plan_destroy_f32(mofft_plan_f32 *plan) {
  plan_release(plan ? &plan->c : NULL);
}

void mofft_plan_destroy_f64(mofft_plan_f64 *plan) {
  plan_release(plan ? &plan->c : NULL);
}

/* ------------------------------------------------------------------ */
/* plan description (host only)                                         */
/* ------------------------------------------------------------------ */

static void describe_schedule(const mofft_schedule *s,
                              mofft_placement placement,
                              mofft_plan_description *d) {
  memset(d, 0, sizeof(*d));
  d->length = s->length;
  d->batch = s->batch;
  d->precision_bits = s->precision_bits;
  d->radix_count = (size_t)s->depth;
  memcpy(d->radices, s->factors, (size_t)s->depth * sizeof(int));
  d->pass_mask = s->pass_mask;
  d->pass_count = s->pass_count;
  for (int i = 0; i < s->pass_count && i < MOFFT_MAX_PASSES; ++i) {
    d->pass_length[i] = s->passes[i].length;
    d->pass_stride[i] = s->passes[i].stride;
    d->pass_span[i] = s->passes[i].span;
    d->pass_kernel[i] = s->passes[i].kernel;
  }
  d->chunk_transforms = s->chunk;
  d->workspace_bytes = schedule_workspace(s, placement);
  d->twiddle_bytes = s->twiddle_count * complex_bytes(s->precision_bits);
}

mofft_status mofft_plan_describe(size_t length, size_t batch,
                                 int precision_bits, const int *radices,
                                 size_t radix_count, uint64_t pass_mask,
                                 mofft_plan_description *description) {
  if (!description || batch == 0 ||
      (precision_bits != 32 && precision_bits != 64))
    return MOFFT_INVALID_ARGUMENT;
  int chain[MOFFT_MAX_PLAN_DEPTH];
  size_t depth = radix_count;
  uint64_t mask = pass_mask;
  unsigned int bank = 0;
  if (!radices) {
    if (length < 2) return MOFFT_UNSUPPORTED_SIZE;
    size_t d2 = 0;
    int direct = 0, layout = 0;
    const int w = wisdom_lookup(length, precision_bits, chain, &d2, &mask,
                                &direct, &layout, &bank);
    if (w < 0) return MOFFT_WISDOM_NOT_FOUND;
    if (w == 0) {
      mofft_status st = model_chain(PASS_KERNELS, length, precision_bits, chain,
                                    &d2, &mask);
      if (st != MOFFT_SUCCESS) return st;
      bank = model_bank(length, precision_bits, chain, d2, mask);
    }
    depth = d2;
  } else {
    if (radix_count == 0 || radix_count > MOFFT_MAX_PLAN_DEPTH)
      return MOFFT_INVALID_ARGUMENT;
    memcpy(chain, radices, radix_count * sizeof(int));
  }
  mofft_schedule s;
  mofft_status st = build_schedule(PASS_KERNELS, length, batch, precision_bits,
                                   chain, depth, mask, bank, &s);
  if (st != MOFFT_SUCCESS) return st;
  describe_schedule(&s, MOFFT_IN_PLACE, description);
  return MOFFT_SUCCESS;
}

double mofft_estimated_schedule_seconds(size_t length, size_t batch,
                                        int precision_bits, const int *radices,
                                        size_t radix_count, uint64_t pass_mask,
                                        unsigned int kernel_bank) {
  if (!radices || radix_count == 0 || radix_count > MOFFT_MAX_PLAN_DEPTH ||
      batch == 0 || (precision_bits != 32 && precision_bits != 64) ||
      kernel_bank >= kernel_bank_count())
    return DBL_MAX;
  mofft_schedule s;
  if (build_schedule(PASS_KERNELS, length, batch, precision_bits, radices,
                     radix_count, pass_mask, kernel_bank, &s) != MOFFT_SUCCESS)
    return DBL_MAX;
  return schedule_seconds(&s);
}

int mofft_estimated_pass_costs_measured(void) { return pass_costs_measured(); }

mofft_status mofft_plan_get_description_f32(const mofft_plan_f32 *plan,
                                            mofft_plan_description *out) {
  if (!plan || !out) return MOFFT_INVALID_ARGUMENT;
  describe_schedule(&plan->c.s, plan->c.placement, out);
  return MOFFT_SUCCESS;
}

mofft_status mofft_plan_get_description_f64(const mofft_plan_f64 *plan,
                                            mofft_plan_description *out) {
  if (!plan || !out) return MOFFT_INVALID_ARGUMENT;
  describe_schedule(&plan->c.s, plan->c.placement, out);
  return MOFFT_SUCCESS;
}

/* ------------------------------------------------------------------ */
/* expert setters (include/mofft.h:146-176)                             */
/* ------------------------------------------------------------------ */

/* Changing the pass grouping rebuilds the schedule and twiddle tables. */
static mofft_status regroup_layout(mofft_plan_common *p, uint64_t mask,
                                   unsigned int bank,
                                   mofft_stage_layout layout) {
  mofft_schedule s;
  mofft_status st = build_schedule_layout(
      PASS_KERNELS, p->s.length, p->s.batch, p->s.precision_bits, p->s.factors,
      (size_t)p->s.depth, mask, bank, layout == MOFFT_STAGE_LAYOUT_SECTION, &s);
  if (st != MOFFT_SUCCESS) return st;
  int same = s.pass_count == p->s.pass_count &&
             s.factored_twiddles == p->s.factored_twiddles;
  for (int i = 0; same && i < s.pass_count; ++i)
    same = s.passes[i].kernel == p->s.passes[i].kernel;
  if (same) {
    p->s.pass_mask = s.pass_mask;
    p->s.bank = bank;
    p->kernel_bank = bank;
    p->stage_layout = layout;
    return MOFFT_SUCCESS;
  }
  /* graphs captured earlier keep their reference to the old tables */
  st = make_tables(p, &s);
  if (st != MOFFT_SUCCESS) return st;
  p->s = s;
  p->kernel_bank = bank;
  p->stage_layout = layout;
  return MOFFT_SUCCESS;
}

static mofft_status regroup(mofft_plan_common *p, uint64_t mask,
                            unsigned int bank) {
  return regroup_layout(p, mask, bank, p->stage_layout);
}

static mofft_status set_mask(mofft_plan_common *p, uint64_t mask) {
  if (!p) return MOFFT_INVALID_ARGUMENT;
  const int transpose_count = p->s.depth > 1 ? p->s.depth - 1 : 0;
  if (transpose_count < 64 && (mask >> transpose_count)) return MOFFT_INVALID_ARGUMENT;
  return regroup(p, mask, p->kernel_bank);
}

static mofft_status set_strategy(mofft_plan_common *p,
                                 mofft_transpose_strategy strategy) {
  if (!p || (strategy != MOFFT_TRANSPOSE_LINEAR &&
             strategy != MOFFT_TRANSPOSE_BLOCKED))
    return MOFFT_INVALID_ARGUMENT;
  const uint64_t all = p->s.depth <= 1 ? 0
                                       : ((uint64_t)1 << (p->s.depth - 1)) - 1;
  return regroup(p, strategy == MOFFT_TRANSPOSE_BLOCKED ? all : 0,
                 p->kernel_bank);
}

mofft_status mofft_plan_set_transpose_strategy_f32(
    mofft_plan_f32 *plan, mofft_transpose_strategy strategy) {
  return set_strategy(plan ? &plan->c : NULL, strategy);
}

mofft_status mofft_plan_set_transpose_strategy_f64(
    mofft_plan_f64 *plan, mofft_transpose_strategy strategy) {
  return set_strategy(plan ? &plan->c : NULL, strategy);
}

mofft_status mofft_plan_set_blocked_transpose_mask_f32(mofft_plan_f32 *plan,
                                                       uint64_t levels) {
  return set_mask(plan ? &plan->c : NULL, levels);
}

mofft_status mofft_plan_set_blocked_transpose_mask_f64(mofft_plan_f64 *plan,
                                                       uint64_t levels) {
  return set_mask(plan ? &plan->c : NULL, levels);
}

/* direct_input and stage_layout select CPU data layouts in the reference
 * (src/mofft.c:1946-2009).  On the GPU direct_input is validated, kept,
 * reported and exported only (both input layouts are the same LDS pass);
 * stage_layout selects the layout of the pass twiddles: TRANSPOSE (the
 * default) a full table per later pass up to a size limit (one coalesced
 * load per point), SECTION the two-level factored table for every later
 * pass (two small-table loads and a complex product per point; faster for
 * some lengths, the autotuner records which). */
mofft_status mofft_plan_set_direct_input_f32(mofft_plan_f32 *plan, int enabled) {
  if (!plan || (enabled != 0 && enabled != 1)) return MOFFT_INVALID_ARGUMENT;
  plan->c.direct_input = enabled;
  return MOFFT_SUCCESS;
}

mofft_status mofft_plan_set_direct_input_f64(mofft_plan_f64 *plan, int enabled) {
  if (!plan || (enabled != 0 && enabled != 1)) return MOFFT_INVALID_ARGUMENT;
  plan->c.direct_input = enabled;
  return MOFFT_SUCCESS;
}

mofft_status mofft_plan_set_stage_layout_f32(mofft_plan_f32 *plan,
                                             mofft_stage_layout layout) {
  if (!plan || (layout != MOFFT_STAGE_LAYOUT_TRANSPOSE &&
                layout != MOFFT_STAGE_LAYOUT_SECTION))
    return MOFFT_INVALID_ARGUMENT;
  if (layout == plan->c.stage_layout) return MOFFT_SUCCESS;
  return regroup_layout(&plan->c, plan->c.s.pass_mask, plan->c.kernel_bank,
                        layout);
}

mofft_status mofft_plan_set_stage_layout_f64(mofft_plan_f64 *plan,
                                             mofft_stage_layout layout) {
  if (!plan || (layout != MOFFT_STAGE_LAYOUT_TRANSPOSE &&
                layout != MOFFT_STAGE_LAYOUT_SECTION))
    return MOFFT_INVALID_ARGUMENT;
  if (layout == plan->c.stage_layout) return MOFFT_SUCCESS;
  return regroup_layout(&plan->c, plan->c.s.pass_mask, plan->c.kernel_bank,
                        layout);
}

/* A kernel bank is a tile-shape family (see kernel_bank_count); changing
 * it rebuilds the schedule with the bank's kernels. */
mofft_status mofft_plan_set_kernel_bank_f32(mofft_plan_f32 *plan,
                                            unsigned int bank) {
  if (!plan || bank >= kernel_bank_count()) return MOFFT_INVALID_ARGUMENT;
  return regroup(&plan->c, plan->c.s.pass_mask, bank);
}

mofft_status mofft_plan_set_kernel_bank_f64(mofft_plan_f64 *plan,
                                            unsigned int bank) {
  if (!plan || bank >= kernel_bank_count()) return MOFFT_INVALID_ARGUMENT;
  return regroup(&plan->c, plan->c.s.pass_mask, bank);
}

mofft_status mofft_plan_get_kernel_bank_f32(const mofft_plan_f32 *plan,
                                            unsigned int *bank) {
  if (!plan || !bank) return MOFFT_INVALID_ARGUMENT;
  *bank = plan->c.kernel_bank;
  return MOFFT_SUCCESS;
}

mofft_status mofft_plan_get_kernel_bank_f64(const mofft_plan_f64 *plan,
                                            unsigned int *bank) {
  if (!plan || !bank) return MOFFT_INVALID_ARGUMENT;
  *bank = plan->c.kernel_bank;
  return MOFFT_SUCCESS;
}

/* ------------------------------------------------------------------ */
/* plan recipes (format of src/mofft.c:1769-1789, parser rules 1675-1758) */
/* ------------------------------------------------------------------ */

const char *mofft_build_id(void) { return MOFFT_BINARY_ID; }

static int recipe_id_valid(const char *id) {
  if (!id || strlen(id) != 64) return 0;
  for (size_t i = 0; i < 64; ++i)
    if (!((id[i] >= '0' && id[i] <= '9') || (id[i] >= 'a' && id[i] <= 'f')))
      return 0;
  return 1;
}

static int device_family(uint32_t *family) {
  if (!mofft_hip_device_available()) return 0;
  *family = MOFFT_GPU_FAMILY;
  return 1;
}

static int parse_uint(const char *token, int base, uint64_t *value) {
  if (!token || !*token) return 0;
  uint64_t v = 0;
  for (const char *c = token; *c; ++c) {
    int digit;
    if (*c >= '0' && *c <= '9') digit = *c - '0';
    else if (base == 16 && *c >= 'a' && *c <= 'f') digit = *c - 'a' + 10;
    else if (base == 16 && *c >= 'A' && *c <= 'F') digit = *c - 'A' + 10;
    else return 0;
    if (v > (UINT64_MAX - (uint64_t)digit) / (uint64_t)base) return 0;
    v = v * (uint64_t)base + (uint64_t)digit;
  }
  *value = v;
  return 1;
}

typedef struct {
  size_t length;
  int direction;
  mofft_placement placement;
  unsigned int bank;
  uint64_t mask;
  int direct_input;
  mofft_stage_layout layout;
  size_t depth;
  int radices[MOFFT_MAX_PLAN_DEPTH];
} recipe_fields;

static mofft_status parse_recipe(const char *recipe, size_t size,
                                 int precision_bits, recipe_fields *f) {
  if (!recipe || size < 2 || size > MOFFT_RECIPE_MAX_BYTES ||
      recipe[size - 1] != '\0' || recipe[size - 2] != '\n' ||
      memchr(recipe, '\0', size - 1) != NULL)
    return MOFFT_INVALID_ARGUMENT;
  char text[MOFFT_RECIPE_MAX_BYTES];
  for (size_t i = 0; i + 2 < size; ++i) {
    const unsigned char ch = (unsigned char)recipe[i];
    if (ch < 32 || ch > 126) return MOFFT_INVALID_ARGUMENT;
    text[i] = (char)ch;
  }
  text[size - 2] = '\0';
  if (size > 2 && (text[0] == ' ' || (size > 3 && text[size - 3] == ' ')))
    return MOFFT_INVALID_ARGUMENT;
  char *tokens[12 + MOFFT_MAX_PLAN_DEPTH];
  size_t n = 0;
  char *cur = text;
  while (*cur) {
    if (n == 12 + MOFFT_MAX_PLAN_DEPTH) return MOFFT_INVALID_ARGUMENT;
    tokens[n++] = cur;
    while (*cur && *cur != ' ') ++cur;
    if (*cur == ' ') {
      *cur++ = '\0';
      if (*cur == ' ') return MOFFT_INVALID_ARGUMENT;
    }
  }
  if (n < 13 || strcmp(tokens[0], "MOFFT_PLAN_V1") != 0 ||
      !recipe_id_valid(MOFFT_BINARY_ID) || strlen(tokens[1]) != 64 ||
      strcmp(tokens[1], MOFFT_BINARY_ID) != 0)
    return MOFFT_INCOMPATIBLE_RECIPE;
  if (strlen(tokens[2]) != 8 || strlen(tokens[8]) != 16)
    return MOFFT_INVALID_ARGUMENT;
  uint64_t family, length, placement, bank, mask, direct, layout, depth;
  if (!parse_uint(tokens[2], 16, &family) || family > UINT32_MAX ||
      !parse_uint(tokens[4], 10, &length) ||
      !parse_uint(tokens[6], 10, &placement) ||
      !parse_uint(tokens[7], 10, &bank) ||
      !parse_uint(tokens[8], 16, &mask) ||
      !parse_uint(tokens[9], 10, &direct) ||
      !parse_uint(tokens[10], 10, &layout) ||
      !parse_uint(tokens[11], 10, &depth))
    return MOFFT_INVALID_ARGUMENT;
  if ((precision_bits == 32 && strcmp(tokens[3], "f32") != 0) ||
      (precision_bits == 64 && strcmp(tokens[3], "f64") != 0))
    return MOFFT_INCOMPATIBLE_RECIPE;
  if ((strcmp(tokens[5], "-1") != 0 && strcmp(tokens[5], "1") != 0) ||
      (uint64_t)(size_t)length != length || length < 2 || placement > 1 ||
      bank >= kernel_bank_count() || direct > 1 || layout > 1 || depth == 0 ||
      depth > MOFFT_MAX_PLAN_DEPTH || n != 12 + (size_t)depth)
    return MOFFT_INVALID_ARGUMENT;
  if (depth - 1 < 64 && (mask >> (depth - 1)) != 0) return MOFFT_INVALID_ARGUMENT;
  f->length = (size_t)length;
  f->direction = strcmp(tokens[5], "-1") == 0 ? MOFFT_FORWARD : MOFFT_BACKWARD;
  f->placement = (mofft_placement)placement;
  f->bank = (unsigned int)bank;
  f->mask = mask;
  f->direct_input = (int)direct;
  f->layout = (mofft_stage_layout)layout;
  f->depth = (size_t)depth;
  size_t remaining = f->length;
  for (size_t i = 0; i < f->depth; ++i) {
    uint64_t r = 0;
    if (!parse_uint(tokens[12 + i], 10, &r) || r > INT32_MAX || r < 2 ||
        remaining % (size_t)r != 0)
      return MOFFT_INVALID_ARGUMENT;
    f->radices[i] = (int)r;
    remaining /= (size_t)r;
  }
  if (remaining != 1) return MOFFT_INVALID_ARGUMENT;
  uint32_t current = 0;
  if (!device_family(&current) || current != (uint32_t)family)
    return MOFFT_INCOMPATIBLE_RECIPE;
  return MOFFT_SUCCESS;
}

static mofft_status export_plan(const mofft_plan_common *p, int precision_bits,
                                char *buffer, size_t capacity,
                                size_t *required) {
  if (!p || !required) return MOFFT_INVALID_ARGUMENT;
  uint32_t family = 0;
  if (!device_family(&family) || !recipe_id_valid(MOFFT_BINARY_ID))
    return MOFFT_UNSUPPORTED_HARDWARE;
  char recipe[MOFFT_RECIPE_MAX_BYTES];
  int written = snprintf(recipe, sizeof(recipe),
                         "MOFFT_PLAN_V1 %s %08x f%d %zu %d %d %u %016llx %d %d %d",
                         MOFFT_BINARY_ID, family, precision_bits, p->s.length,
                         p->direction, (int)p->placement, p->kernel_bank,
                         (unsigned long long)p->s.pass_mask, p->direct_input,
                         (int)p->stage_layout, p->s.depth);
  if (written < 0 || (size_t)written >= sizeof(recipe)) return MOFFT_INTERNAL_ERROR;
  size_t used = (size_t)written;
  for (int i = 0; i < p->s.depth; ++i) {
    written = snprintf(recipe + used, sizeof(recipe) - used, " %d", p->s.factors[i]);
    if (written < 0 || (size_t)written >= sizeof(recipe) - used)
      return MOFFT_INTERNAL_ERROR;
    used += (size_t)written;
  }
  if (used + 2 > sizeof(recipe)) return MOFFT_INTERNAL_ERROR;
  recipe[used++] = '\n';
  recipe[used] = '\0';
  *required = used + 1;
  if (!buffer) return capacity == 0 ? MOFFT_SUCCESS : MOFFT_INVALID_ARGUMENT;
  if (capacity < *required) return MOFFT_BUFFER_TOO_SMALL;
  memcpy(buffer, recipe, *required);
  return MOFFT_SUCCESS;
}

static mofft_status import_plan(mofft_plan_common **out, const char *recipe,
                                size_t size, int precision_bits,
                                mofft_plan_diagnostics *diagnostics) {
  if (!out) return MOFFT_INVALID_ARGUMENT;
  *out = NULL;
  if (diagnostics) memset(diagnostics, 0, sizeof(*diagnostics));
  recipe_fields f;
  mofft_status st = parse_recipe(recipe, size, precision_bits, &f);
  if (st != MOFFT_SUCCESS) return st;
  if (validate_chain(f.length, f.radices, f.depth) != MOFFT_SUCCESS)
    return MOFFT_INVALID_ARGUMENT;
  st = plan_from_chain_layout(out, f.length, 1, precision_bits, f.direction,
                              f.placement, f.radices, f.depth, f.mask, f.bank,
                              f.layout);
  if (st != MOFFT_SUCCESS) { *out = NULL; return st; }
  (*out)->direct_input = f.direct_input;
  if (diagnostics) {
    diagnostics->used_wisdom = 1;
    fill_diagnostics(diagnostics, *out);
  }
  return MOFFT_SUCCESS;
}

mofft_status mofft_plan_export_f32(const mofft_plan_f32 *plan, char *buffer,
                                   size_t capacity, size_t *required) {
  return export_plan(plan ? &plan->c : NULL, 32, buffer, capacity, required);
}

mofft_status mofft_plan_export_f64(const mofft_plan_f64 *plan, char *buffer,
                                   size_t capacity, size_t *required) {
  return export_plan(plan ? &plan->c : NULL, 64, buffer, capacity, required);
}

mofft_status mofft_plan_import_f32(mofft_plan_f32 **plan, const char *recipe,
                                   size_t recipe_size,
                                   mofft_plan_diagnostics *diagnostics) {
  return import_plan((mofft_plan_common **)(void *)plan, recipe, recipe_size,
                     32, diagnostics);
}

mofft_status mofft_plan_import_f64(mofft_plan_f64 **plan, const char *recipe,
                                   size_t recipe_size,
                                   mofft_plan_diagnostics *diagnostics) {
  return import_plan((mofft_plan_common **)(void *)plan, recipe, recipe_size,
                     64, diagnostics);
}

/* ------------------------------------------------------------------ */
/* analytical cost model (src/mofft.c:133-325) re-fitted to MI355X       */
/* ------------------------------------------------------------------ */

/* Cost unit: nanoseconds on one MI355X.  A stage's kernel cost is the
 * codelet flop count of one butterfly (plus the 6(r-1) twiddle flops of a
 * non-first stage) over `lanes` butterflies at the FP32 (157.3 TFLOP/s) or
 * FP64 (78.6 TFLOP/s) vector peak.  Memory tiers: LDS 160 KiB/CU, L2 4 MiB
 * per XCD, Infinity Cache 256 MiB, HBM3E 8 TB/s. */
#define MODEL_LANES 64
#define MODEL_CACHE_LINE 128

/* flops of one radix-r butterfly: the generated codelet's operation count
 * (adds + multiplies of its straight-line code, emit.py codelet_flops) */
static double codelet_flops(int radix) {
  if (radix < 2 || radix > 64 || !is_supported_radix(radix)) return -1;
  const double f = mofft_codelet_flops[radix];
  return f > 0 ? f : -1;
}

double mofft_estimated_radix_cost(int radix, int precision_bits,
                                  int first_stage) {
  const double f = codelet_flops(radix);
  if (f < 0 || (precision_bits != 32 && precision_bits != 64)) return 1.0e300;
  const double flops = f + (first_stage ? 0.0 : 6.0 * (radix - 1));
  const double peak = precision_bits == 32 ? 157.3e3 : 78.6e3; /* flop/ns */
  return flops * MODEL_LANES / peak;
}

/* Memory tiers of the reference's cost profile (profiles/apple-m5-
 * bootstrap.json tiers, src/mofft.c:133-168) re-measured on one MI355X with
 * a 16-byte-per-lane copy / read / write probe (B/ns): HBM3E 5900 read,
 * 5700 write, 5300 copy (1 GiB buffers); 128 MiB buffers (a copy's source
 * and destination within the 256 MB Infinity Cache) 7600 copy.  LDS and L2
 * are nominal (aggregate LDS 128 B/clk/CU x 256 CUs x 2.4 GHz; L2 twice
 * the Infinity Cache). */
static double model_memory_cost(size_t working_set, size_t read, size_t write) {
  static const struct { size_t cap; double read_bw, write_bw, mixed_bw, lat; } tier[] = {
      {(size_t)160 << 10, 78643.0, 78643.0, 78643.0, 0.05},   /* LDS, all CUs */
      {(size_t)32 << 20, 15200.0, 15200.0, 15200.0, 0.5},     /* 8 x 4 MiB L2 */
      {(size_t)256 << 20, 7600.0, 7600.0, 7600.0, 1.0},       /* Infinity Cache */
      {SIZE_MAX, 5900.0, 5700.0, 5300.0, 2.0},                /* HBM3E, B/ns */
  };
  for (size_t i = 0; i < sizeof(tier) / sizeof(tier[0]); ++i) {
    if (working_set <= tier[i].cap) {
      const double r = (double)read / tier[i].read_bw;
      const double w = (double)write / tier[i].write_bw;
      const double m = (double)(read + write) / tier[i].mixed_bw;
      double c = r > w ? r : w;
      if (m > c) c = m;
      return c + tier[i].lat;
    }
  }
  return 1.0e300;
}

typedef struct {
  double kernel_cost, memory_cost, total_cost;
  size_t read, write, ws;
} stage_cost;

static int estimate_stage(int radix, int precision_bits, int first_stage,
                          size_t transform_length, size_t stage_length,
                          size_t extra_read, stage_cost *r) {
  if (radix < 2 || (precision_bits != 32 && precision_bits != 64) ||
      transform_length == 0 || stage_length == 0 ||
      stage_length > transform_length ||
      transform_length % (size_t)radix != 0 || !r)
    return 0;
  const double rc = mofft_estimated_radix_cost(radix, precision_bits, first_stage);
  if (rc >= 1.0e300) return 0;
  const size_t independent = transform_length / (size_t)radix;
  r->kernel_cost = rc * (double)((independent + MODEL_LANES - 1) / MODEL_LANES);
  const size_t cb = complex_bytes(precision_bits);
  const size_t data = sat_mul(transform_length, cb);
  const size_t tw = first_stage ? 0 : sat_mul(stage_length, cb);
  r->read = sat_add(data, tw);
  r->write = data;
  r->ws = sat_add(sat_mul(data, 2), tw);
  r->memory_cost = model_memory_cost(r->ws, sat_add(r->read, extra_read), r->write);
  r->total_cost = r->kernel_cost > r->memory_cost ? r->kernel_cost : r->memory_cost;
  return 1;
}

double mofft_estimated_stage_cost(int radix, int precision_bits,
                                  int first_stage, size_t transform_length,
                                  size_t stage_length) {
  stage_cost c;
  return estimate_stage(radix, precision_bits, first_stage, transform_length,
                        stage_length, 0, &c) ? c.total_cost : DBL_MAX;
}

static size_t overfetch(size_t bytes, size_t useful, size_t line) {
  if (useful >= line) return bytes;
  if (useful == 0 || bytes > SIZE_MAX / line) return SIZE_MAX;
  const size_t product = bytes * line;
  if (product > SIZE_MAX - (useful - 1)) return SIZE_MAX;
  return (product + useful - 1) / useful;
}

/* Same accounting as src/mofft.c:190-304 with the reference's default build
 * options (first-transpose fusion and repeated-twiddle broadcast on): the
 * logical/transferred byte fields are layout invariants of the reference
 * executor, the costs use the MI355X tiers above. */
mofft_status mofft_estimated_plan_cost_breakdown_with_layout(
    size_t length, int precision_bits, const int *radices, size_t radix_count,
    uint64_t blocked_transpose_mask, int direct_input,
    mofft_stage_layout layout, mofft_plan_cost_breakdown *breakdown) {
  if (!radices || radix_count == 0 || radix_count > MOFFT_MAX_PLAN_DEPTH ||
      length < 2 || (precision_bits != 32 && precision_bits != 64) ||
      (direct_input != 0 && direct_input != 1) ||
      (layout != MOFFT_STAGE_LAYOUT_TRANSPOSE &&
       layout != MOFFT_STAGE_LAYOUT_SECTION) ||
      blocked_transpose_mask >> (radix_count - 1) || !breakdown)
    return MOFFT_INVALID_ARGUMENT;
  memset(breakdown, 0, sizeof(*breakdown));
  const size_t lanes = precision_bits == 32 ? 16 : 8;
  const size_t cb = complex_bytes(precision_bits);
  const size_t data = sat_mul(length, cb);
  size_t n = length;
  for (size_t level = 0; level < radix_count; ++level) {
    const int radix = radices[level];
    if (!is_supported_radix(radix) || n % (size_t)radix != 0)
      return MOFFT_INVALID_ARGUMENT;
    const size_t rem = n / (size_t)radix;
    size_t extra = 0;
    const int fused = radix_count == 2 && level == 0;
    if (rem != 1 && layout == MOFFT_STAGE_LAYOUT_SECTION) {
      const size_t t = overfetch(data, sat_mul(rem, cb), MODEL_CACHE_LINE);
      extra = t > data ? t - data : 0;
    } else if (rem != 1 && !fused) {
      const size_t parent = length / n;
      if (direct_input && parent >= lanes / 2) {
        const size_t t = overfetch(data, sat_mul(parent, cb), MODEL_CACHE_LINE);
        extra = t > data ? t - data : 0;
      } else {
        const int blocked = (int)((blocked_transpose_mask >> level) & 1u);
        const size_t ws = sat_mul(data, 2);
        const size_t reads = blocked ? data
                                     : overfetch(data, sat_mul(parent, cb),
                                                 MODEL_CACHE_LINE);
        breakdown->layout_cost += model_memory_cost(ws, reads, data);
        breakdown->logical_read_bytes = sat_add(breakdown->logical_read_bytes, data);
        breakdown->logical_write_bytes = sat_add(breakdown->logical_write_bytes, data);
        breakdown->transferred_read_bytes = sat_add(breakdown->transferred_read_bytes, reads);
        breakdown->transferred_write_bytes = sat_add(breakdown->transferred_write_bytes, data);
        if (ws > breakdown->peak_working_set_bytes) breakdown->peak_working_set_bytes = ws;
      }
    }
    stage_cost sc;
    if (!estimate_stage(radix, precision_bits, rem == 1, length, n, extra, &sc))
      return MOFFT_INVALID_ARGUMENT;
    breakdown->kernel_cost += sc.kernel_cost;
    breakdown->memory_cost += sc.memory_cost;
    breakdown->total_cost += sc.total_cost;
    breakdown->logical_read_bytes = sat_add(breakdown->logical_read_bytes, sc.read);
    breakdown->logical_write_bytes = sat_add(breakdown->logical_write_bytes, sc.write);
    breakdown->transferred_read_bytes = sat_add(breakdown->transferred_read_bytes,
                                                sat_add(sc.read, extra));
    breakdown->transferred_write_bytes = sat_add(breakdown->transferred_write_bytes, sc.write);
    if (sc.ws > breakdown->peak_working_set_bytes) breakdown->peak_working_set_bytes = sc.ws;
    n = rem;
  }
  if (n != 1) return MOFFT_INVALID_ARGUMENT;
  breakdown->total_cost += breakdown->layout_cost;
  return MOFFT_SUCCESS;
}

double mofft_estimated_plan_cost_with_layout(
    size_t length, int precision_bits, const int *radices, size_t radix_count,
    uint64_t blocked_transpose_mask, int direct_input,
    mofft_stage_layout layout) {
  mofft_plan_cost_breakdown b;
  return mofft_estimated_plan_cost_breakdown_with_layout(
             length, precision_bits, radices, radix_count,
             blocked_transpose_mask, direct_input, layout, &b) == MOFFT_SUCCESS
             ? b.total_cost : DBL_MAX;
}

double mofft_estimated_plan_cost(size_t length, int precision_bits,
                                 const int *radices, size_t radix_count,
                                 uint64_t blocked_transpose_mask,
                                 int direct_input) {
  return mofft_estimated_plan_cost_with_layout(
      length, precision_bits, radices, radix_count, blocked_transpose_mask,
      direct_input, MOFFT_STAGE_LAYOUT_TRANSPOSE);
}

/* ------------------------------------------------------------------ */
/* strings                                                              */
/* ------------------------------------------------------------------ */

const char *mofft_status_string(mofft_status status) {
  switch (status) {
    case MOFFT_SUCCESS: return "success";
    case MOFFT_INVALID_ARGUMENT: return "invalid argument";
    case MOFFT_UNSUPPORTED_SIZE: return "unsupported size";
    case MOFFT_ALLOCATION_FAILURE: return "allocation failure";
    case MOFFT_UNSUPPORTED_HARDWARE: return "unsupported hardware";
    case MOFFT_INTERNAL_ERROR: return "internal error";
    case MOFFT_WISDOM_NOT_FOUND: return "compiled wisdom not found";
    case MOFFT_INCOMPATIBLE_RECIPE: return "incompatible plan recipe";
    case MOFFT_BUFFER_TOO_SMALL: return "output buffer too small";
    default: return "unknown status";
  }
}

const char *mofft_version(void) { return "0.4.0"; }

/* ------------------------------------------------------------------ */
/* internal entry points for the other runtime files (mofft_internal.h) */
/* ------------------------------------------------------------------ */

mofft_status mofft_internal_build_schedule(const mofft_kernel_set *ks,
                                           size_t length, size_t batch,
                                           int precision_bits,
                                           const int *radices, size_t depth,
                                           uint64_t mask, unsigned int bank,
                                           mofft_schedule *s) {
  return build_schedule(ks, length, batch, precision_bits, radices, depth,
                        mask, bank, s);
}

mofft_status mofft_internal_model_chain(const mofft_kernel_set *ks,
                                        size_t length, int precision_bits,
                                        int *radices, size_t *depth,
                                        uint64_t *mask) {
  return model_chain(ks, length, precision_bits, radices, depth, mask);
}

int mofft_internal_wisdom_lookup(size_t length, int precision_bits,
                                 int *radices, size_t *depth, uint64_t *mask) {
  return wisdom_lookup(length, precision_bits, radices, depth, mask, NULL,
                       NULL, NULL);
}

mofft_status mofft_internal_fill_twiddles(const mofft_schedule *s,
                                          void **twiddles) {
  return fill_schedule_twiddles(s, twiddles);
}

/* One device twiddle table, generated by the plan-time kernels and copied
 * to `host` (count complex values), for a bitwise check against correctly
 * rounded values.  kind 0: out[i] = w_a^((i * b) mod a), i < count (the
 * two-level pass tables); kind 1: the sub-stage table of ns = a, radix b
 * (count = a * (b - 1)); kind 2: the full pass table of ns = a, L = b
 * (count = a * b). */
mofft_status mofft_internal_twiddle_table(int kind, int precision_bits,
                                          int64_t a, int64_t b, void *host,
                                          size_t count) {
  if (!host || count == 0 || a < 1 || b < 1 ||
      (precision_bits != 32 && precision_bits != 64))
    return MOFFT_INVALID_ARGUMENT;
  if ((kind == 1 && count != (size_t)(a * (b - 1))) ||
      (kind == 2 && count != (size_t)(a * b)) || kind < 0 || kind > 2)
    return MOFFT_INVALID_ARGUMENT;
  if (!mofft_hip_device_available()) return MOFFT_UNSUPPORTED_HARDWARE;
  const size_t bytes = count * complex_bytes(precision_bits);
  void *dev = NULL;
  if (mofft_hip_malloc(&dev, bytes) != 0) return MOFFT_ALLOCATION_FAILURE;
  int rc;
  if (kind == 0)
    rc = mofft_hip_fill_powers(dev, precision_bits, (int64_t)count, a, b, NULL);
  else if (kind == 1)
    rc = mofft_hip_fill_stage(dev, precision_bits, a, (int)b, NULL);
  else
    rc = mofft_hip_fill_pass(dev, precision_bits, a, b, NULL);
  if (rc == 0) rc = mofft_hip_synchronize(NULL);
  if (rc == 0) rc = mofft_hip_copy(host, dev, bytes);
  mofft_hip_free(dev);
  return rc == 0 ? MOFFT_SUCCESS : MOFFT_INTERNAL_ERROR;
}

size_t mofft_internal_complex_bytes(int precision_bits) {
  return complex_bytes(precision_bits);
}

void mofft_internal_describe(const mofft_schedule *s, size_t workspace_bytes,
                             mofft_plan_description *d) {
  describe_schedule(s, MOFFT_IN_PLACE, d);
  d->workspace_bytes = workspace_bytes;
}

size_t mofft_internal_workspace_count(const void *plan, size_t count) {
  const mofft_plan_common *p = (const mofft_plan_common *)plan;
  if (!p || !schedule_workspace(&p->s, p->placement)) return 0;
  const size_t c = count < p->s.chunk ? count : p->s.chunk;
  return sat_mul(sat_mul(c, p->s.length), complex_bytes(p->s.precision_bits));
}

mofft_status mofft_internal_execute_count(const void *plan, const void *in,
                                          void *out, void *work, size_t count,
                                          void *stream) {
  const mofft_plan_common *p = (const mofft_plan_common *)plan;
  if (!p || !in || !out || count == 0 || count > p->s.batch)
    return MOFFT_INVALID_ARGUMENT;
  if ((p->placement == MOFFT_IN_PLACE) != (in == out))
    return MOFFT_INVALID_ARGUMENT;
  if (!work) work = p->workspace;
  if (mofft_internal_workspace_count(plan, count) && !work)
    return MOFFT_INVALID_ARGUMENT;
  return launch_schedule_count(p, in, out, work, stream, count) == 0
             ? MOFFT_SUCCESS
             : MOFFT_INTERNAL_ERROR;
}

/* Execute once with a timestamp between consecutive pass launches; the
 * per-launch durations are summed per pass index into pass_seconds. */
static mofft_status execute_timed(const mofft_plan_common *p, const void *in,
                                  void *out, void *stream,
                                  double *pass_seconds, size_t capacity) {
  if (!p || !in || !out || !pass_seconds ||
      capacity < (size_t)p->s.pass_count)
    return MOFFT_INVALID_ARGUMENT;
  if (!mofft_hip_is_device_pointer(in) || !mofft_hip_is_device_pointer(out))
    return MOFFT_INVALID_ARGUMENT;
  if ((p->placement == MOFFT_IN_PLACE) != (in == out))
    return MOFFT_INVALID_ARGUMENT;
  const size_t chunks = (p->s.batch + p->s.chunk - 1) / p->s.chunk;
  const size_t n_events = 1 + chunks * (size_t)p->s.pass_count;
  void **events = (void **)calloc(n_events, sizeof(void *));
  if (!events) return MOFFT_ALLOCATION_FAILURE;
  mofft_status st = MOFFT_SUCCESS;
  for (size_t k = 0; k < n_events && st == MOFFT_SUCCESS; ++k)
    if (mofft_hip_timing_event_create(&events[k]) != 0) st = MOFFT_INTERNAL_ERROR;
  launch_timing t = {events, n_events, 0};
  if (st == MOFFT_SUCCESS &&
      launch_schedule_timed(p, in, out, p->workspace, stream, p->s.batch, &t) != 0)
    st = MOFFT_INTERNAL_ERROR;
  if (st == MOFFT_SUCCESS && mofft_hip_synchronize(stream) != 0)
    st = MOFFT_INTERNAL_ERROR;
  for (int i = 0; i < p->s.pass_count; ++i) pass_seconds[i] = 0.0;
  for (size_t k = 0; k + 1 <= t.used && st == MOFFT_SUCCESS; ++k) {
    double sec = 0.0;
    if (mofft_hip_event_elapsed(events[k], events[k + 1], &sec) != 0)
      st = MOFFT_INTERNAL_ERROR;
    pass_seconds[k % (size_t)p->s.pass_count] += sec;
  }
  for (size_t k = 0; k < n_events; ++k) mofft_hip_event_destroy(events[k]);
  free(events);
  return st;
}

mofft_status mofft_execute_timed_f32(const mofft_plan_f32 *plan,
                                     const mofft_complex_f32 *input,
                                     mofft_complex_f32 *output, void *stream,
                                     double *pass_seconds, size_t capacity) {
  return execute_timed(plan ? &plan->c : NULL, input, output, stream,
                       pass_seconds, capacity);
}

mofft_status mofft_execute_timed_f64(const mofft_plan_f64 *plan,
                                     const mofft_complex_f64 *input,
                                     mofft_complex_f64 *output, void *stream,
                                     double *pass_seconds, size_t capacity) {
  return execute_timed(plan ? &plan->c : NULL, input, output, stream,
                       pass_seconds, capacity);
}

/* ------------------------------------------------------------------ */
/* captured plans (HIP graphs)                                          */
/* ------------------------------------------------------------------ */

struct mofft_graph {
  void *exec;
  mofft_device_tables *tables;  /* the plan tables the captured launches read */
};

static mofft_status graph_create(const mofft_plan_common *p, const void *in,
                                 void *out, mofft_graph **graph) {
  if (!graph) return MOFFT_INVALID_ARGUMENT;
  *graph = NULL;
  if (!p || !in || !out) return MOFFT_INVALID_ARGUMENT;
  if ((p->placement == MOFFT_IN_PLACE) != (in == out))
    return MOFFT_INVALID_ARGUMENT;
  if (!mofft_hip_is_device_pointer(in) || !mofft_hip_is_device_pointer(out))
    return MOFFT_INVALID_ARGUMENT;
  /* kernel attributes are set outside the capture */
  for (int i = 0; i < p->s.pass_count; ++i)
    if (mofft_hip_launch_pass(p->s.passes[i].kernel, NULL, NULL) != 0)
      return MOFFT_INTERNAL_ERROR;
  mofft_graph *g = (mofft_graph *)calloc(1, sizeof(*g));
  if (!g) return MOFFT_ALLOCATION_FAILURE;
  void *cs = NULL;
  if (mofft_hip_capture_begin(&cs) != 0) {
    free(g);
    return MOFFT_INTERNAL_ERROR;
  }
  const int rc = launch_schedule(p, in, out, p->workspace, cs);
  const int ec = mofft_hip_capture_end(cs, &g->exec);
  if (rc != 0 || ec != 0 || !g->exec) {
    mofft_hip_graph_destroy(g->exec);
    free(g);
    return MOFFT_INTERNAL_ERROR;
  }
  g->tables = p->tables;
  if (g->tables) g->tables->refs++;
  *graph = g;
  return MOFFT_SUCCESS;
}

mofft_status mofft_plan_graph_create_f32(const mofft_plan_f32 *plan,
                                         const mofft_complex_f32 *input,
                                         mofft_complex_f32 *output,
                                         mofft_graph **graph) {
  return graph_create(plan ? &plan->c : NULL, input, output, graph);
}

mofft_status mofft_plan_graph_create_f64(const mofft_plan_f64 *plan,
                                         const mofft_complex_f64 *input,
                                         mofft_complex_f64 *output,
                                         mofft_graph **graph) {
  return graph_create(plan ? &plan->c : NULL, input, output, graph);
}

mofft_status mofft_graph_launch(const mofft_graph *graph, void *stream) {
  if (!graph) return MOFFT_INVALID_ARGUMENT;
  return mofft_hip_graph_launch(graph->exec, stream) == 0 ? MOFFT_SUCCESS
                                                          : MOFFT_INTERNAL_ERROR;
}

void mofft_graph_destroy(mofft_graph *graph) {
  if (!graph) return;
  mofft_hip_graph_destroy(graph->exec);
  tables_drop(graph->tables);
  free(graph);
}
