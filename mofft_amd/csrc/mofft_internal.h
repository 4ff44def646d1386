/*
 * mofft_internal.h - host-side planning types shared by the runtime files
 * (mofft.c: single plans; mofft_batch.c: persistent uneven batches).
 * Not installed; nothing here is part of the public ABI.
 */
#ifndef MOFFT_INTERNAL_H
#define MOFFT_INTERNAL_H

#include "mofft_abi.h"
#include "../../include/mofft.h"
#include "../../include/mofft_hip.h"

#include <stddef.h>
#include <stdint.h>

/* A family of compiled kernels a schedule may use: the pass-kernel
 * registry (one launch per pass) or the persistent batch tile kinds. */
typedef struct {
  const mofft_registry_entry *entries;
  const int *count;
} mofft_kernel_set;

extern const mofft_kernel_set mofft_pass_kernel_set;
extern const mofft_kernel_set mofft_batch_kernel_set;

typedef struct {
  int kernel;            /* registry index */
  size_t length;         /* points per column in LDS */
  size_t stride;         /* n / length */
  size_t span;           /* product of earlier pass lengths (ns) */
  size_t stage_offset;   /* element offset of this pass's sub-stage table */
  size_t stage_count;
  size_t pass_offset;    /* element offset of the full pass twiddle table */
  int pass_table;        /* 1: the full table exists for this pass */
} mofft_pass;

/* Passes after the first multiply by w_(ns*L)^((j%ns)*k).  Up to this many
 * bytes per pass the factors come from a full table laid out so adjacent
 * columns read adjacent entries (one coalesced load per point, L2/Infinity
 * Cache resident); beyond it from a two-level table (two gathered loads and
 * a complex product per point).  Measured on MI355X (2^24 points per
 * launch): FP32 N = 2^16..2^20 run 3-7 % faster on full tables up to 8 MiB;
 * FP64 N = 2^19 runs 3.6 % slower on its 8 MiB table than on the two-level
 * one, so FP64 keeps full tables to 4 MiB. */
#define MOFFT_PASS_TABLE_BYTES ((size_t)8 << 20)
#define MOFFT_PASS_TABLE_BYTES_F64 ((size_t)4 << 20)

typedef struct {
  const mofft_kernel_set *ks;          /* kernel indices refer to this set */
  size_t length;
  size_t batch;
  int precision_bits;
  unsigned int bank;
  int depth;
  int factors[MOFFT_MAX_PLAN_DEPTH];   /* outer -> inner */
  uint64_t pass_mask;
  int pass_count;
  mofft_pass passes[MOFFT_MAX_PLAN_DEPTH];
  size_t chunk;
  size_t workspace_bytes;
  uint32_t split_shift;
  size_t tw_hi_offset, tw_hi_count;
  size_t tw_lo_offset, tw_lo_count;
  size_t twiddle_count;                /* complex elements */
  /* 1: every later pass reads its pass twiddles from the two-level
   * (factored) table, whatever the full table's size (the plan's
   * MOFFT_STAGE_LAYOUT_SECTION) */
  int factored_twiddles;
} mofft_schedule;

/* Measured data of the analytical planner (generated mofft_costs.c):
 * flops of one radix-r codelet (index r), and ns per byte moved of every
 * pass-kernel registry entry as a single pass (column 0), a strided first
 * pass at stride 16 / 64 / 256 / 1024 (columns 1-4) and a strided pass
 * with pass twiddles at span 16 / 64 / 256 / 1024 (columns 5-8);
 * 0 = not measured. */
#define MOFFT_ROLE_ROWS 0
#define MOFFT_ROLE_FIRST 1
#define MOFFT_ROLE_LATER 2
extern const double mofft_codelet_flops[65];
extern const double mofft_pass_cost_ns[][9];
extern const char *const mofft_pass_cost_source;

/* Chain + pass mask -> passes over kernels of `ks` (greedy longest match
 * from the innermost stage; boundaries in `mask` always split). */
mofft_status mofft_internal_build_schedule(const mofft_kernel_set *ks,
                                           size_t length, size_t batch,
                                           int precision_bits,
                                           const int *radices, size_t depth,
                                           uint64_t mask, unsigned int bank,
                                           mofft_schedule *s);
/* Analytical chain over kernels of `ks`: one pass when a row kernel covers
 * the length, else the fewest passes (largest smallest pass). */
mofft_status mofft_internal_model_chain(const mofft_kernel_set *ks,
                                        size_t length, int precision_bits,
                                        int *radices, size_t *depth,
                                        uint64_t *mask);
/* Compiled wisdom: 1 hit, 0 miss, -1 stale entry. */
int mofft_internal_wisdom_lookup(size_t length, int precision_bits,
                                 int *radices, size_t *depth, uint64_t *mask);
/* Allocate the schedule's twiddle tables on the device and fill them. */
mofft_status mofft_internal_fill_twiddles(const mofft_schedule *s,
                                          void **twiddles);
size_t mofft_internal_complex_bytes(int precision_bits);
/* A device-generated twiddle table copied to the host (kind 0: powers
 * w_a^((i*b) mod a); 1: sub-stage table of ns = a, radix b; 2: full pass
 * table of ns = a, L = b). */
mofft_status mofft_internal_twiddle_table(int kind, int precision_bits,
                                          int64_t a, int64_t b, void *host,
                                          size_t count);
/* MEASURE's result check on host copies of `count` complex values: 0 when
 * all are finite and (ref != NULL) within relative L2 `bound` of ref, 1 when
 * a value is not finite, 2 when they disagree. */
int mofft_internal_check_output(const void *out, const void *ref,
                                size_t count, int precision_bits,
                                double bound);
void mofft_internal_describe(const mofft_schedule *s, size_t workspace_bytes,
                             mofft_plan_description *d);
/* Tile order of the pass launches that follow (tests): 1 XCD-aware, 0
 * linear, -1 the default (MOFFT_TILE_ORDER).  Results never depend on it;
 * not for use while other threads launch. */
void mofft_internal_set_tile_order(int xcd);

/* Enqueue the first `count` (<= the plan's batch) transforms of a batched
 * plan (a mofft_plan_f32 * or mofft_plan_f64 *) on `stream`, device
 * buffers; `work` (NULL: the plan's own) holds
 * mofft_internal_workspace_count(plan, count) bytes. */
size_t mofft_internal_workspace_count(const void *plan, size_t count);
mofft_status mofft_internal_execute_count(const void *plan, const void *in,
                                          void *out, void *work, size_t count,
                                          void *stream);

#endif
