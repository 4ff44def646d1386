/*
 * mofft_abi.h - the thin C ABI between the C runtime (mofft.c) and the HIP
 * layer (hip/device.hip + generated kernels).
 *
 * This is the MI355X replacement for the reference's streaming-mode entry
 * (src/streaming_entry.S:4-38) and generated dispatchers
 * (compiler/mofft_compiler/emitter.py:557-620): the runtime decides the
 * schedule on the host and calls these plain-C functions to allocate device
 * memory, fill twiddle tables and launch pass kernels.  No HIP or C++ type
 * crosses this boundary; streams travel as void *.
 */
#ifndef MOFFT_ABI_H
#define MOFFT_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arguments of one pass-kernel launch.  A pass computes, for every column
 * (b, j) with b < transforms and j < stride (= n / L), the L-point DFT of
 *   x_k = in[b*n + j + k*stride] * w_N^((j % ns) * k * tw_scale),  k < L,
 * and writes result o to out[b*n + (j/ns)*ns*L + (j%ns) + o*ns]
 * (one radix-L Stockham step of the whole n-point transform). */
typedef struct {
  const void *in;
  void *out;
  const void *stage_tw;   /* sub-stage twiddles of the L-point kernel */
  const void *tw_hi;      /* w_N^(h << split_shift) */
  const void *tw_lo;      /* w_N^l, l < 2^split_shift */
  const void *tw_pass;    /* full pass table: [k * ns + (j % ns)] =
                             w_(ns*L)^((j % ns) * k), used with
                             MOFFT_PASS_TWIDDLE_TABLE */
  int64_t n;
  int64_t stride;
  int64_t ns;
  int64_t tw_scale;       /* n / (ns * L) */
  int64_t columns;        /* transforms * stride */
  uint32_t split_shift;
  uint32_t flags;         /* MOFFT_PASS_* */
  /* Division by the launch-invariant divisors stride, ns and span = ns * L
   * as a multiply-high (Granlund and Montgomery, "Division by invariant
   * integers using multiplication", PLDI 1994, fig. 4.1): for a divisor
   * d >= 2, shift = ceil(log2 d) and magic = floor(2^32 (2^shift - d) / d)
   * + 1; then for every 32-bit x, t = mulhi(x, magic) and
   * x / d = (t + ((x - t) >> 1)) >> (shift - 1).  d = 1 is magic = shift =
   * 0.  Filled by mofft_pass_args_set_divisors(). */
  uint32_t stride_magic;
  uint32_t stride_shift;
  uint32_t ns_magic;
  uint32_t ns_shift;
  uint32_t span_magic;
  uint32_t span_shift;
} mofft_pass_args;

/* Element offsets inside one launch are 32-bit: columns * L (the elements
 * a launch touches) and n stay below this bound. */
#define MOFFT_PASS_MAX_ELEMENTS ((int64_t)1 << 31)

static inline void mofft_div_magic(uint64_t d, uint32_t *magic,
                                   uint32_t *shift) {
  uint32_t l = 0;
  if (d <= 1) {
    *magic = 0;
    *shift = 0;
    return;
  }
  while (((uint64_t)1 << l) < d) ++l;
  *magic = (uint32_t)(((((uint64_t)1 << l) - d) << 32) / d + 1);
  *shift = l;
}

/* Fill the divisor fields from stride, ns and the pass length L. */
static inline void mofft_pass_args_set_divisors(mofft_pass_args *a,
                                                int64_t L) {
  mofft_div_magic((uint64_t)a->stride, &a->stride_magic, &a->stride_shift);
  mofft_div_magic((uint64_t)a->ns, &a->ns_magic, &a->ns_shift);
  mofft_div_magic((uint64_t)(a->ns * L), &a->span_magic, &a->span_shift);
}

#define MOFFT_PASS_SWAP_IN 1u   /* exchange re/im on load (backward) */
#define MOFFT_PASS_SWAP_OUT 2u  /* exchange re/im on store (backward) */
#define MOFFT_PASS_TWIDDLE 4u   /* apply the w_N pass twiddle on load */
#define MOFFT_PASS_TWIDDLE_TABLE 8u  /* ... read from the coalesced full
                                        table instead of two-level lookups */
#define MOFFT_PASS_XCD_ORDER 16u /* strided tiles narrower than a cache
                                    line: workgroups on one XCD take the
                                    tiles that share lines */

/* One compiled pass kernel (host-side copy of the generated registry). */
typedef struct {
  int precision_bits;
  int length;
  int rows;               /* 1: columns are whole contiguous transforms */
  int threads_per_column;
  int columns_per_block;
  int stage_count;
  int stages[8];          /* execution order */
  int lds_bytes;
  int variant;            /* tile-shape family; plans select it as their
                             kernel bank */
} mofft_registry_entry;

extern const mofft_registry_entry mofft_registry[];
extern const int mofft_registry_count;

/* Device layer.  Return 0 on success, nonzero HIP error otherwise. */
int mofft_hip_device_available(void);
int mofft_hip_malloc(void **ptr, size_t bytes);
int mofft_hip_free(void *ptr);
int mofft_hip_is_device_pointer(const void *ptr);
int mofft_hip_copy(void *dst, const void *src, size_t bytes);
int mofft_hip_copy_async(void *dst, const void *src, size_t bytes,
                         void *stream);
int mofft_hip_synchronize(void *stream);
/* Bytes of the per-device scratch array behind the allocation-free
 * execute, and its device address on the current device (resolved once
 * per device; no allocation). */
#define MOFFT_SCRATCH_BYTES ((size_t)64 << 20)
int mofft_hip_scratch(void **ptr, size_t *bytes);
/* out[i] = w_M^((i * step) mod M) for i < count, rounded once from FP64. */
int mofft_hip_fill_powers(void *out, int precision_bits, int64_t count,
                          int64_t M, int64_t step, void *stream);
/* Stockham sub-stage table: out[a*(r-1) + k-1] = w_(ns*r)^(a*k),
 * a < ns, 1 <= k < r, rounded once from FP64. */
int mofft_hip_fill_stage(void *out, int precision_bits, int64_t ns, int r,
                         void *stream);
/* Pass table: out[k*ns + a] = w_(ns*L)^(a*k), a < ns, k < L, rounded once
 * from FP64. */
int mofft_hip_fill_pass(void *out, int precision_bits, int64_t ns, int64_t L,
                        void *stream);
/* args == NULL only sets the kernel's launch attributes (large dynamic
 * LDS), as the first launch would; nothing is enqueued. */
int mofft_hip_launch_pass(int kernel, const mofft_pass_args *args,
                          void *stream);
/* Device-side interval timer on `stream` (HIP events): begin records the
 * start, end records the stop, waits for it and returns seconds. */
int mofft_hip_timer_begin(void *stream);
int mofft_hip_timer_end(void *stream, double *seconds);

/* ---- persistent uneven batch (one launch, device work queue) ---- */

/* One transform of a batch whose whole DFT is a single row tile
 * (gathered: adjacent columns of a tile belong to different jobs). */
typedef struct {
  const void *in;
  void *out;
  uint32_t flags;         /* MOFFT_PASS_SWAP_IN | MOFFT_PASS_SWAP_OUT */
  uint32_t reserved;
} mofft_batch_job;

/* One pass description of the batch queue.  Three uses:
 *   - a gathered row segment (jobs != NULL): column c of its tiles is the
 *     whole transform jobs[c] (args.columns = job count, in/out unused);
 *     one queue item per tile;
 *   - a shared pass of a multi-pass transform (jobs == NULL, next < 0):
 *     one queue item per `group` consecutive tiles, items run by different
 *     workgroups; with dep >= 0 an item may start only once segment dep has
 *     finished dep_target items, and signal != 0 makes its own items count;
 *   - an owned transform (next >= 0 on its first pass): ONE queue item; the
 *     workgroup that takes it runs every tile of this pass, then of segment
 *     `next`, and so on, with a workgroup barrier between passes.
 * Only queue segments (the first segment_count of the array) carry queue
 * items; continuation passes of owned transforms follow them. */
typedef struct {
  mofft_pass_args args;
  int64_t first_item;     /* queue index of this segment's first item */
  const mofft_batch_job *jobs;
  int32_t kind;           /* mofft_batch_registry index */
  int32_t dep;
  uint32_t dep_target;
  uint32_t signal;
  int32_t next;
  uint32_t tiles;
  uint32_t group;         /* tiles per queue item (gathered, shared) */
  uint32_t reserved;
} mofft_batch_segment;

/* Kernel argument of the persistent batch kernel.  Every counter has a
 * 128-byte line of its own (MOFFT_BATCH_COUNTER_WORDS words apart), so the
 * queue head that every dequeue hits, the error flag that waits read and
 * the per-segment counters that producers release never share a line:
 * word MOFFT_BATCH_HEAD is the queue head, MOFFT_BATCH_ERROR the error flag
 * (nonzero: a wait exceeded spin_limit polls and the launch gave up),
 * MOFFT_BATCH_DONE(s) the finished items of segment s (s < counter_count).
 * The array holds MOFFT_BATCH_COUNTER_ARRAY(counter_count) words; the
 * launch clears all of them. */
#define MOFFT_BATCH_COUNTER_WORDS 32
#define MOFFT_BATCH_HEAD 0
#define MOFFT_BATCH_ERROR MOFFT_BATCH_COUNTER_WORDS
#define MOFFT_BATCH_DONE(s) (((int64_t)(s) + 2) * MOFFT_BATCH_COUNTER_WORDS)
#define MOFFT_BATCH_COUNTER_ARRAY(count) MOFFT_BATCH_DONE(count)
typedef struct {
  const mofft_batch_segment *segments;
  unsigned int *counters;
  int64_t segment_count;  /* queue segments */
  int64_t total_items;
  int64_t counter_count;
  uint64_t spin_limit;
} mofft_batch_header;

extern const mofft_registry_entry mofft_batch_registry[];
extern const int mofft_batch_registry_count;

/* Resident workgroups per CU of the batch kernel with `lds_bytes` of
 * dynamic LDS, and the CU count of the device. */
int mofft_hip_batch_occupancy(int precision_bits, int lds_bytes,
                              int *blocks_per_cu, int *cu_count);
/* Clear the counters, then run the persistent kernel with `grid`
 * workgroups (two kernel launches on `stream`, no memset node). */
int mofft_hip_batch_launch(int precision_bits, const mofft_batch_header *h,
                           int grid, int lds_bytes, void *stream);
/* Capture mofft_hip_batch_launch into an instantiated HIP graph. */
int mofft_hip_batch_graph(int precision_bits, const mofft_batch_header *h,
                          int grid, int lds_bytes, void **graph_exec);
int mofft_hip_graph_launch(void *graph_exec, void *stream);
/* Begin capturing a private stream; end the capture and instantiate. */
int mofft_hip_capture_begin(void **stream);
int mofft_hip_capture_end(void *stream, void **graph_exec);
int mofft_hip_graph_destroy(void *graph_exec);

/* ---- streams, events and device-side inputs (serving loop) ---- */

/* Streams are created with the default priority. */
int mofft_hip_stream_create(void **stream);
int mofft_hip_stream_destroy(void *stream);
int mofft_hip_event_create(void **event);
int mofft_hip_event_destroy(void *event);
int mofft_hip_event_record(void *event, void *stream);
/* 1 when the event has completed, 0 when not yet, negative on error. */
int mofft_hip_event_query(void *event);
/* An event that records timestamps, and the seconds between two of them
 * (both complete). */
int mofft_hip_timing_event_create(void **event);
int mofft_hip_event_elapsed(void *start, void *end, double *seconds);
int mofft_hip_event_synchronize(void *event);
/* Interleaved complex values from a seed: component c (0 = real,
 * 1 = imaginary) of element i is 2u - 1 with u the top 53 bits of
 * splitmix64(seed + (2i + c + 1) * 0x9E3779B97F4A7C15) times 2^-53, rounded
 * to the precision.  count = elements. */
int mofft_hip_fill_uniform(void *out, int precision_bits, int64_t count,
                           uint64_t seed, void *stream);

#ifdef __cplusplus
}
#endif
#endif
