/*
 * mofft_hip.h - additive MI355X extensions to the MoFFT C ABI.
 *
 * The reference has one transform per plan and no device memory
 * (include/mofft.h:107-208); nothing here replaces a reference entry point.
 * These calls add what a GPU caller needs on top of mofft.h:
 *   - batched plans: `batch` transforms of `length` points stored back to
 *     back (transform b starts at element b*length);
 *   - stream-ordered execution on device pointers with no host sync;
 *   - a caller-supplied workspace so one plan can run on several streams;
 *   - host-only plan description (no device needed) for tooling and tests;
 *   - persistent uneven batches: many transforms of different lengths and
 *     directions in one launch fed by a device work queue.
 * Streams are passed as `void *` holding a hipStream_t (NULL = the null
 * stream), so this header needs no HIP include.
 */
#ifndef MOFFT_HIP_H
#define MOFFT_HIP_H

#include "mofft.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MOFFT_MAX_PASSES 16

/* Batched plan creation.  `batch` >= 1; batch == 1 is the mofft.h plan.
 * options->flags may add MOFFT_PLAN_MFMA (FP32, PRETUNED, no other flag):
 * the radix-16 sub-stages run on the matrix cores as split-precision FP16
 * products accumulated in FP32 (kernel bank 3; error stated per length by
 * the tests and the benchmark), with the analytical radix chain.  Radix-32
 * matrix-core sub-stages exist for the whole-transform chains 1024 =
 * 32 x 32, 2048 = 4 x 16 x 32, 4096 = 4 x 32 x 32, 8192 = 16 x 16 x 32 and
 * 16384 = 16 x 32 x 32 (radices outermost first): a plan made with that
 * chain (mofft_plan_create_batched_with_radices_f32) on kernel bank 3
 * (mofft_plan_set_kernel_bank_f32) runs them. */
#define MOFFT_PLAN_MFMA 0x100u
mofft_status mofft_plan_create_batched_f32(
    mofft_plan_f32 **plan, size_t length, size_t batch,
    mofft_direction direction, mofft_placement placement,
    const mofft_plan_options *options, mofft_plan_diagnostics *diagnostics);
mofft_status mofft_plan_create_batched_f64(
    mofft_plan_f64 **plan, size_t length, size_t batch,
    mofft_direction direction, mofft_placement placement,
    const mofft_plan_options *options, mofft_plan_diagnostics *diagnostics);

/* Batched plan with an explicit radix chain (outermost first). */
mofft_status mofft_plan_create_batched_with_radices_f32(
    mofft_plan_f32 **plan, size_t length, size_t batch,
    mofft_direction direction, mofft_placement placement,
    const int *radices, size_t radix_count);
mofft_status mofft_plan_create_batched_with_radices_f64(
    mofft_plan_f64 **plan, size_t length, size_t batch,
    mofft_direction direction, mofft_placement placement,
    const int *radices, size_t radix_count);

mofft_status mofft_plan_get_batch_f32(const mofft_plan_f32 *plan,
                                      size_t *batch);
mofft_status mofft_plan_get_batch_f64(const mofft_plan_f64 *plan,
                                      size_t *batch);

/* Enqueue the transform on `stream`.  Input and output must be device
 * pointers (MOFFT_INVALID_ARGUMENT otherwise).  Returns after enqueue; the
 * plan's own workspace is used, so executions of one plan must be ordered on
 * one stream. */
mofft_status mofft_execute_async_f32(const mofft_plan_f32 *plan,
                                     const mofft_complex_f32 *input,
                                     mofft_complex_f32 *output, void *stream);
mofft_status mofft_execute_async_f64(const mofft_plan_f64 *plan,
                                     const mofft_complex_f64 *input,
                                     mofft_complex_f64 *output, void *stream);

/* Bytes of device workspace an execution needs (0 when the schedule needs
 * none), and an execution that uses caller-owned workspace instead of the
 * plan's, so one plan may run concurrently on several streams. */
mofft_status mofft_plan_workspace_size_f32(const mofft_plan_f32 *plan,
                                           size_t *bytes);
mofft_status mofft_plan_workspace_size_f64(const mofft_plan_f64 *plan,
                                           size_t *bytes);
mofft_status mofft_execute_async_workspace_f32(
    const mofft_plan_f32 *plan, const mofft_complex_f32 *input,
    mofft_complex_f32 *output, void *workspace, size_t workspace_bytes,
    void *stream);
mofft_status mofft_execute_async_workspace_f64(
    const mofft_plan_f64 *plan, const mofft_complex_f64 *input,
    mofft_complex_f64 *output, void *workspace, size_t workspace_bytes,
    void *stream);

/* Allocation-free execute.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The execution of
 * mofft_execute_async_workspace_* with the library's own per-device scratch
 * array as the workspace: a static device array of MOFFT_SCRATCH_MAX_BYTES
 * that each GPU the library runs on holds until the process ends (its
 * address is resolved on a device's first call; no device allocation,
 * free or synchronisation happens in this call).  Device buffers only.  A
 * batch whose workspace exceeds the scratch runs in chunks that fit it; a
 * plan whose single transform needs more workspace than the scratch (in
 * place or three-pass plans of N x complex bytes > 64 MiB: FP64 N > 2^22,
 * FP32 N > 2^23) is rejected with MOFFT_INVALID_ARGUMENT before anything
 * runs.  Every call on one GPU shares that GPU's scratch, which keeps the
 * last call's contents (nothing in it needs to start at zero): calls that
 * overlap on one GPU - on different streams or from different host
 * threads - race; order them on one stream. */
#define MOFFT_SCRATCH_MAX_BYTES ((size_t)64 << 20)
mofft_status mofft_execute_async_scratch_f32(const mofft_plan_f32 *plan,
                                             const mofft_complex_f32 *input,
                                             mofft_complex_f32 *output,
                                             void *stream);
mofft_status mofft_execute_async_scratch_f64(const mofft_plan_f64 *plan,
                                             const mofft_complex_f64 *input,
                                             mofft_complex_f64 *output,
                                             void *stream);

/* A plan's launches on fixed device buffers captured into a HIP graph: one
 * replay per execution, for launch-bound loops.  The graph holds a
 * reference to the plan's twiddle tables and workspace as they were at
 * capture: a later setter that regroups the plan's passes gives the plan
 * new tables while the graph keeps replaying the captured schedule on the
 * old ones, and the graph stays valid after the plan is destroyed (the
 * tables are freed with the last of the two).  Replays of one graph must be
 * ordered on one stream. */
typedef struct mofft_graph mofft_graph;
mofft_status mofft_plan_graph_create_f32(const mofft_plan_f32 *plan,
                                         const mofft_complex_f32 *input,
                                         mofft_complex_f32 *output,
                                         mofft_graph **graph);
mofft_status mofft_plan_graph_create_f64(const mofft_plan_f64 *plan,
                                         const mofft_complex_f64 *input,
                                         mofft_complex_f64 *output,
                                         mofft_graph **graph);
mofft_status mofft_graph_launch(const mofft_graph *graph, void *stream);
void mofft_graph_destroy(mofft_graph *graph);

/* Execute once (device buffers) with a device timestamp between
 * consecutive pass launches, wait, and return in pass_seconds[p] the time of
 * pass p summed over the batch chunks (capacity >= the plan's pass count).
 * For profiling; the kernels are those of mofft_execute_async_*. */
mofft_status mofft_execute_timed_f32(const mofft_plan_f32 *plan,
                                     const mofft_complex_f32 *input,
                                     mofft_complex_f32 *output, void *stream,
                                     double *pass_seconds, size_t capacity);
mofft_status mofft_execute_timed_f64(const mofft_plan_f64 *plan,
                                     const mofft_complex_f64 *input,
                                     mofft_complex_f64 *output, void *stream,
                                     double *pass_seconds, size_t capacity);

/* The GPU schedule of a plan: the chain grouped into HBM passes.  Pass 0 is
 * executed first.  Each pass is one kernel launch per batch chunk that keeps
 * `pass_length` points of every column in LDS. */
typedef struct {
  size_t length;
  size_t batch;
  int precision_bits;
  size_t radix_count;
  int radices[MOFFT_MAX_PLAN_DEPTH];      /* outermost first */
  uint64_t pass_mask;                     /* bit l: HBM pass boundary
                                             before level l */
  int pass_count;
  size_t pass_length[MOFFT_MAX_PASSES];   /* points per column, pass order */
  size_t pass_stride[MOFFT_MAX_PASSES];   /* length / pass_length */
  size_t pass_span[MOFFT_MAX_PASSES];     /* product of earlier passes */
  int pass_kernel[MOFFT_MAX_PASSES];      /* kernel registry index */
  size_t chunk_transforms;                /* transforms per launch */
  size_t workspace_bytes;
  size_t twiddle_bytes;
} mofft_plan_description;

/* Plan on the host only: no device is touched, so this runs anywhere.  The
 * result is the schedule mofft_plan_create_batched_* would build with
 * PRETUNED options (radices == NULL) or with the given chain and mask. */
mofft_status mofft_plan_describe(size_t length, size_t batch,
                                 int precision_bits, const int *radices,
                                 size_t radix_count, uint64_t pass_mask,
                                 mofft_plan_description *description);
mofft_status mofft_plan_get_description_f32(const mofft_plan_f32 *plan,
                                            mofft_plan_description *out);
mofft_status mofft_plan_get_description_f64(const mofft_plan_f64 *plan,
                                            mofft_plan_description *out);

/* The analytical planner's cost model (MI355X counterpart of the
 * reference's generated cost functions, compiler/mofft_compiler/
 * emitter.py:621-725): predicted device seconds of `batch` transforms under
 * the schedule a chain, pass mask and kernel bank give - every pass moves
 * the batch once through HBM at its kernel's measured cost in its role
 * (single pass, first strided pass, strided pass with twiddles; the table
 * `mofft_tune --pass-costs` writes, compiled in), plus the generated
 * codelets' arithmetic at the VALU peak.  DBL_MAX for an invalid plan.  The
 * PRETUNED model plan (lengths without wisdom) and MEASURE's candidate
 * order are the lowest predictions. */
double mofft_estimated_schedule_seconds(size_t length, size_t batch,
                                        int precision_bits, const int *radices,
                                        size_t radix_count, uint64_t pass_mask,
                                        unsigned int kernel_bank);
/* 1 when the compiled pass-cost table holds measurements (else the planner
 * uses nominal pass weights). */
int mofft_estimated_pass_costs_measured(void);

/* Kernel registry introspection (host only). */
typedef struct {
  int precision_bits;
  int length;          /* points per column kept in LDS */
  int rows;            /* 1: columns are whole contiguous transforms */
  int threads_per_column;
  int columns_per_block;
  int stage_count;
  int stages[8];       /* execution order: stage 0 first (innermost) */
  int variant;         /* tile-shape family = the kernel bank selecting it */
} mofft_kernel_info;
int mofft_kernel_count(void);
mofft_status mofft_kernel_get_info(int index, mofft_kernel_info *info);

/* 1 when a gfx950 device is present and usable, else 0. */
int mofft_device_available(void);

/* ---- persistent uneven batch ------------------------------------------
 *
 * No counterpart in the reference, which executes one transform per call
 * (reference src/mofft.c:2169-2201).  A batch plan binds `count` transforms
 * of one precision, each with its own length, direction and device
 * buffers (in-place when input == output), and runs all of them in ONE
 * persistent kernel launch fed by a device-side work queue.  Plans,
 * twiddles, workspace and the queue stay resident on the device; by default
 * the launch is captured once into a HIP graph that every execution
 * replays.  Lengths: powers of two >= 2 (one row tile up to 4096 points, two
 * strided passes up to 2^18, three beyond).  A multi-pass transform is run
 * whole by one workgroup (owned) unless it is larger than `owned_tiles`;
 * then its passes are split into items of `item_tiles` tiles spread over
 * workgroups that synchronise through device counters (shared).  The
 * buffers of different transforms must not overlap.  The radix chain of each length
 * is the compiled wisdom's when its passes exist as batch tile kinds, else
 * the analytical planner's over those kinds. */
typedef struct mofft_batch_plan mofft_batch_plan;

#define MOFFT_BATCH_NO_GRAPH 1u   /* launch directly, no captured graph */
#define MOFFT_BATCH_SHARE_ALL 2u  /* every multi-pass transform is shared */

typedef struct {
  uint32_t flags;        /* MOFFT_BATCH_* */
  uint32_t grid;         /* workgroups; 0 = every resident slot */
  uint64_t spin_limit;   /* polls before a dependency wait gives up and
                            the execution reports MOFFT_INTERNAL_ERROR;
                            0 = default (about a second) */
  uint32_t owned_tiles;  /* largest multi-pass transform (in tiles) run
                            whole by one workgroup; 0 = automatic */
  uint32_t item_tiles;   /* tiles per queue item of shared passes;
                            0 = automatic */
} mofft_batch_options;

typedef struct {
  size_t job_count;
  size_t distinct_lengths;
  size_t multi_pass_jobs;
  size_t shared_jobs;         /* multi-pass transforms run as shared tiles */
  size_t owned_tiles;         /* the owned/shared threshold in effect */
  size_t item_tiles;          /* tiles per item of shared passes */
  size_t wisdom_lengths;      /* lengths whose chain came from wisdom */
  size_t segment_count;
  size_t item_count;          /* queue items of one execution */
  size_t grid;
  size_t blocks_per_cu;
  size_t lds_bytes;
  size_t workspace_bytes;
  size_t twiddle_bytes;
  double flops;               /* sum of 5 N log2 N */
  double algorithmic_bytes;   /* sum of 2 N sizeof(complex) */
  int precision_bits;
  int graph;                  /* 1: executions replay a captured graph */
} mofft_batch_description;

void mofft_batch_options_init(mofft_batch_options *options);
/* directions[i] is MOFFT_FORWARD or MOFFT_BACKWARD; inputs/outputs are
 * device pointers of lengths[i] complex values. */
mofft_status mofft_batch_plan_create(mofft_batch_plan **plan,
                                     int precision_bits, size_t count,
                                     const size_t *lengths,
                                     const int *directions,
                                     const void *const *inputs,
                                     void *const *outputs,
                                     const mofft_batch_options *options);
/* Rebind the buffers (same placements); no execution may be in flight. */
mofft_status mofft_batch_set_buffers(mofft_batch_plan *plan,
                                     const void *const *inputs,
                                     void *const *outputs);
/* Enqueue every transform on `stream`.  Executions of one plan share its
 * queue and must be ordered on one stream. */
mofft_status mofft_batch_execute_async(const mofft_batch_plan *plan,
                                       void *stream);
/* After the stream's executions completed: MOFFT_INTERNAL_ERROR if one of
 * them gave up on a dependency wait, else MOFFT_SUCCESS. */
mofft_status mofft_batch_check(const mofft_batch_plan *plan);
/* Execute on the null stream, wait, and check. */
mofft_status mofft_batch_execute(const mofft_batch_plan *plan);
mofft_status mofft_batch_describe(const mofft_batch_plan *plan,
                                  mofft_batch_description *description);
/* Host only: the schedule a batch plan uses for one length (pass_kernel
 * holds batch tile-kind indices). */
mofft_status mofft_batch_describe_length(int precision_bits, size_t length,
                                         mofft_plan_description *description);
void mofft_batch_plan_destroy(mofft_batch_plan *plan);

/* ---- serving loop ------------------------------------------------------
 *
 * No counterpart in the reference.  A server admits 1-D C2C requests of
 * mixed length, batch and precision against a device-memory budget and
 * runs them with continuous batching: requests that arrive while others
 * are on the GPU are admitted (FIFO) as soon as an in-flight slot and
 * budget free up, and run concurrently on several streams.  When a request
 * is admitted, later pending requests of the same length, precision and
 * direction whose input the server generates join it, up to `max_batch`
 * transforms: one arena block and one batched execution serve them all,
 * each still completed and released on its own.  Plans (and
 * their twiddles) stay resident per (length, precision, direction);
 * request buffers come from one device arena of `budget_bytes` allocated
 * at creation, so serving calls no device allocator.  Results are
 * out-of-place in server-owned buffers that stay valid until released.
 * A server is driven from one host thread at a time. */
typedef struct mofft_server mofft_server;

typedef struct {
  size_t budget_bytes;    /* device arena for request buffers */
  uint32_t streams;       /* device streams; 0 = 4 */
  uint32_t max_inflight;  /* requests on the device at once; 0 = 4 * streams */
  uint32_t max_batch;     /* largest batch of one request or execution;
                             0 = 64 */
  uint32_t flags;         /* MOFFT_SERVER_* */
} mofft_server_options;

/* Run every request as its own execution (no coalescing). */
#define MOFFT_SERVER_NO_COALESCE 1u

typedef struct {
  size_t length;
  size_t batch;
  int precision_bits;
  int direction;
  const void *input;      /* device buffer of length * batch values, or
                             NULL: generated on the device from `seed` by
                             the counter-based generator of the library
                             (mofft_server_generate) */
  uint64_t seed;
  uint64_t tag;           /* returned in the completion */
} mofft_request;

typedef struct {
  uint64_t ticket;
  uint64_t tag;
  int status;             /* mofft_status of the request */
  double submit_seconds;  /* host monotonic clock */
  double start_seconds;   /* admitted and enqueued on a stream */
  double done_seconds;    /* completion observed */
  const void *output;     /* result (device), valid until released */
  size_t bytes;
} mofft_completion;

typedef struct {
  size_t submitted;
  size_t completed;
  size_t pending;
  size_t inflight;
  size_t resident_plans;
  size_t bytes_in_use;
  size_t peak_bytes_in_use;
  size_t budget_bytes;
  size_t peak_inflight;
  double flops_completed;   /* 5 N log2 N per completed transform */
  double bytes_completed;   /* 2 N sizeof(complex) per completed transform */
  size_t executions;        /* batched executions enqueued (one or more
                               requests each) */
} mofft_server_stats;

void mofft_server_options_init(mofft_server_options *options);
mofft_status mofft_server_create(mofft_server **server,
                                 const mofft_server_options *options);
/* Queue a request (never blocks).  MOFFT_ALLOCATION_FAILURE if it can never
 * fit the budget, MOFFT_UNSUPPORTED_SIZE / MOFFT_INVALID_ARGUMENT as plan
 * creation reports them. */
mofft_status mofft_server_submit(mofft_server *server,
                                 const mofft_request *request,
                                 uint64_t *ticket);
/* Retire finished work, admit pending requests, and return up to
 * `capacity` new completions.  With wait != 0, block until at least one
 * completion is available or nothing is pending or in flight.  When the
 * pending head cannot be admitted while nothing is in flight - the arena
 * space it needs is held by completions the caller has not released - a
 * wait could never end: the call then returns MOFFT_ALLOCATION_FAILURE with
 * *count = 0, and the caller releases completions and polls again.
 * mofft_server_run_trace fails such a head (status
 * MOFFT_ALLOCATION_FAILURE in its completion) instead of stalling. */
mofft_status mofft_server_poll(mofft_server *server, mofft_completion *out,
                               size_t capacity, size_t *count, int wait);
/* Return a completed request's buffers to the arena. */
mofft_status mofft_server_release(mofft_server *server, uint64_t ticket);
mofft_status mofft_server_get_stats(const mofft_server *server,
                                    mofft_server_stats *stats);
/* Replay a trace in native code: request i is submitted `arrival_seconds[i]`
 * after the start (wall clock); each result is released as soon as it is
 * observed; completions[i] receives request i's record. */
mofft_status mofft_server_run_trace(mofft_server *server,
                                    const mofft_request *requests,
                                    const double *arrival_seconds,
                                    size_t count,
                                    mofft_completion *completions);
/* The server's input generator on a caller buffer (for checking results):
 * `count` complex values from `seed`, as for requests with input == NULL. */
mofft_status mofft_server_generate(void *out, int precision_bits,
                                   size_t count, uint64_t seed, void *stream);
void mofft_server_destroy(mofft_server *server);

#ifdef __cplusplus
}
#endif
#endif
