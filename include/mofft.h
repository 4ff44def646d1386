/*
 * mofft.h - public C ABI of the MI355X (gfx950) MoFFT library.
 *
 * Drop-in for the reference header include/mofft.h:1-236: every type, enum
 * value, struct layout and function signature below matches it, so code
 * written against the reference compiles and links unchanged.  The
 * implementation behind it runs on one AMD Instinct MI355X through HIP.
 *
 * Transforms are unnormalised 1-D complex DFTs,
 *     X[k] = sum_j x[j] * exp(direction * 2*pi*i * j*k / N),
 * so forward followed by backward returns N*x.  Data is interleaved
 * {real, imag}.  Input and output pointers may be host memory (staged through
 * the device by the library) or device memory from hipMalloc (used in place).
 * Passing the same pointer as input and output selects in-place operation.
 */
#ifndef MOFFT_H
#define MOFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/mofft.h:11-12 */
typedef struct { float real, imag; } mofft_complex_f32;
typedef struct { double real, imag; } mofft_complex_f64;

/* include/mofft.h:14-24 - status codes, same numeric values. */
typedef enum {
  MOFFT_SUCCESS = 0,
  MOFFT_INVALID_ARGUMENT = 1,
  MOFFT_UNSUPPORTED_SIZE = 2,
  MOFFT_ALLOCATION_FAILURE = 3,
  MOFFT_UNSUPPORTED_HARDWARE = 4,
  MOFFT_INTERNAL_ERROR = 5,
  MOFFT_WISDOM_NOT_FOUND = 6,
  MOFFT_INCOMPATIBLE_RECIPE = 7,
  MOFFT_BUFFER_TOO_SMALL = 8
} mofft_status;

/* include/mofft.h:26-35 */
typedef enum { MOFFT_FORWARD = -1, MOFFT_BACKWARD = 1 } mofft_direction;
typedef enum { MOFFT_OUT_OF_PLACE = 0, MOFFT_IN_PLACE = 1 } mofft_placement;
typedef enum {
  MOFFT_TRANSPOSE_LINEAR = 0,
  MOFFT_TRANSPOSE_BLOCKED = 1
} mofft_transpose_strategy;
typedef enum {
  MOFFT_STAGE_LAYOUT_TRANSPOSE = 0,
  MOFFT_STAGE_LAYOUT_SECTION = 1
} mofft_stage_layout;

/* include/mofft.h:37-44.  PRETUNED takes the autotuner's compiled wisdom for
 * this length and precision when there is an entry, and the analytical GPU
 * planner otherwise.  MEASURE times a bounded set of complete plans built
 * from the kernels compiled into this library, with HIP events on the GPU. */
typedef enum {
  MOFFT_PLANNER_PRETUNED = 0,
  MOFFT_PLANNER_MEASURE = 1
} mofft_planner_mode;

/* include/mofft.h:46-50.  PRETUNED only: fail with MOFFT_WISDOM_NOT_FOUND
 * rather than fall back to the analytical planner. */
enum {
  MOFFT_PLAN_STRICT_WISDOM = 1u << 0
};

/* include/mofft.h:52-59.  Zero selects the defaults (16 candidates, 5
 * samples); the limits are 64 candidates and 31 samples. */
typedef struct {
  mofft_planner_mode mode;
  uint32_t flags;
  size_t candidate_limit;
  uint32_t timing_samples;
} mofft_plan_options;

/* include/mofft.h:61-81 */
#define MOFFT_MAX_PLAN_DEPTH 64
typedef struct {
  int measured;
  int used_wisdom;
  int used_model_fallback;
  size_t candidates_measured;
  size_t radix_count;
  int radices[MOFFT_MAX_PLAN_DEPTH];
  unsigned int kernel_bank;
  unsigned int kernel_bank_count;
  uint64_t blocked_transpose_mask;
  int direct_input;
  mofft_stage_layout stage_layout;
  double planning_seconds;
  double selected_median_seconds;
  size_t length;
  int precision_bits;
  mofft_direction direction;
  mofft_placement placement;
} mofft_plan_diagnostics;

/* include/mofft.h:83-90 - flags of the bound-array constructors. */
enum {
  MOFFT_FLAG_PRETUNED = 0,
  MOFFT_FLAG_MEASURE = 1u << 0,
  MOFFT_FLAG_WISDOM_ONLY = 1u << 1
};

/* include/mofft.h:92-102 */
typedef struct {
  double kernel_cost;
  double memory_cost;
  double layout_cost;
  double total_cost;
  size_t logical_read_bytes;
  size_t logical_write_bytes;
  size_t transferred_read_bytes;
  size_t transferred_write_bytes;
  size_t peak_working_set_bytes;
} mofft_plan_cost_breakdown;

/* include/mofft.h:104-105 - opaque plans. */
typedef struct mofft_plan_f32 mofft_plan_f32;
typedef struct mofft_plan_f64 mofft_plan_f64;

/* include/mofft.h:107-112.  One transform of `length` points. */
mofft_status mofft_plan_create_f32(mofft_plan_f32 **plan, size_t length,
                                   mofft_direction direction,
                                   mofft_placement placement);
mofft_status mofft_plan_create_f64(mofft_plan_f64 **plan, size_t length,
                                   mofft_direction direction,
                                   mofft_placement placement);

/* include/mofft.h:114-124.  NULL options means the PRETUNED defaults. */
void mofft_plan_options_init(mofft_plan_options *options);
mofft_status mofft_plan_create_with_options_f32(
    mofft_plan_f32 **plan, size_t length, mofft_direction direction,
    mofft_placement placement, const mofft_plan_options *options,
    mofft_plan_diagnostics *diagnostics);
mofft_status mofft_plan_create_with_options_f64(
    mofft_plan_f64 **plan, size_t length, mofft_direction direction,
    mofft_placement placement, const mofft_plan_options *options,
    mofft_plan_diagnostics *diagnostics);

/* include/mofft.h:126-136.  Returns NULL on any failure.  The plan keeps the
 * two buffer addresses for mofft_execute_plan_*. */
mofft_plan_f32 *mofft_plan_dft_1d_f32(
    size_t length, const mofft_complex_f32 *input,
    mofft_complex_f32 *output, mofft_direction direction, uint32_t flags);
mofft_plan_f64 *mofft_plan_dft_1d_f64(
    size_t length, const mofft_complex_f64 *input,
    mofft_complex_f64 *output, mofft_direction direction, uint32_t flags);
mofft_status mofft_execute_plan_f32(const mofft_plan_f32 *plan);
mofft_status mofft_execute_plan_f64(const mofft_plan_f64 *plan);

/* include/mofft.h:138-145.  Radices run outermost first; the last entry is
 * the first stage executed. */
mofft_status mofft_plan_create_with_radices_f32(
    mofft_plan_f32 **plan, size_t length, mofft_direction direction,
    mofft_placement placement, const int *radices, size_t radix_count);
mofft_status mofft_plan_create_with_radices_f64(
    mofft_plan_f64 **plan, size_t length, mofft_direction direction,
    mofft_placement placement, const int *radices, size_t radix_count);

/* include/mofft.h:146-176.  Plan-search controls; call before executing a
 * plan from several threads.  On the GPU a set bit `level` of the blocked
 * transpose mask makes the data round-trip through HBM before that
 * outer-to-inner level (a pass boundary); clear bits let the stages fuse in
 * LDS where a compiled kernel covers them. */
mofft_status mofft_plan_set_transpose_strategy_f32(
    mofft_plan_f32 *plan, mofft_transpose_strategy strategy);
mofft_status mofft_plan_set_transpose_strategy_f64(
    mofft_plan_f64 *plan, mofft_transpose_strategy strategy);
mofft_status mofft_plan_set_blocked_transpose_mask_f32(
    mofft_plan_f32 *plan, uint64_t blocked_levels);
mofft_status mofft_plan_set_blocked_transpose_mask_f64(
    mofft_plan_f64 *plan, uint64_t blocked_levels);
mofft_status mofft_plan_set_direct_input_f32(mofft_plan_f32 *plan,
                                              int enabled);
mofft_status mofft_plan_set_direct_input_f64(mofft_plan_f64 *plan,
                                              int enabled);
mofft_status mofft_plan_set_stage_layout_f32(mofft_plan_f32 *plan,
                                              mofft_stage_layout layout);
mofft_status mofft_plan_set_stage_layout_f64(mofft_plan_f64 *plan,
                                              mofft_stage_layout layout);
mofft_status mofft_plan_set_kernel_bank_f32(mofft_plan_f32 *plan,
                                             unsigned int bank);
mofft_status mofft_plan_set_kernel_bank_f64(mofft_plan_f64 *plan,
                                             unsigned int bank);
mofft_status mofft_plan_get_kernel_bank_f32(const mofft_plan_f32 *plan,
                                             unsigned int *bank);
mofft_status mofft_plan_get_kernel_bank_f64(const mofft_plan_f64 *plan,
                                             unsigned int *bank);

/* include/mofft.h:178-201.  Text plan recipes, scoped to the library binary
 * (its build ID) and to the GPU architecture.  Export with buffer=NULL, capacity=0 to query the
 * size (including the final NUL). */
const char *mofft_build_id(void);
mofft_status mofft_plan_export_f32(const mofft_plan_f32 *plan, char *buffer,
                                   size_t capacity, size_t *required);
mofft_status mofft_plan_export_f64(const mofft_plan_f64 *plan, char *buffer,
                                   size_t capacity, size_t *required);
mofft_status mofft_plan_import_f32(mofft_plan_f32 **plan, const char *recipe,
                                   size_t recipe_size,
                                   mofft_plan_diagnostics *diagnostics);
mofft_status mofft_plan_import_f64(mofft_plan_f64 **plan, const char *recipe,
                                   size_t recipe_size,
                                   mofft_plan_diagnostics *diagnostics);

/* include/mofft.h:203-208.  Synchronous: returns once the output holds the
 * result. */
mofft_status mofft_execute_f32(const mofft_plan_f32 *plan,
                               const mofft_complex_f32 *input,
                               mofft_complex_f32 *output);
mofft_status mofft_execute_f64(const mofft_plan_f64 *plan,
                               const mofft_complex_f64 *input,
                               mofft_complex_f64 *output);

/* include/mofft.h:210-231 */
void mofft_plan_destroy_f32(mofft_plan_f32 *plan);
void mofft_plan_destroy_f64(mofft_plan_f64 *plan);
double mofft_estimated_radix_cost(int radix, int precision_bits,
                                  int first_stage);
double mofft_estimated_stage_cost(int radix, int precision_bits,
                                  int first_stage, size_t transform_length,
                                  size_t stage_length);
double mofft_estimated_plan_cost(size_t length, int precision_bits,
                                 const int *radices, size_t radix_count,
                                 uint64_t blocked_transpose_mask,
                                 int direct_input);
double mofft_estimated_plan_cost_with_layout(
    size_t length, int precision_bits, const int *radices, size_t radix_count,
    uint64_t blocked_transpose_mask, int direct_input,
    mofft_stage_layout layout);
mofft_status mofft_estimated_plan_cost_breakdown_with_layout(
    size_t length, int precision_bits, const int *radices, size_t radix_count,
    uint64_t blocked_transpose_mask, int direct_input,
    mofft_stage_layout layout, mofft_plan_cost_breakdown *breakdown);
int mofft_wisdom_is_formal(void);
const char *mofft_wisdom_source(void);
const char *mofft_status_string(mofft_status status);
const char *mofft_version(void);

#ifdef __cplusplus
}
#endif
#endif
