/*
 * mofft_wisdom.h - the compiled PRETUNED wisdom table (generated from the
 * autotuner's JSON by mofft_compiler.wisdom, like the reference's
 * tools/compile_wisdom.py:98-153).
 */
#ifndef MOFFT_WISDOM_H
#define MOFFT_WISDOM_H
#include <stddef.h>
#include <stdint.h>

/* Exact (length, precision) lookup.  Returns the radix count written to
 * `radices` (outermost first), or 0 on a miss. */
size_t mofft_generated_wisdom_lookup(size_t length, int precision_bits,
                                     int *radices, size_t capacity,
                                     uint64_t *blocked_transpose_mask,
                                     int *direct_input, int *stage_layout,
                                     unsigned int *kernel_bank);
int mofft_generated_wisdom_is_formal(void);
const char *mofft_generated_wisdom_source(void);
#endif
