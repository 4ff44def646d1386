/*
 * mofft_oracle.h - CPU oracle for the MoFFT 1-D complex FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or as the timed CPU restatement; the product library never links
 * or calls it.
 *
 * Arrays are interleaved complex ({re, im} pairs).  `sign` is the DFT
 * exponent sign: -1 forward, +1 backward (reference include/mofft.h:26).
 */
#ifndef MOFFT_ORACLE_H
#define MOFFT_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* exp(sign*2*pi*i*m/M) evaluated in long double after an exact integer
 * octant reduction, returned as a double-double (hi, lo) pair per part. */
void oracle_omega_dd(uint64_t m, uint64_t M, int sign, double out_re[2],
                     double out_im[2]);
/* Same value rounded once to double / to float. */
void oracle_omega_f64(uint64_t m, uint64_t M, int sign, double out[2]);
void oracle_omega_f32(uint64_t m, uint64_t M, int sign, float out[2]);
/* out[i] = omega(m[i], M, sign) rounded once to FP32 or FP64, i < count. */
void oracle_omega_many(const uint64_t *m, size_t count, uint64_t M, int sign,
                       int precision_bits, void *out);

/* Direct O(N^2) DFT with double-double accumulation and double-double
 * twiddles; result rounded once to double. */
int oracle_dft_direct(size_t n, size_t batch, int sign, const double *in,
                      double *out);
/* Compensated FFT: mixed-radix Stockham in double-double arithmetic for any
 * N whose prime factors are <= 61; result rounded once to double. */
int oracle_fft_dd(size_t n, size_t batch, int sign, const double *in,
                  double *out);
/* The oracle entry: direct summation for N <= 4096, compensated FFT above. */
int oracle_dft(size_t n, size_t batch, int sign, const double *in,
               double *out);

/* Restatement of the reference executor (src/mofft.c:1932-2045, transpose
 * layout with linear transposes) with its twiddle fill
 * (src/mofft.c:1082-1140) and generated-kernel contract
 * (compiler/mofft_compiler/emitter.py:56-65, validate.py:115-127), computed
 * in the working precision.  `radices` run outermost first. */
int oracle_ref_execute_f32(size_t n, int sign, const int *radices,
                           size_t depth, const float *in, float *out);
int oracle_ref_execute_f64(size_t n, int sign, const int *radices,
                           size_t depth, const double *in, double *out);

/* Byte accounting of mofft_estimated_plan_cost_breakdown_with_layout
 * (src/mofft.c:190-304) with the reference build defaults
 * (MOFFT_FUSE_FIRST_TRANSPOSE and MOFFT_BROADCAST_REPEATED_TWIDDLES on,
 * CMakeLists.txt:14-17).  out[0..4] = logical read, logical write,
 * transferred read, transferred write, peak working set.  Returns 0 on
 * success, 1 on invalid input. */
int oracle_plan_bytes(size_t length, int precision_bits, const int *radices,
                      size_t radix_count, uint64_t blocked_mask,
                      int direct_input, int section_layout,
                      size_t cache_line_bytes, uint64_t out[5]);

#ifdef __cplusplus
}
#endif
#endif
