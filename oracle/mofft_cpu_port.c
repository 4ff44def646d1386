/*
 * mofft_cpu_port.c - the reference's executor on a host core, for timing.
 *
 * TEST AND MEASUREMENT INFRASTRUCTURE ONLY.  bench.py's cpu_baseline leg
 * times it (and tests/ check it against the oracle); the product library
 * never links or calls it.
 *
 * The reference runs a plan as a chain of radix-r stages, each an r x r
 * DFT-matrix contraction over a batch of columns, with a materialising
 * transpose between stages (src/mofft.c:1932-2045; appendix dataflow):
 *
 *   r0 = f[d-1], M = N/r0
 *   cur[k*M + b] = sum_j x[j*M + b] * w_r0^(j*k)                  (k<r0, b<M)
 *   for l = d-2 .. 0:  r = f[l], m = n_l / r, P = N / n_l, S = P*m
 *     T[j*S + k*P + b] = cur[k*(P*r) + j*P + b]                   (transpose)
 *     y[q*S + k*P + b] = sum_j T[j*S + k*P + b] * w_(n_l)^(j*k) * w_r^(j*q)
 *     cur = y
 *
 * As in the reference, everything that does not depend on the data is
 * computed once when the plan is made: the DFT-matrix coefficients of each
 * radix (the values the reference generator bakes into its kernels as
 * literals, compiler/mofft_compiler/emitter.py:41-65) and the stage twiddles
 * (fill_twiddles_*, src/mofft.c:1082-1140, cos/sin in double rounded once).
 * The execute only moves and multiplies data.  The contraction is written in
 * the reference's outer-product order (emitter.py:283-466: for each input
 * row j, Out[k, :] += W[j, k] * In[j, :] over the batch of columns), on
 * split real/imaginary planes so the batch loop is contiguous and the
 * compiler vectorises it (FMA over AVX2 lanes where the SME kernels use
 * streaming-vector outer products).
 */
#include "mofft_cpu_port.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PORT_PI 3.141592653589793238462643383279502884
#define PORT_MAX_DEPTH 64

static int port_radix_ok(int r) {
  static const int k[] = {64, 32, 16, 15, 14, 13, 12, 11, 10,
                          9, 8, 7, 6, 5, 4, 3, 2};
  for (size_t i = 0; i < sizeof(k) / sizeof(k[0]); ++i)
    if (k[i] == r) return 1;
  return 0;
}

#define DEFINE_PORT(SUF, T)                                                    \
  struct port_plan_##SUF {                                                    \
    size_t n;                                                                 \
    int sign;                                                                 \
    size_t depth;                                                             \
    int radices[PORT_MAX_DEPTH];                                              \
    size_t lengths[PORT_MAX_DEPTH];                                           \
    T *coef_re[PORT_MAX_DEPTH]; /* [j*r + k] = Re w_r^(j k) */                \
    T *coef_im[PORT_MAX_DEPTH];                                               \
    T *tw_re[PORT_MAX_DEPTH];   /* [j*S + q]: w_(n_l)^(j * k(q)), q = k*P+b */ \
    T *tw_im[PORT_MAX_DEPTH];                                                 \
    T *a_re, *a_im, *b_re, *b_im;                                             \
  };                                                                          \
                                                                              \
  void port_destroy_##SUF(port_plan_##SUF *p) {                               \
    if (!p) return;                                                           \
    for (size_t l = 0; l < p->depth; ++l) {                                   \
      free(p->coef_re[l]); free(p->coef_im[l]);                               \
      free(p->tw_re[l]); free(p->tw_im[l]);                                   \
    }                                                                         \
    free(p->a_re); free(p->a_im); free(p->b_re); free(p->b_im);               \
    free(p);                                                                  \
  }                                                                           \
                                                                              \
  port_plan_##SUF *port_create_##SUF(size_t n, int sign, const int *radices,  \
                                     size_t depth) {                          \
    if (n < 2 || !radices || depth == 0 || depth > PORT_MAX_DEPTH ||          \
        (sign != 1 && sign != -1))                                            \
      return NULL;                                                            \
    port_plan_##SUF *p = (port_plan_##SUF *)calloc(1, sizeof(*p));            \
    if (!p) return NULL;                                                      \
    p->n = n;                                                                 \
    p->sign = sign;                                                           \
    p->depth = depth;                                                         \
    size_t remaining = n;                                                     \
    for (size_t l = 0; l < depth; ++l) {                                      \
      const int r = radices[l];                                               \
      if (!port_radix_ok(r) || remaining % (size_t)r) {                       \
        port_destroy_##SUF(p);                                                \
        return NULL;                                                          \
      }                                                                       \
      p->radices[l] = r;                                                      \
      p->lengths[l] = remaining;                                              \
      remaining /= (size_t)r;                                                 \
      p->coef_re[l] = (T *)malloc((size_t)r * r * sizeof(T));                 \
      p->coef_im[l] = (T *)malloc((size_t)r * r * sizeof(T));                 \
      if (!p->coef_re[l] || !p->coef_im[l]) {                                 \
        port_destroy_##SUF(p);                                                \
        return NULL;                                                          \
      }                                                                       \
      for (int j = 0; j < r; ++j)                                             \
        for (int k = 0; k < r; ++k) {                                         \
          /* exact integer reduction of j*k mod r before the angle */         \
          const double a = (double)sign * 2.0 * PORT_PI *                     \
                           (double)((j * k) % r) / (double)r;                 \
          double c = cos(a), s = sin(a);                                      \
          if (fabs(c) < 5e-16) c = 0.0;                                       \
          if (fabs(s) < 5e-16) s = 0.0;                                       \
          p->coef_re[l][j * r + k] = (T)c;                                    \
          p->coef_im[l][j * r + k] = (T)s;                                    \
        }                                                                     \
    }                                                                         \
    if (remaining != 1) {                                                     \
      port_destroy_##SUF(p);                                                  \
      return NULL;                                                            \
    }                                                                         \
    for (size_t l = 0; l + 1 < depth; ++l) {                                  \
      const size_t nl = p->lengths[l];                                        \
      const int r = p->radices[l];                                            \
      const size_t m = nl / (size_t)r, P = n / nl, S = P * m;                 \
      p->tw_re[l] = (T *)malloc(n * sizeof(T));                               \
      p->tw_im[l] = (T *)malloc(n * sizeof(T));                               \
      if (!p->tw_re[l] || !p->tw_im[l]) {                                     \
        port_destroy_##SUF(p);                                                \
        return NULL;                                                          \
      }                                                                       \
      for (int j = 0; j < r; ++j)                                             \
        for (size_t k = 0; k < m; ++k) {                                      \
          const double a = (double)sign * 2.0 * PORT_PI *                     \
                           (double)(((size_t)j * k) % nl) / (double)nl;       \
          const T c = (T)cos(a), s = (T)sin(a);                               \
          for (size_t b = 0; b < P; ++b) {                                    \
            p->tw_re[l][(size_t)j * S + k * P + b] = c;                       \
            p->tw_im[l][(size_t)j * S + k * P + b] = s;                       \
          }                                                                   \
        }                                                                     \
    }                                                                         \
    p->a_re = (T *)malloc(n * sizeof(T));                                     \
    p->a_im = (T *)malloc(n * sizeof(T));                                     \
    p->b_re = (T *)malloc(n * sizeof(T));                                     \
    p->b_im = (T *)malloc(n * sizeof(T));                                     \
    if (!p->a_re || !p->a_im || !p->b_re || !p->b_im) {                       \
      port_destroy_##SUF(p);                                                  \
      return NULL;                                                            \
    }                                                                         \
    return p;                                                                 \
  }                                                                           \
                                                                              \
  /* Out[k*S + q] = sum_j W[j,k] * In[j*S + q] for q < S, in the outer-      \
   * product order of the reference kernels: the batch of columns is cut   \
   * into blocks of QB (as the SME kernels cut it into vector-length        \
   * blocks), and each block's r x QB output tile is accumulated by one     \
   * rank-1 update per input row while it stays in L1.  (Measured on a      \
   * Xeon core: 256-byte blocks 2.2 / 1.3 GFLOP/s FP32 / FP64 at N = 2^16,  \
   * unblocked 2.0 / 1.1, register-accumulated 8-lane blocks 0.9 / 1.2.) */ \
  static void port_contract_##SUF(int r, size_t S, const T *wr, const T *wi,  \
                                  const T *restrict in_re,                    \
                                  const T *restrict in_im,                    \
                                  T *restrict out_re, T *restrict out_im) {   \
    enum { QB = 256 / (int)sizeof(T) };                                       \
    for (size_t q0 = 0; q0 < S; q0 += QB) {                                   \
      const size_t qn = S - q0 < (size_t)QB ? S - q0 : (size_t)QB;            \
      for (int k = 0; k < r; ++k) {                                           \
        T *restrict ore = out_re + (size_t)k * S + q0;                        \
        T *restrict oim = out_im + (size_t)k * S + q0;                        \
        const T c0 = wr[k], s0 = wi[k]; /* row j = 0 */                       \
        const T *restrict ire0 = in_re + q0, *restrict iim0 = in_im + q0;     \
        for (size_t q = 0; q < qn; ++q) {                                     \
          ore[q] = c0 * ire0[q] - s0 * iim0[q];                               \
          oim[q] = c0 * iim0[q] + s0 * ire0[q];                               \
        }                                                                     \
        for (int j = 1; j < r; ++j) {                                         \
          const T c = wr[j * r + k], s = wi[j * r + k];                       \
          const T *restrict ire = in_re + (size_t)j * S + q0;                 \
          const T *restrict iim = in_im + (size_t)j * S + q0;                 \
          for (size_t q = 0; q < qn; ++q) {                                   \
            ore[q] += c * ire[q] - s * iim[q];                                \
            oim[q] += c * iim[q] + s * ire[q];                                \
          }                                                                   \
        }                                                                     \
      }                                                                       \
    }                                                                         \
  }                                                                           \
                                                                              \
  int port_execute_##SUF(const port_plan_##SUF *p, const T *in, T *out) {     \
    if (!p || !in || !out) return 1;                                          \
    const size_t n = p->n, d = p->depth;                                      \
    T *cre = p->a_re, *cim = p->a_im, *sre = p->b_re, *sim = p->b_im;         \
    /* planes of the input */                                                 \
    for (size_t i = 0; i < n; ++i) {                                          \
      sre[i] = in[2 * i];                                                     \
      sim[i] = in[2 * i + 1];                                                 \
    }                                                                         \
    /* first stage: radix f[d-1] over the M = N / r0 strided columns */      \
    const int r0 = p->radices[d - 1];                                         \
    port_contract_##SUF(r0, n / (size_t)r0, p->coef_re[d - 1],                \
                        p->coef_im[d - 1], sre, sim, cre, cim);               \
    for (size_t level = d - 1; level-- > 0;) {                                \
      const size_t nl = p->lengths[level];                                    \
      const int r = p->radices[level];                                        \
      const size_t m = nl / (size_t)r, P = n / nl, S = P * m;                 \
      /* transpose (src/mofft.c:2014-2020) with the twiddle multiply */       \
      const T *twr = p->tw_re[level], *twi = p->tw_im[level];                 \
      for (int j = 0; j < r; ++j)                                             \
        for (size_t k = 0; k < m; ++k) {                                      \
          const size_t dst = (size_t)j * S + k * P;                           \
          const size_t src = k * (P * (size_t)r) + (size_t)j * P;             \
          for (size_t b = 0; b < P; ++b) {                                    \
            const T xr = cre[src + b], xi = cim[src + b];                     \
            const T tr = twr[dst + b], ti = twi[dst + b];                     \
            sre[dst + b] = xr * tr - xi * ti;                                 \
            sim[dst + b] = xr * ti + xi * tr;                                 \
          }                                                                   \
        }                                                                     \
      port_contract_##SUF(r, S, p->coef_re[level], p->coef_im[level], sre,    \
                          sim, cre, cim);                                     \
    }                                                                         \
    for (size_t i = 0; i < n; ++i) {                                          \
      out[2 * i] = cre[i];                                                    \
      out[2 * i + 1] = cim[i];                                                \
    }                                                                         \
    return 0;                                                                 \
  }

DEFINE_PORT(f32, float)
DEFINE_PORT(f64, double)
