/*
 * mofft_oracle.c - CPU oracle for the MoFFT 1-D complex FFT path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker (or as the timed CPU
 * restatement).  The product library never links or calls anything here.
 *
 * Contents
 *   1. exp(2*pi*i*m/M) in long double after an exact integer octant
 *      reduction (the twiddle values the reference fills in
 *      src/mofft.c:1082-1140, evaluated more accurately).
 *   2. The FP64 DFT oracle: direct summation in double-double for N <= 4096
 *      and a compensated (double-double) mixed-radix Stockham FFT above
 *      that.  Definition of the transform: tests/test_fft.c:20-26 and
 *      include/mofft.h:26 (unnormalised, exponent sign = direction).
 *   3. A restatement of the reference executor dataflow
 *      (src/mofft.c:1932-2045, transpose layout, linear transposes) in the
 *      working precision, for timing as the CPU path and for dataflow
 *      checks.
 *   4. The byte accounting of the reference plan-cost breakdown
 *      (src/mofft.c:133-304).
 *
 * Build: see oracle/Makefile (-O2 -ffp-contract=off; the double-double
 * error-free transformations need IEEE evaluation without contraction).
 */
#include "mofft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_PI_L 3.141592653589793238462643383279502884L

/* ------------------------------------------------------------------ */
/* 1. Twiddles                                                          */
/* ------------------------------------------------------------------ */

/* cos and sin of 2*pi*m/M.  The angle is folded into [0, pi/4] with exact
 * integer arithmetic on p/Q = 8m/8M before any rounding happens, so the only
 * rounding is in p/Q and in cosl/sinl of a small argument. */
static void omega_ld(uint64_t m, uint64_t M, long double *c, long double *s) {
  m %= M;
  unsigned __int128 Q = (unsigned __int128)M * 8u;
  unsigned __int128 p = (unsigned __int128)m * 8u;
  int neg_s = 0, neg_c = 0, swap = 0;
  if (p > Q / 2) { p = Q - p; neg_s = 1; }
  if (p > Q / 4) { p = Q / 2 - p; neg_c = 1; }
  if (p > Q / 8) { p = Q / 4 - p; swap = 1; }
  const long double theta =
      2.0L * ORACLE_PI_L * ((long double)p / (long double)Q);
  long double cv = cosl(theta), sv = sinl(theta);
  if (swap) { long double t = cv; cv = sv; sv = t; }
  if (neg_c) cv = -cv;
  if (neg_s) sv = -sv;
  *c = cv;
  *s = sv;
}

void oracle_omega_dd(uint64_t m, uint64_t M, int sign, double out_re[2],
                     double out_im[2]) {
  long double c, s;
  omega_ld(m, M, &c, &s);
  if (sign < 0) s = -s;
  out_re[0] = (double)c;
  out_re[1] = (double)(c - (long double)out_re[0]);
  out_im[0] = (double)s;
  out_im[1] = (double)(s - (long double)out_im[0]);
}

void oracle_omega_f64(uint64_t m, uint64_t M, int sign, double out[2]) {
  long double c, s;
  omega_ld(m, M, &c, &s);
  out[0] = (double)c;
  out[1] = (double)(sign < 0 ? -s : s);
}

void oracle_omega_f32(uint64_t m, uint64_t M, int sign, float out[2]) {
  long double c, s;
  omega_ld(m, M, &c, &s);
  out[0] = (float)c;
  out[1] = (float)(sign < 0 ? -s : s);
}

void oracle_omega_many(const uint64_t *m, size_t count, uint64_t M, int sign,
                       int precision_bits, void *out) {
  for (size_t i = 0; i < count; ++i) {
    if (precision_bits == 32) oracle_omega_f32(m[i], M, sign, (float *)out + 2 * i);
    else oracle_omega_f64(m[i], M, sign, (double *)out + 2 * i);
  }
}

/* ------------------------------------------------------------------ */
/* Double-double arithmetic (Dekker / Knuth error-free transforms)       */
/* ------------------------------------------------------------------ */

typedef struct { double hi, lo; } dd;
typedef struct { dd re, im; } cdd;

static inline dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  dd r = {s, e};
  return r;
}

static inline dd quick_two_sum(double a, double b) {
  const double s = a + b;
  dd r = {s, b - (s - a)};
  return r;
}

static inline dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}

static inline dd dd_neg(dd a) {
  dd r = {-a.hi, -a.lo};
  return r;
}

static inline dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }

static inline dd dd_mul(dd a, dd b) {
  const double p = a.hi * b.hi;
  const double e = fma(a.hi, b.hi, -p);
  return quick_two_sum(p, e + (a.hi * b.lo + a.lo * b.hi));
}

static inline dd dd_from(double a) {
  dd r = {a, 0.0};
  return r;
}

static inline dd dd_from_ld(long double v) {
  dd r;
  r.hi = (double)v;
  r.lo = (double)(v - (long double)r.hi);
  return r;
}

static inline cdd cdd_add(cdd a, cdd b) {
  cdd r = {dd_add(a.re, b.re), dd_add(a.im, b.im)};
  return r;
}

static inline cdd cdd_sub(cdd a, cdd b) {
  cdd r = {dd_sub(a.re, b.re), dd_sub(a.im, b.im)};
  return r;
}

static inline cdd cdd_mul(cdd a, cdd b) {
  cdd r = {dd_sub(dd_mul(a.re, b.re), dd_mul(a.im, b.im)),
           dd_add(dd_mul(a.re, b.im), dd_mul(a.im, b.re))};
  return r;
}

/* multiply by sign*i: (a+ib)*(i*sign) = -sign*b + i*sign*a */
static inline cdd cdd_mul_i(cdd a, int sign) {
  cdd r;
  if (sign < 0) { r.re = a.im; r.im = dd_neg(a.re); }
  else { r.re = dd_neg(a.im); r.im = a.re; }
  return r;
}

static cdd omega_cdd(uint64_t m, uint64_t M, int sign) {
  long double c, s;
  omega_ld(m, M, &c, &s);
  if (sign < 0) s = -s;
  cdd r = {dd_from_ld(c), dd_from_ld(s)};
  return r;
}

/* ------------------------------------------------------------------ */
/* 2a. Direct DFT                                                        */
/* ------------------------------------------------------------------ */

int oracle_dft_direct(size_t n, size_t batch, int sign, const double *in,
                      double *out) {
  if (n == 0 || !in || !out || (sign != 1 && sign != -1)) return 1;
  cdd *table = (cdd *)malloc(n * sizeof(*table));
  cdd *acc = (cdd *)malloc(n * sizeof(*acc));
  if (!table || !acc) { free(table); free(acc); return 1; }
  for (size_t m = 0; m < n; ++m) table[m] = omega_cdd(m, n, sign);
  for (size_t b = 0; b < batch; ++b) {
    const double *x = in + 2 * n * b;
    for (size_t k = 0; k < n; ++k) {
      cdd sum = {dd_from(0.0), dd_from(0.0)};
      size_t idx = 0; /* (j*k) mod n, advanced incrementally */
      for (size_t j = 0; j < n; ++j) {
        const cdd w = table[idx];
        cdd xv = {dd_from(x[2 * j]), dd_from(x[2 * j + 1])};
        sum = cdd_add(sum, cdd_mul(xv, w));
        idx += k;
        if (idx >= n) idx -= n;
      }
      acc[k] = sum;
    }
    double *y = out + 2 * n * b;
    for (size_t k = 0; k < n; ++k) {
      y[2 * k] = acc[k].re.hi + acc[k].re.lo;
      y[2 * k + 1] = acc[k].im.hi + acc[k].im.lo;
    }
  }
  free(table);
  free(acc);
  return 0;
}

/* ------------------------------------------------------------------ */
/* 2b. Compensated mixed-radix Stockham FFT in double-double             */
/* ------------------------------------------------------------------ */

static int factor_chain(size_t n, int *radix, int cap) {
  int count = 0;
  while (n % 4 == 0 && count < cap) { radix[count++] = 4; n /= 4; }
  while (n % 2 == 0 && count < cap) { radix[count++] = 2; n /= 2; }
  for (size_t p = 3; p <= 61 && n > 1; p += 2) {
    while (n % p == 0 && count < cap) { radix[count++] = (int)p; n /= p; }
  }
  return n == 1 ? count : -1;
}

/* Twiddle w_n^e for e < n from two long-double tables: w_n^(h*B) * w_n^l. */
typedef struct {
  size_t n, block;
  long double *hi_re, *hi_im, *lo_re, *lo_im;
} twiddle2;

static int twiddle2_init(twiddle2 *t, size_t n, int sign) {
  size_t block = 1;
  while (block * block < n) ++block;
  const size_t hcount = (n + block - 1) / block;
  t->n = n;
  t->block = block;
  t->hi_re = (long double *)malloc(hcount * sizeof(long double));
  t->hi_im = (long double *)malloc(hcount * sizeof(long double));
  t->lo_re = (long double *)malloc(block * sizeof(long double));
  t->lo_im = (long double *)malloc(block * sizeof(long double));
  if (!t->hi_re || !t->hi_im || !t->lo_re || !t->lo_im) return 1;
  for (size_t h = 0; h < hcount; ++h) {
    omega_ld((uint64_t)h * block, n, &t->hi_re[h], &t->hi_im[h]);
    if (sign < 0) t->hi_im[h] = -t->hi_im[h];
  }
  for (size_t l = 0; l < block; ++l) {
    omega_ld(l, n, &t->lo_re[l], &t->lo_im[l]);
    if (sign < 0) t->lo_im[l] = -t->lo_im[l];
  }
  return 0;
}

static void twiddle2_free(twiddle2 *t) {
  free(t->hi_re); free(t->hi_im); free(t->lo_re); free(t->lo_im);
}

static cdd twiddle2_get(const twiddle2 *t, size_t e) {
  const size_t h = e / t->block, l = e % t->block;
  const long double ar = t->hi_re[h], ai = t->hi_im[h];
  const long double br = t->lo_re[l], bi = t->lo_im[l];
  cdd r = {dd_from_ld(ar * br - ai * bi), dd_from_ld(ar * bi + ai * br)};
  return r;
}

static void small_dft(cdd *v, int r, int sign, const cdd *wr, cdd *tmp) {
  if (r == 2) {
    const cdd a = v[0], b = v[1];
    v[0] = cdd_add(a, b);
    v[1] = cdd_sub(a, b);
    return;
  }
  if (r == 4) {
    const cdd s0 = cdd_add(v[0], v[2]), d0 = cdd_sub(v[0], v[2]);
    const cdd s1 = cdd_add(v[1], v[3]), d1 = cdd_sub(v[1], v[3]);
    const cdd d1i = cdd_mul_i(d1, sign);
    v[0] = cdd_add(s0, s1);
    v[2] = cdd_sub(s0, s1);
    v[1] = cdd_add(d0, d1i);
    v[3] = cdd_sub(d0, d1i);
    return;
  }
  for (int k = 0; k < r; ++k) {
    cdd sum = v[0];
    for (int j = 1; j < r; ++j)
      sum = cdd_add(sum, cdd_mul(v[j], wr[(j * k) % r]));
    tmp[k] = sum;
  }
  memcpy(v, tmp, (size_t)r * sizeof(*v));
}

int oracle_fft_dd(size_t n, size_t batch, int sign, const double *in,
                  double *out) {
  if (n == 0 || !in || !out || (sign != 1 && sign != -1)) return 1;
  int radix[64];
  const int count = factor_chain(n, radix, 64);
  if (count < 0) return 1;
  cdd *a = (cdd *)malloc(n * sizeof(*a));
  cdd *b = (cdd *)malloc(n * sizeof(*b));
  twiddle2 tw;
  memset(&tw, 0, sizeof(tw));
  if (!a || !b || twiddle2_init(&tw, n, sign)) {
    free(a); free(b); twiddle2_free(&tw);
    return 1;
  }
  for (size_t t = 0; t < batch; ++t) {
    const double *x = in + 2 * n * t;
    for (size_t i = 0; i < n; ++i) {
      a[i].re = dd_from(x[2 * i]);
      a[i].im = dd_from(x[2 * i + 1]);
    }
    cdd *src = a, *dst = b;
    size_t ns = 1;
    for (int s = 0; s < count; ++s) {
      const int r = radix[s];
      const size_t span = n / (size_t)r;
      cdd wr[64], v[64], tmp[64];
      for (int m = 0; m < r; ++m) wr[m] = omega_cdd((uint64_t)m, (uint64_t)r, sign);
      const size_t scale = n / (ns * (size_t)r);
      for (size_t j = 0; j < span; ++j) {
        const size_t jm = j % ns;
        for (int k = 0; k < r; ++k) {
          cdd value = src[j + (size_t)k * span];
          const size_t e = jm * (size_t)k;
          if (e != 0) value = cdd_mul(value, twiddle2_get(&tw, e * scale));
          v[k] = value;
        }
        small_dft(v, r, sign, wr, tmp);
        const size_t base = (j / ns) * ns * (size_t)r + jm;
        for (int k = 0; k < r; ++k) dst[base + (size_t)k * ns] = v[k];
      }
      cdd *swap = src; src = dst; dst = swap;
      ns *= (size_t)r;
    }
    double *y = out + 2 * n * t;
    for (size_t i = 0; i < n; ++i) {
      y[2 * i] = src[i].re.hi + src[i].re.lo;
      y[2 * i + 1] = src[i].im.hi + src[i].im.lo;
    }
  }
  free(a); free(b); twiddle2_free(&tw);
  return 0;
}

int oracle_dft(size_t n, size_t batch, int sign, const double *in,
               double *out) {
  return n <= 4096 ? oracle_dft_direct(n, batch, sign, in, out)
                   : oracle_fft_dd(n, batch, sign, in, out);
}

/* ------------------------------------------------------------------ */
/* 3. Reference executor restatement                                     */
/* ------------------------------------------------------------------ */

#define REF_PI 3.141592653589793238462643383279502884

static int ref_supported_radix(int r) {
  static const int k[] = {64, 32, 16, 15, 14, 13, 12, 11, 10,
                          9, 8, 7, 6, 5, 4, 3, 2};
  for (size_t i = 0; i < sizeof(k) / sizeof(k[0]); ++i)
    if (k[i] == r) return 1;
  return 0;
}

/* DFT-matrix coefficient as the reference generator computes it
 * (compiler/mofft_compiler/emitter.py:41-65): double cos/sin of
 * sign*2*pi*j*k/r, magnitudes below 5e-16 flushed to zero. */
static double ref_coefficient(int sign, int j, int k, int r, int imag) {
  const double angle = (double)sign * 2.0 * REF_PI * (double)j * (double)k /
                       (double)r;
  double v = imag ? sin(angle) : cos(angle);
  if (fabs(v) < 5e-16) v = 0.0;
  return v;
}

#define DEFINE_REF_EXECUTE(SUFFIX, T, FMA)                                    \
  int oracle_ref_execute_##SUFFIX(size_t n, int sign, const int *radices,     \
                                  size_t depth, const T *in, T *out) {        \
    if (n < 2 || !radices || depth == 0 || depth > 64 || !in || !out ||       \
        (sign != 1 && sign != -1))                                            \
      return 1;                                                               \
    size_t lengths[64], remaining = n;                                        \
    for (size_t l = 0; l < depth; ++l) {                                      \
      if (!ref_supported_radix(radices[l]) || remaining % radices[l]) return 1; \
      lengths[l] = remaining;                                                 \
      remaining /= (size_t)radices[l];                                        \
    }                                                                         \
    if (remaining != 1) return 1;                                             \
    T *cur = (T *)malloc(2 * n * sizeof(T));                                  \
    T *spare = (T *)malloc(2 * n * sizeof(T));                                \
    T *tw = (T *)malloc(2 * n * sizeof(T));                                   \
    if (!cur || !spare || !tw) { free(cur); free(spare); free(tw); return 1; } \
    /* first stage: Out[k*M+b] = sum_j W[j,k] In[j*M+b], M = N/r0 */          \
    const int r0 = radices[depth - 1];                                        \
    const size_t M = n / (size_t)r0;                                          \
    for (int k = 0; k < r0; ++k) {                                            \
      for (size_t b = 0; b < M; ++b) {                                        \
        T re = 0, im = 0;                                                     \
        for (int j = 0; j < r0; ++j) {                                        \
          const T wr = (T)ref_coefficient(sign, j, k, r0, 0);                 \
          const T wi = (T)ref_coefficient(sign, j, k, r0, 1);                 \
          const T xr = in[2 * ((size_t)j * M + b)];                           \
          const T xi = in[2 * ((size_t)j * M + b) + 1];                       \
          re = FMA(wr, xr, re); re = FMA(-wi, xi, re);                        \
          im = FMA(wr, xi, im); im = FMA(wi, xr, im);                         \
        }                                                                     \
        cur[2 * ((size_t)k * M + b)] = re;                                    \
        cur[2 * ((size_t)k * M + b) + 1] = im;                                \
      }                                                                       \
    }                                                                         \
    for (size_t level = depth - 1; level-- > 0;) {                            \
      const size_t nl = lengths[level];                                       \
      const int r = radices[level];                                           \
      const size_t m = nl / (size_t)r, P = n / nl, S = P * m;                 \
      /* linear transpose src/mofft.c:2014-2020 */                            \
      for (int j = 0; j < r; ++j)                                             \
        for (size_t k = 0; k < m; ++k)                                        \
          for (size_t b = 0; b < P; ++b) {                                    \
            const size_t d = (size_t)j * S + k * P + b;                       \
            const size_t s = k * (P * (size_t)r) + (size_t)j * P + b;         \
            spare[2 * d] = cur[2 * s];                                        \
            spare[2 * d + 1] = cur[2 * s + 1];                                \
          }                                                                   \
      /* full twiddle table, src/mofft.c:1090-1107 */                         \
      for (int k1 = 0; k1 < r; ++k1)                                          \
        for (size_t b = 0; b < m; ++b) {                                      \
          const double a = (double)sign * 2.0 * REF_PI * (double)k1 *         \
                           (double)b / (double)nl;                            \
          const T c = (T)cos(a), s = (T)sin(a);                               \
          for (size_t t = 0; t < P; ++t) {                                    \
            const size_t index = (size_t)k1 * S + b * P + t;                  \
            tw[2 * index] = c;                                                \
            tw[2 * index + 1] = s;                                            \
          }                                                                   \
        }                                                                     \
      T *dst = level == 0 ? out : cur;                                        \
      T *xs = (T *)malloc(2 * (size_t)r * sizeof(T));                         \
      if (!xs) { free(cur); free(spare); free(tw); return 1; }                \
      for (size_t q = 0; q < S; ++q) {                                        \
        for (int j = 0; j < r; ++j) {                                         \
          const size_t i = (size_t)j * S + q;                                 \
          const T xr = spare[2 * i], xi = spare[2 * i + 1];                   \
          const T tr = tw[2 * i], ti = tw[2 * i + 1];                         \
          xs[2 * j] = xr * tr - xi * ti;                                      \
          xs[2 * j + 1] = xr * ti + xi * tr;                                  \
        }                                                                     \
        for (int k = 0; k < r; ++k) {                                         \
          T re = 0, im = 0;                                                   \
          for (int j = 0; j < r; ++j) {                                       \
            const T wr = (T)ref_coefficient(sign, j, k, r, 0);                \
            const T wi = (T)ref_coefficient(sign, j, k, r, 1);                \
            re = FMA(wr, xs[2 * j], re); re = FMA(-wi, xs[2 * j + 1], re);    \
            im = FMA(wr, xs[2 * j + 1], im); im = FMA(wi, xs[2 * j], im);     \
          }                                                                   \
          dst[2 * ((size_t)k * S + q)] = re;                                  \
          dst[2 * ((size_t)k * S + q) + 1] = im;                              \
        }                                                                     \
      }                                                                       \
      free(xs);                                                               \
    }                                                                         \
    if (depth == 1) memcpy(out, cur, 2 * n * sizeof(T));                      \
    free(cur); free(spare); free(tw);                                         \
    return 0;                                                                 \
  }

DEFINE_REF_EXECUTE(f32, float, fmaf)
DEFINE_REF_EXECUTE(f64, double, fma)

/* ------------------------------------------------------------------ */
/* 4. Plan byte accounting (src/mofft.c:133-168, 190-304)                */
/* ------------------------------------------------------------------ */

static size_t sat_mul(size_t a, size_t b) {
  return a != 0 && b > SIZE_MAX / a ? SIZE_MAX : a * b;
}

static size_t sat_add(size_t a, size_t b) {
  return b > SIZE_MAX - a ? SIZE_MAX : a + b;
}

static size_t overfetch(size_t bytes, size_t useful, size_t line) {
  if (useful >= line) return bytes;
  if (useful == 0 || bytes > SIZE_MAX / line) return SIZE_MAX;
  const size_t product = bytes * line;
  if (product > SIZE_MAX - (useful - 1)) return SIZE_MAX;
  return (product + useful - 1) / useful;
}

int oracle_plan_bytes(size_t length, int precision_bits, const int *radices,
                      size_t radix_count, uint64_t blocked_mask,
                      int direct_input, int section_layout,
                      size_t cache_line_bytes, uint64_t out[5]) {
  if (!radices || radix_count == 0 || radix_count > 64 || length < 2 ||
      (precision_bits != 32 && precision_bits != 64) || !out ||
      (direct_input != 0 && direct_input != 1) ||
      (blocked_mask >> (radix_count - 1)))
    return 1;
  const size_t lanes = precision_bits == 32 ? 16 : 8;
  const size_t cb = precision_bits == 32 ? 8 : 16;
  const size_t data = sat_mul(length, cb);
  size_t lr = 0, lw = 0, tr = 0, tw = 0, peak = 0;
  size_t n = length;
  for (size_t level = 0; level < radix_count; ++level) {
    const int radix = radices[level];
    if (!ref_supported_radix(radix) || n % (size_t)radix) return 1;
    const size_t rem = n / (size_t)radix;
    size_t extra = 0;
    const int fused = radix_count == 2 && level == 0;
    if (rem != 1 && section_layout) {
      const size_t t = overfetch(data, sat_mul(rem, cb), cache_line_bytes);
      extra = t > data ? t - data : 0;
    } else if (rem != 1 && !fused) {
      const size_t parent = length / n;
      const int use_direct = direct_input && parent >= lanes / 2;
      if (use_direct) {
        const size_t t = overfetch(data, sat_mul(parent, cb), cache_line_bytes);
        extra = t > data ? t - data : 0;
      } else {
        const int blocked = (int)((blocked_mask >> level) & 1u);
        lr = sat_add(lr, data);
        lw = sat_add(lw, data);
        tr = sat_add(tr, blocked ? data
                                 : overfetch(data, sat_mul(parent, cb),
                                             cache_line_bytes));
        tw = sat_add(tw, data);
        if (sat_mul(data, 2) > peak) peak = sat_mul(data, 2);
      }
    }
    /* stage traffic: every stage reads and writes N points; later stages
     * also read a twiddle table of n points */
    const size_t twb = rem == 1 ? 0 : sat_mul(n, cb);
    const size_t sr = sat_add(data, twb);
    lr = sat_add(lr, sr);
    lw = sat_add(lw, data);
    tr = sat_add(tr, sat_add(sr, extra));
    tw = sat_add(tw, data);
    const size_t ws = sat_add(sat_mul(data, 2), twb);
    if (ws > peak) peak = ws;
    n = rem;
  }
  if (n != 1) return 1;
  out[0] = lr; out[1] = lw; out[2] = tr; out[3] = tw; out[4] = peak;
  return 0;
}
