/*
 * test_dropin.c - the reference's C-level correctness contract, exercised
 * through include/mofft.h only, with host buffers as the reference tests
 * use them.  Each check names the reference test it mirrors; the inputs,
 * sizes and tolerances are the reference's, the code is this project's.
 *
 *   tests/test_fft.c:11-99     direct-DFT parity, in/out of place, both signs
 *   tests/test_fft.c:101-145   forward/backward round trips
 *   tests/test_fft.c:147-195   model plan vs explicit {32,8}; bad radix
 *   tests/test_fft.c:197-251   every radix as a two-stage {2, r} plan
 *   tests/test_fft.c:253-306   linear vs blocked mask (on the GPU the mask
 *                              moves HBM pass boundaries, which changes
 *                              where four-step twiddles apply: agreement to
 *                              the layout tolerance, not bit identity)
 *   tests/test_fft.c:308-406   section vs transpose layout
 *   tests/test_fft.c:408-487   direct-input vs materialised, {13, r, 2}
 *   tests/test_fft.c:493-560   cost-model byte known answers
 *   tests/test_fft.c:574-587   unsupported 17, 27 extended sizes plan
 *   tests/test_binary_banks.c  closed-form three-tone spectra, every bank
 *   tests/test_planning.c      MEASURE, STRICT wisdom, bound arrays
 *   tests/test_plan_recipe.c   export/import round trip and rejections
 *
 * Exit status 0 when every check passes; failures are printed.
 */
#include "mofft.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int failures = 0;
#define FAIL(...)                                  \
  do {                                             \
    fprintf(stderr, "FAIL %s:%d: ", __func__, __LINE__); \
    fprintf(stderr, __VA_ARGS__);                  \
    fputc('\n', stderr);                           \
    ++failures;                                    \
  } while (0)

static const double kPi = 3.14159265358979323846264338327950288;

/* O(N^2) reference in double: X[k] = sum_j x[j] exp(sign 2 pi i jk/N) */
static void dft(const double *x, double *y, size_t n, int sign) {
  for (size_t k = 0; k < n; ++k) {
    double re = 0, im = 0;
    for (size_t j = 0; j < n; ++j) {
      const double a = sign * 2.0 * kPi * (double)((j * k) % n) / (double)n;
      re += x[2 * j] * cos(a) - x[2 * j + 1] * sin(a);
      im += x[2 * j] * sin(a) + x[2 * j + 1] * cos(a);
    }
    y[2 * k] = re;
    y[2 * k + 1] = im;
  }
}

static double max_dist(const double *a, const double *b, size_t n) {
  double e = 0;
  for (size_t i = 0; i < n; ++i) {
    const double d = hypot(a[2 * i] - b[2 * i], a[2 * i + 1] - b[2 * i + 1]);
    if (!(d <= e)) e = d;  /* also catches NaN */
  }
  return e;
}

/* Run one transform of length n through the public API.  radices == NULL
 * selects the default planner.  Returns status; output in y (double). */
static mofft_status run(int bits, size_t n, int sign, int in_place,
                        const int *radices, size_t depth, uint64_t mask,
                        int set_mask, const double *x, double *y) {
  mofft_status st;
  const mofft_placement pl = in_place ? MOFFT_IN_PLACE : MOFFT_OUT_OF_PLACE;
  if (bits == 32) {
    mofft_plan_f32 *p = NULL;
    mofft_complex_f32 *a = calloc(n, sizeof(*a)), *b = calloc(n, sizeof(*b));
    for (size_t i = 0; i < n; ++i) { a[i].real = (float)x[2 * i]; a[i].imag = (float)x[2 * i + 1]; }
    st = radices ? mofft_plan_create_with_radices_f32(&p, n, sign, pl, radices, depth)
                 : mofft_plan_create_f32(&p, n, sign, pl);
    if (st == MOFFT_SUCCESS && set_mask) st = mofft_plan_set_blocked_transpose_mask_f32(p, mask);
    if (st == MOFFT_SUCCESS) {
      if (in_place) { memcpy(b, a, n * sizeof(*a)); st = mofft_execute_f32(p, b, b); }
      else st = mofft_execute_f32(p, a, b);
    }
    for (size_t i = 0; i < n; ++i) { y[2 * i] = b[i].real; y[2 * i + 1] = b[i].imag; }
    mofft_plan_destroy_f32(p); free(a); free(b);
  } else {
    mofft_plan_f64 *p = NULL;
    mofft_complex_f64 *a = calloc(n, sizeof(*a)), *b = calloc(n, sizeof(*b));
    for (size_t i = 0; i < n; ++i) { a[i].real = x[2 * i]; a[i].imag = x[2 * i + 1]; }
    st = radices ? mofft_plan_create_with_radices_f64(&p, n, sign, pl, radices, depth)
                 : mofft_plan_create_f64(&p, n, sign, pl);
    if (st == MOFFT_SUCCESS && set_mask) st = mofft_plan_set_blocked_transpose_mask_f64(p, mask);
    if (st == MOFFT_SUCCESS) {
      if (in_place) { memcpy(b, a, n * sizeof(*a)); st = mofft_execute_f64(p, b, b); }
      else st = mofft_execute_f64(p, a, b);
    }
    for (size_t i = 0; i < n; ++i) { y[2 * i] = b[i].real; y[2 * i + 1] = b[i].imag; }
    mofft_plan_destroy_f64(p); free(a); free(b);
  }
  return st;
}

static void fill(double *x, size_t n, double fa, double fb, int ma, int mb,
                 double sa, double sb) {
  for (size_t i = 0; i < n; ++i) {
    x[2 * i] = sin((double)i * fa) + (double)(i % ma) * sa;
    x[2 * i + 1] = cos((double)i * fb) - (double)(i % mb) * sb;
  }
}

static void round_to_float(double *x, size_t n) {
  for (size_t i = 0; i < 2 * n; ++i) x[i] = (float)x[i];
}

static void check_direct(void) {
  const size_t sizes[] = {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15,
                          16, 32, 45, 64, 144, 225, 256};
  for (size_t s = 0; s < sizeof(sizes) / sizeof(sizes[0]); ++s) {
    const size_t n = sizes[s];
    double *x = malloc(16 * n), *r = malloc(16 * n), *y = malloc(16 * n);
    for (int bits = 32; bits <= 64; bits += 32)
      for (int sign = -1; sign <= 1; sign += 2)
        for (int ip = 0; ip <= 1; ++ip) {
          fill(x, n, 0.17, 0.11, 7, 5, 0.03, 0.02);
          if (bits == 32) round_to_float(x, n);
          dft(x, r, n, sign);
          const mofft_status st = run(bits, n, sign, ip, NULL, 0, 0, 0, x, y);
          const double e = max_dist(y, r, n);
          const double tol = bits == 32 ? 32.0 * FLT_EPSILON * n : 96.0 * DBL_EPSILON * n;
          if (st != MOFFT_SUCCESS || e > tol)
            FAIL("f%d n=%zu sign=%d in_place=%d status=%s err=%g tol=%g", bits, n,
                 sign, ip, mofft_status_string(st), e, tol);
        }
    free(x); free(r); free(y);
  }
}

static void check_roundtrip(void) {
  const size_t sizes[] = {169, 196, 1024, 1728, 2197, 38416, 248832};
  for (size_t s = 0; s < sizeof(sizes) / sizeof(sizes[0]); ++s) {
    const size_t n = sizes[s];
    double *x = malloc(16 * n), *y = malloc(16 * n), *z = malloc(16 * n);
    for (int bits = 32; bits <= 64; bits += 32) {
      fill(x, n, 0.017, 0.013, 17, 13, 0.003, 0.002);
      if (bits == 32) round_to_float(x, n);
      mofft_status st = run(bits, n, -1, 0, NULL, 0, 0, 0, x, y);
      if (st == MOFFT_SUCCESS) st = run(bits, n, 1, 0, NULL, 0, 0, 0, y, z);
      for (size_t i = 0; i < 2 * n; ++i) z[i] /= (double)n;
      const double e = max_dist(z, x, n);
      const double stages = fmax(1.0, log2((double)n));
      const double tol = bits == 32 ? 128.0 * FLT_EPSILON * stages : 512.0 * DBL_EPSILON * stages;
      if (st != MOFFT_SUCCESS || e > tol)
        FAIL("roundtrip f%d n=%zu status=%s err=%g tol=%g", bits, n, mofft_status_string(st), e, tol);
    }
    free(x); free(y); free(z);
  }
}

static void check_specific(void) {
  const size_t n = 256;
  const int chain[] = {32, 8};
  double *x = malloc(16 * n), *a = malloc(16 * n), *b = malloc(16 * n);
  for (size_t i = 0; i < n; ++i) { x[2 * i] = (float)sin(.03 * i); x[2 * i + 1] = (float)cos(.02 * i); }
  for (int bits = 32; bits <= 64; bits += 32) {
    mofft_status s1 = run(bits, n, -1, 0, NULL, 0, 0, 0, x, a);
    mofft_status s2 = run(bits, n, -1, 0, chain, 2, 0, 0, x, b);
    const double tol = bits == 32 ? 64.0 * FLT_EPSILON * n : 256.0 * DBL_EPSILON * n;
    if (s1 || s2 || max_dist(a, b, n) > tol) FAIL("f%d {32,8} vs default", bits);
  }
  const int bad[] = {17, 16};
  mofft_plan_f32 *p = NULL;
  if (mofft_plan_create_with_radices_f32(&p, n, MOFFT_FORWARD, MOFFT_OUT_OF_PLACE, bad, 2) !=
      MOFFT_INVALID_ARGUMENT || p != NULL)
    FAIL("radix 17 accepted");
  free(x); free(a); free(b);
}

static const int kRadices[] = {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 32, 64};

static void check_two_stage_radices(void) {
  for (size_t r = 0; r < sizeof(kRadices) / sizeof(kRadices[0]); ++r) {
    const int chain[] = {2, kRadices[r]};
    const size_t n = 2 * (size_t)kRadices[r];
    double *x = malloc(16 * n), *ref = malloc(16 * n), *y = malloc(16 * n);
    for (size_t i = 0; i < n; ++i) {
      x[2 * i] = sin(.13 * i) + .01 * (i % 3);
      x[2 * i + 1] = cos(.07 * i) - .02 * (i % 5);
    }
    dft(x, ref, n, -1);
    for (int bits = 32; bits <= 64; bits += 32) {
      const mofft_status st = run(bits, n, -1, 0, chain, 2, 0, 0, x, y);
      const double tol = bits == 32 ? 64.0 * FLT_EPSILON * n : 256.0 * DBL_EPSILON * n;
      if (st || max_dist(y, ref, n) > tol) FAIL("f%d chain {2,%d}", bits, kRadices[r]);
    }
    free(x); free(ref); free(y);
  }
}

static void check_mask_grouping(void) {
  const size_t n = 4096;
  const int chain[] = {16, 16, 16};
  double *x = malloc(16 * n), *a = malloc(16 * n), *b = malloc(16 * n);
  for (size_t i = 0; i < n; ++i) { x[2 * i] = sin(.017 * i); x[2 * i + 1] = cos(.013 * i); }
  for (int bits = 32; bits <= 64; bits += 32) {
    mofft_status s1 = run(bits, n, -1, 0, chain, 3, 0, 1, x, a);
    mofft_status s2 = run(bits, n, -1, 0, chain, 3, bits == 32 ? 1 : 2, 1, x, b);
    const double e = max_dist(a, b, n);
    const double tol = bits == 32 ? 128.0 * FLT_EPSILON * n : 512.0 * DBL_EPSILON * n;
    if (s1 || s2 || e > tol) FAIL("f%d pass grouping changed the result by %g", bits, e);
  }
  free(x); free(a); free(b);
}

static void check_layout_setters(void) {
  const size_t n = 4096;
  const int cases[][6] = {{8, 16, 32}, {8, 8, 64}, {8, 8, 8, 8}, {4, 4, 16, 16}, {4, 4, 4, 4, 4, 4}};
  const size_t depth[] = {3, 3, 4, 4, 6};
  double *x = malloc(16 * n);
  for (size_t i = 0; i < n; ++i) { x[2 * i] = sin(.017 * i); x[2 * i + 1] = cos(.013 * i); }
  for (size_t c = 0; c < 5; ++c)
    for (int sign = -1; sign <= 1; sign += 2) {
      mofft_plan_f32 *a = NULL, *b = NULL;
      if (mofft_plan_create_with_radices_f32(&a, n, sign, MOFFT_OUT_OF_PLACE, cases[c], depth[c]) ||
          mofft_plan_create_with_radices_f32(&b, n, sign, MOFFT_OUT_OF_PLACE, cases[c], depth[c]) ||
          mofft_plan_set_stage_layout_f32(b, MOFFT_STAGE_LAYOUT_SECTION) ||
          mofft_plan_set_direct_input_f32(b, 0)) {
        FAIL("layout setters case %zu", c);
      } else {
        mofft_complex_f32 *in = calloc(n, 8), *ya = calloc(n, 8), *yb = calloc(n, 8);
        for (size_t i = 0; i < n; ++i) { in[i].real = (float)x[2 * i]; in[i].imag = (float)x[2 * i + 1]; }
        if (mofft_execute_f32(a, in, ya) || mofft_execute_f32(b, in, yb)) FAIL("execute case %zu", c);
        double e = 0;
        for (size_t i = 0; i < n; ++i)
          e = fmax(e, hypot((double)ya[i].real - yb[i].real, (double)ya[i].imag - yb[i].imag));
        if (e > 128.0 * FLT_EPSILON * n) FAIL("section vs transpose case %zu err %g", c, e);
        free(in); free(ya); free(yb);
      }
      mofft_plan_destroy_f32(a); mofft_plan_destroy_f32(b);
    }
  if (mofft_plan_set_stage_layout_f32(NULL, MOFFT_STAGE_LAYOUT_SECTION) != MOFFT_INVALID_ARGUMENT)
    FAIL("NULL plan accepted");
  free(x);
}

static void check_tones_every_bank(void) {
  mofft_plan_f32 *probe = NULL;
  mofft_plan_options options;
  mofft_plan_options_init(&options);
  mofft_plan_diagnostics info;
  if (mofft_plan_create_with_options_f32(&probe, 16, MOFFT_FORWARD, MOFFT_OUT_OF_PLACE,
                                         &options, &info) != MOFFT_SUCCESS) {
    FAIL("probe plan");
    return;
  }
  mofft_plan_destroy_f32(probe);
  const size_t sizes[] = {121, 169, 196, 225, 256, 1024, 1728, 2197, 2744, 3375, 20736,
                          28561, 38416, 50625, 65536, 248832, 262144, 371293, 537824, 759375};
  for (unsigned bank = 0; bank < info.kernel_bank_count; ++bank)
    for (size_t s = 0; s < sizeof(sizes) / sizeof(sizes[0]); ++s)
      for (int sign = -1; sign <= 1; sign += 2) {
        const size_t n = sizes[s];
        mofft_complex_f64 *x = calloc(n, 16), *y = calloc(n, 16);
        for (size_t j = 0; j < n; ++j) {
          const double a = 2 * kPi * 3 * (double)j / n, b = 2 * kPi * 5 * (double)j / n;
          x[j].real = 0.25 + cos(a) + 0.5 * cos(b);
          x[j].imag = sin(a) - 0.5 * sin(b);
        }
        mofft_plan_f64 *p = NULL;
        mofft_status st = mofft_plan_create_f64(&p, n, sign, MOFFT_OUT_OF_PLACE);
        if (st == MOFFT_SUCCESS) st = mofft_plan_set_kernel_bank_f64(p, bank);
        if (st == MOFFT_SUCCESS) st = mofft_execute_f64(p, x, y);
        double e = 0;
        for (size_t k = 0; k < n && st == MOFFT_SUCCESS; ++k) {
          double want = k == 0 ? 0.25 : 0.0;
          if (k == (sign < 0 ? 3 : n - 3)) want += 1.0;
          if (k == (sign < 0 ? n - 5 : 5)) want += 0.5;
          e = fmax(e, hypot(y[k].real / n - want, y[k].imag / n));
        }
        if (st != MOFFT_SUCCESS || !(e <= 2e-11))
          FAIL("tones n=%zu bank=%u sign=%d status=%s err=%g", n, bank, sign, mofft_status_string(st), e);
        mofft_plan_destroy_f64(p); free(x); free(y);
      }
}

static void check_sizes(void) {
  mofft_plan_f32 *bad = NULL;
  if (mofft_plan_create_f32(&bad, 17, MOFFT_FORWARD, MOFFT_OUT_OF_PLACE) != MOFFT_UNSUPPORTED_SIZE)
    FAIL("17 accepted");
  const size_t ext[] = {256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072, 262144,
                        144, 169, 196, 225, 1728, 2197, 2744, 3375, 20736, 28561, 38416, 50625,
                        248832, 371293, 537824, 759375};
  for (size_t i = 0; i < sizeof(ext) / sizeof(ext[0]); ++i) {
    mofft_plan_f32 *a = NULL;
    mofft_plan_f64 *b = NULL;
    if (mofft_plan_create_f32(&a, ext[i], MOFFT_FORWARD, MOFFT_OUT_OF_PLACE) != MOFFT_SUCCESS ||
        mofft_plan_create_f64(&b, ext[i], MOFFT_FORWARD, MOFFT_OUT_OF_PLACE) != MOFFT_SUCCESS)
      FAIL("plan n=%zu", ext[i]);
    mofft_plan_destroy_f32(a);
    mofft_plan_destroy_f64(b);
  }
}

static void check_planning(void) {
  enum { N = 30 };
  mofft_complex_f32 in[N], out[N];
  mofft_complex_f64 in64[N], out64[N];
  double x[2 * N], ref[2 * N];
  for (size_t i = 0; i < N; ++i) {
    x[2 * i] = (float)(sin(i * 0.31) + (i % 5) * 0.07);
    x[2 * i + 1] = (float)(cos(i * 0.23) - (i % 7) * 0.03);
    in[i].real = (float)x[2 * i]; in[i].imag = (float)x[2 * i + 1];
    in64[i].real = x[2 * i]; in64[i].imag = x[2 * i + 1];
  }
  mofft_plan_options opt;
  mofft_plan_options_init(&opt);
  opt.mode = MOFFT_PLANNER_MEASURE;
  opt.candidate_limit = 4;
  opt.timing_samples = 3;
  for (int sign = -1; sign <= 1; sign += 2)
    for (int ip = 0; ip <= 1; ++ip) {
      mofft_plan_diagnostics d32, d64;
      mofft_plan_f32 *p32 = NULL;
      mofft_plan_f64 *p64 = NULL;
      const mofft_placement pl = ip ? MOFFT_IN_PLACE : MOFFT_OUT_OF_PLACE;
      if (mofft_plan_create_with_options_f32(&p32, N, sign, pl, &opt, &d32) ||
          mofft_plan_create_with_options_f64(&p64, N, sign, pl, &opt, &d64)) {
        FAIL("measured planner");
        mofft_plan_destroy_f32(p32); mofft_plan_destroy_f64(p64);
        continue;
      }
      if (!d32.measured || d32.candidates_measured == 0 || d32.candidates_measured > 4 ||
          !d64.measured || d64.kernel_bank >= d64.kernel_bank_count || d32.length != N ||
          d64.precision_bits != 64 || d32.direction != sign || d32.placement != pl)
        FAIL("measured diagnostics");
      dft(x, ref, N, sign);
      memcpy(out, in, sizeof(in));
      memcpy(out64, in64, sizeof(in64));
      mofft_status s1 = ip ? mofft_execute_f32(p32, out, out) : mofft_execute_f32(p32, in, out);
      mofft_status s2 = ip ? mofft_execute_f64(p64, out64, out64) : mofft_execute_f64(p64, in64, out64);
      double e32 = 0, e64 = 0;
      for (size_t k = 0; k < N; ++k) {
        e32 = fmax(e32, hypot(out[k].real - ref[2 * k], out[k].imag - ref[2 * k + 1]));
        e64 = fmax(e64, hypot(out64[k].real - ref[2 * k], out64[k].imag - ref[2 * k + 1]));
      }
      if (s1 || s2 || e32 > 2e-4 * N || e64 > 2e-11 * N) FAIL("measured plan result %g %g", e32, e64);
      mofft_plan_destroy_f32(p32); mofft_plan_destroy_f64(p64);
    }
  /* STRICT wisdom: no compiled entry for 17 or for 30 */
  mofft_plan_options strict;
  mofft_plan_options_init(&strict);
  strict.flags = MOFFT_PLAN_STRICT_WISDOM;
  mofft_plan_f32 *p = (mofft_plan_f32 *)(uintptr_t)1;
  if (mofft_plan_create_with_options_f32(&p, 30, MOFFT_FORWARD, MOFFT_OUT_OF_PLACE, &strict, NULL) !=
          MOFFT_WISDOM_NOT_FOUND || p != NULL)
    FAIL("strict wisdom miss");
  /* bound arrays */
  mofft_plan_f32 *b = mofft_plan_dft_1d_f32(N, in, out, MOFFT_FORWARD, MOFFT_FLAG_PRETUNED);
  if (!b || mofft_execute_plan_f32(b) != MOFFT_SUCCESS) FAIL("bound array plan");
  mofft_plan_destroy_f32(b);
  if (mofft_plan_dft_1d_f32(N, in, out, MOFFT_FORWARD, MOFFT_FLAG_MEASURE | MOFFT_FLAG_WISDOM_ONLY))
    FAIL("contradictory flags accepted");
  b = mofft_plan_dft_1d_f32(N, in, out, MOFFT_FORWARD, MOFFT_FLAG_MEASURE);
  if (!b || mofft_execute_plan_f32(b) != MOFFT_SUCCESS) FAIL("bound measure plan");
  mofft_plan_destroy_f32(b);
  /* placement must match the pointers */
  mofft_plan_f32 *q = NULL;
  if (mofft_plan_create_f32(&q, N, MOFFT_FORWARD, MOFFT_IN_PLACE) ||
      mofft_execute_f32(q, in, out) != MOFFT_INVALID_ARGUMENT)
    FAIL("in-place plan ran out of place");
  mofft_plan_destroy_f32(q);
}

static void check_recipes(void) {
  const size_t n = 96;
  mofft_plan_f64 *p = NULL, *q = NULL;
  const int chain[] = {4, 8, 3};
  if (mofft_plan_create_with_radices_f64(&p, n, MOFFT_BACKWARD, MOFFT_IN_PLACE, chain, 3) ||
      mofft_plan_set_blocked_transpose_mask_f64(p, 1)) {
    FAIL("recipe source plan");
    return;
  }
  size_t need = 0;
  if (mofft_plan_export_f64(p, NULL, 0, &need) != MOFFT_SUCCESS || need < 20) FAIL("export size");
  char small[8];
  memset(small, 'x', sizeof(small));
  size_t need2 = 0;
  if (mofft_plan_export_f64(p, small, sizeof(small), &need2) != MOFFT_BUFFER_TOO_SMALL ||
      need2 != need || small[0] != 'x')
    FAIL("undersized export");
  char *text = malloc(need);
  if (mofft_plan_export_f64(p, text, need, &need2) != MOFFT_SUCCESS) FAIL("export");
  if (strncmp(text, "MOFFT_PLAN_V1 ", 14) != 0 || strlen(text) + 1 != need) FAIL("recipe text");
  mofft_plan_diagnostics d;
  if (mofft_plan_import_f64(&q, text, need, &d) != MOFFT_SUCCESS || !q) FAIL("import");
  if (d.length != n || d.radix_count != 3 || d.radices[2] != 3 || d.direction != MOFFT_BACKWARD ||
      d.placement != MOFFT_IN_PLACE || !d.used_wisdom)
    FAIL("import diagnostics");
  mofft_complex_f64 *x = calloc(n, 16), *y = calloc(n, 16);
  for (size_t i = 0; i < n; ++i) { x[i].real = sin(i * .3); x[i].imag = cos(i * .7); y[i] = x[i]; }
  if (mofft_execute_f64(p, x, x) || mofft_execute_f64(q, y, y)) FAIL("replay");
  for (size_t i = 0; i < n; ++i)
    if (x[i].real != y[i].real || x[i].imag != y[i].imag) { FAIL("replay differs at %zu", i); break; }
  mofft_plan_f32 *r = (mofft_plan_f32 *)(uintptr_t)1;
  if (mofft_plan_import_f32(&r, text, need, NULL) != MOFFT_INCOMPATIBLE_RECIPE || r) FAIL("precision mismatch");
  text[need - 2] = ' ';
  if (mofft_plan_import_f64(&q, text, need, NULL) != MOFFT_INVALID_ARGUMENT) FAIL("malformed accepted");
  mofft_plan_destroy_f64(p); mofft_plan_destroy_f64(q);
  free(text); free(x); free(y);
}

static void check_cost_model(void) {
  const int r[] = {8, 8, 8, 8};
  mofft_plan_cost_breakdown s, l, b;
  if (mofft_estimated_plan_cost_breakdown_with_layout(4096, 64, r, 4, 0, 0, MOFFT_STAGE_LAYOUT_SECTION, &s) ||
      mofft_estimated_plan_cost_breakdown_with_layout(4096, 64, r, 4, 0, 0, MOFFT_STAGE_LAYOUT_TRANSPOSE, &l) ||
      mofft_estimated_plan_cost_breakdown_with_layout(4096, 64, r, 4, 7, 0, MOFFT_STAGE_LAYOUT_TRANSPOSE, &b))
    FAIL("breakdown status");
  if (s.logical_read_bytes != 336896 || s.logical_write_bytes != 262144 ||
      s.transferred_read_bytes != 336896 || s.transferred_write_bytes != 262144 ||
      s.peak_working_set_bytes != 196608 || s.layout_cost != 0.0 ||
      l.logical_read_bytes != 533504 || l.logical_write_bytes != 458752 ||
      l.transferred_read_bytes != 992256 || l.transferred_write_bytes != 458752 ||
      b.transferred_read_bytes != 533504 || !(l.layout_cost > 0.0))
    FAIL("byte known answers");
}

int main(void) {
  check_cost_model();
  check_sizes();
  check_direct();
  check_roundtrip();
  check_specific();
  check_two_stage_radices();
  check_mask_grouping();
  check_layout_setters();
  check_tones_every_bank();
  check_planning();
  check_recipes();
  if (failures) {
    fprintf(stderr, "%d drop-in checks failed\n", failures);
    return 1;
  }
  printf("all drop-in checks passed\n");
  return 0;
}
