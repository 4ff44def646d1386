/*
 * mofft_tune - GPU autotuner that writes the wisdom MOFFT_FLAG_PRETUNED reads.
 *
 * Role of the reference's tools/plan_search.cpp (model shortlist, measured
 * finalists, JSON wisdom, plan_search.cpp:100-639, 782-882) for the MI355X
 * build: for every (length, precision) it runs the library's MEASURE planner
 * over the GPU candidate space (pass decompositions x pass boundaries x
 * kernel banks / tile shapes, timed with HIP events on a batch of transforms),
 * re-times the winner and the PRETUNED plan (the compiled wisdom's entry,
 * else the analytical model's plan) in an interleaved steady
 * phase, and records the faster one.  The JSON goes through
 * mofft_compiler.wisdom into the library at the next build.
 *
 * MEASURE times at most 64 candidates ranked by the cost model (the option
 * limit of the reference); the model can rank a fast kernel low.  So the
 * tuner also sweeps the kernel registry itself: every single-pass row
 * kernel of the length, every split into two column-kernel passes whose
 * lengths multiply to it (and, above 2^20 points, into three passes of at
 * least 64 points), each under every plain tile bank, timed the same way;
 * the fastest of the MEASURE winner, the sweep winner and the PRETUNED plan is
 * recorded, in the faster of its two pass-twiddle layouts (stage_layout:
 * full per-pass tables or the two-level table), after its output for a
 * seeded input has been checked against the PRETUNED plan's (relative L2
 * within four times the precision's bound).
 *
 * usage: mofft_tune [--precision 32|64|both] [--lengths N,N,...]
 *                   [--min-log2 A --max-log2 B] [--points P]
 *                   [--candidates C] [--samples S] [--no-sweep]
 *                   -o wisdom.json
 *        mofft_tune --pass-costs passes.json [--precision ...] [--points P]
 *
 * --pass-costs measures the analytical planner's cost table instead: every
 * compiled pass kernel timed in the roles it plays in plans, strided ones
 * at four strides / spans (the planner's
 * counterpart of the reference's calibrated cost profile,
 * compiler/mofft_compiler/emitter.py:621-725, profiles/apple-m5-bootstrap.json).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mofft_hip.h"

namespace {

struct Result {
  size_t length;
  int precision;
  std::vector<int> radices;
  uint64_t mask;
  unsigned bank;
  int layout;  /* mofft_stage_layout: the pass-twiddle table layout */
  double seconds;
  double model_seconds;
  size_t batch;
  size_t candidates;
};

void *plan_create(int bits, size_t n, size_t batch, const mofft_plan_options *o,
                  mofft_plan_diagnostics *d, mofft_status *st) {
  void *p = nullptr;
  if (bits == 32)
    *st = mofft_plan_create_batched_f32((mofft_plan_f32 **)&p, n, batch, MOFFT_FORWARD,
                                        MOFFT_OUT_OF_PLACE, o, d);
  else
    *st = mofft_plan_create_batched_f64((mofft_plan_f64 **)&p, n, batch, MOFFT_FORWARD,
                                        MOFFT_OUT_OF_PLACE, o, d);
  return p;
}

void plan_destroy(int bits, void *p) {
  if (bits == 32) mofft_plan_destroy_f32((mofft_plan_f32 *)p);
  else mofft_plan_destroy_f64((mofft_plan_f64 *)p);
}

int pass_count(int bits, void *p) {
  mofft_plan_description d;
  const mofft_status st =
      bits == 32 ? mofft_plan_get_description_f32((const mofft_plan_f32 *)p, &d)
                 : mofft_plan_get_description_f64((const mofft_plan_f64 *)p, &d);
  return st == MOFFT_SUCCESS ? (int)d.pass_count : 0;
}

bool set_layout(int bits, void *p, mofft_stage_layout layout) {
  const mofft_status st =
      bits == 32 ? mofft_plan_set_stage_layout_f32((mofft_plan_f32 *)p, layout)
                 : mofft_plan_set_stage_layout_f64((mofft_plan_f64 *)p, layout);
  return st == MOFFT_SUCCESS;
}

mofft_status run(int bits, void *p, void *in, void *out, hipStream_t s) {
  if (bits == 32)
    return mofft_execute_async_f32((const mofft_plan_f32 *)p, (const mofft_complex_f32 *)in,
                                   (mofft_complex_f32 *)out, s);
  return mofft_execute_async_f64((const mofft_plan_f64 *)p, (const mofft_complex_f64 *)in,
                                 (mofft_complex_f64 *)out, s);
}

/* median device time of `reps` executions, each bracketed by events */
double steady(int bits, void *p, void *in, void *out, hipStream_t s, int reps) {
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  std::vector<double> t;
  for (int i = 0; i < 3; ++i) run(bits, p, in, out, s);
  for (int i = 0; i < reps; ++i) {
    hipEventRecord(a, s);
    run(bits, p, in, out, s);
    hipEventRecord(b, s);
    hipEventSynchronize(b);
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    t.push_back(ms * 1e-3);
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

/* Relative L2 distance between two plans' outputs for one seeded input
 * (at most the first 2^20 points of the batch), so that a timing never
 * stands in for a result: a plan is recorded only when it agrees with the
 * PRETUNED plan. */
double disagreement(int bits, void *a, void *b, void *in, void *out, void *out2,
                    size_t n, size_t batch, hipStream_t s) {
  const size_t cb = bits == 32 ? 8 : 16;
  if (mofft_server_generate(in, bits, n * batch, 2027, s) != MOFFT_SUCCESS ||
      run(bits, a, in, out, s) != MOFFT_SUCCESS ||
      run(bits, b, in, out2, s) != MOFFT_SUCCESS ||
      hipStreamSynchronize(s) != hipSuccess)
    return INFINITY;
  const size_t count = std::min(n * batch, std::max(n, (size_t)1 << 20));
  std::vector<double> x(2 * count), y(2 * count);
  std::vector<unsigned char> hx(count * cb), hy(count * cb);
  if (hipMemcpy(hx.data(), out, count * cb, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hy.data(), out2, count * cb, hipMemcpyDeviceToHost) != hipSuccess)
    return INFINITY;
  for (size_t i = 0; i < 2 * count; ++i) {
    x[i] = bits == 32 ? ((const float *)hx.data())[i] : ((const double *)hx.data())[i];
    y[i] = bits == 32 ? ((const float *)hy.data())[i] : ((const double *)hy.data())[i];
  }
  double num = 0, den = 0;
  for (size_t i = 0; i < 2 * count; ++i) {
    num += (x[i] - y[i]) * (x[i] - y[i]);
    den += y[i] * y[i];
  }
  return den > 0 ? std::sqrt(num / den) : INFINITY;
}

/* A chain (outer -> inner) and pass mask built from registry kernels. */
struct Chain {
  std::vector<int> radices;
  uint64_t mask;
  bool operator==(const Chain &o) const {
    return radices == o.radices && mask == o.mask;
  }
};

struct Kernel {
  int length;
  std::vector<int> stages;  /* execution order */
};

/* Distinct (length, stages) kernels of one precision and kind. */
std::vector<Kernel> registry_kernels(int bits, bool rows) {
  std::vector<Kernel> out;
  for (int i = 0; i < mofft_kernel_count(); ++i) {
    mofft_kernel_info k;
    if (mofft_kernel_get_info(i, &k) != MOFFT_SUCCESS) continue;
    if (k.precision_bits != bits || (k.rows != 0) != rows) continue;
    Kernel e{k.length, std::vector<int>(k.stages, k.stages + k.stage_count)};
    bool seen = false;
    for (const Kernel &o : out) seen |= o.length == e.length && o.stages == e.stages;
    if (!seen) out.push_back(e);
  }
  return out;
}

/* Passes in execution order -> chain: the last pass's stages first, each
 * reversed; a boundary bit after every pass but the first. */
Chain chain_of(const std::vector<const Kernel *> &passes) {
  Chain c{{}, 0};
  for (size_t p = passes.size(); p-- > 0;) {
    const std::vector<int> &st = passes[p]->stages;
    for (size_t i = st.size(); i-- > 0;) c.radices.push_back(st[i]);
    if (p > 0) c.mask |= (uint64_t)1 << (c.radices.size() - 1);
  }
  return c;
}

std::vector<Chain> registry_chains(int bits, size_t n) {
  std::vector<Chain> out;
  auto add = [&](const Chain &c) {
    if (std::find(out.begin(), out.end(), c) == out.end()) out.push_back(c);
  };
  for (const Kernel &k : registry_kernels(bits, true))
    if ((size_t)k.length == n) add(chain_of({&k}));
  const std::vector<Kernel> cols = registry_kernels(bits, false);
  for (const Kernel &a : cols) {
    if (n % (size_t)a.length) continue;
    for (const Kernel &b : cols)
      if ((size_t)b.length == n / (size_t)a.length) add(chain_of({&a, &b}));
  }
  /* three passes only where two would need tiles of more than 1024
   * points: passes of >= 64 points, the first two stage chains of each
   * pass length */
  if (n <= ((size_t)1 << 20)) return out;
  std::vector<const Kernel *> c3;
  for (const Kernel &k : cols) {
    if (k.length < 64) continue;
    int same = 0;
    for (const Kernel *o : c3) same += o->length == k.length;
    if (same < 2) c3.push_back(&k);
  }
  for (const Kernel *a : c3) {
    if (n % (size_t)a->length) continue;
    const size_t r = n / (size_t)a->length;
    for (const Kernel *b : c3) {
      if (r % (size_t)b->length) continue;
      for (const Kernel *c : c3)
        if ((size_t)c->length == r / (size_t)b->length) add(chain_of({a, b, c}));
    }
  }
  return out;
}

void *plan_with_chain(int bits, size_t n, size_t batch, const Chain &c,
                      unsigned bank) {
  void *p = nullptr;
  mofft_status st;
  if (bits == 32)
    st = mofft_plan_create_batched_with_radices_f32(
        (mofft_plan_f32 **)&p, n, batch, MOFFT_FORWARD, MOFFT_OUT_OF_PLACE,
        c.radices.data(), c.radices.size());
  else
    st = mofft_plan_create_batched_with_radices_f64(
        (mofft_plan_f64 **)&p, n, batch, MOFFT_FORWARD, MOFFT_OUT_OF_PLACE,
        c.radices.data(), c.radices.size());
  if (st != MOFFT_SUCCESS) return nullptr;
  if (bits == 32)
    st = mofft_plan_set_blocked_transpose_mask_f32((mofft_plan_f32 *)p, c.mask);
  else
    st = mofft_plan_set_blocked_transpose_mask_f64((mofft_plan_f64 *)p, c.mask);
  if (st == MOFFT_SUCCESS)
    st = bits == 32 ? mofft_plan_set_kernel_bank_f32((mofft_plan_f32 *)p, bank)
                    : mofft_plan_set_kernel_bank_f64((mofft_plan_f64 *)p, bank);
  /* regroup under the bank: a chain whose kernel exists only there (the
   * radix-32 matrix-core whole-transform kernels) groups into fewer passes */
  if (st == MOFFT_SUCCESS && bank != 0)
    st = bits == 32
        ? mofft_plan_set_blocked_transpose_mask_f32((mofft_plan_f32 *)p, c.mask)
        : mofft_plan_set_blocked_transpose_mask_f64((mofft_plan_f64 *)p, c.mask);
  if (st != MOFFT_SUCCESS) {
    plan_destroy(bits, p);
    return nullptr;
  }
  return p;
}

/* Per-pass device seconds of one execution (median over `reps`), from
 * the library's timed execute (events between pass launches). */
bool pass_times(int bits, void *p, void *in, void *out, hipStream_t s, int reps,
                std::vector<double> *secs) {
  const int cap = 16;
  std::vector<std::vector<double>> t(cap);
  for (int i = 0; i < 3 + reps; ++i) {
    double v[cap];
    const mofft_status st =
        bits == 32 ? mofft_execute_timed_f32((const mofft_plan_f32 *)p, (const mofft_complex_f32 *)in,
                                             (mofft_complex_f32 *)out, s, v, cap)
                   : mofft_execute_timed_f64((const mofft_plan_f64 *)p, (const mofft_complex_f64 *)in,
                                             (mofft_complex_f64 *)out, s, v, cap);
    if (st != MOFFT_SUCCESS) return false;
    if (i >= 3)
      for (int k = 0; k < cap; ++k) t[k].push_back(v[k]);
  }
  secs->clear();
  for (int k = 0; k < cap; ++k) {
    std::sort(t[k].begin(), t[k].end());
    secs->push_back(t[k][t[k].size() / 2]);
  }
  return true;
}

/* The pass-cost table of the analytical planner: every registry kernel
 * timed as the pass it is in plans, on `points` points (one HBM read and
 * write of them per pass) - a whole-transform kernel as the single pass of
 * N = L ("rows"); a strided kernel as the first pass of N = L x S with
 * stride S ("first", no pass twiddles) and as the second pass of N = S x L
 * with span S ("later", pass twiddles), for S = 16, 64, 256 and 1024 (the
 * planner interpolates in log2 S), the other pass a bank-0 strided kernel
 * of length S.  Each plan's description must name the kernel (else the
 * case is skipped).  Writes JSON that mofft_compiler compiles into the
 * library. */
int pass_costs(const std::vector<int> &precisions, size_t points, int reps,
               const char *path) {
  struct Row {
    int bits, rows, length, variant, span;
    std::vector<int> stages;
    const char *role;
    double bytes, seconds;
  };
  std::vector<Row> rows;
  hipStream_t stream = nullptr;
  const int spans[] = {16, 64, 256, 1024};
  for (int bits : precisions) {
    const size_t cb = bits == 32 ? 8 : 16;
    const size_t total = bits == 32 ? points : points / 2;
    void *in = nullptr, *out = nullptr;
    if (hipMalloc(&in, total * cb) != hipSuccess || hipMalloc(&out, total * cb) != hipSuccess) {
      fprintf(stderr, "pass costs: allocation failed\n");
      return 1;
    }
    if (mofft_server_generate(in, bits, total, 2026, stream) != MOFFT_SUCCESS ||
        hipStreamSynchronize(stream) != hipSuccess)
      return 1;
    /* the strided pass of each span length the kernels are paired with */
    std::vector<std::vector<int>> pair(4);
    for (int q = 0; q < 4; ++q)
      for (int i = 0; i < mofft_kernel_count(); ++i) {
        mofft_kernel_info k;
        if (mofft_kernel_get_info(i, &k) == MOFFT_SUCCESS && k.precision_bits == bits &&
            !k.rows && k.length == spans[q] && k.variant == 0) {
          pair[q].assign(k.stages, k.stages + k.stage_count);
          break;
        }
      }
    for (int i = 0; i < mofft_kernel_count(); ++i) {
      mofft_kernel_info k;
      if (mofft_kernel_get_info(i, &k) != MOFFT_SUCCESS || k.precision_bits != bits) continue;
      std::vector<int> st(k.stages, k.stages + k.stage_count);
      std::vector<int> inner(st.rbegin(), st.rend());     /* outer -> inner */
      struct Case { const char *role; size_t n; Chain c; int slot; int span; };
      std::vector<Case> cases;
      if (k.rows) {
        cases.push_back({"rows", (size_t)k.length, Chain{inner, 0}, 0, 1});
      } else {
        for (int q = 0; q < 4; ++q) {
          if (pair[q].empty()) continue;
          std::vector<int> pr(pair[q].rbegin(), pair[q].rend());
          Chain first{pr, 0};
          first.radices.insert(first.radices.end(), inner.begin(), inner.end());
          first.mask = (uint64_t)1 << (pr.size() - 1);
          Chain later{inner, 0};
          later.radices.insert(later.radices.end(), pr.begin(), pr.end());
          later.mask = (uint64_t)1 << (inner.size() - 1);
          const size_t n = (size_t)k.length * (size_t)spans[q];
          if (n > total) continue;
          /* the wide bank widens only the passes after the first */
          if (k.variant != 4) cases.push_back({"first", n, first, 0, spans[q]});
          cases.push_back({"later", n, later, 1, spans[q]});
        }
      }
      for (const Case &c : cases) {
        const size_t batch = std::max<size_t>(1, total / c.n);
        void *p = plan_with_chain(bits, c.n, batch, c.c, (unsigned)k.variant);
        if (!p) continue;
        mofft_plan_description d;
        const mofft_status ds = bits == 32
            ? mofft_plan_get_description_f32((const mofft_plan_f32 *)p, &d)
            : mofft_plan_get_description_f64((const mofft_plan_f64 *)p, &d);
        std::vector<double> secs;
        if (ds == MOFFT_SUCCESS && d.pass_kernel[c.slot] == i &&
            pass_times(bits, p, in, out, stream, reps, &secs)) {
          rows.push_back({bits, k.rows, k.length, k.variant, c.span, st, c.role,
                          2.0 * (double)(c.n * batch * cb), secs[c.slot]});
        }
        plan_destroy(bits, p);
      }
      if (i % 100 == 0) {
        printf("pass costs: f%d kernel %d of %d\n", bits, i, mofft_kernel_count());
        fflush(stdout);
      }
    }
    hipFree(in);
    hipFree(out);
  }
  FILE *f = fopen(path, "w");
  if (!f) return 1;
  fprintf(f, "{\n \"schema_version\": 2,\n \"target\": \"gfx950\",\n");
  fprintf(f, " \"source\": \"mofft_tune --pass-costs\",\n \"points\": %zu,\n", points);
  fprintf(f, " \"build_id\": \"%s\",\n \"entries\": [\n", mofft_build_id());
  for (size_t i = 0; i < rows.size(); ++i) {
    const Row &r = rows[i];
    fprintf(f, "  {\"precision\": %d, \"rows\": %d, \"length\": %d, \"stages\": [",
            r.bits, r.rows, r.length);
    for (size_t k = 0; k < r.stages.size(); ++k) fprintf(f, "%s%d", k ? ", " : "", r.stages[k]);
    fprintf(f, "], \"variant\": %d, \"role\": \"%s\", \"span\": %d, \"bytes\": %.0f, "
               "\"seconds\": %.9g}%s\n",
            r.variant, r.role, r.span, r.bytes, r.seconds, i + 1 < rows.size() ? "," : "");
  }
  fprintf(f, " ]\n}\n");
  fclose(f);
  printf("wrote %zu pass costs to %s\n", rows.size(), path);
  return 0;
}

std::vector<size_t> parse_list(const char *s) {
  std::vector<size_t> v;
  while (*s) {
    char *end = nullptr;
    v.push_back(strtoull(s, &end, 10));
    s = *end == ',' ? end + 1 : end;
    if (end == s && *s) break;
  }
  return v;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<int> precisions = {32, 64};
  std::vector<size_t> lengths;
  int lo = 4, hi = 20;
  size_t points = (size_t)1 << 24;
  size_t candidates = 32;
  unsigned samples = 5;
  bool sweep = true;
  const char *out_path = "wisdom.json";
  const char *pass_costs_path = nullptr;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() { return i + 1 < argc ? argv[++i] : (char *)""; };
    if (a == "--precision") {
      std::string v = next();
      precisions = v == "32" ? std::vector<int>{32} : v == "64" ? std::vector<int>{64} : std::vector<int>{32, 64};
    } else if (a == "--lengths") lengths = parse_list(next());
    else if (a == "--min-log2") lo = atoi(next());
    else if (a == "--max-log2") hi = atoi(next());
    else if (a == "--points") points = strtoull(next(), nullptr, 10);
    else if (a == "--candidates") candidates = strtoull(next(), nullptr, 10);
    else if (a == "--samples") samples = (unsigned)atoi(next());
    else if (a == "--no-sweep") sweep = false;
    else if (a == "--pass-costs") pass_costs_path = next();
    else if (a == "-o") out_path = next();
    else { fprintf(stderr, "unknown argument %s\n", a.c_str()); return 2; }
  }
  if (lengths.empty())
    for (int k = lo; k <= hi; ++k) lengths.push_back((size_t)1 << k);
  if (!mofft_device_available()) {
    fprintf(stderr, "mofft_tune: no gfx950 device\n");
    return 1;
  }
  if (pass_costs_path) return pass_costs(precisions, points, (int)samples * 2, pass_costs_path);
  hipStream_t stream = nullptr;
  std::vector<Result> results;
  size_t mismatches = 0;
  for (int bits : precisions) {
    const size_t cb = bits == 32 ? 8 : 16;
    for (size_t n : lengths) {
      const size_t batch = std::max<size_t>(1, points / n);
      void *in = nullptr, *out = nullptr, *out2 = nullptr;
      if (hipMalloc(&in, n * batch * cb) != hipSuccess ||
          hipMalloc(&out, n * batch * cb) != hipSuccess ||
          hipMalloc(&out2, n * batch * cb) != hipSuccess) {
        fprintf(stderr, "allocation failed for N=%zu\n", n);
        return 1;
      }
      /* seeded uniform input, as the transforms will see: all-zero
       * buffers time 10-15 % faster on MI355X for multi-pass plans */
      if (mofft_server_generate(in, bits, n * batch, 2026, stream) != MOFFT_SUCCESS ||
          hipStreamSynchronize(stream) != hipSuccess) {
        fprintf(stderr, "input generation failed for N=%zu\n", n);
        return 1;
      }
      mofft_plan_options o;
      mofft_plan_options_init(&o);
      o.mode = MOFFT_PLANNER_MEASURE;
      o.candidate_limit = candidates;
      o.timing_samples = samples;
      mofft_plan_diagnostics d;
      mofft_status st;
      void *measured = plan_create(bits, n, batch, &o, &d, &st);
      if (st != MOFFT_SUCCESS) {
        fprintf(stderr, "N=%zu f%d: %s\n", n, bits, mofft_status_string(st));
        hipFree(in); hipFree(out); hipFree(out2);
        continue;
      }
      mofft_plan_diagnostics dm;
      void *model = plan_create(bits, n, batch, nullptr, &dm, &st);
      /* registry sweep: every kernel-backed chain under every plain bank */
      void *swept = nullptr;
      Chain swept_chain{{}, 0};
      unsigned swept_bank = 0;
      size_t swept_count = 0;
      if (sweep) {
        double best = 0;
        for (const Chain &c : registry_chains(bits, n)) {
          for (unsigned bank : {0u, 1u, 2u, 4u, 5u}) {
            void *p = plan_with_chain(bits, n, batch, c, bank);
            if (!p) continue;
            ++swept_count;
            const double t = steady(bits, p, in, out, stream, samples);
            if (!swept || t < best) {
              if (swept) plan_destroy(bits, swept);
              swept = p;
              swept_chain = c;
              swept_bank = bank;
              best = t;
            } else {
              plan_destroy(bits, p);
            }
          }
        }
      }
      /* interleaved steady timing of the finalists (plan_search.cpp:612-622) */
      double tw = 0, tm = 0, tsw = 0;
      for (int round = 0; round < 3; ++round) {
        tw += steady(bits, measured, in, out, stream, 15);
        if (model) tm += steady(bits, model, in, out, stream, 15);
        if (swept) tsw += steady(bits, swept, in, out, stream, 15);
      }
      Result r;
      r.length = n;
      r.precision = bits;
      r.batch = batch;
      r.candidates = d.candidates_measured + swept_count;
      r.model_seconds = tm / 3;
      /* the PRETUNED plan ("model": wisdom entry or analytical plan) unless
       * a measured one is at least 1% faster */
      const mofft_plan_diagnostics *win = model ? &dm : &d;
      double tbest = model ? tm : tw;
      const char *which = model ? "model" : "measured";
      void *chosen = model ? model : measured;
      if (!model || tw <= tbest * 0.99) {
        win = &d;
        tbest = tw;
        which = "measured";
        chosen = measured;
      }
      if (swept && tsw <= tbest * 0.99) {
        tbest = tsw;
        which = "swept";
        chosen = swept;
        r.radices = swept_chain.radices;
        r.mask = swept_chain.mask;
        r.bank = swept_bank;
      } else {
        r.radices.assign(win->radices, win->radices + win->radix_count);
        r.mask = win->blocked_transpose_mask;
        r.bank = win->kernel_bank;
      }
      r.seconds = tbest / 3;
      /* the winner's pass-twiddle layout (mofft_plan_set_stage_layout:
       * full per-pass tables or the two-level table) against the other,
       * interleaved; the other is kept when it is at least 1% faster */
      r.layout = chosen == model ? (int)dm.stage_layout
                                 : chosen == measured ? (int)d.stage_layout
                                                      : (int)MOFFT_STAGE_LAYOUT_TRANSPOSE;
      if (pass_count(bits, chosen) > 1) {
        const mofft_stage_layout cur = (mofft_stage_layout)r.layout;
        const mofft_stage_layout alt = cur == MOFFT_STAGE_LAYOUT_SECTION
                                           ? MOFFT_STAGE_LAYOUT_TRANSPOSE
                                           : MOFFT_STAGE_LAYOUT_SECTION;
        double tc = 0, ta = 0;
        bool ok = true;
        for (int round = 0; ok && round < 3; ++round) {
          ok = set_layout(bits, chosen, cur);
          if (ok) tc += steady(bits, chosen, in, out, stream, 15);
          ok = ok && set_layout(bits, chosen, alt);
          if (ok) ta += steady(bits, chosen, in, out, stream, 15);
        }
        if (ok && ta <= tc * 0.99) {
          r.layout = (int)alt;
          r.seconds = ta / 3;
        }
        if (!set_layout(bits, chosen, (mofft_stage_layout)r.layout)) {
          fprintf(stderr, "N=%zu f%d: stage layout could not be set\n", n, bits);
          return 1;
        }
      }
      if (model && chosen != model) {
        const double tol = (bits == 32 ? 1e-6 : 1e-15) * std::max(1.0, std::log2((double)n));
        const double dis = disagreement(bits, chosen, model, in, out, out2, n, batch, stream);
        if (!(dis <= 4 * tol)) {
          printf("f%d N=%zu: the %s plan disagrees with the model plan (relative L2 "
                 "%.3e > %.1e); no entry recorded\n", bits, n, which, dis, 4 * tol);
          ++mismatches;
          plan_destroy(bits, measured);
          plan_destroy(bits, model);
          if (swept) plan_destroy(bits, swept);
          hipFree(in); hipFree(out); hipFree(out2);
          continue;
        }
      }
      results.push_back(r);
      const double gbs = 2.0 * n * batch * cb / r.seconds / 1e9;
      printf("f%d N=%zu batch=%zu: %s plan, %.3f us (model %.3f us, measured %.3f us, "
             "swept %.3f us), %.0f GB/s, %zu candidates\n",
             bits, n, batch, which, r.seconds * 1e6, r.model_seconds * 1e6, tw / 3 * 1e6,
             tsw / 3 * 1e6, gbs, r.candidates);
      fflush(stdout);
      plan_destroy(bits, measured);
      if (model) plan_destroy(bits, model);
      if (swept) plan_destroy(bits, swept);
      hipFree(in);
      hipFree(out);
      hipFree(out2);
    }
  }
  FILE *f = fopen(out_path, "w");
  if (!f) return 1;
  fprintf(f, "{\n \"schema_version\": 1,\n \"target\": \"gfx950\",\n");
  fprintf(f, " \"formal_measurement\": false,\n \"source\": \"mofft_tune\",\n");
  fprintf(f, " \"build_id\": \"%s\",\n \"entries\": [\n", mofft_build_id());
  for (size_t i = 0; i < results.size(); ++i) {
    const Result &r = results[i];
    fprintf(f, "  {\"length\": %zu, \"precision\": \"fp%d\", \"radices\": [", r.length, r.precision);
    for (size_t k = 0; k < r.radices.size(); ++k)
      fprintf(f, "%s%d", k ? ", " : "", r.radices[k]);
    fprintf(f, "], \"blocked_transpose_mask\": %llu, \"kernel_bank\": %u, "
               "\"direct_input\": false, \"stage_layout\": \"%s\", "
               "\"median_seconds\": %.9g, \"model_seconds\": %.9g, \"batch\": %zu}%s\n",
            (unsigned long long)r.mask, r.bank,
            r.layout == (int)MOFFT_STAGE_LAYOUT_SECTION ? "section" : "transpose",
            r.seconds, r.model_seconds, r.batch, i + 1 < results.size() ? "," : "");
  }
  fprintf(f, " ]\n}\n");
  fclose(f);
  printf("wrote %zu entries to %s\n", results.size(), out_path);
  if (mismatches) printf("%zu plans disagreed with the model plan\n", mismatches);
  return mismatches ? 3 : 0;
}
