/*
 * mofft_batch.c - host side of the persistent uneven batch (mofft_hip.h).
 *
 * A batch plan turns `count` independent transforms (each with its own
 * length, direction and buffers) into one queue of tiles for the persistent
 * kernel of csrc/hip/batch.h:
 *
 *   - every distinct length gets one schedule over the batch tile kinds
 *     (wisdom chain when its passes exist as kinds, else the analytical
 *     chain) and one resident twiddle table, shared by all its transforms
 *     and by both directions (backward runs as swap(F(swap(x))));
 *   - transforms whose schedule is one row tile are gathered: a segment per
 *     length whose tile columns are different transforms, one queue item
 *     per tile;
 *   - a multi-pass transform gets one segment per pass; in-place and
 *     three-or-more-pass transforms get a private slice of the workspace;
 *   - a multi-pass transform of at most twice the fair share of tiles per
 *     workgroup of the grid is OWNED (one queue item, all passes run by one
 *     workgroup); a larger one is SHARED (items of a few tiles, later passes
 *     wait for earlier ones through counters);
 *   - queue order: shared first passes, owned transforms longest first,
 *     gathered segments longest first, shared later passes level by level
 *     (the level-major order is what makes the dependency waits finite).
 *
 * Reference counterpart: none (the reference executes one transform per
 * call, src/mofft.c:2169-2201).  The per-transform buffer rules are the
 * ones of the single-plan executor in mofft.c.
 */
#include "mofft_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Polls of one dependency wait before the launch gives up.  A poll sleeps
 * about 0.2 us, so the default is roughly a second - orders of magnitude
 * above any tile's duration. */
#define MOFFT_BATCH_DEFAULT_SPIN_LIMIT ((uint64_t)1 << 22)

typedef struct {
  size_t length;
  mofft_schedule s;
  void *twiddles;
  int from_wisdom;
  size_t gather_first;  /* first slot in the job table (row schedules) */
  size_t gather_count;
} batch_length;

struct mofft_batch_plan {
  int precision_bits;
  size_t count;
  int *directions;
  unsigned char *in_place;
  int *length_index;          /* job -> batch_length */
  size_t *work_offset;        /* job -> byte offset in workspace, or SIZE_MAX */
  batch_length *bl;
  size_t bl_count;
  mofft_batch_segment *segs;  /* host copy */
  size_t *seg_job;            /* job of a pass segment, SIZE_MAX if gathered */
  int *seg_pass;
  size_t seg_count;
  size_t queue_count;         /* segments that carry queue items */
  unsigned char *shared;      /* job -> 1 if its passes are shared tiles */
  mofft_batch_job *jobs;      /* host copy of the gathered job table */
  size_t *slot_job;
  size_t job_slots;
  int64_t total_items;
  void *d_segs;
  void *d_jobs;
  void *d_counters;
  void *workspace;
  size_t workspace_bytes;
  size_t twiddle_bytes;
  int grid;
  int blocks_per_cu;
  int lds_bytes;
  void *graph;
  mofft_batch_header header;
  mofft_batch_options options;
  size_t multi_pass_jobs;
  size_t shared_jobs;
  int64_t owned_max;
  int64_t item_tiles;
  size_t wisdom_lengths;
  double flops;
  double bytes;
};

void mofft_batch_options_init(mofft_batch_options *options) {
  if (options) memset(options, 0, sizeof(*options));
}

static int is_pow2(size_t n) { return n >= 2 && (n & (n - 1)) == 0; }

/* The schedule one length uses in a batch: the wisdom chain when the batch
 * tile kinds run it with the same passes as the tuned plan and no more than
 * the analytical chain over those kinds, else the analytical chain. */
static mofft_status batch_schedule(int precision_bits, size_t length,
                                   mofft_schedule *s, int *from_wisdom) {
  if (!is_pow2(length)) return MOFFT_UNSUPPORTED_SIZE;
  int radices[MOFFT_MAX_PLAN_DEPTH];
  size_t depth = 0;
  uint64_t mask = 0;
  *from_wisdom = 0;
  mofft_status st = mofft_internal_model_chain(&mofft_batch_kernel_set, length,
                                               precision_bits, radices, &depth,
                                               &mask);
  if (st != MOFFT_SUCCESS) return st;
  st = mofft_internal_build_schedule(&mofft_batch_kernel_set, length, 1,
                                     precision_bits, radices, depth, mask, 0, s);
  if (st != MOFFT_SUCCESS) return st;
  /* the tuned chain, where the batch kinds reproduce its pass grouping */
  mofft_schedule w, tuned;
  if (mofft_internal_wisdom_lookup(length, precision_bits, radices, &depth,
                                   &mask) > 0 &&
      mofft_internal_build_schedule(&mofft_batch_kernel_set, length, 1,
                                    precision_bits, radices, depth, mask, 0,
                                    &w) == MOFFT_SUCCESS &&
      mofft_internal_build_schedule(&mofft_pass_kernel_set, length, 1,
                                    precision_bits, radices, depth, mask, 0,
                                    &tuned) == MOFFT_SUCCESS &&
      w.pass_count == tuned.pass_count && w.pass_count <= s->pass_count) {
    *s = w;
    *from_wisdom = 1;
  }
  return MOFFT_SUCCESS;
}

static int needs_work(const mofft_schedule *s, int in_place) {
  return s->pass_count >= 2 && (in_place || s->pass_count >= 3);
}

static const mofft_registry_entry *kind_of(const mofft_schedule *s, int pass) {
  return &mofft_batch_kernel_set.entries[s->passes[pass].kernel];
}

static void plan_free(mofft_batch_plan *p) {
  if (!p) return;
  mofft_hip_graph_destroy(p->graph);
  for (size_t i = 0; i < p->bl_count; ++i) mofft_hip_free(p->bl[i].twiddles);
  mofft_hip_free(p->d_segs);
  mofft_hip_free(p->d_jobs);
  mofft_hip_free(p->d_counters);
  mofft_hip_free(p->workspace);
  free(p->directions);
  free(p->in_place);
  free(p->length_index);
  free(p->work_offset);
  free(p->bl);
  free(p->segs);
  free(p->seg_job);
  free(p->seg_pass);
  free(p->shared);
  free(p->jobs);
  free(p->slot_job);
  free(p);
}

/* Arguments of pass `i` of job `t`: the last pass writes the output;
 * in-place transforms go through their workspace slice first; out-of-place
 * ones through the output (the last pass of a Stockham schedule reads and
 * writes the same elements of each column, so it may run in place). */
static void pass_args(const mofft_batch_plan *p, size_t t, const void *in,
                      void *out, int i, mofft_pass_args *a) {
  const batch_length *b = &p->bl[p->length_index[t]];
  const mofft_schedule *s = &b->s;
  const size_t cb = mofft_internal_complex_bytes(p->precision_bits);
  const int P = s->pass_count;
  void *work = p->work_offset[t] == SIZE_MAX
                   ? NULL
                   : (char *)p->workspace + p->work_offset[t];
  const void *src = in;
  void *dst = out;
  for (int k = 0; k <= i; ++k) {
    if (k == P - 1) dst = out;
    else if (p->in_place[t]) dst = (k % 2 == 0) ? work : out;
    else dst = ((P - 2 - k) % 2 == 0) ? out : work;
    if (k < i) src = dst;
  }
  const mofft_pass *ps = &s->passes[i];
  const char *tw = (const char *)b->twiddles;
  memset(a, 0, sizeof(*a));
  a->in = src;
  a->out = dst;
  a->stage_tw = tw ? tw + ps->stage_offset * cb : NULL;
  a->tw_hi = tw ? tw + s->tw_hi_offset * cb : NULL;
  a->tw_lo = tw ? tw + s->tw_lo_offset * cb : NULL;
  a->tw_pass = (tw && ps->pass_table) ? tw + ps->pass_offset * cb : NULL;
  a->n = (int64_t)s->length;
  a->stride = (int64_t)ps->stride;
  a->ns = (int64_t)ps->span;
  a->tw_scale = (int64_t)(s->length / (ps->span * ps->length));
  a->columns = (int64_t)ps->stride;
  a->split_shift = s->split_shift;
  mofft_pass_args_set_divisors(a, (int64_t)ps->length);
  const int backward = p->directions[t] == MOFFT_BACKWARD;
  if (backward && i == 0) a->flags |= MOFFT_PASS_SWAP_IN;
  if (backward && i == P - 1) a->flags |= MOFFT_PASS_SWAP_OUT;
  if (ps->span > 1) a->flags |= MOFFT_PASS_TWIDDLE;
  if (ps->span > 1 && ps->pass_table) a->flags |= MOFFT_PASS_TWIDDLE_TABLE;
}

/* Segment arguments and the gathered job table from the buffers (the
 * queue layout is fixed at plan creation), uploaded to the device. */
static mofft_status bind_buffers(mofft_batch_plan *p, const void *const *inputs,
                                 void *const *outputs) {
  if (!inputs || !outputs) return MOFFT_INVALID_ARGUMENT;
  for (size_t t = 0; t < p->count; ++t) {
    if (!inputs[t] || !outputs[t]) return MOFFT_INVALID_ARGUMENT;
    if ((inputs[t] == outputs[t]) != (p->in_place[t] != 0))
      return MOFFT_INVALID_ARGUMENT;
    if (!mofft_hip_is_device_pointer(inputs[t]) ||
        !mofft_hip_is_device_pointer(outputs[t]))
      return MOFFT_INVALID_ARGUMENT;
  }
  for (size_t k = 0; k < p->job_slots; ++k) {
    const size_t t = p->slot_job[k];
    mofft_batch_job *j = &p->jobs[k];
    j->in = inputs[t];
    j->out = outputs[t];
    j->flags = p->directions[t] == MOFFT_BACKWARD
                   ? (MOFFT_PASS_SWAP_IN | MOFFT_PASS_SWAP_OUT)
                   : 0u;
    j->reserved = 0;
  }
  for (size_t g = 0; g < p->seg_count; ++g) {
    const size_t t = p->seg_job[g];
    if (t == SIZE_MAX) continue;
    pass_args(p, t, inputs[t], outputs[t], p->seg_pass[g], &p->segs[g].args);
  }
  int rc = 0;
  if (p->job_slots)
    rc |= mofft_hip_copy(p->d_jobs, p->jobs, p->job_slots * sizeof(mofft_batch_job));
  rc |= mofft_hip_copy(p->d_segs, p->segs, p->seg_count * sizeof(mofft_batch_segment));
  return rc == 0 ? MOFFT_SUCCESS : MOFFT_INTERNAL_ERROR;
}

static mofft_status add_segment(mofft_batch_plan *p, size_t *cap,
                                const mofft_batch_segment *sg, size_t job,
                                int pass) {
  if (p->seg_count == *cap) {
    const size_t nc = *cap ? 2 * *cap : 64;
    mofft_batch_segment *a = (mofft_batch_segment *)realloc(p->segs, nc * sizeof(*a));
    if (!a) return MOFFT_ALLOCATION_FAILURE;
    p->segs = a;
    size_t *b = (size_t *)realloc(p->seg_job, nc * sizeof(*b));
    if (!b) return MOFFT_ALLOCATION_FAILURE;
    p->seg_job = b;
    int *c = (int *)realloc(p->seg_pass, nc * sizeof(*c));
    if (!c) return MOFFT_ALLOCATION_FAILURE;
    p->seg_pass = c;
    *cap = nc;
  }
  p->segs[p->seg_count] = *sg;
  p->seg_job[p->seg_count] = job;
  p->seg_pass[p->seg_count] = pass;
  p->seg_count++;
  return MOFFT_SUCCESS;
}

typedef struct {
  size_t length;
  size_t job;
} length_key;

/* longest first, then job order */
static int by_length_desc(const void *l, const void *r) {
  const length_key *a = (const length_key *)l, *b = (const length_key *)r;
  if (a->length != b->length) return a->length > b->length ? -1 : 1;
  return a->job < b->job ? -1 : (a->job > b->job);
}

static int64_t pass_tiles(const mofft_schedule *s, int pass) {
  const mofft_registry_entry *e = kind_of(s, pass);
  return ((int64_t)s->passes[pass].stride + e->columns_per_block - 1) /
         e->columns_per_block;
}

static int64_t transform_tiles(const mofft_schedule *s) {
  int64_t n = 0;
  for (int i = 0; i < s->pass_count; ++i) n += pass_tiles(s, i);
  return n;
}

static void pass_segment(const mofft_batch_plan *p, size_t t, int pass,
                         int64_t first_item, mofft_batch_segment *sg) {
  const batch_length *b = &p->bl[p->length_index[t]];
  memset(sg, 0, sizeof(*sg));
  sg->first_item = first_item;
  sg->kind = b->s.passes[pass].kernel;
  sg->dep = -1;
  sg->next = -1;
  sg->tiles = (uint32_t)pass_tiles(&b->s, pass);
  sg->group = sg->tiles;
}

static int64_t seg_items(const mofft_batch_segment *sg) {
  return ((int64_t)sg->tiles + sg->group - 1) / sg->group;
}

/* Queue layout (see the file comment): shared first passes, owned
 * transforms (longest first), gathered row segments, shared later passes
 * level by level; then the continuation passes of owned transforms. */
static mofft_status build_queue(mofft_batch_plan *p, const size_t *order,
                                int64_t workers) {
  size_t cap = 0;
  mofft_status st = MOFFT_SUCCESS;
  size_t *last = (size_t *)malloc(p->count * sizeof(size_t));
  p->shared = (unsigned char *)calloc(p->count, 1);
  if (!last || !p->shared) { free(last); return MOFFT_ALLOCATION_FAILURE; }
  /* fair share of tiles per workgroup of the grid */
  int64_t all = 0;
  for (size_t i = 0; i < p->bl_count; ++i) {
    const batch_length *b = &p->bl[i];
    if (b->s.pass_count == 1) {
      const mofft_registry_entry *e = kind_of(&b->s, 0);
      all += ((int64_t)b->gather_count + e->columns_per_block - 1) / e->columns_per_block;
    }
  }
  for (size_t t = 0; t < p->count; ++t) {
    const mofft_schedule *s = &p->bl[p->length_index[t]].s;
    if (s->pass_count > 1) all += transform_tiles(s);
  }
  const int64_t share = workers > 0 ? all / workers : 0;
  /* Owned up to twice a fair share; shared passes in items of an eighth
   * of a share, so each item's cross-workgroup synchronisation (an L2
   * write-back and invalidate on gfx950) is amortised over several tiles.
   * Measured on 4096 FP32 transforms of N = 2^6..2^18 (MI355X): every
   * transform owned 2.47 ms; transforms above 64 tiles shared in items of
   * 16 / 4 / 1 tiles 2.53 / 3.09 / 5.69 ms. */
  int64_t owned_max = p->options.owned_tiles ? (int64_t)p->options.owned_tiles
                                             : 2 * share;
  if (owned_max < 8 && !p->options.owned_tiles) owned_max = 8;
  int64_t group = p->options.item_tiles ? (int64_t)p->options.item_tiles
                                        : share / 8;
  if (group < 1) group = 1;
  p->owned_max = owned_max;
  p->item_tiles = group;
  int max_passes = 1;
  for (size_t t = 0; t < p->count; ++t) {
    const mofft_schedule *s = &p->bl[p->length_index[t]].s;
    if (s->pass_count < 2) continue;
    if (s->pass_count > max_passes) max_passes = s->pass_count;
    p->shared[t] = (p->options.flags & MOFFT_BATCH_SHARE_ALL) ||
                   transform_tiles(s) > owned_max;
    p->shared_jobs += p->shared[t];
  }
  int64_t items = 0;
  mofft_batch_segment sg;
  /* shared first passes */
  for (size_t k = 0; k < p->count && st == MOFFT_SUCCESS; ++k) {
    const size_t t = order[k];
    if (!p->shared[t]) continue;
    pass_segment(p, t, 0, items, &sg);
    if (sg.group > group) sg.group = (uint32_t)group;
    items += seg_items(&sg);
    st = add_segment(p, &cap, &sg, t, 0);
    last[t] = p->seg_count - 1;
  }
  /* owned transforms: one item each (next is linked below) */
  for (size_t k = 0; k < p->count && st == MOFFT_SUCCESS; ++k) {
    const size_t t = order[k];
    if (p->shared[t] || p->bl[p->length_index[t]].s.pass_count < 2) continue;
    pass_segment(p, t, 0, items, &sg);
    items += 1;
    st = add_segment(p, &cap, &sg, t, 0);
    last[t] = p->seg_count - 1;
  }
  /* gathered row segments, longest first */
  const size_t cb = mofft_internal_complex_bytes(p->precision_bits);
  for (size_t i = 0; i < p->bl_count && st == MOFFT_SUCCESS; ++i) {
    const batch_length *b = &p->bl[i];
    if (b->s.pass_count != 1) continue;
    const mofft_registry_entry *e = kind_of(&b->s, 0);
    const char *tw = (const char *)b->twiddles;
    memset(&sg, 0, sizeof(sg));
    sg.first_item = items;
    sg.kind = b->s.passes[0].kernel;
    sg.dep = -1;
    sg.next = -1;
    sg.args.stage_tw = tw ? tw + b->s.passes[0].stage_offset * cb : NULL;
    sg.args.n = (int64_t)b->length;
    sg.args.stride = 1;
    sg.args.ns = 1;
    sg.args.tw_scale = 1;
    sg.args.columns = (int64_t)b->gather_count;
    mofft_pass_args_set_divisors(&sg.args, (int64_t)b->length);
    sg.jobs = (const mofft_batch_job *)p->d_jobs + b->gather_first;
    sg.tiles = (uint32_t)(((int64_t)b->gather_count + e->columns_per_block - 1) /
                          e->columns_per_block);
    sg.group = 1;
    items += sg.tiles;
    st = add_segment(p, &cap, &sg, SIZE_MAX, 0);
  }
  /* shared later passes, level-major */
  for (int level = 1; level < max_passes && st == MOFFT_SUCCESS; ++level) {
    for (size_t k = 0; k < p->count && st == MOFFT_SUCCESS; ++k) {
      const size_t t = order[k];
      if (!p->shared[t] || level >= p->bl[p->length_index[t]].s.pass_count) continue;
      pass_segment(p, t, level, items, &sg);
      if (sg.group > group) sg.group = (uint32_t)group;
      sg.dep = (int32_t)last[t];
      sg.dep_target = (uint32_t)seg_items(&p->segs[last[t]]);
      p->segs[last[t]].signal = 1;
      items += seg_items(&sg);
      st = add_segment(p, &cap, &sg, t, level);
      last[t] = p->seg_count - 1;
    }
  }
  p->queue_count = p->seg_count;
  /* continuation passes of owned transforms */
  for (size_t k = 0; k < p->count && st == MOFFT_SUCCESS; ++k) {
    const size_t t = order[k];
    const int P = p->bl[p->length_index[t]].s.pass_count;
    if (p->shared[t] || P < 2) continue;
    size_t prev = last[t];
    for (int i = 1; i < P && st == MOFFT_SUCCESS; ++i) {
      pass_segment(p, t, i, -1, &sg);
      st = add_segment(p, &cap, &sg, t, i);
      p->segs[prev].next = (int32_t)(p->seg_count - 1);
      prev = p->seg_count - 1;
    }
  }
  if (st == MOFFT_SUCCESS &&
      (items > (int64_t)INT32_MAX - workers || p->seg_count > (size_t)INT32_MAX))
    st = MOFFT_UNSUPPORTED_SIZE;
  p->total_items = items;
  free(last);
  return st;
}

mofft_status mofft_batch_plan_create(mofft_batch_plan **plan,
                                     int precision_bits, size_t count,
                                     const size_t *lengths,
                                     const int *directions,
                                     const void *const *inputs,
                                     void *const *outputs,
                                     const mofft_batch_options *options) {
  if (!plan) return MOFFT_INVALID_ARGUMENT;
  *plan = NULL;
  if ((precision_bits != 32 && precision_bits != 64) || count == 0 ||
      !lengths || !directions || !inputs || !outputs ||
      count > (size_t)INT32_MAX)
    return MOFFT_INVALID_ARGUMENT;
  for (size_t t = 0; t < count; ++t) {
    if (directions[t] != MOFFT_FORWARD && directions[t] != MOFFT_BACKWARD)
      return MOFFT_INVALID_ARGUMENT;
    if (lengths[t] < 2) return MOFFT_UNSUPPORTED_SIZE;
    if (!is_pow2(lengths[t])) return MOFFT_UNSUPPORTED_SIZE;
  }
  if (!mofft_hip_device_available()) return MOFFT_UNSUPPORTED_HARDWARE;
  mofft_batch_plan *p = (mofft_batch_plan *)calloc(1, sizeof(*p));
  if (!p) return MOFFT_ALLOCATION_FAILURE;
  p->precision_bits = precision_bits;
  p->count = count;
  if (options) p->options = *options;
  else mofft_batch_options_init(&p->options);
  if (p->options.spin_limit == 0)
    p->options.spin_limit = MOFFT_BATCH_DEFAULT_SPIN_LIMIT;
  const size_t cb = mofft_internal_complex_bytes(precision_bits);
  mofft_status st = MOFFT_SUCCESS;
  size_t *order = (size_t *)malloc(count * sizeof(size_t));
  p->directions = (int *)malloc(count * sizeof(int));
  p->in_place = (unsigned char *)malloc(count);
  p->length_index = (int *)malloc(count * sizeof(int));
  p->work_offset = (size_t *)malloc(count * sizeof(size_t));
  p->bl = (batch_length *)calloc(count, sizeof(batch_length));
  if (!order || !p->directions || !p->in_place || !p->length_index ||
      !p->work_offset || !p->bl) {
    st = MOFFT_ALLOCATION_FAILURE;
    goto fail;
  }
  {
    length_key *keys = (length_key *)malloc(count * sizeof(length_key));
    if (!keys) { st = MOFFT_ALLOCATION_FAILURE; goto fail; }
    for (size_t t = 0; t < count; ++t) {
      keys[t].length = lengths[t];
      keys[t].job = t;
      p->directions[t] = directions[t];
      p->in_place[t] = inputs[t] == outputs[t];
    }
    qsort(keys, count, sizeof(length_key), by_length_desc);
    for (size_t t = 0; t < count; ++t) order[t] = keys[t].job;
    free(keys);
  }
  /* distinct lengths (order is longest first) */
  for (size_t k = 0; k < count; ++k) {
    const size_t t = order[k];
    if (p->bl_count == 0 || p->bl[p->bl_count - 1].length != lengths[t]) {
      batch_length *b = &p->bl[p->bl_count++];
      b->length = lengths[t];
      st = batch_schedule(precision_bits, b->length, &b->s, &b->from_wisdom);
      if (st != MOFFT_SUCCESS) goto fail;
      if (b->s.pass_count > MOFFT_MAX_PASSES) { st = MOFFT_UNSUPPORTED_SIZE; goto fail; }
      p->wisdom_lengths += (size_t)b->from_wisdom;
    }
    p->length_index[t] = (int)(p->bl_count - 1);
  }
  /* twiddles, workspace slices, gathered slots */
  for (size_t i = 0; i < p->bl_count; ++i) {
    st = mofft_internal_fill_twiddles(&p->bl[i].s, &p->bl[i].twiddles);
    if (st != MOFFT_SUCCESS) goto fail;
    p->twiddle_bytes += p->bl[i].s.twiddle_count * cb;
  }
  for (size_t k = 0; k < count; ++k) {
    const size_t t = order[k];
    batch_length *b = &p->bl[p->length_index[t]];
    const double n = (double)b->length;
    p->flops += 5.0 * n * log2(n);
    p->bytes += 2.0 * n * (double)cb;
    p->work_offset[t] = SIZE_MAX;
    if (b->s.pass_count == 1) {
      if (b->gather_count == 0) b->gather_first = p->job_slots;
      b->gather_count++;
      p->job_slots++;
    } else {
      p->multi_pass_jobs++;
      if (needs_work(&b->s, p->in_place[t])) {
        p->work_offset[t] = p->workspace_bytes;
        /* 256-byte aligned slices */
        p->workspace_bytes += (b->length * cb + 255) & ~(size_t)255;
      }
    }
  }
  if (p->job_slots) {
    p->jobs = (mofft_batch_job *)calloc(p->job_slots, sizeof(mofft_batch_job));
    p->slot_job = (size_t *)malloc(p->job_slots * sizeof(size_t));
    if (!p->jobs || !p->slot_job) { st = MOFFT_ALLOCATION_FAILURE; goto fail; }
    size_t *fill = (size_t *)calloc(p->bl_count, sizeof(size_t));
    if (!fill) { st = MOFFT_ALLOCATION_FAILURE; goto fail; }
    for (size_t k = 0; k < count; ++k) {
      const size_t t = order[k];
      const int li = p->length_index[t];
      if (p->bl[li].s.pass_count != 1) continue;
      p->slot_job[p->bl[li].gather_first + fill[li]++] = t;
    }
    free(fill);
    if (mofft_hip_malloc(&p->d_jobs, p->job_slots * sizeof(mofft_batch_job)) != 0) {
      st = MOFFT_ALLOCATION_FAILURE;
      goto fail;
    }
  }
  if (p->workspace_bytes &&
      mofft_hip_malloc(&p->workspace, p->workspace_bytes) != 0) {
    st = MOFFT_ALLOCATION_FAILURE;
    goto fail;
  }
  /* launch shape: 256-thread workgroups, at most as many per CU as fit
   * with the largest LDS tile among the kinds this batch uses */
  for (size_t i = 0; i < p->bl_count; ++i)
    for (int k = 0; k < p->bl[i].s.pass_count; ++k) {
      const mofft_registry_entry *e = kind_of(&p->bl[i].s, k);
      if (e->lds_bytes > p->lds_bytes) p->lds_bytes = e->lds_bytes;
    }
  int bpc = 0, cus = 0;
  if (mofft_hip_batch_occupancy(precision_bits, p->lds_bytes, &bpc, &cus) != 0 ||
      bpc <= 0 || cus <= 0) {
    st = MOFFT_INTERNAL_ERROR;
    goto fail;
  }
  p->blocks_per_cu = bpc;
  /* two workgroups per CU unless fewer fit: measured on 4096 FP32
   * transforms of N = 2^6..2^18 (MI355X, 4 resident per CU), grids of
   * 1024 / 768 / 512 / 384 workgroups ran 1.55 / 1.40 / 1.30 / 1.43 ms;
   * FP64 (2 resident per CU) is flat from 512 to 640 */
  int64_t workers = (int64_t)(bpc < 2 ? bpc : 2) * cus;
  if (p->options.grid) workers = (int64_t)p->options.grid;
  st = build_queue(p, order, workers);
  if (st != MOFFT_SUCCESS) goto fail;
  if (mofft_hip_malloc(&p->d_segs, p->seg_count * sizeof(mofft_batch_segment)) != 0 ||
      mofft_hip_malloc(&p->d_counters,
                       (size_t)MOFFT_BATCH_COUNTER_ARRAY(p->queue_count) *
                           sizeof(unsigned int)) != 0) {
    st = MOFFT_ALLOCATION_FAILURE;
    goto fail;
  }
  st = bind_buffers(p, inputs, outputs);
  if (st != MOFFT_SUCCESS) goto fail;
  int64_t grid = workers;
  if (grid > p->total_items) grid = p->total_items;
  if (grid < 1) grid = 1;
  p->grid = (int)grid;
  p->header.segments = (const mofft_batch_segment *)p->d_segs;
  p->header.counters = (unsigned int *)p->d_counters;
  p->header.segment_count = (int64_t)p->queue_count;
  p->header.total_items = p->total_items;
  p->header.counter_count = (int64_t)p->queue_count;
  p->header.spin_limit = p->options.spin_limit;
  if (!(p->options.flags & MOFFT_BATCH_NO_GRAPH) &&
      mofft_hip_batch_graph(precision_bits, &p->header, p->grid, p->lds_bytes,
                            &p->graph) != 0) {
    st = MOFFT_INTERNAL_ERROR;
    goto fail;
  }
  free(order);
  *plan = p;
  return MOFFT_SUCCESS;
fail:
  free(order);
  plan_free(p);
  return st;
}

mofft_status mofft_batch_set_buffers(mofft_batch_plan *plan,
                                     const void *const *inputs,
                                     void *const *outputs) {
  if (!plan) return MOFFT_INVALID_ARGUMENT;
  return bind_buffers(plan, inputs, outputs);
}

mofft_status mofft_batch_execute_async(const mofft_batch_plan *plan,
                                       void *stream) {
  if (!plan) return MOFFT_INVALID_ARGUMENT;
  int rc;
  if (plan->graph) rc = mofft_hip_graph_launch(plan->graph, stream);
  else rc = mofft_hip_batch_launch(plan->precision_bits, &plan->header,
                                   plan->grid, plan->lds_bytes, stream);
  return rc == 0 ? MOFFT_SUCCESS : MOFFT_INTERNAL_ERROR;
}

mofft_status mofft_batch_check(const mofft_batch_plan *plan) {
  if (!plan) return MOFFT_INVALID_ARGUMENT;
  unsigned int flag = 0;
  if (mofft_hip_copy(&flag, (unsigned int *)plan->d_counters + MOFFT_BATCH_ERROR,
                     sizeof(flag)) != 0)
    return MOFFT_INTERNAL_ERROR;
  return flag == 0 ? MOFFT_SUCCESS : MOFFT_INTERNAL_ERROR;
}

mofft_status mofft_batch_execute(const mofft_batch_plan *plan) {
  mofft_status st = mofft_batch_execute_async(plan, NULL);
  if (st != MOFFT_SUCCESS) return st;
  if (mofft_hip_synchronize(NULL) != 0) return MOFFT_INTERNAL_ERROR;
  return mofft_batch_check(plan);
}

mofft_status mofft_batch_describe(const mofft_batch_plan *plan,
                                  mofft_batch_description *d) {
  if (!plan || !d) return MOFFT_INVALID_ARGUMENT;
  memset(d, 0, sizeof(*d));
  d->job_count = plan->count;
  d->distinct_lengths = plan->bl_count;
  d->multi_pass_jobs = plan->multi_pass_jobs;
  d->shared_jobs = plan->shared_jobs;
  d->owned_tiles = (size_t)plan->owned_max;
  d->item_tiles = (size_t)plan->item_tiles;
  d->wisdom_lengths = plan->wisdom_lengths;
  d->segment_count = plan->seg_count;
  d->item_count = (size_t)plan->total_items;
  d->grid = (size_t)plan->grid;
  d->blocks_per_cu = (size_t)plan->blocks_per_cu;
  d->lds_bytes = (size_t)plan->lds_bytes;
  d->workspace_bytes = plan->workspace_bytes;
  d->twiddle_bytes = plan->twiddle_bytes;
  d->flops = plan->flops;
  d->algorithmic_bytes = plan->bytes;
  d->precision_bits = plan->precision_bits;
  d->graph = plan->graph != NULL;
  return MOFFT_SUCCESS;
}

mofft_status mofft_batch_describe_length(int precision_bits, size_t length,
                                         mofft_plan_description *d) {
  if (!d || (precision_bits != 32 && precision_bits != 64))
    return MOFFT_INVALID_ARGUMENT;
  mofft_schedule s;
  int w = 0;
  mofft_status st = batch_schedule(precision_bits, length, &s, &w);
  if (st != MOFFT_SUCCESS) return st;
  mofft_internal_describe(&s, needs_work(&s, 1)
                                  ? length * mofft_internal_complex_bytes(precision_bits)
                                  : 0,
                          d);
  return MOFFT_SUCCESS;
}

void mofft_batch_plan_destroy(mofft_batch_plan *plan) { plan_free(plan); }
