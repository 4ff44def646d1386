/*
 * mofft_server.cpp - the serving loop of mofft_hip.h (host C++).
 *
 * Continuous batching of 1-D C2C requests on one GPU:
 *
 *   submit   -> pending FIFO (never blocks)
 *   admit    -> while the head request fits (in-flight slots free and the
 *               arena has room for its input, output and workspace), carve
 *               one block from the arena for it and for the later pending
 *               requests of the same plan (length, precision, direction)
 *               with generated input that fit in `max_batch` transforms,
 *               pick the least-loaded stream and enqueue [generate each
 *               input] -> one batched plan execution -> event
 *   retire   -> an execution whose event completed yields one completion
 *               per request; a block returns to the arena when the last of
 *               its requests is released, which lets the next head in
 *
 * Requests therefore join while others are running and overlap with them
 * on the device (several streams), small requests of one plan share
 * launches, and admission is bounded by the budget - the arena is the only
 * device memory request data ever uses, so the serving path calls no
 * device allocator and never synchronises the device.  Plans are resident
 * per (length, precision, direction), built for `max_batch` transforms and
 * executed for each group's total batch; a transform's result does not
 * depend on the transforms executed beside it.
 *
 * Reference counterpart: none (the reference has no serving layer).
 */
#include <sched.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <cmath>
#include <deque>
#include <map>
#include <new>
#include <vector>

extern "C" {
#include "mofft_internal.h"
}

namespace {

double now_seconds() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* First-fit allocator over one device arena, 256-byte granules; free
 * blocks coalesce with their neighbours. */
class Arena {
 public:
  static constexpr size_t kAlign = 256;
  void reset(size_t bytes) {
    free_.clear();
    size_ = bytes & ~(kAlign - 1);
    if (size_) free_[0] = size_;
    used_ = 0;
  }
  bool alloc(size_t bytes, size_t *offset) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *offset = it->first;
        const size_t rest = it->second - bytes;
        const size_t at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t offset, size_t bytes) {
    bytes = round(bytes);
    used_ -= bytes;
    auto next = free_.lower_bound(offset);
    if (next != free_.end() && offset + bytes == next->first) {
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == offset) {
        prev->second += bytes;
        return;
      }
    }
    free_[offset] = bytes;
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }
  static size_t round(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0;
  size_t used_ = 0;
};

struct Job {
  uint64_t ticket;
  mofft_request req;
  double submit;
  double start = 0;
  size_t out_off = 0;     /* this request's output inside its block */
  size_t out_bytes = 0;
  uint64_t block = 0;     /* arena block shared with its execution group */
  void *plan = nullptr;
};

/* Arena ranges of one execution: input, output and workspace of all its
 * requests; freed when the last of them is released. */
struct Block {
  size_t in_off = 0, out_off = 0, work_off = 0;
  size_t in_bytes = 0, out_bytes = 0, work_bytes = 0;
  size_t refs = 0;
};

/* Requests executed together: one stream, one event. */
struct Group {
  std::vector<Job> jobs;
  int stream = -1;
  void *event = nullptr;
};

struct PlanKey {
  int bits;
  int direction;
  size_t length;
  bool operator<(const PlanKey &o) const {
    if (bits != o.bits) return bits < o.bits;
    if (direction != o.direction) return direction < o.direction;
    return length < o.length;
  }
};

size_t complex_bytes(int bits) { return bits == 32 ? 8 : 16; }

}  // namespace

struct mofft_server {
  mofft_server_options opt;
  char *arena_base = nullptr;
  Arena arena;
  std::vector<void *> streams;
  std::vector<int> stream_load;
  std::vector<void *> free_events;
  std::map<PlanKey, void *> plans;
  std::deque<Job> pending;
  std::deque<Group> inflight;
  size_t inflight_jobs = 0;
  std::map<uint64_t, Job> done;     /* completed, buffers held until release */
  std::map<uint64_t, Block> blocks;
  std::deque<mofft_completion> ready;
  uint64_t next_ticket = 1, next_block = 1;
  size_t submitted = 0, completed = 0, peak_bytes = 0, peak_inflight = 0;
  size_t executions = 0;
  double flops = 0, bytes = 0;
};

namespace {

void destroy_plan(int bits, void *p) {
  if (bits == 32) mofft_plan_destroy_f32((mofft_plan_f32 *)p);
  else mofft_plan_destroy_f64((mofft_plan_f64 *)p);
}

mofft_status get_plan(mofft_server *s, const mofft_request &r, void **plan) {
  const PlanKey key{r.precision_bits, r.direction, r.length};
  auto it = s->plans.find(key);
  if (it != s->plans.end()) {
    *plan = it->second;
    return MOFFT_SUCCESS;
  }
  void *p = nullptr;
  mofft_status st;
  const mofft_direction dir = (mofft_direction)r.direction;
  if (r.precision_bits == 32)
    st = mofft_plan_create_batched_f32((mofft_plan_f32 **)&p, r.length,
                                       s->opt.max_batch, dir, MOFFT_OUT_OF_PLACE,
                                       nullptr, nullptr);
  else
    st = mofft_plan_create_batched_f64((mofft_plan_f64 **)&p, r.length,
                                       s->opt.max_batch, dir, MOFFT_OUT_OF_PLACE,
                                       nullptr, nullptr);
  if (st != MOFFT_SUCCESS) return st;
  s->plans[key] = p;
  *plan = p;
  return MOFFT_SUCCESS;
}

/* Arena bytes of `batch` transforms of request r's plan: generated input,
 * output, and the schedule's workspace (three or more passes), each
 * 256-byte aligned. */
size_t group_bytes(const mofft_request &r, size_t batch, const void *plan,
                   size_t *in_bytes, size_t *out_bytes, size_t *work_bytes) {
  const size_t b = r.length * batch * complex_bytes(r.precision_bits);
  *out_bytes = b;
  *in_bytes = r.input ? 0 : b;
  *work_bytes = mofft_internal_workspace_count(plan, batch);
  return Arena::round(*in_bytes) + Arena::round(*out_bytes) +
         Arena::round(*work_bytes);
}

mofft_completion completion_of(const mofft_server *s, const Job &j, double t,
                               int status) {
  mofft_completion c;
  memset(&c, 0, sizeof(c));
  c.ticket = j.ticket;
  c.tag = j.req.tag;
  c.status = status;
  c.submit_seconds = j.submit;
  c.start_seconds = j.start;
  c.done_seconds = t;
  c.output = status == MOFFT_SUCCESS ? s->arena_base + j.out_off : nullptr;
  c.bytes = j.out_bytes;
  return c;
}

/* Drop one request's reference to its block; the last one frees it. */
void release_job(mofft_server *s, const Job &j) {
  auto it = s->blocks.find(j.block);
  if (it == s->blocks.end()) return;
  Block &b = it->second;
  if (--b.refs > 0) return;
  if (b.in_bytes) s->arena.release(b.in_off, b.in_bytes);
  if (b.out_bytes) s->arena.release(b.out_off, b.out_bytes);
  if (b.work_bytes) s->arena.release(b.work_off, b.work_bytes);
  s->blocks.erase(it);
}

void finish_job(mofft_server *s, const Job &j, double t, int status) {
  s->completed++;
  if (status == MOFFT_SUCCESS) {
    const double n = (double)j.req.length;
    s->flops += 5.0 * n * std::log2(n) * (double)j.req.batch;
    s->bytes += 2.0 * n * (double)complex_bytes(j.req.precision_bits) *
                (double)j.req.batch;
  }
  s->ready.push_back(completion_of(s, j, t, status));
  s->done[j.ticket] = j;
}

/* Move completed executions' requests to the completion queue. */
void retire(mofft_server *s) {
  const double t = now_seconds();
  for (auto it = s->inflight.begin(); it != s->inflight.end();) {
    const int q = mofft_hip_event_query(it->event);
    if (q == 0) { ++it; continue; }
    const int status = q > 0 ? MOFFT_SUCCESS : MOFFT_INTERNAL_ERROR;
    Group g = std::move(*it);
    it = s->inflight.erase(it);
    s->free_events.push_back(g.event);
    s->stream_load[g.stream]--;
    s->inflight_jobs -= g.jobs.size();
    for (const Job &j : g.jobs) finish_job(s, j, t, status);
  }
}

/* Carve the block of a group of `batch` transforms; false if it does not
 * fit the arena now. */
bool alloc_block(mofft_server *s, const mofft_request &r, size_t batch,
                 const void *plan, Block *b) {
  group_bytes(r, batch, plan, &b->in_bytes, &b->out_bytes, &b->work_bytes);
  if (b->in_bytes && !s->arena.alloc(b->in_bytes, &b->in_off)) return false;
  if (!s->arena.alloc(b->out_bytes, &b->out_off)) {
    if (b->in_bytes) s->arena.release(b->in_off, b->in_bytes);
    return false;
  }
  if (b->work_bytes && !s->arena.alloc(b->work_bytes, &b->work_off)) {
    if (b->in_bytes) s->arena.release(b->in_off, b->in_bytes);
    s->arena.release(b->out_off, b->out_bytes);
    return false;
  }
  return true;
}

/* How far past the head admission looks for requests to join it. */
constexpr size_t kCoalesceLookahead = 256;

/* Admit pending requests in FIFO order while slots and budget allow; each
 * admitted head takes same-plan pending requests with generated input
 * along into its execution. */
void admit(mofft_server *s) {
  const bool coalesce = (s->opt.flags & MOFFT_SERVER_NO_COALESCE) == 0;
  while (!s->pending.empty() && s->inflight_jobs < s->opt.max_inflight) {
    const Job &h = s->pending.front();
    std::vector<size_t> members{0};
    size_t total = h.req.batch;
    if (coalesce && !h.req.input) {
      const size_t end = std::min(s->pending.size(), kCoalesceLookahead);
      for (size_t k = 1; k < end; ++k) {
        if (s->inflight_jobs + members.size() >= s->opt.max_inflight) break;
        const Job &c = s->pending[k];
        if (c.plan != h.plan || c.req.input ||
            total + c.req.batch > s->opt.max_batch)
          continue;
        members.push_back(k);
        total += c.req.batch;
      }
    }
    Block b;
    if (!alloc_block(s, h.req, total, h.plan, &b)) {
      if (members.size() == 1) return;  /* the head waits for memory */
      members.resize(1);
      total = h.req.batch;
      b = Block();
      if (!alloc_block(s, h.req, total, h.plan, &b)) return;
    }
    b.refs = members.size();
    const uint64_t block_id = s->next_block++;
    s->blocks[block_id] = b;
    if (s->arena.used() > s->peak_bytes) s->peak_bytes = s->arena.used();
    int best = 0;
    for (size_t k = 1; k < s->streams.size(); ++k)
      if (s->stream_load[k] < s->stream_load[best]) best = (int)k;
    void *stream = s->streams[best];
    const size_t cb = complex_bytes(h.req.precision_bits);
    const size_t n = h.req.length;
    Group g;
    g.stream = best;
    int rc = 0;
    size_t offset = 0;  /* transforms before this member */
    const double start = now_seconds();
    for (size_t m : members) {
      Job j = s->pending[m];
      j.block = block_id;
      j.start = start;
      j.out_off = b.out_off + offset * n * cb;
      j.out_bytes = j.req.batch * n * cb;
      if (!j.req.input && rc == 0)
        rc = mofft_hip_fill_uniform(s->arena_base + b.in_off + offset * n * cb,
                                    j.req.precision_bits,
                                    (int64_t)(j.req.batch * n), j.req.seed,
                                    stream);
      offset += j.req.batch;
      g.jobs.push_back(j);
    }
    /* (h refers into `pending`: not used past this point) */
    for (size_t k = members.size(); k-- > 0;)
      s->pending.erase(s->pending.begin() + (ptrdiff_t)members[k]);
    const Job &head = g.jobs[0];
    const void *in = head.req.input ? head.req.input : s->arena_base + b.in_off;
    void *out = s->arena_base + b.out_off;
    void *work = b.work_bytes ? s->arena_base + b.work_off : nullptr;
    mofft_status st = rc == 0 ? mofft_internal_execute_count(
                                    head.plan, in, out, work, total, stream)
                              : MOFFT_INTERNAL_ERROR;
    if (st == MOFFT_SUCCESS && s->free_events.empty()) {
      void *ev = nullptr;
      if (mofft_hip_event_create(&ev) != 0) st = MOFFT_INTERNAL_ERROR;
      else s->free_events.push_back(ev);
    }
    if (st == MOFFT_SUCCESS) {
      g.event = s->free_events.back();
      s->free_events.pop_back();
      if (mofft_hip_event_record(g.event, stream) != 0) {
        s->free_events.push_back(g.event);
        st = MOFFT_INTERNAL_ERROR;
      }
    }
    s->executions++;
    if (st != MOFFT_SUCCESS) {
      const double t = now_seconds();
      for (const Job &j : g.jobs) finish_job(s, j, t, st);
      continue;
    }
    s->stream_load[best]++;
    s->inflight_jobs += g.jobs.size();
    s->inflight.push_back(std::move(g));
    if (s->inflight_jobs > s->peak_inflight) s->peak_inflight = s->inflight_jobs;
  }
}

mofft_status do_submit(mofft_server *s, const mofft_request *r, double t,
                       uint64_t *ticket) {
  if (!s || !r) return MOFFT_INVALID_ARGUMENT;
  if ((r->precision_bits != 32 && r->precision_bits != 64) || r->batch == 0 ||
      r->batch > s->opt.max_batch ||
      (r->direction != MOFFT_FORWARD && r->direction != MOFFT_BACKWARD))
    return MOFFT_INVALID_ARGUMENT;
  if (r->length < 2) return MOFFT_UNSUPPORTED_SIZE;
  if (r->input && !mofft_hip_is_device_pointer(r->input))
    return MOFFT_INVALID_ARGUMENT;
  Job j;
  j.req = *r;
  j.submit = t;
  mofft_status st = get_plan(s, *r, &j.plan);
  if (st != MOFFT_SUCCESS) return st;
  size_t in_b = 0, out_b = 0, work_b = 0;
  if (group_bytes(*r, r->batch, j.plan, &in_b, &out_b, &work_b) > s->arena.size())
    return MOFFT_ALLOCATION_FAILURE;
  j.ticket = s->next_ticket++;
  s->pending.push_back(j);
  s->submitted++;
  if (ticket) *ticket = j.ticket;
  return MOFFT_SUCCESS;
}

/* True when pending requests exist, nothing is in flight and admission
 * (just run) admitted none of them: the head waits for arena space that no
 * running execution will return. */
bool admission_stalled(const mofft_server *s) {
  return !s->pending.empty() && s->inflight.empty() && s->ready.empty();
}

/* Fail the pending head with `status` (its completion goes to `ready`). */
void fail_head(mofft_server *s, int status) {
  Job j = s->pending.front();
  s->pending.pop_front();
  j.block = 0;
  j.start = now_seconds();
  finish_job(s, j, j.start, status);
}

}  // namespace

extern "C" {

void mofft_server_options_init(mofft_server_options *o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->budget_bytes = (size_t)1 << 30;
}

mofft_status mofft_server_create(mofft_server **server,
                                 const mofft_server_options *options) {
  if (!server) return MOFFT_INVALID_ARGUMENT;
  *server = nullptr;
  mofft_server_options o;
  if (options) o = *options;
  else mofft_server_options_init(&o);
  if (o.budget_bytes < Arena::kAlign || (o.flags & ~MOFFT_SERVER_NO_COALESCE) != 0)
    return MOFFT_INVALID_ARGUMENT;
  if (o.streams == 0) o.streams = 4;
  if (o.max_inflight == 0) o.max_inflight = 4 * o.streams;
  if (o.max_batch == 0) o.max_batch = 64;
  if (o.streams > 64) return MOFFT_INVALID_ARGUMENT;
  if (!mofft_hip_device_available()) return MOFFT_UNSUPPORTED_HARDWARE;
  mofft_server *s = new (std::nothrow) mofft_server;
  if (!s) return MOFFT_ALLOCATION_FAILURE;
  s->opt = o;
  void *base = nullptr;
  if (mofft_hip_malloc(&base, o.budget_bytes) != 0) {
    delete s;
    return MOFFT_ALLOCATION_FAILURE;
  }
  s->arena_base = (char *)base;
  s->arena.reset(o.budget_bytes);
  for (uint32_t k = 0; k < o.streams; ++k) {
    void *st = nullptr;
    if (mofft_hip_stream_create(&st) != 0) {
      mofft_server_destroy(s);
      return MOFFT_INTERNAL_ERROR;
    }
    s->streams.push_back(st);
    s->stream_load.push_back(0);
  }
  *server = s;
  return MOFFT_SUCCESS;
}

mofft_status mofft_server_submit(mofft_server *s, const mofft_request *r,
                                 uint64_t *ticket) {
  mofft_status st = do_submit(s, r, now_seconds(), ticket);
  if (st == MOFFT_SUCCESS) admit(s);
  return st;
}

mofft_status mofft_server_poll(mofft_server *s, mofft_completion *out,
                               size_t capacity, size_t *count, int wait) {
  if (!s || (capacity && !out)) return MOFFT_INVALID_ARGUMENT;
  size_t n = 0;
  for (;;) {
    retire(s);
    admit(s);
    while (n < capacity && !s->ready.empty()) {
      out[n++] = s->ready.front();
      s->ready.pop_front();
    }
    if (n > 0 || !wait || (s->pending.empty() && s->inflight.empty())) break;
    /* Nothing in flight and the head still not admitted: only the arena
     * blocks it, and only a release of held completions can free arena
     * space, so waiting would never end. */
    if (admission_stalled(s)) {
      if (count) *count = 0;
      return MOFFT_ALLOCATION_FAILURE;
    }
    sched_yield();
  }
  if (count) *count = n;
  return MOFFT_SUCCESS;
}

mofft_status mofft_server_release(mofft_server *s, uint64_t ticket) {
  if (!s) return MOFFT_INVALID_ARGUMENT;
  auto it = s->done.find(ticket);
  if (it == s->done.end()) return MOFFT_INVALID_ARGUMENT;
  release_job(s, it->second);
  s->done.erase(it);
  admit(s);
  return MOFFT_SUCCESS;
}

mofft_status mofft_server_get_stats(const mofft_server *s,
                                    mofft_server_stats *st) {
  if (!s || !st) return MOFFT_INVALID_ARGUMENT;
  memset(st, 0, sizeof(*st));
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->pending = s->pending.size();
  st->inflight = s->inflight_jobs;
  st->resident_plans = s->plans.size();
  st->bytes_in_use = s->arena.used();
  st->peak_bytes_in_use = s->peak_bytes;
  st->budget_bytes = s->arena.size();
  st->peak_inflight = s->peak_inflight;
  st->flops_completed = s->flops;
  st->bytes_completed = s->bytes;
  st->executions = s->executions;
  return MOFFT_SUCCESS;
}

mofft_status mofft_server_run_trace(mofft_server *s,
                                    const mofft_request *requests,
                                    const double *arrival, size_t count,
                                    mofft_completion *completions) {
  if (!s || (count && (!requests || !arrival || !completions)))
    return MOFFT_INVALID_ARGUMENT;
  std::map<uint64_t, size_t> index;   /* ticket -> request */
  std::vector<mofft_completion> batch(256);
  const double t0 = now_seconds();
  size_t next = 0, finished = 0;
  while (finished < count) {
    const double t = now_seconds() - t0;
    bool progress = false;
    while (next < count && arrival[next] <= t) {
      uint64_t ticket = 0;
      const mofft_status st = do_submit(s, &requests[next], t0 + arrival[next], &ticket);
      if (st == MOFFT_SUCCESS) {
        index[ticket] = next;
      } else {
        mofft_completion &c = completions[next];
        memset(&c, 0, sizeof(c));
        c.tag = requests[next].tag;
        c.status = st;
        c.submit_seconds = c.done_seconds = t0 + arrival[next];
        finished++;
      }
      next++;
      progress = true;
    }
    retire(s);
    admit(s);
    while (!s->ready.empty()) {
      const mofft_completion c = s->ready.front();
      s->ready.pop_front();
      auto it = index.find(c.ticket);
      if (it != index.end()) {
        completions[it->second] = c;
        completions[it->second].output = nullptr;
        index.erase(it);
      }
      mofft_server_release(s, c.ticket);
      finished++;
      progress = true;
    }
    /* A head that cannot be admitted with nothing in flight waits for
     * memory held by completions the caller took before this call; none of
     * the trace's own requests can free it, so it fails instead of
     * stalling the replay. */
    if (!progress && admission_stalled(s)) {
      fail_head(s, MOFFT_ALLOCATION_FAILURE);
      progress = true;
    }
    if (!progress) sched_yield();
  }
  return MOFFT_SUCCESS;
}

mofft_status mofft_server_generate(void *out, int precision_bits, size_t count,
                                   uint64_t seed, void *stream) {
  if (!out || (precision_bits != 32 && precision_bits != 64))
    return MOFFT_INVALID_ARGUMENT;
  if (!mofft_hip_device_available()) return MOFFT_UNSUPPORTED_HARDWARE;
  return mofft_hip_fill_uniform(out, precision_bits, (int64_t)count, seed,
                                stream) == 0
             ? MOFFT_SUCCESS
             : MOFFT_INTERNAL_ERROR;
}

void mofft_server_destroy(mofft_server *s) {
  if (!s) return;
  for (void *st : s->streams) mofft_hip_synchronize(st);
  for (auto &kv : s->plans) destroy_plan(kv.first.bits, kv.second);
  for (auto &g : s->inflight) mofft_hip_event_destroy(g.event);
  for (void *ev : s->free_events) mofft_hip_event_destroy(ev);
  for (void *st : s->streams) mofft_hip_stream_destroy(st);
  mofft_hip_free(s->arena_base);
  delete s;
}

}  // extern "C"
