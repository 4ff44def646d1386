/*
 * mofft_server.cpp - the serving loop of mofft_hip.h (host C++).
 *
 * Continuous batching of 1-D C2C requests on one GPU:
 *
 *   submit   -> pending FIFO (never blocks)
 *   admit    -> while the head request fits (an in-flight slot is free and
 *               the arena has room for its input and output), carve its
 *               buffers from the arena, pick the least-loaded stream, enqueue
 *               [generate input] -> plan execution -> event
 *   retire   -> events that completed become completions; buffers return
 *               to the arena on release, which lets the next head in
 *
 * Requests therefore join while others are running and overlap with them
 * on the device (several streams), and admission is bounded by the budget
 * - the arena is the only device memory request data ever uses, so the
 * serving path calls no device allocator and never synchronises the device.
 * Plans are resident per (length, precision, direction), built for
 * `max_batch` transforms and executed for each request's batch.
 *
 * Reference counterpart: none (the reference has no serving layer).
 */
#include <sched.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

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
  size_t in_off = 0, out_off = 0, work_off = 0;
  size_t in_bytes = 0, out_bytes = 0, work_bytes = 0;
  int stream = -1;
  void *event = nullptr;
  void *plan = nullptr;
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
  std::deque<Job> inflight;
  std::map<uint64_t, Job> done;     /* completed, buffers held until release */
  std::deque<mofft_completion> ready;
  uint64_t next_ticket = 1;
  size_t submitted = 0, completed = 0, peak_bytes = 0, peak_inflight = 0;
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

/* Arena bytes of a request: generated input, output, and the schedule's
 * workspace (three or more passes), each 256-byte aligned. */
size_t request_bytes(const mofft_request &r, const void *plan, size_t *in_bytes,
                     size_t *out_bytes, size_t *work_bytes) {
  const size_t b = r.length * r.batch * complex_bytes(r.precision_bits);
  *out_bytes = b;
  *in_bytes = r.input ? 0 : b;
  *work_bytes = mofft_internal_workspace_count(plan, r.batch);
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

void release_job(mofft_server *s, const Job &j) {
  if (j.in_bytes) s->arena.release(j.in_off, j.in_bytes);
  if (j.out_bytes) s->arena.release(j.out_off, j.out_bytes);
  if (j.work_bytes) s->arena.release(j.work_off, j.work_bytes);
}

/* Move completed in-flight requests to the completion queue. */
void retire(mofft_server *s) {
  const double t = now_seconds();
  for (auto it = s->inflight.begin(); it != s->inflight.end();) {
    const int q = mofft_hip_event_query(it->event);
    if (q == 0) { ++it; continue; }
    const int status = q > 0 ? MOFFT_SUCCESS : MOFFT_INTERNAL_ERROR;
    Job j = *it;
    it = s->inflight.erase(it);
    s->free_events.push_back(j.event);
    j.event = nullptr;
    s->stream_load[j.stream]--;
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
}

/* Admit pending requests in FIFO order while slots and budget allow. */
void admit(mofft_server *s) {
  while (!s->pending.empty() && s->inflight.size() < s->opt.max_inflight) {
    Job &j = s->pending.front();
    request_bytes(j.req, j.plan, &j.in_bytes, &j.out_bytes, &j.work_bytes);
    size_t in_off = 0, out_off = 0, work_off = 0;
    if (j.in_bytes && !s->arena.alloc(j.in_bytes, &in_off)) return;
    if (!s->arena.alloc(j.out_bytes, &out_off)) {
      if (j.in_bytes) s->arena.release(in_off, j.in_bytes);
      return;
    }
    if (j.work_bytes && !s->arena.alloc(j.work_bytes, &work_off)) {
      if (j.in_bytes) s->arena.release(in_off, j.in_bytes);
      s->arena.release(out_off, j.out_bytes);
      return;
    }
    j.in_off = in_off;
    j.out_off = out_off;
    j.work_off = work_off;
    if (s->arena.used() > s->peak_bytes) s->peak_bytes = s->arena.used();
    int best = 0;
    for (size_t k = 1; k < s->streams.size(); ++k)
      if (s->stream_load[k] < s->stream_load[best]) best = (int)k;
    j.stream = best;
    void *stream = s->streams[best];
    const void *in = j.req.input ? j.req.input : s->arena_base + j.in_off;
    void *out = s->arena_base + j.out_off;
    const size_t count = j.req.length * j.req.batch;
    int rc = 0;
    if (!j.req.input)
      rc = mofft_hip_fill_uniform((void *)in, j.req.precision_bits,
                                  (int64_t)count, j.req.seed, stream);
    void *work = j.work_bytes ? s->arena_base + j.work_off : nullptr;
    mofft_status st = rc == 0 ? mofft_internal_execute_count(
                                    j.plan, in, out, work, j.req.batch, stream)
                              : MOFFT_INTERNAL_ERROR;
    if (s->free_events.empty()) {
      void *ev = nullptr;
      if (mofft_hip_event_create(&ev) != 0) st = MOFFT_INTERNAL_ERROR;
      else s->free_events.push_back(ev);
    }
    j.start = now_seconds();
    if (st == MOFFT_SUCCESS) {
      j.event = s->free_events.back();
      s->free_events.pop_back();
      if (mofft_hip_event_record(j.event, stream) != 0) st = MOFFT_INTERNAL_ERROR;
    }
    if (st != MOFFT_SUCCESS) {
      if (j.event) s->free_events.push_back(j.event);
      j.event = nullptr;
      s->completed++;
      s->ready.push_back(completion_of(s, j, now_seconds(), st));
      s->done[j.ticket] = j;
      s->pending.pop_front();
      continue;
    }
    s->stream_load[best]++;
    s->inflight.push_back(j);
    if (s->inflight.size() > s->peak_inflight) s->peak_inflight = s->inflight.size();
    s->pending.pop_front();
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
  if (request_bytes(*r, j.plan, &in_b, &out_b, &work_b) > s->arena.size())
    return MOFFT_ALLOCATION_FAILURE;
  j.ticket = s->next_ticket++;
  s->pending.push_back(j);
  s->submitted++;
  if (ticket) *ticket = j.ticket;
  return MOFFT_SUCCESS;
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
  if (o.budget_bytes < Arena::kAlign || o.flags != 0) return MOFFT_INVALID_ARGUMENT;
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
  st->inflight = s->inflight.size();
  st->resident_plans = s->plans.size();
  st->bytes_in_use = s->arena.used();
  st->peak_bytes_in_use = s->peak_bytes;
  st->budget_bytes = s->arena.size();
  st->peak_inflight = s->peak_inflight;
  st->flops_completed = s->flops;
  st->bytes_completed = s->bytes;
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
  for (auto &j : s->inflight) mofft_hip_event_destroy(j.event);
  for (void *ev : s->free_events) mofft_hip_event_destroy(ev);
  for (void *st : s->streams) mofft_hip_stream_destroy(st);
  mofft_hip_free(s->arena_base);
  delete s;
}

}  // extern "C"
