/*
 * batch.h - persistent uneven-batch execution for gfx950.
 *
 * Many transforms of different lengths, directions and placements run in
 * ONE launch.  The host lays the work out as a queue of tiles grouped into
 * segments (mofft_abi.h: mofft_batch_segment); a fixed grid of resident
 * 256-thread workgroups pulls tile indices from a device atomic (the queue
 * head) until the queue is empty.  Every tile kind is a pass-kernel tile of
 * fft_pass.h re-instantiated with 256 threads, so one kernel can run any of
 * them; which one is a switch over the segment's kind (the generated
 * dispatcher of mofft_batch_f32.hip / mofft_batch_f64.hip).
 *
 * Dependencies: pass p+1 of a multi-pass transform needs every tile of pass
 * p.  A transform no larger than a fair share of the launch is OWNED: one
 * queue item, all of its passes run by the workgroup that takes it, and a
 * workgroup barrier (workgroup-scope release/acquire) orders the passes -
 * no cross-workgroup synchronisation at all.  Only a transform larger than
 * that is SHARED: its pass tiles are separate items, and a tile of pass p+1
 * waits for a counter of pass p.  Shared passes are queued level-major (all
 * shared first passes before everything else, later passes after it), so a
 * tile only ever waits for tiles dequeued before it, by workgroups that are
 * running or done: the wait always ends.  It is bounded anyway: after
 * spin_limit polls the launch raises the error flag and every workgroup
 * leaves the loop.
 *
 * Memory ordering of shared passes: a producer tile ends with an
 * agent-scope release fence in every thread, a barrier, and one release
 * increment of its segment's counter; a consumer polls with agent-scope
 * acquire loads and then fences (acquire) in every thread.  On gfx950 those
 * fences write back and invalidate the L2 of the XCD, which is why owned
 * execution is preferred (measured: 3.5x faster on a 4096-transform batch).
 *
 * Reference counterpart: none - the reference executes one transform per
 * call (src/mofft.c:2169-2201); this is the batched/serving execution layer
 * built on the same pass kernels.
 */
#pragma once

#include "fft_pass.h"

/* Build-time option: a minimum of waves per SIMD for the batch kernels
 * (caps their registers; default: the compiler's choice). */
#ifdef MOFFT_BATCH_MIN_WAVES_F32
#define MOFFT_BATCH_WAVES_F32 \
  __attribute__((amdgpu_waves_per_eu(MOFFT_BATCH_MIN_WAVES_F32)))
#else
#define MOFFT_BATCH_WAVES_F32
#endif
#ifdef MOFFT_BATCH_MIN_WAVES_F64
#define MOFFT_BATCH_WAVES_F64 \
  __attribute__((amdgpu_waves_per_eu(MOFFT_BATCH_MIN_WAVES_F64)))
#else
#define MOFFT_BATCH_WAVES_F64
#endif

namespace mofft {

constexpr int kBatchThreads = 256;

/* Gathered row tile: column c of tile `block` is the whole transform
 * jobs[block * C + c], with its own buffers and direction. */
template <typename T, int L, int TPT, int C>
__device__ __forceinline__ void setup_gather(Ctx<T> &x,
                                             const mofft_pass_args &p,
                                             const mofft_batch_job *jobs,
                                             int64_t block, T *lds, int tid) {
  x.lds = lds;
  x.tid = tid;
  x.c = tid / TPT;
  x.t = tid % TPT;
  x.col0 = (uint32_t)block * C;
  const uint32_t col = x.col0 + x.c;
  x.live = col < (uint32_t)p.columns;
  x.contiguous = false; /* each column its own job's buffer */
  const void *in = nullptr;
  void *out = nullptr;
  uint32_t flags = 0;
  if (x.live) {
    in = jobs[col].in;
    out = jobs[col].out;
    flags = jobs[col].flags;
  }
  x.stride = 1;
  x.ns = 1;
  x.src = (const cx<T> *)in;
  x.dst = (cx<T> *)out;
  x.stw = (const cx<T> *)p.stage_tw;
  x.thi = nullptr;
  x.tlo = nullptr;
  x.tpass = nullptr;
  x.obase = 0;
  x.jm_scale = 0;
  x.split_shift = 0;
  x.split_mask = 0;
  x.swap_in = (flags & MOFFT_PASS_SWAP_IN) != 0;
  x.swap_out = (flags & MOFFT_PASS_SWAP_OUT) != 0;
  x.pass_tw = false;
  x.pass_table = false;
}

/* One queue tile of kind <T, L, TPT, C, ROWS, R...>.  Row kinds are always
 * gathered segments; strided kinds are passes of one transform. */
template <typename T, int L, int TPT, int C, bool ROWS, int... R>
__device__ __forceinline__ void batch_tile(const mofft_batch_segment &sg,
                                           int64_t block, T *lds) {
  static_assert(TPT * C == kBatchThreads, "batch tiles are 256 threads");
  /* The thread index is loop-invariant in the persistent loop, and so is
   * every LDS and twiddle address computed from it; hidden behind an empty
   * asm the compiler recomputes them per tile instead of hoisting dozens of
   * them out of the loop into registers held across every tile. */
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));
  Ctx<T> x;
  if constexpr (ROWS)
    setup_gather<T, L, TPT, C>(x, sg.args, sg.jobs, block, lds, tid);
  else
    setup_tile<T, L, TPT, C, ROWS>(x, sg.args, block, lds, tid);
  run_tile<T, L, TPT, C, ROWS, R...>(x, sg.args);
}

/* Poll the finished-tile counter of segment `dep` until it reaches
 * `target`.  False when the error flag is (or gets) raised.  The polls are
 * relaxed atomic loads (no cache maintenance per poll); the caller issues
 * one agent-scope acquire fence after a successful wait, which pairs with
 * the producers' release and orders every later load of their data. */
__device__ __forceinline__ bool batch_wait(const mofft_batch_header &h,
                                           int dep, unsigned target) {
  unsigned int *done = h.counters + MOFFT_BATCH_DONE(dep);
  unsigned int *error = h.counters + MOFFT_BATCH_ERROR;
  uint64_t polls = 0;
  while (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
         target) {
    if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return false;
    if (++polls > h.spin_limit) {
      __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(8);
  }
  return true;
}

/* A segment descriptor in scalar registers.  The kernel stores to global
 * memory, so the compiler cannot prove the descriptor array unchanged and
 * would keep every field (and all address arithmetic derived from it) in
 * vector registers; reading it word by word through readfirstlane states
 * that the values are uniform. */
__device__ __forceinline__ mofft_batch_segment
load_segment(const mofft_batch_segment *p) {
  static_assert(sizeof(mofft_batch_segment) % 4 == 0, "word copy");
  mofft_batch_segment s;
  const int *src = reinterpret_cast<const int *>(p);
  int *dst = reinterpret_cast<int *>(&s);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(s) / 4); ++i)
    dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return s;
}

/* The persistent loop.  D::run(kind, segment, block, lds) dispatches a
 * tile to its kind; it has ONE call site (the loop is a small state
 * machine), so the kernel's register budget is that of its largest kind. */
template <typename T, typename D>
__device__ __forceinline__ void batch_loop(const mofft_batch_header &h) {
  extern __shared__ __align__(16) unsigned char mofft_lds[];
  __shared__ int s_item;
  __shared__ int s_seg;
  T *lds = reinterpret_cast<T *>(mofft_lds);
  mofft_batch_segment c;
  int seg = -1;
  int64_t block = 0, end = 0;
  bool active = false;  /* inside a run of tiles [block, end) of segment c */
  bool owned = false;   /* ... that continues with the transform's next pass */
  for (;;) {
    if (!active) {
      if (threadIdx.x == 0) {
        int sseg = -1;
        const int64_t t = (int64_t)__hip_atomic_fetch_add(
            h.counters + MOFFT_BATCH_HEAD, 1u, __ATOMIC_RELAXED,
            __HIP_MEMORY_SCOPE_AGENT);
        if (t < h.total_items) {
          int lo = 0, hi = (int)h.segment_count - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (h.segments[mid].first_item <= t) lo = mid;
            else hi = mid - 1;
          }
          sseg = lo;
          const int dep = h.segments[sseg].dep;
          if (dep >= 0 && !batch_wait(h, dep, h.segments[sseg].dep_target))
            sseg = -1;
        }
        s_item = (int)t;
        s_seg = sseg;
      }
      __syncthreads();
      /* uniform across the workgroup: keep them (and everything derived
       * from the segment) in scalar registers */
      seg = __builtin_amdgcn_readfirstlane(s_seg);
      const int64_t item = __builtin_amdgcn_readfirstlane(s_item);
      if (seg < 0) break;
      c = load_segment(h.segments + seg);
      owned = c.next >= 0;
      if (owned) {
        block = 0;
        end = c.tiles;
      } else {
        block = (item - c.first_item) * (int64_t)c.group;
        end = block + c.group < (int64_t)c.tiles ? block + c.group : c.tiles;
        if (c.dep >= 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      active = true;
    }
    D::run(c.kind, c, block, lds);
    const bool done = ++block == end;
    const bool signal = done && !owned && c.signal;
    if (signal) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    /* LDS reuse by the next tile, and (owned) visibility of this pass's
     * stores to the whole workgroup before the next pass loads them */
    __syncthreads();
    if (done) {
      if (owned && c.next >= 0) {
        c = load_segment(h.segments + c.next);
        block = 0;
        end = c.tiles;
      } else {
        if (signal && threadIdx.x == 0)
          __hip_atomic_fetch_add(h.counters + MOFFT_BATCH_DONE(seg), 1u,
                                 __ATOMIC_RELEASE,
                                 __HIP_MEMORY_SCOPE_AGENT);
        active = false;
      }
    }
  }
}

/* Clears the queue head, the error flag and the segment counters (a
 * kernel rather than a memset so that captured graphs replay it safely). */
template <int UNUSED>
__global__ void batch_reset_kernel(unsigned int *counters, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    counters[i] = 0u;
}

/* Host side of one precision's batch kernel (instantiated by the generated
 * file that defines K). */
template <void (*K)(mofft_batch_header)>
struct BatchLauncher {
  /* The LDS opt-in covers the largest tile any batch may use, so a later
   * batch with a larger one needs no second call on that device. */
  static int allow(int lds_bytes) {
    static std::atomic<unsigned long long> configured{0};
    if (lds_bytes > kLdsCap) return (int)hipErrorInvalidValue;
    if (lds_bytes <= 65536) return 0;
    return allow_lds(configured, reinterpret_cast<const void *>(K), kLdsCap);
  }
  static int occupancy(int lds_bytes, int *blocks_per_cu, int *cu_count) {
    const int rc = allow(lds_bytes);
    if (rc != 0) return rc;
    int n = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, reinterpret_cast<const void *>(K), kBatchThreads, lds_bytes);
    if (e != hipSuccess) return (int)e;
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    *blocks_per_cu = n;
    *cu_count = cus;
    return 0;
  }
  static int launch(const mofft_batch_header *h, int grid, int lds_bytes,
                    hipStream_t s) {
    const int rc = allow(lds_bytes);
    if (rc != 0) return rc;
    const int64_t words = MOFFT_BATCH_COUNTER_ARRAY(h->counter_count);
    const unsigned rb = (unsigned)((words + 255) / 256);
    hipLaunchKernelGGL(batch_reset_kernel<0>, dim3(rb), dim3(256), 0, s,
                       h->counters, words);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    if (h->total_items == 0) return 0;
    hipLaunchKernelGGL(K, dim3((unsigned)grid), dim3(kBatchThreads),
                       lds_bytes, s, *h);
    return (int)hipGetLastError();
  }
};

}  // namespace mofft
