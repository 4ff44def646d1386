/*
 * fused.h - XCD-resident two-pass execution for gfx950 (mofft_abi.h:
 * mofft_fused_args describes the protocol).
 *
 * MI355X has eight XCDs, each with its own 4 MB L2; HBM traffic of a
 * two-pass transform is halved when the array between the passes stays in
 * the L2 of the XCD that wrote it.  One persistent launch runs both passes
 * of every transform of the batch with the column tiles of fft_pass.h; the
 * tile arithmetic is that of the two pass-kernel launches it replaces, so
 * the results are bit-identical to them.
 *
 * Teams.  A workgroup's team is the XCD it runs on (s_getreg XCC_ID), so
 * every member of a team shares one L2.  A team's queue (one counter) is a
 * sequence of rounds; round r holds the `tiles` pass-0 tiles of the team's
 * slot r followed by the pass-1 tiles of slot r - 1.  The first pass-0 tile
 * of a slot takes the slot's transform from the launch-wide counter, after
 * the previous slot has taken its own, and publishes it; so a team's
 * transforms increase with its slots, a transform >= batch marks the slot
 * and every later one void, and a workgroup that draws a pass-1 tile of a
 * void slot leaves.
 *
 * Waits (bounded by spin_limit, then the error flag ends the launch): a
 * pass-1 tile waits for all pass-0 tiles of its slot; a pass-0 tile of slot
 * s waits until pass 1 of slot s - slots has read the workspace it reuses;
 * every tile waits for its slot's transform to be published.  Each wait is
 * for a tile drawn EARLIER from the same queue, by a workgroup that is
 * running and waits only for tiles drawn before its own, so every wait
 * ends.
 *
 * Ordering without cache maintenance: data stores are write-through to the
 * XCD's L2; a producer waits for its stores to complete (vmcnt 0), meets
 * its workgroup at a barrier, and one thread adds to the slot's counter.
 * The consumer polls the counter, meets its workgroup, and loads the slot
 * with L1-missing loads (Ctx::coherent) that find the producer's lines in
 * the shared L2 (or, once evicted, written back to memory).  Counters and
 * flags are only ever accessed by atomics (sync_read).
 * Nothing crosses XCDs inside the launch: slots are team-private and each
 * transform belongs to one team; kernel boundaries order the rest.
 *
 * Reference counterpart: none - the reference streams every stage through
 * memory (src/mofft.c:1932-2160); this is the GPU execution of the same
 * two Stockham passes.
 */
#pragma once

#include "fft_pass.h"

namespace mofft {

/* gfx940-family hardware register XCC_ID (bits 3:0: the XCD). */
__device__ __forceinline__ unsigned xcd_id() {
  return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11));
}

/* Current value of a synchronisation word, read by an atomic.  Device-scope
 * atomics are performed past the L2 and do not refresh a line the L2
 * already holds, so a plain load can keep returning a stale value; every
 * access to these words is therefore an atomic (here a compare-and-swap
 * that never changes the value). */
__device__ __forceinline__ unsigned sync_read(unsigned *word) {
  unsigned v = 0u;
  __hip_atomic_compare_exchange_strong(word, &v, 0u, __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return v;
}

/* Poll `word` until it is at least `target` (or, with target == 0, until it
 * is non-zero); returns the value, or 0 when the error flag is (or gets)
 * raised. */
__device__ __forceinline__ unsigned fused_wait(unsigned *word, unsigned target,
                                               unsigned *error,
                                               uint64_t spin_limit) {
  uint64_t polls = 0;
  for (;;) {
    const unsigned v = sync_read(word);
    if (target ? v >= target : v != 0) return v;
    if (sync_read(error)) return 0;
    if (++polls > spin_limit) {
      __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
      return 0;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

enum : int { kFusedRun = 0, kFusedSkip = 1, kFusedExit = 2 };

template <typename T, int L, int TPT, int C, int... R>
__global__ void __launch_bounds__(TPT * C)
fused_two_pass_kernel(mofft_fused_args fa) {
  extern __shared__ __align__(16) unsigned char mofft_lds[];
  __shared__ int s_state, s_pass, s_tile;
  __shared__ unsigned s_slot;
  __shared__ long long s_t;
  T *lds = reinterpret_cast<T *>(mofft_lds);
  unsigned *const error = fa.counters + 1;
  const unsigned team = xcd_id() % fa.teams;
  unsigned *const head = fa.counters + 2 + team;
  unsigned *const published =
      fa.counters + 2 + fa.teams + (size_t)team * fa.max_slots * 3;
  unsigned *const done = published + fa.max_slots;
  const uint32_t T2 = 2 * fa.tiles;
  const int64_t n = fa.pass[0].n;
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));
  for (;;) {
    if (tid == 0) {
      int state = kFusedRun, pass = 0, tile = 0;
      unsigned slot = 0;
      long long t = 0;
      const unsigned i =
          __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
      const unsigned r = i / T2, h = i - r * T2;
      if (h < fa.tiles) {
        pass = 0;
        slot = r;
        tile = (int)h;
      } else {
        pass = 1;
        tile = (int)(h - fa.tiles);
        if (r == 0) state = kFusedSkip;
        slot = r - 1;
      }
      if (state == kFusedRun && slot >= fa.max_slots) state = kFusedExit;
      if (state == kFusedRun) {
        unsigned pub;
        if (pass == 0 && tile == 0) {
          /* take the next transform only after the previous slot has its
           * own: a team's transforms then increase with its slots, so the
           * first void slot is followed by void slots only */
          pub = slot == 0 ? 1u
                          : fused_wait(published + slot - 1, 0u, error,
                                       fa.spin_limit);
          if (pub != 0) {
            const unsigned g = __hip_atomic_fetch_add(
                fa.counters, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pub = g + 1u;
            __hip_atomic_exchange(published + slot, pub, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
          }
        } else {
          pub = fused_wait(published + slot, 0u, error, fa.spin_limit);
        }
        if (pub == 0) {
          state = kFusedExit;
        } else {
          t = (long long)pub - 1;
          if (t >= fa.batch) state = pass == 1 ? kFusedExit : kFusedSkip;
        }
      }
      if (state == kFusedRun) {
        bool ok = true;
        if (pass == 1)
          ok = fused_wait(done + 2 * slot, fa.tiles, error, fa.spin_limit) != 0;
        else if (slot >= fa.slots)
          ok = fused_wait(done + 2 * (slot - fa.slots) + 1, fa.tiles, error,
                          fa.spin_limit) != 0;
        if (!ok) state = kFusedExit;
      }
      s_state = state;
      s_pass = pass;
      s_tile = tile;
      s_slot = slot;
      s_t = t;
    }
    __syncthreads();
    const int state = __builtin_amdgcn_readfirstlane(s_state);
    if (state == kFusedExit) break;
    if (state == kFusedSkip) continue;
    const int pass = __builtin_amdgcn_readfirstlane(s_pass);
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    const unsigned slot = __builtin_amdgcn_readfirstlane(s_slot);
    const long long t = s_t;
    cx<T> *ws = reinterpret_cast<cx<T> *>(fa.work) +
                ((size_t)team * fa.slots + slot % fa.slots) * (size_t)n;
    mofft_pass_args a = fa.pass[pass];
    if (pass == 0) {
      a.in = reinterpret_cast<const cx<T> *>(fa.in) + t * n;
      a.out = ws;
    } else {
      a.in = ws;
      a.out = reinterpret_cast<cx<T> *>(fa.out) + t * n;
    }
    Ctx<T> x;
    setup_tile<T, L, TPT, C, false>(x, a, (int64_t)tile, lds, tid);
    x.coherent = pass == 1;
    run_tile<T, L, TPT, C, false, R...>(x, a);
    /* this thread's stores have reached the L2 ... */
    __builtin_amdgcn_s_waitcnt(0x0F70); /* vmcnt(0) */
    __syncthreads();
    /* ... and so have the workgroup's: count the tile */
    if (tid == 0)
      __hip_atomic_fetch_add(done + 2 * slot + pass, 1u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int UNUSED>
__global__ void fused_reset_kernel(unsigned int *counters, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    counters[i] = 0u;
}

template <typename T, int L, int TPT, int C, int... R>
struct FusedKernel {
  static constexpr int THREADS = TPT * C;
  static constexpr int LDS = Tile<T, L, TPT, C, false>::BYTES;
  static const void *fn() {
    return reinterpret_cast<const void *>(
        &fused_two_pass_kernel<T, L, TPT, C, R...>);
  }
  static int configure() {
    static bool configured = false;
    if (!configured) {
      if (LDS > 65536) {
        hipError_t e = hipFuncSetAttribute(
            fn(), hipFuncAttributeMaxDynamicSharedMemorySize, LDS);
        if (e != hipSuccess) return (int)e;
      }
      configured = true;
    }
    return 0;
  }
  static int occupancy(int *blocks_per_cu, int *cu_count) {
    int rc = configure();
    if (rc) return rc;
    int nb = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn(),
                                                                THREADS, LDS);
    if (e != hipSuccess) return (int)e;
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    *blocks_per_cu = nb;
    *cu_count = cus;
    return 0;
  }
  static int launch(const mofft_fused_args *a, int grid, hipStream_t s) {
    int rc = configure();
    if (rc || !a) return rc;
    if (grid <= 0 || a->teams == 0 || a->slots < 2 || a->tiles == 0)
      return (int)hipErrorInvalidValue;
    const int64_t words = mofft_fused_counter_words(a->teams, a->max_slots);
    const unsigned rb = (unsigned)((words + 255) / 256);
    hipLaunchKernelGGL(fused_reset_kernel<0>, dim3(rb), dim3(256), 0, s,
                       a->counters, words);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL((fused_two_pass_kernel<T, L, TPT, C, R...>),
                       dim3((unsigned)grid), dim3(THREADS), LDS, s, *a);
    return (int)hipGetLastError();
  }
};

}  // namespace mofft
