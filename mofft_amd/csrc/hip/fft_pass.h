/*
 * fft_pass.h - the LDS-resident pass kernel for gfx950.
 *
 * One launch performs one radix-L step of an n-point Stockham FFT over a
 * batch (the arguments are described in mofft_abi.h).  Inside the launch the
 * L-point DFT of every column is itself a chain of radix-r Stockham
 * sub-stages (the template parameter pack R, executed left to right).  Each
 * sub-stage is a batch of r x r DFT-matrix contractions evaluated by a
 * generated straight-line codelet (mofft_codelets.h) on data held in VGPRs;
 * sub-stages exchange data through LDS.
 *
 * Work mapping
 *   - A workgroup owns C columns and TPT threads per column.
 *   - ROWS: each column is a whole contiguous transform (stride == 1);
 *     threads of a column are adjacent lanes, so loads and stores of the
 *     first and last sub-stage are contiguous runs of the transform.
 *   - COLS: columns are adjacent elements of a strided transform; the column
 *     index is the fastest-varying lane index, so every global access is a
 *     run of C consecutive complex values (128 B for C = 16 FP32 or 8 FP64).
 *     When the pass output interleaves fewer than C columns (ns < C), the
 *     results go through LDS and leave as one contiguous block.
 *
 * LDS holds one real array: an exchange moves the real parts, then the
 * imaginary parts, through the same buffer.  That halves LDS per thread and
 * doubles the workgroups a CU can keep resident (LDS, not VGPRs, bounds
 * occupancy for these kernels); the price is two extra barriers per
 * exchange.
 *
 * Reference counterpart: the per-stage kernels `mofft_kernel_r*_first/other`
 * (compiler/mofft_compiler/emitter.py:283-466) and the executor's transposes
 * (src/mofft.c:1987-2044); here several stages fuse into one HBM pass.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "../mofft_abi.h"
#include "mofft_codelets.h"

/* Build-time kernel options (defaults are the measured best on MI355X):
 *   MOFFT_NT_LOAD / MOFFT_NT_STORE  non-temporal hints on the streaming HBM
 *                                   input / output accesses (off: they cost
 *                                   up to 4x bandwidth on gfx950)
 *   MOFFT_SPLIT_EXCHANGE            exchange real and imaginary parts through
 *                                   one LDS array in turn (half the LDS)
 *   MOFFT_VEC_SMALL                 16-byte accesses when one thread owns a
 *                                   whole FP32 transform */
#ifndef MOFFT_NT_LOAD
#define MOFFT_NT_LOAD 0
#endif
#ifndef MOFFT_NT_STORE
#define MOFFT_NT_STORE 0
#endif
#ifndef MOFFT_SPLIT_EXCHANGE
#define MOFFT_SPLIT_EXCHANGE 1
#endif
#ifndef MOFFT_VEC_SMALL
#define MOFFT_VEC_SMALL 1
#endif

namespace mofft {

template <typename T> struct cx_of;
template <> struct cx_of<float> { using type = float2; };
template <> struct cx_of<double> { using type = double2; };
template <typename T> using cx = typename cx_of<T>::type;

/* LDS tile geometry (one array of T).  The LDS serves 128 bytes per cycle
 * (32 banks of 4 bytes; the gfx950 bank-conflict counters match this
 * model): G = 32 FP32 or 16 FP64 lanes of one access.
 * ROWS tiles pad one element every G: a run of G consecutive elements then
 * meets G different banks, and so do the stride-R writes of a first
 * sub-stage (R a power of two up to 16).  (One element every 16 for FP32
 * made even contiguous runs span 34 words, every access 2-way conflicted:
 * FP32 8192 / 16384 single pass 7-8 % slower.  A per-exchange XOR swizzle
 * also clears the middle sub-stages' writes, but its addresses are not
 * affine in the element index, so the compiler holds one per element: up
 * to 18 more VGPRs per thread, no faster.)  COLS tiles (interleaved
 * columns) pad one word every 32. */
template <typename T, int L, int TPT, int C, bool ROWS>
struct Tile {
  static constexpr int NT = TPT * C;
  static constexpr int GSH = sizeof(T) == 4 ? 5 : 4;  /* log2 G */
  static constexpr int LP = L + (L >> GSH) + 1;
  static constexpr int ROW_WORDS = C * LP;
  static constexpr int COL_WORDS = L * C + ((L * C) >> 5) + 1;
  static constexpr int WORDS = ROWS ? ROW_WORDS
                                    : (ROW_WORDS > COL_WORDS ? ROW_WORDS
                                                             : COL_WORDS);
  static constexpr int ARRAYS = MOFFT_SPLIT_EXCHANGE ? 1 : 2;
  static constexpr int BYTES = ARRAYS * WORDS * (int)sizeof(T);

  static __device__ __forceinline__ int row_addr(int c, int e) {
    return c * LP + e + (e >> GSH);
  }
  static __device__ __forceinline__ int stage_addr(int c, int e) {
    if constexpr (ROWS) {
      return row_addr(c, e);
    } else {
      const int q = e * C + c;
      return q + (q >> 5);
    }
  }

  /* stage_addr(c, e0 + K) == stage_addr(c, e0) + step(K) whenever
   * (e0 mod 2^SH) + (K mod 2^SH) < 2^SH.  A row tile's pad term is
   * e >> GSH; a column tile with C a power of two has q >> 5 =
   * e >> (5 - log2 C) below 32 columns (c < C) and e * C / 32 + (c >> 5)
   * from 32 columns on (SH = 0: linear in e).  The exchanges add the
   * step as a compile-time offset instead of forming every element's
   * padded address. */
  static constexpr bool AFFINE = ROWS || (C & (C - 1)) == 0;
  static constexpr int SH = ROWS ? GSH
                                 : (C >= 32 ? 0
                                            : 5 - (C >= 16 ? 4 : C >= 8 ? 3
                                                   : C >= 4 ? 2 : C >= 2 ? 1 : 0));
  static constexpr int step(int k) {
    return ROWS ? k + (k >> GSH)
                : (C >= 32 ? k * C + ((k * C) >> 5) : k * C + (k >> SH));
  }
  /* e0 = jj (any), K = k * NB */
  static constexpr bool read_affine(int nb) {
    return AFFINE && (SH == 0 || nb % (1 << SH) == 0);
  }
  /* e0 = (jj / ns) * ns * r + jj % ns, K = k * ns (k < r): exact when ns
   * is a multiple of 2^SH, or a power of two dividing 2^SH with ns * r a
   * multiple of 2^SH (then e0 mod 2^SH = jj % ns and K mod 2^SH <=
   * 2^SH - ns) */
  static constexpr bool write_affine(int ns, int r) {
    return AFFINE &&
           (SH == 0 || ns % (1 << SH) == 0 ||
            ((ns & (ns - 1)) == 0 && (1 << SH) % ns == 0 &&
             (ns * r) % (1 << SH) == 0));
  }
  /* row_addr(c, e0 + K) == row_addr(c, e0) + row_step(K) for K a
   * multiple of 2^GSH */
  static constexpr int row_step(int k) { return k + (k >> GSH); }
};

template <typename T>
struct Ctx {
  T *lds;
  const cx<T> *src;   /* element e of this thread's column: src[e * stride] */
  cx<T> *dst;         /* output base of this thread's transform */
  const cx<T> *stw;
  const cx<T> *thi;
  const cx<T> *tlo;
  const cx<T> *tpass; /* &tw_pass[j % ns] when the full table is used */
  /* Element offsets are 32-bit: a launch touches fewer than
   * MOFFT_PASS_MAX_ELEMENTS elements (checked on the host). */
  uint32_t stride;
  uint32_t ns;
  uint32_t obase;     /* (j / ns) * ns * L + j % ns */
  uint32_t col0;      /* first column of this tile */
  uint32_t tvalid;    /* contiguous row tile: elements it covers */
  const cx<T> *tsrc;  /* ... its first input element */
  cx<T> *tdst;        /* ... its first output element */
  uint32_t jm_scale;  /* (j % ns) * tw_scale */
  uint32_t split_mask;
  uint32_t split_shift;
  int tid;            /* thread index in the workgroup */
  int c;
  int t;
  bool live;
  bool contiguous;    /* row tile over consecutive transforms of one buffer */
  bool swap_in;
  bool swap_out;
  bool pass_tw;
  bool pass_table;
};

/* x / d for a launch-invariant divisor d given by its multiplier and
 * shift (mofft_div_magic in mofft_abi.h): a multiply-high, two adds and two
 * shifts instead of a full integer division per lane. */
__device__ __forceinline__ uint32_t div_by(uint32_t x, uint32_t magic,
                                           uint32_t shift) {
  if (shift == 0) return x; /* d == 1 */
  const uint32_t t = __umulhi(x, magic);
  return (t + ((x - t) >> 1)) >> (shift - 1);
}

/* Native clang vectors for the non-temporal load/store builtins. */
typedef float nf2 __attribute__((ext_vector_type(2)));
typedef float nf4 __attribute__((ext_vector_type(4)));
typedef double nd2 __attribute__((ext_vector_type(2)));
template <typename T> struct nvec2;
template <> struct nvec2<float> { using type = nf2; };
template <> struct nvec2<double> { using type = nd2; };

template <typename T>
__device__ __forceinline__ cx<T> ld_stream(const cx<T> *p) {
  using V = typename nvec2<T>::type;
#if MOFFT_NT_LOAD
  const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
#else
  const V v = *reinterpret_cast<const V *>(p);
#endif
  cx<T> r;
  r.x = v.x;
  r.y = v.y;
  return r;
}

template <typename T>
__device__ __forceinline__ void st_stream(cx<T> *p, T a, T b) {
  using V = typename nvec2<T>::type;
  V v;
  v.x = a;
  v.y = b;
#if MOFFT_NT_STORE
  __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
#else
  *reinterpret_cast<V *>(p) = v;
#endif
}

/* (a.x + i a.y)(b.x + i b.y) with the fused multiply-adds written out, so
 * every kernel rounds a complex product the same way whatever the compiler
 * would contract in its context (a batch tile and a single-plan tile of the
 * same chain stay bit-identical). */
__device__ __forceinline__ float fused(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fused(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

template <typename T>
__device__ __forceinline__ cx<T> cprod(const cx<T> a, const cx<T> b) {
  cx<T> r;
  r.x = fused(a.x, b.x, -(a.y * b.y));
  r.y = fused(a.x, b.y, a.y * b.x);
  return r;
}

template <typename T>
__device__ __forceinline__ void cmul(T &re, T &im, const cx<T> w) {
  cx<T> a;
  a.x = re;
  a.y = im;
  const cx<T> r = cprod<T>(a, w);
  re = r.x;
  im = r.y;
}

template <typename T>
__device__ __forceinline__ void store_global(cx<T> *p, T a, T b, bool swap) {
  st_stream<T>(p, swap ? b : a, swap ? a : b);
}

/* Cooperative store of the whole tile, used when the pass output of
 * adjacent columns is not contiguous (ns < C, C a multiple of ns: the tile
 * is C / ns whole groups, each one contiguous block).  Each thread owns the tile
 * elements q = tid + i*NT; real and imaginary parts pass through LDS in
 * turn (row layout). */
template <typename T, int L, int TPT, int C, bool ROWS, int IT, int R>
__device__ __forceinline__ void store_tile(const Ctx<T> &x,
                                           const mofft_pass_args &p,
                                           T (&vr)[IT][R], T (&vi)[IT][R]) {
  using TL = Tile<T, L, TPT, C, ROWS>;
  constexpr int NB = L / R;
  constexpr bool FULL = IT * TPT == NB;
  constexpr int PER = (C * L) / TL::NT;
  static_assert((C * L) % TL::NT == 0, "every thread stores PER elements");
  const int tid = x.tid;
  const uint32_t ns = (uint32_t)p.ns;
  const uint32_t span = ns * (uint32_t)L;
  int ad[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const uint32_t q = (uint32_t)(tid + i * TL::NT);
    uint32_t cc, o;
    if (ns == 1) {
      cc = q / L;
      o = q - cc * L;
    } else {
      const uint32_t g = div_by(q, p.span_magic, p.span_shift);
      const uint32_t rem = q - g * span;
      o = div_by(rem, p.ns_magic, p.ns_shift);
      cc = g * ns + (rem - o * ns);
    }
    ad[i] = TL::row_addr((int)cc, (int)o);
  }
  T outr[PER], outi[PER];
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    if (half) __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int jj = x.t + it * TPT;
      if (FULL || jj < NB) {
        if constexpr (NB % (1 << TL::GSH) == 0) {
          T *a = x.lds + TL::row_addr(x.c, jj);
#pragma unroll
          for (int k = 0; k < R; ++k)
            a[TL::row_step(k * NB)] = half ? vi[it][k] : vr[it][k];
        } else {
#pragma unroll
          for (int k = 0; k < R; ++k)
            x.lds[TL::row_addr(x.c, jj + k * NB)] = half ? vi[it][k] : vr[it][k];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (half) outi[i] = x.lds[ad[i]];
      else outr[i] = x.lds[ad[i]];
    }
  }
  cx<T> *out = (cx<T> *)p.out;
  const uint32_t col0 = x.col0;
  const uint32_t stride = (uint32_t)p.stride;
  const uint32_t n = (uint32_t)p.n;
  const uint32_t b0 = div_by(col0, p.stride_magic, p.stride_shift);
  const uint32_t j0 = col0 - b0 * stride;
  const uint32_t jq0 = div_by(j0, p.ns_magic, p.ns_shift);
  const uint32_t cq = div_by((uint32_t)C, p.ns_magic, p.ns_shift);
  const bool swap = (p.flags & MOFFT_PASS_SWAP_OUT) != 0;
  const bool aligned = (j0 + C <= stride) && (cq * ns == (uint32_t)C) &&
                       (jq0 * ns == j0) && (col0 + C <= (uint32_t)p.columns);
  if (aligned) {
    cx<T> *base = out + (b0 * n + jq0 * span);
#pragma unroll
    for (int i = 0; i < PER; ++i)
      store_global<T>(base + tid + i * TL::NT, outr[i], outi[i], swap);
  } else {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const uint32_t q = (uint32_t)(tid + i * TL::NT);
      uint32_t cc, o;
      if (ns == 1) {
        cc = q / L;
        o = q - cc * L;
      } else {
        const uint32_t g = div_by(q, p.span_magic, p.span_shift);
        const uint32_t rem = q - g * span;
        o = div_by(rem, p.ns_magic, p.ns_shift);
        cc = g * ns + (rem - o * ns);
      }
      const uint32_t col = col0 + cc;
      if (col >= (uint32_t)p.columns) continue;
      const uint32_t b = div_by(col, p.stride_magic, p.stride_shift);
      const uint32_t j = col - b * stride;
      const uint32_t jq = div_by(j, p.ns_magic, p.ns_shift);
      const uint32_t pos = b * n + jq * span + (j - jq * ns) + o * ns;
      store_global<T>(out + pos, outr[i], outi[i], swap);
    }
  }
}

/* The cooperative tile store enumerates whole output groups of ns adjacent
 * columns (one contiguous run of ns*L elements each): it needs ns < C and C
 * a multiple of ns, so that a tile is whole groups.  Otherwise (ns = 12
 * under 256 columns, say) each thread stores its own column, runs of ns
 * elements across adjacent lanes. */
template <int C>
__device__ __forceinline__ bool tile_store(const mofft_pass_args &p) {
  const uint32_t ns = (uint32_t)p.ns;
  return ns < (uint32_t)C &&
         div_by((uint32_t)C, p.ns_magic, p.ns_shift) * ns == (uint32_t)C;
}

/* Radix values that mark a radix-16 / radix-32 sub-stage evaluated on the
 * matrix cores (MStage below) instead of by a VALU codelet.  In a kernel's
 * chain the matrix-core sub-stages come first (the compiler's registry only
 * builds such chains), so one scale per column covers all of them. */
constexpr int kMfma16 = -16;
constexpr int kMfma32 = -32;
constexpr int radix_value(int r) { return r < 0 ? -r : r; }

/* Sub-stage twiddle table of a kernel: entries OFF + (jj % NS) * (R - 1) + k
 * for every sub-stage with NS > 1, in execution order.  When it is small
 * next to the tile (at most a quarter of its LDS, and within the LDS cap)
 * each tile copies it into LDS behind the tile area, loading it together
 * with its data, so the sub-stages read their twiddles at LDS latency.
 * Otherwise the sub-stages read it from global memory (L2), each issuing
 * the next sub-stage's twiddle loads before the exchange that precedes it.
 * The registry (mofft_compiler/registry.py, Entry.lds_bytes) sizes LDS by
 * the same rule. */
constexpr int kLdsCap = 150 * 1024;

/* entries of the sub-stages R... when the first of them has stride ns */
template <int NS0, int... R>
constexpr int stage_tw_entries() {
  constexpr int rs[] = {radix_value(R)...};
  int ns = NS0, count = 0;
  for (int i = 0; i < (int)sizeof...(R); ++i) {
    if (ns > 1) count += ns * (rs[i] - 1);
    ns *= rs[i];
  }
  return count;
}

template <typename T, int L, int TPT, int C, bool ROWS>
constexpr bool stage_tw_in_lds(int count) {
  const int tile = Tile<T, L, TPT, C, ROWS>::BYTES;
  const int table = count * (int)sizeof(cx<T>);
  return count > 0 && 4 * table <= tile && (tile + 15) / 16 * 16 + table <= kLdsCap;
}

template <typename T, int L, int TPT, int C, bool ROWS, int NS, int OFF,
          int R, int... REST>
struct Stage;
template <typename T, int L, int TPT, int C, bool ROWS, int NS, int OFF,
          int RM, int... REST>
struct MStage;

template <typename T, int L, int TPT, int C, bool ROWS, int NS, int OFF,
          int R, int... REST>
struct StageSel {
  using type = Stage<T, L, TPT, C, ROWS, NS, OFF, R, REST...>;
};
template <typename T, int L, int TPT, int C, bool ROWS, int NS, int OFF,
          int... REST>
struct StageSel<T, L, TPT, C, ROWS, NS, OFF, kMfma16, REST...> {
  using type = MStage<T, L, TPT, C, ROWS, NS, OFF, 16, REST...>;
};
template <typename T, int L, int TPT, int C, bool ROWS, int NS, int OFF,
          int... REST>
struct StageSel<T, L, TPT, C, ROWS, NS, OFF, kMfma32, REST...> {
  using type = MStage<T, L, TPT, C, ROWS, NS, OFF, 32, REST...>;
};

template <typename T, int L, int TPT, int C, bool ROWS, int NS, int OFF,
          int R, int... REST>
struct Stage {
  using TL = Tile<T, L, TPT, C, ROWS>;
  static constexpr int RAD = R;
  static constexpr int NB = L / R;
  static constexpr int IT = (NB + TPT - 1) / TPT;
  static constexpr bool FULL = IT * TPT == NB;
  static constexpr bool LAST = sizeof...(REST) == 0;
  /* the kernel's sub-stage table is in LDS (run_tile copies it there) */
  static constexpr bool TW_IN_LDS = stage_tw_in_lds<T, L, TPT, C, ROWS>(
      OFF + stage_tw_entries<NS, R, REST...>());
  static constexpr bool PREFETCH_TW = true;
  /* FP32 row tiles of 16384 points and more keep their sub-stage table
   * (about L entries) in global memory.  Their radix-16 / radix-8
   * sub-stages load the twiddles of k = 1, 2, 3, 4, 8, 12 only and form
   * each other one as the product w^(a j) w^(b j), k = a + b, a < 4, b a
   * multiple of 4, of two correctly rounded table values (one rounding
   * more).  The smaller set also fits the prefetch into the 64 VGPRs that
   * two 1024-thread workgroups per CU allow: FP32 16384 single pass 55.3
   * -> 49.9 us per 2^24 points. */
  static constexpr bool SPARSE_TW =
      sizeof(T) == 4 && ROWS && L >= 16384 && (R == 8 || R == 16);
  static constexpr bool tw_loaded(int k) {
    return !SPARSE_TW || k <= 4 || k % 4 == 0;
  }

  /* this sub-stage's twiddles w_{NS R}^(j k), k = 1..R-1, of every
   * butterfly of the thread (j = jj % NS) */
  static __device__ __forceinline__ void load_tw(const Ctx<T> &x,
                                                 cx<T> (&w)[IT][R - 1]) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int jj = x.t + it * TPT;
#pragma unroll
      for (int k = 0; k < R - 1; ++k) {
        w[it][k].x = T(1);
        w[it][k].y = T(0);
      }
      if (FULL || jj < NB) {
        const cx<T> *src = x.stw + OFF + (jj % NS) * (R - 1);
#pragma unroll
        for (int k = 0; k < R - 1; ++k)
          if (tw_loaded(k + 1)) w[it][k] = src[k];
      }
    }
  }

  /* first sub-stage: element e = jj + k*NB straight from HBM.  Every
   * load (data, then pass twiddles) is issued before any result is used:
   * a use between two loads makes the compiler wait for the first before
   * issuing the second, one memory round trip per element. */
  static __device__ __forceinline__ void load_raw(const Ctx<T> &x,
                                                  cx<T> (&raw)[IT][R]) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int jj = x.t + it * TPT;
      const bool ok = (FULL || jj < NB) && x.live;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        cx<T> v;
        v.x = T(0);
        v.y = T(0);
        if (ok) {
          const cx<T> *q = x.src + (uint32_t)(jj + k * NB) * x.stride;
          v = ld_stream<T>(q);
        }
        raw[it][k] = v;
      }
    }
  }

  /* One thread per whole transform (TPT == 1, R == L): a direct load has
   * lanes a transform apart, 64 cache lines per load instruction.  A
   * contiguous tile is loaded coalesced instead (element g of the tile by
   * thread g % C) and handed to its owners through LDS, real parts then
   * imaginary parts. */
  static constexpr bool STAGED = ROWS && TPT == 1 && IT == 1;

  static __device__ __forceinline__ void load_staged(const Ctx<T> &x,
                                                     T (&vr)[IT][R],
                                                     T (&vi)[IT][R]) {
    constexpr int NT = C;
    cx<T> v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const uint32_t g = (uint32_t)(x.tid + i * NT);
      v[i].x = T(0);
      v[i].y = T(0);
      if (g < x.tvalid) v[i] = ld_stream<T>(x.tsrc + g);
    }
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      if (half) __syncthreads();
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const int g = x.tid + i * NT;
        const int c = g / L;
        const bool re = (half == 0) != x.swap_in;
        x.lds[TL::row_addr(c, g - c * L)] = re ? v[i].x : v[i].y;
      }
      __syncthreads();
#pragma unroll
      for (int e = 0; e < R; ++e) {
        const T w = x.lds[TL::row_addr(x.tid, e)];
        if (half) vi[0][e] = w;
        else vr[0][e] = w;
      }
    }
    __syncthreads();
  }

  static __device__ __forceinline__ void store_staged(const Ctx<T> &x,
                                                      T (&vr)[IT][R],
                                                      T (&vi)[IT][R]) {
    constexpr int NT = C;
    T outr[R], outi[R];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      __syncthreads();
#pragma unroll
      for (int o = 0; o < R; ++o)
        x.lds[TL::row_addr(x.tid, o)] = half ? vi[0][o] : vr[0][o];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const int g = x.tid + i * NT;
        const int c = g / L;
        const T w = x.lds[TL::row_addr(c, g - c * L)];
        if (half) outi[i] = w;
        else outr[i] = w;
      }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const uint32_t g = (uint32_t)(x.tid + i * NT);
      if (g < x.tvalid) store_global<T>(x.tdst + g, outr[i], outi[i], x.swap_out);
    }
  }

  static __device__ __forceinline__ void load_first(const Ctx<T> &x,
                                                    T (&vr)[IT][R],
                                                    T (&vi)[IT][R]) {
    if constexpr (STAGED) {
      if (x.contiguous) {
        load_staged(x, vr, vi);
        return;
      }
    }
    if constexpr (MOFFT_VEC_SMALL && ROWS && TPT == 1 && IT == 1 &&
                  (R % 2 == 0) && sizeof(T) == 4) {
      /* a whole small transform per thread: 16-byte loads */
      const nf4 *s = reinterpret_cast<const nf4 *>(x.src);
      nf4 v[R / 2];
#pragma unroll
      for (int k = 0; k < R / 2; ++k) {
        v[k] = nf4{0.f, 0.f, 0.f, 0.f};
#if MOFFT_NT_LOAD
        if (x.live) v[k] = __builtin_nontemporal_load(s + k);
#else
        if (x.live) v[k] = s[k];
#endif
      }
#pragma unroll
      for (int k = 0; k < R; k += 2) {
        const nf4 w = v[k / 2];
        vr[0][k] = x.swap_in ? w.y : w.x;
        vi[0][k] = x.swap_in ? w.x : w.y;
        vr[0][k + 1] = x.swap_in ? w.w : w.z;
        vi[0][k + 1] = x.swap_in ? w.z : w.w;
      }
    } else {
      cx<T> raw[IT][R];
      load_raw(x, raw);
      /* w_N^(e * (j % ns) * tw_scale) per element (1 where unused) */
      cx<T> w[IT][R];
#pragma unroll
      for (int it = 0; it < IT; ++it)
#pragma unroll
        for (int k = 0; k < R; ++k) {
          w[it][k].x = T(1);
          w[it][k].y = T(0);
        }
      /* whole-transform (row) tiles are single-pass plans: no pass
       * twiddles, and no registers held for them */
      if (!ROWS && x.pass_tw) {
        if (x.pass_table) {
#pragma unroll
          for (int it = 0; it < IT; ++it) {
            const int jj = x.t + it * TPT;
            if (FULL || jj < NB) {
#pragma unroll
              for (int k = 0; k < R; ++k)
                w[it][k] = x.tpass[(uint32_t)(jj + k * NB) * x.ns];
            }
          }
        } else {
          cx<T> h[IT][R], l[IT][R];
#pragma unroll
          for (int it = 0; it < IT; ++it) {
            const int jj = x.t + it * TPT;
#pragma unroll
            for (int k = 0; k < R; ++k) {
              h[it][k] = w[it][k];
              l[it][k] = w[it][k];
              if (FULL || jj < NB) {
                const uint32_t m = (uint32_t)(jj + k * NB) * x.jm_scale;
                h[it][k] = x.thi[m >> x.split_shift];
                l[it][k] = x.tlo[m & x.split_mask];
              }
            }
          }
#pragma unroll
          for (int it = 0; it < IT; ++it)
#pragma unroll
            for (int k = 0; k < R; ++k) {
              const cx<T> a = h[it][k], b = l[it][k];
              w[it][k] = cprod<T>(a, b);
            }
        }
      }
#pragma unroll
      for (int it = 0; it < IT; ++it)
#pragma unroll
        for (int k = 0; k < R; ++k) {
          vr[it][k] = x.swap_in ? raw[it][k].y : raw[it][k].x;
          vi[it][k] = x.swap_in ? raw[it][k].x : raw[it][k].y;
        }
      if (!ROWS && x.pass_tw) {
#pragma unroll
        for (int it = 0; it < IT; ++it)
#pragma unroll
          for (int k = 0; k < R; ++k) cmul<T>(vr[it][k], vi[it][k], w[it][k]);
      }
    }
  }

  static __device__ __forceinline__ void load_lds(const Ctx<T> &x,
                                                  const T *arr,
                                                  T (&v)[IT][R]) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int jj = x.t + it * TPT;
      if constexpr (TL::read_affine(NB)) {
        const T *a = arr + TL::stage_addr(x.c, jj);
#pragma unroll
        for (int k = 0; k < R; ++k)
          v[it][k] = (FULL || jj < NB) ? a[TL::step(k * NB)] : T(0);
      } else {
#pragma unroll
        for (int k = 0; k < R; ++k)
          v[it][k] = (FULL || jj < NB) ? arr[TL::stage_addr(x.c, jj + k * NB)]
                                       : T(0);
      }
    }
  }

  static __device__ __forceinline__ void put_lds(const Ctx<T> &x, T *arr,
                                                 T (&v)[IT][R]) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int jj = x.t + it * TPT;
      if (FULL || jj < NB) {
        const int base = (jj / NS) * NS * R + (jj % NS);
        if constexpr (TL::write_affine(NS, R)) {
          T *a = arr + TL::stage_addr(x.c, base);
#pragma unroll
          for (int k = 0; k < R; ++k) a[TL::step(k * NS)] = v[it][k];
        } else {
#pragma unroll
          for (int k = 0; k < R; ++k)
            arr[TL::stage_addr(x.c, base + k * NS)] = v[it][k];
        }
      }
    }
  }

  static __device__ __forceinline__ void store_direct(const Ctx<T> &x,
                                                      T (&vr)[IT][R],
                                                      T (&vi)[IT][R]) {
    if constexpr (STAGED) {
      if (x.contiguous) {
        store_staged(x, vr, vi);
        return;
      }
    }
    if (!x.live) return;
    if constexpr (MOFFT_VEC_SMALL && ROWS && TPT == 1 && IT == 1 &&
                  (R % 2 == 0) && sizeof(T) == 4) {
      nf4 *d = reinterpret_cast<nf4 *>(x.dst);
#pragma unroll
      for (int k = 0; k < R; k += 2) {
        nf4 v;
        v.x = x.swap_out ? vi[0][k] : vr[0][k];
        v.y = x.swap_out ? vr[0][k] : vi[0][k];
        v.z = x.swap_out ? vi[0][k + 1] : vr[0][k + 1];
        v.w = x.swap_out ? vr[0][k + 1] : vi[0][k + 1];
#if MOFFT_NT_STORE
        __builtin_nontemporal_store(v, d + k / 2);
#else
        d[k / 2] = v;
#endif
      }
    } else {
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int jj = x.t + it * TPT;
        if (FULL || jj < NB) {
#pragma unroll
          for (int k = 0; k < R; ++k) {
            const int o = jj + k * NB;
            cx<T> *ptr = ROWS ? x.dst + o : x.dst + (x.obase + (uint32_t)o * x.ns);
            store_global<T>(ptr, vr[it][k], vi[it][k], x.swap_out);
          }
        }
      }
    }
  }

  /* Twiddle, DFT and hand the results to the next sub-stage (or HBM). */
  static __device__ __forceinline__ void run(const Ctx<T> &x,
                                             const mofft_pass_args &p,
                                             T (&vr)[IT][R], T (&vi)[IT][R]) {
    if constexpr (NS > 1) {
      cx<T> w[IT][R - 1];
      load_tw(x, w);
      run_tw(x, p, vr, vi, w);
    } else {
      cx<T> w[IT][1];
      run_tw(x, p, vr, vi, w);
    }
  }

  /* the same with this sub-stage's twiddles already loaded (w is unused
   * when NS == 1) */
  template <int W>
  static __device__ __forceinline__ void run_tw(const Ctx<T> &x,
                                                const mofft_pass_args &p,
                                                T (&vr)[IT][R], T (&vi)[IT][R],
                                                const cx<T> (&w)[IT][W]) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      if constexpr (NS > 1) {
        const int jj = x.t + it * TPT;
        if (FULL || jj < NB) {
          if constexpr (SPARSE_TW) {
            cx<T> ww[R - 1];
#pragma unroll
            for (int k = 1; k < R; ++k)
              ww[k - 1] = tw_loaded(k) ? w[it][k - 1]
                                       : cprod<T>(w[it][k % 4 - 1],
                                                  w[it][(k & ~3) - 1]);
#pragma unroll
            for (int k = 1; k < R; ++k) cmul<T>(vr[it][k], vi[it][k], ww[k - 1]);
          } else {
#pragma unroll
            for (int k = 1; k < R; ++k) cmul<T>(vr[it][k], vi[it][k], w[it][k - 1]);
          }
        }
      }
      Dft<R>::template run<T>(vr[it], vi[it]);
    }
    if constexpr (!LAST) {
      using Next = typename StageSel<T, L, TPT, C, ROWS, NS * R,
                                     OFF + (NS > 1 ? NS * (R - 1) : 0),
                                     REST...>::type;
      T nr[Next::IT][Next::RAD];
      T ni[Next::IT][Next::RAD];
      /* a global table: the next sub-stage's twiddle loads are in flight
       * during the exchange instead of after it (FP32 N = 2^10..2^12 single
       * pass 4-10 % faster).  In 1024-thread workgroups a thread has at
       * most 128 VGPRs: the held twiddles spill FP64 tiles (N = 2^14 single
       * pass 77 -> 127 us) and cost strided FP32 tiles their second
       * resident workgroup; FP32 whole-transform tiles up to 8192 points
       * keep them, and the 16384-point tile where the next sub-stage loads
       * the reduced set (SPARSE_TW): 60 VGPRs, so two of its 1024-thread
       * workgroups share a CU (with the full set prefetched it needed 80,
       * one workgroup per CU: 63.2 us per 2^24 points against 49.9). */
      constexpr bool PF = !TW_IN_LDS && Next::PREFETCH_TW &&
                          (TPT * C < 1024 ||
                           (ROWS && sizeof(T) == 4 &&
                            (L <= 8192 || (L <= 16384 && Next::SPARSE_TW))));
      using NW = cx<T>[Next::IT][PF ? Next::RAD - 1 : 1];
      NW nw;
      if constexpr (PF) Next::load_tw(x, nw);
#if MOFFT_SPLIT_EXCHANGE
      put_lds(x, x.lds, vr);
      __syncthreads();
      Next::load_lds(x, x.lds, nr);
      __syncthreads();
      put_lds(x, x.lds, vi);
      __syncthreads();
      Next::load_lds(x, x.lds, ni);
      __syncthreads();
#else
      put_lds(x, x.lds, vr);
      put_lds(x, x.lds + TL::WORDS, vi);
      __syncthreads();
      Next::load_lds(x, x.lds, nr);
      Next::load_lds(x, x.lds + TL::WORDS, ni);
      __syncthreads();
#endif
      if constexpr (PF) Next::run_tw(x, p, nr, ni, nw);
      else Next::run(x, p, nr, ni);
    } else {
      if (ROWS || !tile_store<C>(p)) store_direct(x, vr, vi);
      else store_tile<T, L, TPT, C, ROWS, IT, R>(x, p, vr, vi);
    }
  }
};

/* Per-thread context of tile `block` (C columns) of a pass. */
template <typename T, int L, int TPT, int C, bool ROWS>
__device__ __forceinline__ void setup_tile(Ctx<T> &x, const mofft_pass_args &p,
                                           int64_t block, T *lds, int tid) {
  x.lds = lds;
  x.tid = tid;
  x.c = ROWS ? tid / TPT : tid % C;
  x.t = ROWS ? tid % TPT : tid / C;
  x.col0 = (uint32_t)block * C;
  const uint32_t col = x.col0 + x.c;
  x.live = col < (uint32_t)p.columns;
  x.contiguous = ROWS;
  if (ROWS) {
    const uint32_t cols = (uint32_t)p.columns - x.col0;
    x.tvalid = (cols < (uint32_t)C ? cols : (uint32_t)C) * (uint32_t)L;
    x.tsrc = (const cx<T> *)p.in + x.col0 * (uint32_t)p.n;
    x.tdst = (cx<T> *)p.out + x.col0 * (uint32_t)p.n;
  }
  const uint32_t b = ROWS ? col : div_by(col, p.stride_magic, p.stride_shift);
  const uint32_t j = ROWS ? 0u : col - b * (uint32_t)p.stride;
  x.stride = (uint32_t)p.stride;
  x.ns = (uint32_t)p.ns;
  x.src = (const cx<T> *)p.in + (b * (uint32_t)p.n + j);
  x.dst = (cx<T> *)p.out + b * (uint32_t)p.n;
  x.stw = (const cx<T> *)p.stage_tw;
  x.thi = (const cx<T> *)p.tw_hi;
  x.tlo = (const cx<T> *)p.tw_lo;
  const uint32_t jq = div_by(j, p.ns_magic, p.ns_shift);
  const uint32_t jr = j - jq * x.ns;
  x.obase = jq * x.ns * L + jr;
  x.jm_scale = jr * (uint32_t)p.tw_scale;
  x.split_shift = p.split_shift;
  x.split_mask = (1u << p.split_shift) - 1u;
  x.swap_in = (p.flags & MOFFT_PASS_SWAP_IN) != 0;
  x.swap_out = (p.flags & MOFFT_PASS_SWAP_OUT) != 0;
  x.pass_tw = (p.flags & MOFFT_PASS_TWIDDLE) != 0 && jr != 0;
  x.pass_table = (p.flags & MOFFT_PASS_TWIDDLE_TABLE) != 0;
  x.tpass = (const cx<T> *)p.tw_pass + jr;
}

template <typename T, int L, int TPT, int C, bool ROWS, int... R>
struct StageTw {
  static constexpr int COUNT = stage_tw_entries<1, R...>();
  static constexpr int TILE = Tile<T, L, TPT, C, ROWS>::BYTES;
  static constexpr int OFFSET = (TILE + 15) / 16 * 16;
  static constexpr int TABLE = COUNT * (int)sizeof(cx<T>);
  static constexpr bool IN_LDS = stage_tw_in_lds<T, L, TPT, C, ROWS>(COUNT);
  static constexpr int BYTES = IN_LDS ? OFFSET + TABLE : TILE;
};

/* Load, transform and store the tile described by x. */
template <typename T, int L, int TPT, int C, bool ROWS, int R0, int... R>
__device__ __forceinline__ void run_tile(const Ctx<T> &x,
                                         const mofft_pass_args &p) {
  using First = typename StageSel<T, L, TPT, C, ROWS, 1, 0, R0, R...>::type;
  using SW = StageTw<T, L, TPT, C, ROWS, R0, R...>;
  static_assert(SW::COUNT == 0 || First::TW_IN_LDS == SW::IN_LDS,
                "one LDS rule for the sub-stage table");
  T vr[First::IT][First::RAD];
  T vi[First::IT][First::RAD];
  if constexpr (SW::IN_LDS) {
    constexpr int NT = TPT * C;
    constexpr int PER = (SW::COUNT + NT - 1) / NT;
    cx<T> tw[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int q = x.tid + i * NT;
      if (q < SW::COUNT) tw[i] = x.stw[q];
    }
    First::load_first(x, vr, vi);
    cx<T> *table = reinterpret_cast<cx<T> *>(
        reinterpret_cast<unsigned char *>(x.lds) + SW::OFFSET);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int q = x.tid + i * NT;
      if (q < SW::COUNT) table[q] = tw[i];
    }
    /* read only after the first exchange's barriers */
    Ctx<T> y = x;
    y.stw = table;
    First::run(y, p, vr, vi);
  } else {
    First::load_first(x, vr, vi);
    First::run(x, p, vr, vi);
  }
}

/* Tile of workgroup b among `blocks`.  A strided tile's row runs are
 * C * sizeof(complex) bytes; below a 128-byte cache line, S = 128 / run
 * adjacent tiles share every line they touch.  Workgroups are dealt to the
 * eight XCDs in turn (b, b + 8, b + 16, ... run side by side on one XCD,
 * whose L2 is its own), so with MOFFT_PASS_XCD_ORDER the S tiles of a line
 * go to workgroups b, b + 8, ..., b + 8 (S - 1): each line is fetched into
 * one L2 and written back whole from it.  A bijection on every full group
 * of 8 S workgroups; the last partial group keeps its order. */
template <typename T, int C, bool ROWS>
__device__ __forceinline__ uint32_t tile_of_block(uint32_t b, uint32_t blocks,
                                                  uint32_t flags) {
  constexpr int RUN = C * (int)sizeof(cx<T>);
  if constexpr (ROWS || RUN >= 128 || 128 % RUN != 0) {
    return b;
  } else {
    constexpr uint32_t S = 128 / RUN;
    constexpr uint32_t G = 8 * S;
    if (!(flags & MOFFT_PASS_XCD_ORDER)) return b;
    const uint32_t g = b / G;
    if ((g + 1) * G > blocks) return b;
    const uint32_t r = b - g * G;
    return g * G + (r % 8) * S + r / 8;
  }
}

template <typename T, int L, int TPT, int C, bool ROWS, int R0, int... R>
__global__ void __launch_bounds__(TPT * C)
fft_pass_kernel(mofft_pass_args p) {
  extern __shared__ __align__(16) unsigned char mofft_lds[];
  Ctx<T> x;
  const uint32_t block =
      tile_of_block<T, C, ROWS>(blockIdx.x, gridDim.x, p.flags);
  setup_tile<T, L, TPT, C, ROWS>(x, p, (int64_t)block,
                                 reinterpret_cast<T *>(mofft_lds),
                                 (int)threadIdx.x);
  run_tile<T, L, TPT, C, ROWS, R0, R...>(x, p);
}

/* A launch with more than 64 KiB of dynamic LDS needs the kernel's
 * opt-in, once per device (`done` holds one bit per device ordinal). */
inline int allow_lds(std::atomic<unsigned long long> &done, const void *fn,
                     int bytes) {
  if (bytes <= 65536) return 0;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  const unsigned long long bit = dev < 64 ? 1ull << dev : 0ull;
  if (bit && (done.load(std::memory_order_acquire) & bit)) return 0;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return (int)e;
  done.fetch_or(bit, std::memory_order_acq_rel);
  return 0;
}

template <typename T, int L, int TPT, int C, bool ROWS, int... R>
struct PassKernel {
  static constexpr int THREADS = TPT * C;
  static constexpr int STAGES = sizeof...(R);
  static constexpr bool HAS_MFMA = ((R < 0) || ...);
  static constexpr int LDS = StageTw<T, L, TPT, C, ROWS, R...>::BYTES;
  static int launch(const mofft_pass_args *a, hipStream_t s) {
    static std::atomic<unsigned long long> configured{0};
    const int rc = allow_lds(
        configured,
        reinterpret_cast<const void *>(&fft_pass_kernel<T, L, TPT, C, ROWS, R...>),
        LDS);
    if (rc != 0) return rc;
    if (!a) return 0; /* configure only (e.g. before a stream capture) */
    /* row tiles never carry pass twiddles (see Stage::load_first) */
    if (ROWS && (a->flags & MOFFT_PASS_TWIDDLE)) return (int)hipErrorInvalidValue;
    const int64_t blocks = (a->columns + C - 1) / C;
    if (blocks <= 0) return 0;
    if (blocks > 0x7fffffff) return (int)hipErrorInvalidValue;
    /* LDS is only touched by sub-stage exchanges and by the transposed
     * store of strided passes whose output interleaves < C columns */
    const bool needs_lds =
        STAGES > 1 || HAS_MFMA || (!ROWS && a->ns < C) || (ROWS && TPT == 1);
    hipLaunchKernelGGL((fft_pass_kernel<T, L, TPT, C, ROWS, R...>),
                       dim3((unsigned)blocks), dim3(THREADS),
                       needs_lds ? LDS : 0, s, *a);
    return (int)hipGetLastError();
  }
};


/* ------------------------------------------------------------------ */
/* radix-16 / radix-32 sub-stages on the matrix cores (FP32, bank 3)    */
/* ------------------------------------------------------------------ */

typedef _Float16 mf_h2 __attribute__((ext_vector_type(2)));
typedef _Float16 mf_h4 __attribute__((ext_vector_type(4)));
typedef _Float16 mf_h8 __attribute__((ext_vector_type(8)));
typedef float mf_f2 __attribute__((ext_vector_type(2)));
typedef float mf_f16 __attribute__((ext_vector_type(16)));

/* matrix-core sub-stages in a chain */
template <int... RS>
constexpr int mfma_stages() {
  return (0 + ... + (RS < 0 ? 1 : 0));
}

/* log2 of the growth bound G of a chain's matrix-core sub-stages, taken at
 * the first of them (RM, then REST): the inputs of the last one have
 * components at most sqrt(2) times the product of the radices before it
 * times the column's largest input component; G = 2 * that product. */
template <int RM, int... REST>
constexpr int mfma_growth_log2() {
  constexpr int rs[] = {RM, REST..., 1};
  const int n = 1 + mfma_stages<REST...>();
  int lg = 1;
  for (int i = 0; i + 1 < n; ++i)
    for (int r = radix_value(rs[i]); r > 1; r >>= 1) ++lg;
  return lg;
}

/* 2^j for j in [-126, 126] */
__device__ __forceinline__ float exp2i(int j) {
  return __uint_as_float((uint32_t)(j + 127) << 23);
}

/* Scale exponent k for values whose largest component magnitude is mx:
 * mx * 2^k lies in [2^14, 2^15) / 2^lg (FP16 range for the high part,
 * normal range for the low part), clamped so that 2^k and 2^-k are normal
 * FP32 values (zero and subnormal mx give the upper clamp). */
__device__ __forceinline__ int mfma_scale_exp(float mx, int lg) {
  const int e = (int)((__float_as_uint(mx) >> 23) & 0xffu);
  int k = 141 - lg - e;
  k = k > 126 ? 126 : k;
  k = k < -126 ? -126 : k;
  return k;
}

/* a = ah + al, b = bh + bl in FP16, each part rounded to nearest (two
 * values per conversion instruction) */
__device__ __forceinline__ void split2(float a, float b, mf_h2 &hi,
                                       mf_h2 &lo) {
  const mf_f2 v = {a, b};
  hi = __builtin_convertvector(v, mf_h2);
  const mf_f2 h = __builtin_convertvector(hi, mf_f2);
  lo = __builtin_convertvector(v - h, mf_h2);
}

__device__ __forceinline__ mf_h8 cat4(mf_h2 a, mf_h2 b, mf_h2 c, mf_h2 d) {
  const mf_h4 ab = __builtin_shufflevector(a, b, 0, 1, 2, 3);
  const mf_h4 cd = __builtin_shufflevector(c, d, 0, 1, 2, 3);
  return __builtin_shufflevector(ab, cd, 0, 1, 2, 3, 4, 5, 6, 7);
}

/* LDS of a kernel's column scales: one exponent per column of the tile,
 * and the partial maxima they are reduced from (one array per kernel,
 * shared by its matrix-core sub-stages) */
template <int C>
__device__ __forceinline__ int *mfma_colexp() {
  __shared__ int a[C];
  return a;
}
template <int N>
__device__ __forceinline__ float *mfma_slots() {
  __shared__ float a[N];
  return a;
}

/* A radix-R DFT (R = 16 or 32) of 32 butterflies is one real product
 * D = M V: column j of V holds butterfly j's inputs (rows 0..R-1 real
 * parts, R..2R-1 imaginary parts), M is the real 2R x 2R form of the
 * R-point DFT matrix (mofft_mfma16_a / mofft_mfma32_a, generated), column j
 * of D the R outputs.  It runs as v_mfma_f32_32x32x16_f16 products over
 * R / 16 row tiles of 32 and R / 8 K-blocks of 16, in split precision:
 * every operand is a sum hi + lo of two FP16 values and D = Ml Vh + Mh Vl
 * + Mh Vh is accumulated in FP32 (the dropped Ml Vl is 2^-22 relative).
 *
 * Scaling: FP16 needs the inputs in a known range.  A chain with one
 * matrix-core sub-stage scales each butterfly by the power of two that
 * puts its largest input component in [2^14, 2^15) and scales the outputs
 * back.  A chain with several (they lead the chain) scales each column
 * once, at the first of them, by the power of two that keeps the inputs of
 * all of them below 2^15 (the product of the radices before the last one
 * bounds the growth), and scales back after the last.  The scaled column
 * maximum is then at least 2^(14 - lg) (2^3 for the deepest chain,
 * 32 x 32 x 16); values down to 2^-3 keep a normal low part, and below that
 * the low part's absolute error (2^-25) is at most 2^-28 of the column
 * maximum, so the error relative to the column stays that of the
 * per-butterfly form (the power-of-two scalings themselves are exact).
 *
 * Operand layouts (gfx950, verified on the device): A[i][k] in lane
 * i%32 + 32*(k/8) element k%8; B[k][j] in lane j + 32*(k/8) element k%8;
 * D[i][j] in lane j + 32*((i/4)%2) register (i/8)*4 + i%4.  So lane
 * (j, h = lane/32) holds the R/2 inputs m = 16*(r/8) + 8h + r%8 and, after
 * the products, the R/2 outputs q = 8*(r/4) + 4h + r%4 of butterfly j
 * (r < R/2; real parts of the outputs in row tile 0 registers r (R = 32)
 * or r (R = 16, r < 8), imaginary parts in row tile 1 or registers 8 + r).
 *
 * Inputs come from the LDS tile and outputs go back to it like any other
 * sub-stage's.  In a whole-transform (row) tile with at least 32
 * butterflies per column, the first sub-stage loads its inputs from HBM in
 * this layout and a last one stores from it (32 lanes: 32 consecutive
 * elements).  Otherwise the first loads in the thread layout of a VALU
 * radix-R stage and exchanges through LDS, the last exchanges back to that
 * layout and stores from it. */
template <typename T, int L, int TPT, int C, bool ROWS, int NS, int OFF,
          int RM, int... REST>
struct MStage {
  static_assert(sizeof(T) == 4, "matrix-core stages are FP32");
  static_assert(RM == 16 || RM == 32, "matrix-core radix");
  using TL = Tile<T, L, TPT, C, ROWS>;
  static constexpr int R = RM;
  static constexpr int RAD = R / 2;        /* values per lane and group */
  static constexpr int KB = R / 8;         /* K-blocks of 16 */
  static constexpr int RT = R / 16;        /* row tiles of 32 */
  static constexpr int NB = L / R;
  static constexpr int NT = TPT * C;
  static constexpr int WAVES = NT / 64;
  static constexpr int GROUPS = (C * NB) / 32;
  static constexpr int IT = GROUPS / (WAVES > 0 ? WAVES : 1);
  static constexpr bool LAST = sizeof...(REST) == 0;
  static constexpr bool FIRST_M = NS == 1;
  static constexpr bool LAST_M = mfma_stages<REST...>() == 0;
  static constexpr bool PER_BUTTERFLY = FIRST_M && LAST_M;
  /* (not for tiles of 32+ points per thread: the 32768-point tile's
   * direct loads spill 1000+ VGPRs, against 99 through the VALU layout) */
  static constexpr bool DIRECT = ROWS && NB >= 32 && L / NT <= 16;
  static constexpr int SPC = NB >= 32 ? NB / 32 : 1; /* slots per column */
  static constexpr bool PREFETCH_TW = false;
  static constexpr bool SPARSE_TW = false;
  static constexpr bool TW_IN_LDS = stage_tw_in_lds<T, L, TPT, C, ROWS>(
      OFF + stage_tw_entries<NS, RM, REST...>());
  static_assert(NT % 64 == 0 && (C * NB) % 32 == 0 && GROUPS % WAVES == 0,
                "matrix-core tile shape");
  /* the VALU radix-R stage whose thread layout HBM loads / stores use
   * when they do not use this one's */
  using V = Stage<T, L, TPT, C, ROWS, NS, OFF, R>;

  static __device__ __forceinline__ void where(const Ctx<T> &x, int it,
                                               int &c, int &jj, int &h) {
    const int lane = x.tid & 63;
    const int g = ((x.tid >> 6) + it * WAVES) * 32 + (lane & 31);
    c = g / NB;
    jj = g - c * NB;
    h = lane >> 5;
  }
  static __device__ __forceinline__ int in_m(int r, int h) {
    return 16 * (r >> 3) + 8 * h + (r & 7);
  }
  static __device__ __forceinline__ int out_q(int r, int h) {
    return 8 * (r >> 2) + 4 * h + (r & 3);
  }

  /* Exchange addresses as base + compile-time step (Tile::step): the
   * lane half h goes into the base (jj + 8 h NB, resp. the butterfly's
   * first output + 4 h NS), the rest of in_m / out_q into the step.  The
   * output form is exact when no residue mod 2^SH carries, checked here
   * over every j = jj mod NS, h and r. */
  static constexpr bool put_affine() {
    if (!TL::AFFINE) return false;
    if (TL::SH == 0) return true;
    const int m = 1 << TL::SH;
    if ((NS * R) % m != 0 && NS % m != 0) return false;
    for (int j = 0; j < NS; ++j)
      for (int h = 0; h < 2; ++h)
        for (int r = 0; r < RAD; ++r) {
          const int e0 = (NS % m == 0) ? 0 : j + 4 * h * NS;
          const int k = (8 * (r >> 2) + (r & 3)) * NS;
          if (e0 % m + k % m >= m) return false;
        }
    return true;
  }

  static __device__ __forceinline__ void load_lds(const Ctx<T> &x,
                                                  const T *arr,
                                                  T (&v)[IT][RAD]) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      int c, jj, h;
      where(x, it, c, jj, h);
      if constexpr (TL::read_affine(NB)) {
        const T *a = arr + TL::stage_addr(c, jj + 8 * h * NB);
#pragma unroll
        for (int r = 0; r < RAD; ++r)
          v[it][r] = a[TL::step((16 * (r >> 3) + (r & 7)) * NB)];
      } else {
#pragma unroll
        for (int r = 0; r < RAD; ++r)
          v[it][r] = arr[TL::stage_addr(c, jj + in_m(r, h) * NB)];
      }
    }
  }

  static __device__ __forceinline__ void put_lds(const Ctx<T> &x, T *arr,
                                                 T (&v)[IT][RAD]) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      int c, jj, h;
      where(x, it, c, jj, h);
      const int base = (jj / NS) * NS * R + (jj % NS);
      if constexpr (put_affine()) {
        T *a = arr + TL::stage_addr(c, base + 4 * h * NS);
#pragma unroll
        for (int r = 0; r < RAD; ++r)
          a[TL::step((8 * (r >> 2) + (r & 3)) * NS)] = v[it][r];
      } else {
#pragma unroll
        for (int r = 0; r < RAD; ++r)
          arr[TL::stage_addr(c, base + out_q(r, h) * NS)] = v[it][r];
      }
    }
  }

  /* thread-layout values of V at their positions jj + k*NB */
  static __device__ __forceinline__ void put_natural(const Ctx<T> &x, T *arr,
                                                     T (&u)[V::IT][R]) {
#pragma unroll
    for (int it = 0; it < V::IT; ++it) {
      const int jj = x.t + it * TPT;
      if (V::FULL || jj < NB) {
        if constexpr (TL::read_affine(NB)) {
          T *a = arr + TL::stage_addr(x.c, jj);
#pragma unroll
          for (int k = 0; k < R; ++k) a[TL::step(k * NB)] = u[it][k];
        } else {
#pragma unroll
          for (int k = 0; k < R; ++k)
            arr[TL::stage_addr(x.c, jj + k * NB)] = u[it][k];
        }
      }
    }
  }

  static __device__ __forceinline__ void load_first(const Ctx<T> &x,
                                                    T (&vr)[IT][RAD],
                                                    T (&vi)[IT][RAD]) {
    if constexpr (DIRECT) {
      /* every load issued before any is used */
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        int c, jj, h;
        where(x, it, c, jj, h);
        const uint32_t base = (uint32_t)c * L + (uint32_t)jj;
        const bool ok = (uint32_t)c * L < x.tvalid;
#pragma unroll
        for (int r = 0; r < RAD; ++r) {
          cx<T> v;
          v.x = T(0);
          v.y = T(0);
          if (ok) v = ld_stream<T>(x.tsrc + base + (uint32_t)(in_m(r, h) * NB));
          vr[it][r] = x.swap_in ? v.y : v.x;
          vi[it][r] = x.swap_in ? v.x : v.y;
        }
      }
    } else {
      T ur[V::IT][R];
      T ui[V::IT][R];
      V::load_first(x, ur, ui);
      put_natural(x, x.lds, ur);
      __syncthreads();
      load_lds(x, x.lds, vr);
      __syncthreads();
      put_natural(x, x.lds, ui);
      __syncthreads();
      load_lds(x, x.lds, vi);
      __syncthreads();
    }
  }

  /* Column scales at the first of several matrix-core sub-stages: every
   * 32-lane group reduces the largest component magnitude of its columns
   * into LDS slots, then every lane takes its column's maximum, records the
   * exponent for the last sub-stage and scales its inputs. */
  static __device__ __forceinline__ void column_scale(const Ctx<T> &x,
                                                      T (&vr)[IT][RAD],
                                                      T (&vi)[IT][RAD]) {
    constexpr int W = NB >= 32 ? 32 : NB;  /* lanes of a column per half */
    constexpr int LG = mfma_growth_log2<RM, REST...>();
    float *slot = mfma_slots<C * SPC>();
    int *cexp = mfma_colexp<C>();
    const int lane = x.tid & 63;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      int c, jj, h;
      where(x, it, c, jj, h);
      float mx = 0.f;
#pragma unroll
      for (int r = 0; r < RAD; ++r)
        mx = fmaxf(mx, fmaxf(fabsf(vr[it][r]), fabsf(vi[it][r])));
      mx = fmaxf(mx, __shfl_xor(mx, 32));
#pragma unroll
      for (int m = 1; m < W; m <<= 1) mx = fmaxf(mx, __shfl_xor(mx, m));
      if (lane < 32 && jj % W == 0) slot[c * SPC + jj / 32] = mx;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      int c, jj, h;
      where(x, it, c, jj, h);
      float mx = 0.f;
#pragma unroll
      for (int k = 0; k < SPC; ++k) mx = fmaxf(mx, slot[c * SPC + k]);
      const int e = mfma_scale_exp(mx, LG);
      if (jj == 0 && h == 0) cexp[c] = e;
      const float up = exp2i(e);
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        vr[it][r] *= up;
        vi[it][r] *= up;
      }
    }
  }

  static __device__ __forceinline__ void dft(const Ctx<T> &x, int it,
                                             T (&re)[RAD], T (&im)[RAD]) {
    int c, jj, h;
    where(x, it, c, jj, h);
    int k = 0;
    if constexpr (PER_BUTTERFLY) {
      float mx = 0.f;
#pragma unroll
      for (int r = 0; r < RAD; ++r)
        mx = fmaxf(mx, fmaxf(fabsf(re[r]), fabsf(im[r])));
      mx = fmaxf(mx, __shfl_xor(mx, 32));
      k = mfma_scale_exp(mx, 0);
      const float up = exp2i(k);
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        re[r] *= up;
        im[r] *= up;
      }
    }
    const int lane = x.tid & 63;
    const mf_h8 *tab = reinterpret_cast<const mf_h8 *>(
        R == 16 ? mofft_mfma16_a : mofft_mfma32_a);
    mf_f16 acc[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) acc[rt] = mf_f16{};
    /* twiddles and FP16 splits eight inputs at a time (K-blocks u of the
     * real and u + KB/2 of the imaginary parts), then the products one row
     * tile and one K-block at a time: the inputs die as they are split and
     * only one block's matrix operands are live */
    mf_h8 bh[KB], bl[KB];
#pragma unroll
    for (int u = 0; u < KB / 2; ++u) {
      if constexpr (NS > 1) {
        const cx<T> *w = x.stw + OFF + (jj % NS) * (R - 1);
#pragma unroll
        for (int r = 8 * u; r < 8 * u + 8; ++r) {
          const int m = in_m(r, h);
          if (m > 0) cmul<T>(re[r], im[r], w[m - 1]);
        }
      }
#pragma unroll
      for (int part = 0; part < 2; ++part) {
        const int kb = u + part * (KB / 2);
        const T *src = (part == 0 ? re : im) + 8 * u;
        mf_h2 h0, h1, h2, h3, l0, l1, l2, l3;
        split2(src[0], src[1], h0, l0);
        split2(src[2], src[3], h1, l1);
        split2(src[4], src[5], h2, l2);
        split2(src[6], src[7], h3, l3);
        bh[kb] = cat4(h0, h1, h2, h3);
        bl[kb] = cat4(l0, l1, l2, l3);
      }
    }
    /* the small terms (Ml Vh, Mh Vl, 2^-11 of the result) of every K-block
     * first, the large ones last: each FP32 addition rounds at the scale of
     * its partial sum, so only the KB additions of Mh Vh round at the
     * result's scale.  Consecutive products alternate between the row
     * tiles' accumulators, so a radix-32 stage issues two independent
     * chains instead of waiting on one product after another. */
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const mf_h8 ml = tab[((1 * RT + rt) * KB + kb) * 64 + lane];
        acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ml, bh[kb], acc[rt], 0, 0, 0);
      }
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const mf_h8 mh = tab[((0 * RT + rt) * KB + kb) * 64 + lane];
        acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(mh, bl[kb], acc[rt], 0, 0, 0);
      }
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const mf_h8 mh = tab[((0 * RT + rt) * KB + kb) * 64 + lane];
        acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(mh, bh[kb], acc[rt], 0, 0, 0);
      }
    if constexpr (RT == 1) {
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        re[r] = acc[0][r];
        im[r] = acc[0][8 + r];
      }
    } else {
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        re[r] = acc[0][r];
        im[r] = acc[RT - 1][r];
      }
    }
    if constexpr (PER_BUTTERFLY || (LAST_M && !FIRST_M)) {
      if constexpr (!PER_BUTTERFLY) k = mfma_colexp<C>()[c];
      const float down = exp2i(-k);
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        re[r] *= down;
        im[r] *= down;
      }
    }
  }

  static __device__ __forceinline__ void run(const Ctx<T> &x,
                                             const mofft_pass_args &p,
                                             T (&vr)[IT][RAD],
                                             T (&vi)[IT][RAD]) {
    if constexpr (FIRST_M && !PER_BUTTERFLY) column_scale(x, vr, vi);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      dft(x, it, vr[it], vi[it]);
      /* one group's twiddles and operands live at a time */
    }
    if constexpr (!LAST) {
      using Next = typename StageSel<T, L, TPT, C, ROWS, NS * R,
                                     OFF + (NS > 1 ? NS * (R - 1) : 0),
                                     REST...>::type;
      T nr[Next::IT][Next::RAD];
      T ni[Next::IT][Next::RAD];
      put_lds(x, x.lds, vr);
      __syncthreads();
      Next::load_lds(x, x.lds, nr);
      __syncthreads();
      put_lds(x, x.lds, vi);
      __syncthreads();
      Next::load_lds(x, x.lds, ni);
      __syncthreads();
      Next::run(x, p, nr, ni);
    } else if constexpr (DIRECT) {
      /* outputs q of butterfly jj at jj + q*NB (NS == NB here) */
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        int c, jj, h;
        where(x, it, c, jj, h);
        const uint32_t base = (uint32_t)c * L + (uint32_t)jj;
        if ((uint32_t)c * L < x.tvalid) {
#pragma unroll
          for (int r = 0; r < RAD; ++r)
            store_global<T>(x.tdst + base + (uint32_t)(out_q(r, h) * NB),
                            vr[it][r], vi[it][r], x.swap_out);
        }
      }
    } else {
      /* back to V's thread layout (outputs at jj + q*NB), then store */
      T ur[V::IT][R];
      T ui[V::IT][R];
      put_lds(x, x.lds, vr);
      __syncthreads();
      V::load_lds(x, x.lds, ur);
      __syncthreads();
      put_lds(x, x.lds, vi);
      __syncthreads();
      V::load_lds(x, x.lds, ui);
      __syncthreads();
      if (ROWS || !tile_store<C>(p)) V::store_direct(x, ur, ui);
      else store_tile<T, L, TPT, C, ROWS, V::IT, R>(x, p, ur, ui);
    }
  }
};

}  // namespace mofft
