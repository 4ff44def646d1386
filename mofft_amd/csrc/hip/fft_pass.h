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
 * Reference counterpart: the per-stage kernels `mofft_kernel_r*_first/other`
 * (compiler/mofft_compiler/emitter.py:283-466) and the executor's transposes
 * (src/mofft.c:1987-2044); here several stages fuse into one HBM pass.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../mofft_abi.h"
#include "mofft_codelets.h"

namespace mofft {

template <typename T> struct cx_of;
template <> struct cx_of<float> { using type = float2; };
template <> struct cx_of<double> { using type = double2; };
template <typename T> using cx = typename cx_of<T>::type;

/* LDS tile geometry.  Real and imaginary parts live in separate arrays.
 * Column-major (ROWS) tiles pad one word every 16; interleaved (COLS) tiles
 * pad one word every 32. */
template <typename T, int L, int TPT, int C, bool ROWS>
struct Tile {
  static constexpr int NT = TPT * C;
  static constexpr int LP = L + (L >> 4) + 1;
  static constexpr int ROW_WORDS = C * LP;
  static constexpr int COL_WORDS = L * C + ((L * C) >> 5) + 1;
  static constexpr int WORDS = ROWS ? ROW_WORDS
                                    : (ROW_WORDS > COL_WORDS ? ROW_WORDS
                                                             : COL_WORDS);
  static constexpr int BYTES = 2 * WORDS * (int)sizeof(T);

  static __device__ __forceinline__ int row_addr(int c, int e) {
    return c * LP + e + (e >> 4);
  }
  static __device__ __forceinline__ int stage_addr(int c, int e) {
    if constexpr (ROWS) {
      return row_addr(c, e);
    } else {
      const int q = e * C + c;
      return q + (q >> 5);
    }
  }
};

template <typename T>
struct Ctx {
  T *sre;
  T *sim;
  const cx<T> *src;   /* element e of this thread's column: src[e * stride] */
  cx<T> *dst;         /* output base of this thread's transform */
  const cx<T> *stw;
  const cx<T> *thi;
  const cx<T> *tlo;
  int64_t stride;
  int64_t ns;
  int64_t obase;      /* (j / ns) * ns * L + j % ns */
  uint64_t jm_scale;  /* (j % ns) * tw_scale */
  uint64_t split_mask;
  uint32_t split_shift;
  int c;
  int t;
  bool live;
  bool swap_in;
  bool swap_out;
  bool pass_tw;
};

template <typename T>
__device__ __forceinline__ void cmul(T &re, T &im, const cx<T> w) {
  const T r = re * w.x - im * w.y;
  const T i = re * w.y + im * w.x;
  re = r;
  im = i;
}

template <typename T>
__device__ __forceinline__ void load_global(const Ctx<T> &x, int e, T &a,
                                            T &b) {
  const cx<T> v = x.src[(int64_t)e * x.stride];
  a = x.swap_in ? v.y : v.x;
  b = x.swap_in ? v.x : v.y;
  if (x.pass_tw) {
    const uint64_t m = (uint64_t)e * x.jm_scale;
    const cx<T> h = x.thi[m >> x.split_shift];
    const cx<T> l = x.tlo[m & x.split_mask];
    cx<T> w;
    w.x = h.x * l.x - h.y * l.y;
    w.y = h.x * l.y + h.y * l.x;
    cmul<T>(a, b, w);
  }
}

template <typename T>
__device__ __forceinline__ void store_global(cx<T> *p, T a, T b, bool swap) {
  cx<T> v;
  v.x = swap ? b : a;
  v.y = swap ? a : b;
  *p = v;
}

/* Cooperative store of the whole tile, used when the pass output of
 * adjacent columns is not contiguous (ns < C).  The tile holds C columns of
 * L outputs in row layout. */
template <typename T, int L, int TPT, int C, bool ROWS>
__device__ __forceinline__ void store_tile(const Ctx<T> &x,
                                           const mofft_pass_args &p) {
  using TL = Tile<T, L, TPT, C, ROWS>;
  cx<T> *out = (cx<T> *)p.out;
  const int tid = threadIdx.x;
  const int64_t col0 = (int64_t)blockIdx.x * C;
  const int64_t b0 = col0 / p.stride;
  const int64_t j0 = col0 - b0 * p.stride;
  const int64_t ns = p.ns;
  const bool swap = (p.flags & MOFFT_PASS_SWAP_OUT) != 0;
  const bool aligned = (j0 + C <= p.stride) && (C % ns == 0) &&
                       (j0 % ns == 0) && (col0 + C <= p.columns);
  if (aligned) {
    cx<T> *base = out + b0 * p.n + (j0 / ns) * ns * L;
    const int span = (int)ns * L;
#pragma unroll 4
    for (int q = tid; q < C * L; q += TL::NT) {
      const int g = q / span;
      const int rem = q - g * span;
      const int o = rem / (int)ns;
      const int r = rem - o * (int)ns;
      const int cc = g * (int)ns + r;
      const int ad = TL::row_addr(cc, o);
      store_global<T>(base + q, x.sre[ad], x.sim[ad], swap);
    }
  } else {
    for (int q = tid; q < C * L; q += TL::NT) {
      const int cc = q / L;
      const int o = q - cc * L;
      const int64_t col = col0 + cc;
      if (col >= p.columns) continue;
      const int64_t b = col / p.stride;
      const int64_t j = col - b * p.stride;
      const int64_t pos = b * p.n + (j / ns) * ns * L + (j % ns) + (int64_t)o * ns;
      const int ad = TL::row_addr(cc, o);
      store_global<T>(out + pos, x.sre[ad], x.sim[ad], swap);
    }
  }
}

template <typename T, int L, int TPT, int C, bool ROWS, int NS, int OFF,
          bool FIRST, int R, int... REST>
struct Stage {
  using TL = Tile<T, L, TPT, C, ROWS>;

  static __device__ __forceinline__ void run(const Ctx<T> &x,
                                             const mofft_pass_args &p) {
    constexpr int NB = L / R;
    constexpr int IT = (NB + TPT - 1) / TPT;
    constexpr bool FULL = IT * TPT == NB;
    constexpr bool LAST = sizeof...(REST) == 0;
    T vr[IT][R];
    T vi[IT][R];

#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int jj = x.t + it * TPT;
      const bool ok = FULL || jj < NB;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int e = jj + k * NB;
        T a = T(0), b = T(0);
        if constexpr (FIRST) {
          if (ok && x.live) load_global<T>(x, e, a, b);
        } else {
          if (ok) {
            const int ad = TL::stage_addr(x.c, e);
            a = x.sre[ad];
            b = x.sim[ad];
          }
        }
        vr[it][k] = a;
        vi[it][k] = b;
      }
    }
    if constexpr (!FIRST) __syncthreads();

#pragma unroll
    for (int it = 0; it < IT; ++it) {
      if constexpr (NS > 1) {
        const int jj = x.t + it * TPT;
        if (FULL || jj < NB) {
          const cx<T> *w = x.stw + OFF + (jj % NS) * (R - 1);
#pragma unroll
          for (int k = 1; k < R; ++k) cmul<T>(vr[it][k], vi[it][k], w[k - 1]);
        }
      }
      Dft<R>::template run<T>(vr[it], vi[it]);
    }

    if constexpr (!LAST) {
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int jj = x.t + it * TPT;
        if (FULL || jj < NB) {
          const int base = (jj / NS) * NS * R + (jj % NS);
#pragma unroll
          for (int k = 0; k < R; ++k) {
            const int ad = TL::stage_addr(x.c, base + k * NS);
            x.sre[ad] = vr[it][k];
            x.sim[ad] = vi[it][k];
          }
        }
      }
      __syncthreads();
      Stage<T, L, TPT, C, ROWS, NS * R, OFF + (NS > 1 ? NS * (R - 1) : 0),
            false, REST...>::run(x, p);
    } else {
      /* final sub-stage: output o = jj + k * NB of the L-point DFT */
      const bool direct = ROWS || p.ns >= C;
      if (direct) {
        if (x.live) {
#pragma unroll
          for (int it = 0; it < IT; ++it) {
            const int jj = x.t + it * TPT;
            if (FULL || jj < NB) {
#pragma unroll
              for (int k = 0; k < R; ++k) {
                const int64_t o = jj + k * NB;
                cx<T> *ptr = ROWS ? x.dst + o : x.dst + x.obase + o * x.ns;
                store_global<T>(ptr, vr[it][k], vi[it][k], x.swap_out);
              }
            }
          }
        }
      } else {
        if constexpr (FIRST) {
          /* single sub-stage: LDS untouched so far */
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
          const int jj = x.t + it * TPT;
          if (FULL || jj < NB) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
              const int ad = TL::row_addr(x.c, jj + k * NB);
              x.sre[ad] = vr[it][k];
              x.sim[ad] = vi[it][k];
            }
          }
        }
        __syncthreads();
        store_tile<T, L, TPT, C, ROWS>(x, p);
      }
    }
  }
};

template <typename T, int L, int TPT, int C, bool ROWS, int... R>
__global__ void __launch_bounds__(TPT * C)
fft_pass_kernel(mofft_pass_args p) {
  using TL = Tile<T, L, TPT, C, ROWS>;
  extern __shared__ __align__(16) unsigned char mofft_lds[];
  Ctx<T> x;
  x.sre = reinterpret_cast<T *>(mofft_lds);
  x.sim = x.sre + TL::WORDS;
  const int tid = threadIdx.x;
  x.c = ROWS ? tid / TPT : tid % C;
  x.t = ROWS ? tid % TPT : tid / C;
  const int64_t col = (int64_t)blockIdx.x * C + x.c;
  x.live = col < p.columns;
  const int64_t b = ROWS ? col : col / p.stride;
  const int64_t j = ROWS ? 0 : col - b * p.stride;
  x.stride = p.stride;
  x.ns = p.ns;
  x.src = (const cx<T> *)p.in + b * p.n + j;
  x.dst = (cx<T> *)p.out + b * p.n;
  x.stw = (const cx<T> *)p.stage_tw;
  x.thi = (const cx<T> *)p.tw_hi;
  x.tlo = (const cx<T> *)p.tw_lo;
  const int64_t jq = p.ns > 1 ? j / p.ns : j;
  const int64_t jr = j - jq * p.ns;
  x.obase = jq * p.ns * L + jr;
  x.jm_scale = (uint64_t)jr * (uint64_t)p.tw_scale;
  x.split_shift = p.split_shift;
  x.split_mask = (uint64_t(1) << p.split_shift) - 1;
  x.swap_in = (p.flags & MOFFT_PASS_SWAP_IN) != 0;
  x.swap_out = (p.flags & MOFFT_PASS_SWAP_OUT) != 0;
  x.pass_tw = (p.flags & MOFFT_PASS_TWIDDLE) != 0 && jr != 0;
  Stage<T, L, TPT, C, ROWS, 1, 0, true, R...>::run(x, p);
}

template <typename T, int L, int TPT, int C, bool ROWS, int... R>
struct PassKernel {
  static constexpr int THREADS = TPT * C;
  static constexpr int LDS = Tile<T, L, TPT, C, ROWS>::BYTES;
  static int launch(const mofft_pass_args *a, hipStream_t s) {
    static bool configured = false;
    if (!configured) {
      if (LDS > 65536) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void *>(
                &fft_pass_kernel<T, L, TPT, C, ROWS, R...>),
            hipFuncAttributeMaxDynamicSharedMemorySize, LDS);
        if (e != hipSuccess) return (int)e;
      }
      configured = true;
    }
    const int64_t blocks = (a->columns + C - 1) / C;
    if (blocks <= 0) return 0;
    if (blocks > 0x7fffffff) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL((fft_pass_kernel<T, L, TPT, C, ROWS, R...>),
                       dim3((unsigned)blocks), dim3(THREADS), LDS, s, *a);
    return (int)hipGetLastError();
  }
};

}  // namespace mofft
