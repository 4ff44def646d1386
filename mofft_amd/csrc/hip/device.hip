/*
 * device.hip - the HIP side of the C ABI declared in mofft_abi.h.
 *
 * Device discovery, allocation, copies, FP64 twiddle generation and the pass
 * launcher.  This file and the generated kernel parts are the only code that
 * talks to HIP; mofft.c stays plain C.  It replaces the reference's
 * hardware probe and streaming-mode entry (src/mofft.c:89-113,
 * src/streaming_entry.S:4-38).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <mutex>

#include "../mofft_abi.h"

namespace mofft {
constexpr int kMaxDevices = 64;
typedef int (*launch_fn)(const mofft_pass_args *, void *);
extern const launch_fn launch_table[];
extern const int launch_table_size;
/* persistent batch kernels (generated mofft_batch_f32/f64.hip) */
int batch_occupancy_f32(int lds, int *bpc, int *cus);
int batch_occupancy_f64(int lds, int *bpc, int *cus);
int batch_launch_f32(const mofft_batch_header *h, int grid, int lds, void *s);
int batch_launch_f64(const mofft_batch_header *h, int grid, int lds, void *s);

/* exp(-2*pi*i*m/M) in FP64: exact integer octant reduction of 8m/8M, then
 * sincospi of an argument in [0, 1/4]. */
__device__ __forceinline__ void omega_f64(uint64_t m, uint64_t M, double &c,
                                          double &s) {
  m %= M;
  const uint64_t Q = 8 * M;
  uint64_t p = 8 * m;
  bool neg_s = false, neg_c = false, swap = false;
  if (p > Q / 2) { p = Q - p; neg_s = true; }
  if (p > Q / 4) { p = Q / 2 - p; neg_c = true; }
  if (p > Q / 8) { p = Q / 4 - p; swap = true; }
  double sv, cv;
  sincospi(2.0 * ((double)p / (double)Q), &sv, &cv);
  if (swap) { const double t = cv; cv = sv; sv = t; }
  if (neg_c) cv = -cv;
  if (neg_s) sv = -sv;
  c = cv;
  s = -sv; /* forward sign */
}

template <typename T>
__global__ void fill_powers_kernel(T *out, int64_t count, uint64_t M,
                                   uint64_t step) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t m = (uint64_t)(((unsigned __int128)i * step) % M);
  double c, s;
  omega_f64(m, M, c, s);
  out[2 * i] = (T)c;
  out[2 * i + 1] = (T)s;
}

template <typename T>
__global__ void fill_stage_kernel(T *out, int64_t ns, int r) {
  const int64_t count = ns * (r - 1);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t a = i / (r - 1);
  const int64_t k = i - a * (r - 1) + 1;
  double c, s;
  omega_f64((uint64_t)(a * k), (uint64_t)(ns * r), c, s);
  out[2 * i] = (T)c;
  out[2 * i + 1] = (T)s;
}
template <typename T>
__global__ void fill_pass_kernel(T *out, int64_t ns, int64_t L) {
  const int64_t count = ns * L;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t k = i / ns;
  const int64_t a = i - k * ns;
  double c, s;
  omega_f64((uint64_t)(a * k), (uint64_t)(ns * L), c, s);
  out[2 * i] = (T)c;
  out[2 * i + 1] = (T)s;
}
/* Workspace of the allocation-free execute (mofft_execute_async_scratch_*):
 * one array per device for the life of the process.  Host code never uses
 * this name as an address: it is a host-side stand-in for the device
 * symbol; mofft_hip_scratch resolves the device address once per device. */
static __device__ __align__(256) unsigned char scratch[MOFFT_SCRATCH_BYTES];

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T> struct pair_of;
template <> struct pair_of<float> { using type = float2; };
template <> struct pair_of<double> { using type = double2; };

__device__ __forceinline__ double unit_interval(uint64_t key) {
  return 2.0 * ((double)(splitmix64(key) >> 11) * 0x1.0p-53) - 1.0;
}

/* One complex element per thread and iteration (one 8- or 16-byte store);
 * the splitmix64 keys seed + (2e + c + 1) * gamma advance by a constant per
 * iteration instead of a 64-bit multiply per value (the same values:
 * arithmetic mod 2^64).  FP32 output 3.5 -> 3.85 TB/s on MI355X (still
 * bound by the two 64-bit mixes per element); FP64 5.9 TB/s. */
template <typename T>
__global__ void fill_uniform_kernel(T *out, int64_t count, uint64_t seed) {
  using P = typename pair_of<T>::type;
  constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t key = seed + (uint64_t)(2 * e + 1) * kGamma;
  const uint64_t step = (uint64_t)(2 * stride) * kGamma;
  for (; e < count; e += stride, key += step) {
    P v;
    v.x = (T)unit_interval(key);
    v.y = (T)unit_interval(key + kGamma);
    reinterpret_cast<P *>(out)[e] = v;
  }
}
}  // namespace mofft

extern "C" {

int mofft_hip_device_available(void) {
  static int cached = -1;
  if (cached >= 0) return cached;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    cached = 0;
    return 0;
  }
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) { cached = 0; return 0; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { cached = 0; return 0; }
  cached = strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
  return cached;
}

/* Device address of `scratch` on the current device: looked up with
 * hipGetSymbolAddress the first time a device asks (that device current),
 * then served from a per-device cache.  Each device holds its own copy of
 * the array at its own address. */
int mofft_hip_scratch(void **ptr, size_t *bytes) {
  static std::mutex lock;
  static std::atomic<void *> address[mofft::kMaxDevices];
  if (!ptr || !bytes) return (int)hipErrorInvalidValue;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= mofft::kMaxDevices) return (int)hipErrorInvalidDevice;
  void *p = address[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> guard(lock);
    p = address[dev].load(std::memory_order_relaxed);
    if (!p) {
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(mofft::scratch));
      if (e != hipSuccess) return (int)e;
      if (!p) return (int)hipErrorInvalidSymbol;
      address[dev].store(p, std::memory_order_release);
    }
  }
  *ptr = p;
  *bytes = MOFFT_SCRATCH_BYTES;
  return 0;
}

int mofft_hip_malloc(void **ptr, size_t bytes) {
  if (!ptr) return (int)hipErrorInvalidValue;
  *ptr = nullptr;
  if (bytes == 0) return 0;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) { *ptr = nullptr; (void)hipGetLastError(); }
  return (int)e;
}

int mofft_hip_free(void *ptr) {
  if (!ptr) return 0;
  return (int)hipFree(ptr);
}

int mofft_hip_is_device_pointer(const void *ptr) {
  hipPointerAttribute_t attr;
  memset(&attr, 0, sizeof(attr));
  hipError_t e = hipPointerGetAttributes(&attr, ptr);
  if (e != hipSuccess) { (void)hipGetLastError(); return 0; }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int mofft_hip_copy(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return 0;
  return (int)hipMemcpy(dst, src, bytes, hipMemcpyDefault);
}

int mofft_hip_copy_async(void *dst, const void *src, size_t bytes,
                         void *stream) {
  if (bytes == 0) return 0;
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault,
                             (hipStream_t)stream);
}

int mofft_hip_synchronize(void *stream) {
  return (int)hipStreamSynchronize((hipStream_t)stream);
}

int mofft_hip_fill_powers(void *out, int precision_bits, int64_t count,
                          int64_t M, int64_t step, void *stream) {
  if (count <= 0) return 0;
  const unsigned blocks = (unsigned)((count + 255) / 256);
  if (precision_bits == 32)
    hipLaunchKernelGGL(mofft::fill_powers_kernel<float>, dim3(blocks),
                       dim3(256), 0, (hipStream_t)stream, (float *)out, count,
                       (uint64_t)M, (uint64_t)step);
  else
    hipLaunchKernelGGL(mofft::fill_powers_kernel<double>, dim3(blocks),
                       dim3(256), 0, (hipStream_t)stream, (double *)out, count,
                       (uint64_t)M, (uint64_t)step);
  return (int)hipGetLastError();
}

int mofft_hip_fill_stage(void *out, int precision_bits, int64_t ns, int r,
                         void *stream) {
  const int64_t count = ns * (r - 1);
  if (count <= 0) return 0;
  const unsigned blocks = (unsigned)((count + 255) / 256);
  if (precision_bits == 32)
    hipLaunchKernelGGL(mofft::fill_stage_kernel<float>, dim3(blocks),
                       dim3(256), 0, (hipStream_t)stream, (float *)out, ns, r);
  else
    hipLaunchKernelGGL(mofft::fill_stage_kernel<double>, dim3(blocks),
                       dim3(256), 0, (hipStream_t)stream, (double *)out, ns, r);
  return (int)hipGetLastError();
}

int mofft_hip_fill_pass(void *out, int precision_bits, int64_t ns, int64_t L,
                        void *stream) {
  const int64_t count = ns * L;
  if (count <= 0) return 0;
  const unsigned blocks = (unsigned)((count + 255) / 256);
  if (precision_bits == 32)
    hipLaunchKernelGGL(mofft::fill_pass_kernel<float>, dim3(blocks), dim3(256),
                       0, (hipStream_t)stream, (float *)out, ns, L);
  else
    hipLaunchKernelGGL(mofft::fill_pass_kernel<double>, dim3(blocks),
                       dim3(256), 0, (hipStream_t)stream, (double *)out, ns, L);
  return (int)hipGetLastError();
}

static hipEvent_t timer_events[2];
static int timer_ready = 0;

int mofft_hip_timer_begin(void *stream) {
  if (!timer_ready) {
    if (hipEventCreate(&timer_events[0]) != hipSuccess ||
        hipEventCreate(&timer_events[1]) != hipSuccess)
      return (int)hipErrorInvalidValue;
    timer_ready = 1;
  }
  return (int)hipEventRecord(timer_events[0], (hipStream_t)stream);
}

int mofft_hip_timer_end(void *stream, double *seconds) {
  if (!timer_ready || !seconds) return (int)hipErrorInvalidValue;
  hipError_t e = hipEventRecord(timer_events[1], (hipStream_t)stream);
  if (e == hipSuccess) e = hipEventSynchronize(timer_events[1]);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, timer_events[0], timer_events[1]);
  *seconds = (double)ms * 1e-3;
  return (int)e;
}

int mofft_hip_launch_pass(int kernel, const mofft_pass_args *args,
                          void *stream) {
  if (kernel < 0 || kernel >= mofft::launch_table_size)
    return (int)hipErrorInvalidValue;
  return mofft::launch_table[kernel](args, stream);
}

int mofft_hip_batch_occupancy(int precision_bits, int lds_bytes,
                              int *blocks_per_cu, int *cu_count) {
  if (!blocks_per_cu || !cu_count || lds_bytes < 0)
    return (int)hipErrorInvalidValue;
  if (precision_bits == 32)
    return mofft::batch_occupancy_f32(lds_bytes, blocks_per_cu, cu_count);
  if (precision_bits == 64)
    return mofft::batch_occupancy_f64(lds_bytes, blocks_per_cu, cu_count);
  return (int)hipErrorInvalidValue;
}

int mofft_hip_batch_launch(int precision_bits, const mofft_batch_header *h,
                           int grid, int lds_bytes, void *stream) {
  if (!h || grid <= 0 || lds_bytes < 0) return (int)hipErrorInvalidValue;
  if (precision_bits == 32)
    return mofft::batch_launch_f32(h, grid, lds_bytes, stream);
  if (precision_bits == 64)
    return mofft::batch_launch_f64(h, grid, lds_bytes, stream);
  return (int)hipErrorInvalidValue;
}

/* Stream capture on a private stream; the graph holds two kernel nodes
 * (counter reset, persistent kernel) and no memset node. */
int mofft_hip_batch_graph(int precision_bits, const mofft_batch_header *h,
                          int grid, int lds_bytes, void **graph_exec) {
  if (!graph_exec) return (int)hipErrorInvalidValue;
  *graph_exec = nullptr;
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return (int)e;
  e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    const int rc = mofft_hip_batch_launch(precision_bits, h, grid, lds_bytes, s);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    e = rc != 0 ? (hipError_t)rc : ec;
    if (e == hipSuccess) {
      hipGraphExec_t x = nullptr;
      e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      if (e == hipSuccess) *graph_exec = (void *)x;
    }
    if (g) (void)hipGraphDestroy(g);
  }
  (void)hipStreamDestroy(s);
  return (int)e;
}

int mofft_hip_capture_begin(void **stream) {
  if (!stream) return (int)hipErrorInvalidValue;
  *stream = nullptr;
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return (int)e;
  e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(s);
    return (int)e;
  }
  *stream = (void *)s;
  return 0;
}

int mofft_hip_capture_end(void *stream, void **graph_exec) {
  if (!stream || !graph_exec) return (int)hipErrorInvalidValue;
  *graph_exec = nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
  if (e == hipSuccess) {
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e == hipSuccess) *graph_exec = (void *)x;
  }
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy((hipStream_t)stream);
  return (int)e;
}

int mofft_hip_graph_launch(void *graph_exec, void *stream) {
  if (!graph_exec) return (int)hipErrorInvalidValue;
  return (int)hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream);
}

int mofft_hip_graph_destroy(void *graph_exec) {
  if (!graph_exec) return 0;
  return (int)hipGraphExecDestroy((hipGraphExec_t)graph_exec);
}

int mofft_hip_stream_create(void **stream) {
  if (!stream) return (int)hipErrorInvalidValue;
  hipStream_t s = nullptr;
  const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  *stream = e == hipSuccess ? (void *)s : nullptr;
  return (int)e;
}

int mofft_hip_stream_destroy(void *stream) {
  if (!stream) return 0;
  return (int)hipStreamDestroy((hipStream_t)stream);
}

int mofft_hip_event_create(void **event) {
  if (!event) return (int)hipErrorInvalidValue;
  hipEvent_t ev = nullptr;
  const hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  *event = e == hipSuccess ? (void *)ev : nullptr;
  return (int)e;
}

int mofft_hip_timing_event_create(void **event) {
  if (!event) return (int)hipErrorInvalidValue;
  hipEvent_t ev = nullptr;
  const hipError_t e = hipEventCreate(&ev);
  *event = e == hipSuccess ? (void *)ev : nullptr;
  return (int)e;
}

int mofft_hip_event_elapsed(void *start, void *end, double *seconds) {
  if (!seconds) return (int)hipErrorInvalidValue;
  float ms = 0.f;
  const hipError_t e = hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)end);
  *seconds = (double)ms * 1e-3;
  return (int)e;
}

int mofft_hip_event_destroy(void *event) {
  if (!event) return 0;
  return (int)hipEventDestroy((hipEvent_t)event);
}

int mofft_hip_event_record(void *event, void *stream) {
  return (int)hipEventRecord((hipEvent_t)event, (hipStream_t)stream);
}

int mofft_hip_event_query(void *event) {
  const hipError_t e = hipEventQuery((hipEvent_t)event);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return -(int)e;
}

int mofft_hip_event_synchronize(void *event) {
  return (int)hipEventSynchronize((hipEvent_t)event);
}

int mofft_hip_fill_uniform(void *out, int precision_bits, int64_t count,
                           uint64_t seed, void *stream) {
  if (count <= 0) return 0;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (precision_bits == 32)
    hipLaunchKernelGGL(mofft::fill_uniform_kernel<float>, dim3((unsigned)blocks),
                       dim3(256), 0, (hipStream_t)stream, (float *)out, count, seed);
  else
    hipLaunchKernelGGL(mofft::fill_uniform_kernel<double>, dim3((unsigned)blocks),
                       dim3(256), 0, (hipStream_t)stream, (double *)out, count, seed);
  return (int)hipGetLastError();
}

}  // extern "C"
