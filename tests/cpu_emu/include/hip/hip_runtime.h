// CPU emulation of the subset of the HIP runtime used by libhsba, for
// bounds-checked (AddressSanitizer) runs of the real kernel sources on the
// host.  Test infrastructure only.
//
// Model: blocks run one after another; the threads of a block run as OS
// threads of a persistent pool.  __syncthreads() is a block barrier, a
// shuffle exchanges through a per-wave buffer between two wave barriers, and
// a thread leaving the kernel drops out of both barriers.  __shared__ is
// `static` (one block at a time).  Launches issued during stream capture are
// recorded and replayed by hipGraphLaunch.
#pragma once

#include <barrier>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

namespace hipemu {
extern thread_local dim3 threadIdx_, blockIdx_;
extern dim3 blockDim_, gridDim_;
void syncthreads();
void wave_exchange(const void *in, void *out, int bytes, int src_lane);
void wave_sync();
void wave_gather(const void *in, void *all, int bytes);
int lane_id();
using Job = std::function<void()>;
void run_kernel(dim3 grid, dim3 block, const std::function<void()> &body);
extern std::vector<Job> *capture;
}  // namespace hipemu

#define threadIdx (hipemu::threadIdx_)
#define blockIdx (hipemu::blockIdx_)
#define blockDim (hipemu::blockDim_)
#define gridDim (hipemu::gridDim_)

inline void __syncthreads() { hipemu::syncthreads(); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }

template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  T out;
  const int lane = hipemu::lane_id();
  const int base = lane & ~(width - 1);
  hipemu::wave_exchange(&v, &out, (int)sizeof(T), base | ((lane - base) ^ mask));
  return out;
}

using std::fabs;
using std::fmax;
using std::sqrt;
inline int min(int a, int b) { return a < b ? a : b; }
inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int max(int a, int b) { return a > b ? a : b; }
inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
// v_mfma_f64_16x16x4_f64: A[i][k] from lane i + 16 k, B[k][j] from lane
// j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4.
typedef double hipemu_dbl4 __attribute__((ext_vector_type(4)));
inline hipemu_dbl4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, hipemu_dbl4 c, int, int, int) {
  double A[64], B[64];
  hipemu::wave_gather(&a, A, 8);
  hipemu::wave_gather(&b, B, 8);
  const int l = hipemu::lane_id() & 63, j = l & 15, i0 = l >> 4;
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k) c[r] += A[i0 + 4 * r + 16 * k] * B[j + 16 * k];
  return c;
}
// v_mfma_f32_16x16x4f32: operands as above; D[i][j] in lane j + 16 (i / 4),
// element i % 4.
typedef float hipemu_flt4 __attribute__((ext_vector_type(4)));
inline hipemu_flt4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, hipemu_flt4 c, int, int, int) {
  float A[64], B[64];
  hipemu::wave_gather(&a, A, 4);
  hipemu::wave_gather(&b, B, 4);
  const int l = hipemu::lane_id() & 63, j = l & 15, i0 = 4 * (l >> 4);
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k) c[r] += A[i0 + r + 16 * k] * B[j + 16 * k];
  return c;
}
// v_readlane: the value of `lane` of this thread's wave.
inline int __builtin_amdgcn_readlane(int v, int lane) {
  int out;
  hipemu::wave_exchange(&v, &out, 4, (hipemu::lane_id() & ~63) | lane);
  return out;
}
// A wave barrier is a real barrier of the wave's 64 threads here.
inline void __builtin_amdgcn_wave_barrier() { hipemu::wave_sync(); }
inline void __builtin_amdgcn_fence(int, const char *) { std::atomic_thread_fence(std::memory_order_seq_cst); }

typedef enum hipError_t { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 } hipError_t;
typedef struct ihipStream_t *hipStream_t;
struct hipemuGraph { std::vector<hipemu::Job> jobs; };
typedef hipemuGraph *hipGraph_t;
typedef hipemuGraph *hipGraphExec_t;
typedef int hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
#define hipStreamNonBlocking 1

inline const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "success" : "emulated error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
inline hipError_t hipMalloc(void **p, size_t bytes) {
  *p = std::malloc(bytes ? bytes : 1);
  if (!*p) return hipErrorOutOfMemory;
  std::memset(*p, 0xFF, bytes);  // uninitialised device memory reads as NaN / -1
  return hipSuccess;
}
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void **p, size_t bytes, unsigned = 0) {
  *p = std::malloc(bytes ? bytes : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t) {
  if (hipemu::capture) { std::fprintf(stderr, "hipemu: memcpy during capture\n"); std::abort(); }
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) {
  if (hipemu::capture) { std::fprintf(stderr, "hipemu: memset during capture\n"); std::abort(); }
  std::memset(d, v, n);
  return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (hipStream_t)0x1; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  hipemu::capture = new std::vector<hipemu::Job>();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
  *g = new hipemuGraph();
  (*g)->jobs = std::move(*hipemu::capture);
  delete hipemu::capture;
  hipemu::capture = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t) {
  *e = new hipemuGraph(*g);
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto &j : e->jobs) j();
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }

#include <chrono>
typedef struct hipemuEvent { double t; } *hipEvent_t;
inline double hipemu_now() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new hipemuEvent{0.0}; return hipSuccess; }
#define hipEventDisableTiming 2
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
// every stream is the one in-order queue of the emulation: waits are no-ops
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  if (hipemu::capture) hipemu::capture->push_back([e]() { e->t = hipemu_now(); });
  else e->t = hipemu_now();
  return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(b->t - a->t); return hipSuccess; }

template <typename K, typename... Args>
void hipemu_launch(K kernel, dim3 grid, dim3 block, Args... args) {
  hipemu::Job job = [=]() { hipemu::run_kernel(grid, block, [=]() { kernel(args...); }); };
  if (hipemu::capture) hipemu::capture->push_back(job);
  else job();
}
#define hipLaunchKernelGGL(k, g, b, shm, st, ...) hipemu_launch(k, g, b, __VA_ARGS__)
