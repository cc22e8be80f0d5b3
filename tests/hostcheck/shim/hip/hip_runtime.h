// CPU stand-in for the part of the HIP runtime the library uses, so that
// every kernel and launcher of jax_bssn_amd/csrc compiles as ordinary C++
// and runs on host threads (test infrastructure; see cpu_build.py).
//
// A launch runs the grid's blocks one after another; each block runs its
// blockDim threads as real threads, with __syncthreads() as a barrier of
// the block and __shfl_xor() as an exchange inside each 64-lane wave.
// __shared__ arrays become function statics, shared by the threads of the
// running block.  Device pointers are host pointers.  Built with
// AddressSanitizer, every global and LDS access is bounds-checked.
#pragma once

#include <barrier>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

#define __host__
#define __device__
#define __global__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define __builtin_amdgcn_fence(...) ((void)0)
#define __builtin_amdgcn_s_barrier() hipcpu::block_sync()

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace hipcpu {
struct Idx {
  unsigned x = 0, y = 0, z = 0;
};
extern thread_local Idx thread_idx, block_idx;
extern dim3 block_dim, grid_dim;
void block_sync();
unsigned long long wave_exchange(unsigned long long v, int src_lane_xor);
void launch(const std::function<void()> &body, dim3 grid, dim3 block);
}  // namespace hipcpu

#define threadIdx (hipcpu::thread_idx)
#define blockIdx (hipcpu::block_idx)
#define blockDim (hipcpu::block_dim)
#define gridDim (hipcpu::grid_dim)

inline void __syncthreads() { hipcpu::block_sync(); }

template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  static_assert(sizeof(T) <= 8, "shuffle of a register value");
  (void)width;
  unsigned long long b = 0;
  std::memcpy(&b, &v, sizeof(T));
  b = hipcpu::wave_exchange(b, mask);
  T r;
  std::memcpy(&r, &b, sizeof(T));
  return r;
}

inline long long __double_as_longlong(double x) {
  long long r;
  std::memcpy(&r, &x, 8);
  return r;
}
inline double __longlong_as_double(long long x) {
  double r;
  std::memcpy(&r, &x, 8);
  return r;
}
inline int __float_as_int(float x) {
  int r;
  std::memcpy(&r, &x, 4);
  return r;
}
inline float __int_as_float(int x) {
  float r;
  std::memcpy(&r, &x, 4);
  return r;
}

typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;
typedef void *hipStream_t;
typedef struct hipcpu_event *hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };

struct hipcpu_event {
  std::chrono::steady_clock::time_point t;
};

inline const char *hipGetErrorString(hipError_t) { return "hip (cpu build) error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t = nullptr) {
  std::memmove(dst, src, n);
  return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new hipcpu_event;
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  *p = const_cast<void *>(sym);
  return hipSuccess;
}

#define hipLaunchKernelGGL(K, G, B, S, ST, ...) hipcpu::launch([&]() { K(__VA_ARGS__); }, dim3(G), dim3(B))
