// CPU stand-in for <hip/hip_runtime.h>, TEST INFRASTRUCTURE ONLY.
//
// Lets the engine's unchanged HIP sources (gpu-cpg_amd/csrc) compile with a
// host C++ compiler and run on the CPU: every GPU thread is a fiber, the
// 64 fibers of a wave meet at each wave collective (__ballot, __shfl*),
// the fibers of a block meet at __syncthreads, and the blocks of a
// cooperative launch run concurrently on OS threads.  A collective that
// not every live lane of its wave reaches is reported as a deadlock.
// This build is loaded only by CPU tests; the product library is the HIP
// build.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <type_traits>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static thread_local
#define __launch_bounds__(...)
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 0

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
struct uint2 { unsigned x, y; };
struct uint4 { unsigned x, y, z, w; };
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }

namespace simt {

dim3 thread_idx();
dim3 block_idx();
dim3 block_dim();
dim3 grid_dim();
int lane_id();
uint64_t* wave_sync(uint64_t value);
uint64_t wave_present();
void block_sync();
void relax();

template <class F, class... A>
void launch(F kernel, dim3 grid, dim3 block, bool cooperative, A... args);
}  // namespace simt

#define threadIdx (simt::thread_idx())
#define blockIdx (simt::block_idx())
#define blockDim (simt::block_dim())
#define gridDim (simt::grid_dim())

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }

inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
inline unsigned __float_as_uint(float f) { unsigned i; std::memcpy(&i, &f, 4); return i; }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned long long atomicOr(unsigned long long* p, unsigned long long v) {
  return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __lane_id() { return simt::lane_id(); }
inline uint64_t __lanemask_lt() {
  const int l = simt::lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

inline uint64_t __ballot(int pred) {
  const uint64_t* s = simt::wave_sync(pred ? 1u : 0u);
  const uint64_t present = simt::wave_present();
  uint64_t m = 0;
  for (int i = 0; i < 64; ++i)
    if (((present >> i) & 1u) && s[i]) m |= 1ull << i;
  return m;
}

template <class T>
inline T simt_from_bits(uint64_t b) {
  T v;
  std::memcpy(&v, &b, sizeof(T));
  return v;
}
template <class T>
inline uint64_t simt_to_bits(T v) {
  uint64_t b = 0;
  std::memcpy(&b, &v, sizeof(T));
  return b;
}
template <class T>
inline T __shfl(T v, int src, int = 64) {
  const uint64_t* s = simt::wave_sync(simt_to_bits(v));
  return simt_from_bits<T>(s[src & 63]);
}
template <class T>
inline T __shfl_up(T v, unsigned d, int = 64) {
  const uint64_t* s = simt::wave_sync(simt_to_bits(v));
  const int l = simt::lane_id();
  return l >= static_cast<int>(d) ? simt_from_bits<T>(s[l - d]) : v;
}
template <class T>
inline T __shfl_xor(T v, int m, int = 64) {
  const uint64_t* s = simt::wave_sync(simt_to_bits(v));
  return simt_from_bits<T>(s[(simt::lane_id() ^ m) & 63]);
}
inline void __syncthreads() { simt::block_sync(); }

// v_mfma_i32_16x16x64_i8 as documented for CDNA wave64: lane l holds A
// row l % 16 and B column l % 16, bytes 0..15 over k = 16 (l / 16) + j;
// the result of lane l is D[4 (l / 16) + i][l % 16], i = 0..3.
typedef int simt_v4i __attribute__((ext_vector_type(4)));
inline simt_v4i __builtin_amdgcn_mfma_i32_16x16x64_i8(simt_v4i a, simt_v4i b, simt_v4i c, int, int,
                                                      int) {
  uint64_t part[4] = {
      (static_cast<uint64_t>(static_cast<uint32_t>(a.y)) << 32) | static_cast<uint32_t>(a.x),
      (static_cast<uint64_t>(static_cast<uint32_t>(a.w)) << 32) | static_cast<uint32_t>(a.z),
      (static_cast<uint64_t>(static_cast<uint32_t>(b.y)) << 32) | static_cast<uint32_t>(b.x),
      (static_cast<uint64_t>(static_cast<uint32_t>(b.w)) << 32) | static_cast<uint32_t>(b.z)};
  int8_t A[64][16], B[64][16];
  for (int p = 0; p < 4; ++p) {
    const uint64_t* all = simt::wave_sync(part[p]);
    for (int l = 0; l < 64; ++l) {
      for (int j = 0; j < 8; ++j) {
        const int8_t v = static_cast<int8_t>((all[l] >> (8 * j)) & 0xff);
        if (p == 0) A[l][j] = v;
        else if (p == 1) A[l][8 + j] = v;
        else if (p == 2) B[l][j] = v;
        else B[l][8 + j] = v;
      }
    }
  }
  const int l = simt::lane_id();
  simt_v4i d = c;
  for (int i = 0; i < 4; ++i) {
    const int r = 4 * (l / 16) + i, col = l % 16;
    int sum = 0;
    for (int g = 0; g < 4; ++g)
      for (int j = 0; j < 16; ++j) sum += static_cast<int>(A[g * 16 + r][j]) * static_cast<int>(B[g * 16 + col][j]);
    d[i] += sum;
  }
  return d;
}
// every live lane of the wave meets here, as the wave does in lockstep
inline void __builtin_amdgcn_wave_barrier() { (void)simt::wave_sync(0); }

template <class T>
inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline float atomicAdd(float* p, float v) {
  float old = *p;
  for (;;) {
    float nv = old + v;
    if (__atomic_compare_exchange(p, &old, &nv, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) break;
  }
  return old;
}
template <class T>
inline T atomicSub(T* p, T v) { return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicExch(T* p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicMax(T* p, T v) {
  T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST,
                                                 __ATOMIC_SEQ_CST)) {}
  return old;
}
template <class T>
inline T atomicMin(T* p, T v) {
  T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (old > v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST,
                                                 __ATOMIC_SEQ_CST)) {}
  return old;
}
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), (order))
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), (order))
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add((p), (v), (order))
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(order)
#define __builtin_amdgcn_s_sleep(n) simt::relax()
// a __device__ variable is host memory here: its address is its own
#define HIP_SYMBOL(x) (&(x))
#define __align__(n) __attribute__((aligned(n)))
// waiting for a thread's outstanding memory operations: on the CPU they are
// done once issued, so only the compiler's ordering is kept
#define __builtin_amdgcn_s_waitcnt(n) __atomic_thread_fence(__ATOMIC_SEQ_CST)

// ---------------------------------------------------------------- host API
typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;
constexpr hipError_t hipErrorNotReady = 600;
constexpr unsigned hipEventDisableTiming = 2;
typedef struct simt_stream* hipStream_t;
typedef struct simt_event* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum hipDeviceAttribute_t {
  hipDeviceAttributeMultiprocessorCount = 1,
  hipDeviceAttributeCooperativeLaunch = 2
};
constexpr unsigned hipStreamNonBlocking = 1;

const char* hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipGetDevice(int* d);
inline hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int dev);
hipError_t simt_malloc(void** p, size_t bytes);
template <class T>
hipError_t hipMalloc(T** p, size_t bytes) {
  void* q = nullptr;
  const hipError_t e = simt_malloc(&q, bytes);
  *p = static_cast<T*>(q);
  return e;
}
hipError_t hipFree(void* p);
// Pinned host memory is ordinary host memory here.
constexpr unsigned hipHostMallocDefault = 0;
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
  *p = std::malloc(bytes ? bytes : 1);
  return *p ? hipSuccess : 2;
}
inline hipError_t hipHostFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k, hipStream_t s);
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemsetAsync(void* p, int v, size_t bytes, hipStream_t s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
// launches run to completion when made, so every recorded event has fired
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
typedef struct simt_graph* hipGraph_t;
typedef struct simt_graph* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);

template <class F>
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, F, int, size_t) {
  *n = 1;
  return hipSuccess;
}

namespace simt {
template <class... A, size_t... I>
void launch_params(void (*f)(A...), dim3 g, dim3 b, void** p, std::index_sequence<I...>) {
  launch(f, g, b, true, *static_cast<std::decay_t<A>*>(p[I])...);
}
}  // namespace simt

template <class... A>
hipError_t hipLaunchCooperativeKernel(void (*f)(A...), dim3 g, dim3 b, void** params, unsigned,
                                      hipStream_t) {
  simt::launch_params(f, g, b, params, std::index_sequence_for<A...>{});
  return hipSuccess;
}

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  simt::launch(kernel, grid, block, false, __VA_ARGS__)

#include "simt_launch.h"
