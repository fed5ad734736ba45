// Shared definitions for the CPG engine's HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace cpg {

constexpr int kScaleUp = 10000;             // gpucpg/gpucpg.cuh:35
constexpr int kIntMax = 2147483647;
constexpr uint32_t kExpandedBit = 0x80000000u;  // live-entry flag: parent already expanded
constexpr uint32_t kKeyInf = 0xFF800000u;       // orderable key of +inf ("no bound")
constexpr uint32_t kKeyPosZero = 0x80000000u;   // key of +0.0f
constexpr uint32_t kKeyNegZero = 0x7FFFFFFFu;   // key of -0.0f
constexpr int kHistBins = 4096;                 // threshold histograms

#define CPG_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                               " at " __FILE__ ":" + std::to_string(__LINE__)); \
  } while (0)

// Total order on float bits that matches float '<' on non-NaN values.
__host__ __device__ inline uint32_t float_key(float f) {
  const uint32_t b = __builtin_bit_cast(uint32_t, f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float key_float(uint32_t k) {
  if (k == kKeyInf) return __builtin_huge_valf();
  const uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  return __builtin_bit_cast(float, b);
}
__host__ __device__ inline bool key_is_zero(uint32_t k) {
  return k == kKeyPosZero || k == kKeyNegZero;
}

// Histogram bin of a key for the window [klo, klo + (kHistBins << shift)).
__host__ __device__ inline uint32_t key_bin(uint32_t k, uint32_t klo, uint32_t shift) {
  if (k < klo) return 0;
  const uint32_t b = (k - klo) >> shift;
  return b < static_cast<uint32_t>(kHistBins) ? b : static_cast<uint32_t>(kHistBins - 1);
}
// Largest key of bin b (saturating).
__host__ __device__ inline uint32_t bin_upper_key(uint32_t b, uint32_t klo, uint32_t shift) {
  const uint64_t u = static_cast<uint64_t>(klo) + ((static_cast<uint64_t>(b) + 1) << shift) - 1;
  return u > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(u);
}
// Smallest shift whose window of `bins` bins covers [klo, khi].
__host__ __device__ inline uint32_t window_shift_bins(uint32_t klo, uint32_t khi, int bins) {
  uint32_t s = 0;
  const uint64_t span = khi >= klo ? static_cast<uint64_t>(khi - klo) : 0;
  while ((span >> s) >= static_cast<uint64_t>(bins)) ++s;
  return s;
}
__host__ __device__ inline uint32_t window_shift(uint32_t klo, uint32_t khi) {
  return window_shift_bins(klo, khi, kHistBins);
}

__device__ inline uint64_t lanemask_lt() { return __lanemask_lt(); }
// Wave-aggregated reservation of one slot per lane with want != 0 on a
// 64-bit counter; returns the lane's slot (undefined where want == 0).
__device__ inline uint64_t wave_reserve(unsigned long long* counter, bool want) {
  const uint64_t mask = __ballot(want);
  if (mask == 0) return 0;
  const int leader = __ffsll(static_cast<long long>(mask)) - 1;
  unsigned long long base = 0;
  if (__lane_id() == leader) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)));
  base = __shfl(base, leader);
  return base + __popcll(mask & lanemask_lt());
}

// Block-wide reservation: each thread asks for `want` consecutive slots on
// a 64-bit counter with one atomic per block; returns the thread's first
// slot.  Every thread of the block must call it (it contains barriers).
template <int kThreads>
__device__ inline uint64_t block_reserve(unsigned long long* counter, unsigned want) {
  static_assert(kThreads % 64 == 0, "block must be whole waves");
  constexpr int kW = kThreads / 64;
  __shared__ unsigned s_wave[kW];
  __shared__ unsigned long long s_base;
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  unsigned incl = want;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kW; ++w) {
      const unsigned c = s_wave[w];
      s_wave[w] = static_cast<unsigned>(tot);
      tot += c;
    }
    s_base = tot ? atomicAdd(counter, tot) : 0ull;
  }
  __syncthreads();
  const uint64_t r = s_base + s_wave[wave] + (incl - want);
  __syncthreads();
  return r;
}

}  // namespace cpg
