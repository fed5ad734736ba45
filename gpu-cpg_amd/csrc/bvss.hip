// BVSS A_dev discovery: which deviation edges u->v (v != succ[u]) leave a
// frontier of source vertices.
//
// The reference packs, per 8-source interval, the destination rows that
// have a deviation edge from that interval with an 8-bit source mask, and
// tests masks against the frontier byte with binary tensor-core products
// (gpucpg/tc_pfxt_bvss.cuh:117-176 build, :289-439 discovery; unit answers
// unittests/tc_pfxt_bvss.cu:9-72).  This engine's layout is sized for
// wave64: an interval covers 64 sources, one 64-bit word per (interval,
// destination row), rows grouped by interval (CSR over intervals), and a
// frontier is one 64-bit word per interval.  Discovery runs up to F
// frontiers at once and emits (frontier, u, v) triples:
//   bitmap path   one lane per (row entry, frontier): AND of the two words,
//                 one triple per set bit
//   MFMA path     one wave per 16-row x 16-frontier tile of an interval:
//                 the 0/1 int8 product rows(16 x 64) x frontiers(64 x 16)
//                 with int32 accumulation (v_mfma_i32_16x16x64_i8) counts
//                 the shared sources of every (row, frontier) pair, and the
//                 pairs with a count above zero are decoded from the words
// An interval takes the MFMA path when its frontier block is dense enough
// for the product to be a real contraction (mode 0), or always (mode 2);
// mode 1 uses the bitmap path only.  Both paths give the same triples.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace cpg {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kBvssBlock = 256;

// Spreads the 4 bits of x over the 4 bytes of a word (bit i -> byte i).
__device__ inline int nibble_bytes(uint32_t x) {
  return static_cast<int>((x & 1u) | ((x & 2u) << 7) | ((x & 4u) << 14) | ((x & 8u) << 21));
}

// 16 bits of w starting at bit `lo`, as 16 bytes of 0/1.
__device__ inline v4i bits_to_bytes(uint64_t w, int lo) {
  const uint32_t x = static_cast<uint32_t>(w >> lo) & 0xffffu;
  v4i r;
  r.x = nibble_bytes(x & 0xfu);
  r.y = nibble_bytes((x >> 4) & 0xfu);
  r.z = nibble_bytes((x >> 8) & 0xfu);
  r.w = nibble_bytes((x >> 12) & 0xfu);
  return r;
}

// Exclusive wave prefix sum of v; *total receives the wave's sum.
__device__ inline unsigned wave_excl_scan(unsigned v, unsigned* total) {
  unsigned incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned x = __shfl_up(incl, o);
    if (__lane_id() >= o) incl += x;
  }
  *total = __shfl(incl, 63);
  return incl - v;
}

}  // namespace

__global__ void __launch_bounds__(kBvssBlock) bvss_frontier_kernel(BvssArgs a, const int* fptr,
                                                                   const int* fsrc) {
  const int f = blockIdx.y;
  const int b = fptr[f], e = fptr[f + 1];
  for (int i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
    const int u = fsrc[i];
    if (u < 0 || u >= a.n) continue;
    atomicOr(&a.front[static_cast<size_t>(f) * a.n_int + (u >> 6)], 1ull << (u & 63));
  }
}

// Marks the intervals that take the MFMA path: the frontier block (64
// sources x nfront frontiers) holds at least dense_min set bits per
// frontier on average (mode 0), every interval (mode 2) or none (mode 1).
__global__ void __launch_bounds__(kBvssBlock) bvss_density_kernel(BvssArgs a, int mode,
                                                                  int dense_min) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n_int; i += gridDim.x * blockDim.x) {
    int bits = 0;
    for (int f = 0; f < a.nfront; ++f) bits += __popcll(a.front[static_cast<size_t>(f) * a.n_int + i]);
    uint8_t d = 0;
    if (mode == 2) d = 1;
    else if (mode == 0) d = bits >= dense_min * a.nfront ? 1 : 0;
    a.dense[i] = d;
  }
}

// Bitmap path: one lane per (row entry, frontier) of the sparse intervals.
__global__ void __launch_bounds__(kBvssBlock) bvss_bitmap_kernel(BvssArgs a) {
  const uint64_t total = static_cast<uint64_t>(a.n_entries) * a.nfront;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x; base < total;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t t = base + threadIdx.x;
    uint64_t hits = 0;
    int row = 0, f = 0, iv = 0;
    if (t < total) {
      const int e = static_cast<int>(t % a.n_entries);
      f = static_cast<int>(t / a.n_entries);
      iv = a.entry_int[e];
      if (!a.dense[iv]) {
        row = a.rows[e];
        hits = a.masks[e] & a.front[static_cast<size_t>(f) * a.n_int + iv];
      }
    }
    unsigned wtot = 0;
    const unsigned off = wave_excl_scan(static_cast<unsigned>(__popcll(hits)), &wtot);
    unsigned long long wbase = 0;
    if (__lane_id() == 0 && wtot) wbase = atomicAdd(&a.ctl->count, static_cast<unsigned long long>(wtot));
    wbase = __shfl(wbase, 0);
    uint64_t slot = wbase + off;
    while (hits) {
      const int bit = __ffsll(static_cast<long long>(hits)) - 1;
      if (slot < a.cap) a.out[slot] = make_int4(f, iv * 64 + bit, row, 0);
      ++slot;
      hits &= hits - 1;
    }
  }
}

// MFMA path: one wave per (tile of 16 rows of a dense interval, group of 16
// frontiers).  Lane l supplies row l % 16 and frontier l % 16 over sources
// 16 (l / 16) .. 16 (l / 16) + 15 as 0/1 bytes; the accumulator of lane l
// holds the shared-source counts of rows 4 (l / 16) + i (i = 0..3) with
// frontier l % 16.
__global__ void __launch_bounds__(kBvssBlock) bvss_mfma_kernel(BvssArgs a) {
  const int lane = __lane_id();
  const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = static_cast<int>((gridDim.x * blockDim.x) >> 6);
  const int fgroups = (a.nfront + 15) / 16;
  const int64_t work = static_cast<int64_t>(a.n_tiles) * fgroups;
  for (int64_t w = wave; w < work; w += nwaves) {
    const int tile = static_cast<int>(w % a.n_tiles);
    const int fg = static_cast<int>(w / a.n_tiles);
    const int2 td = a.tiles[tile];  // {interval, first row entry}
    const int iv = td.x;
    if (!a.dense[iv]) continue;  // wave-uniform
    const int e0 = td.y, e_end = a.int_ptr[iv + 1];
    const int kg = lane >> 4, c = lane & 15;
    const int ea = e0 + c;
    const uint64_t mrow = ea < e_end ? a.masks[ea] : 0ull;
    const int fc = fg * 16 + c;
    const uint64_t fword = fc < a.nfront ? a.front[static_cast<size_t>(fc) * a.n_int + iv] : 0ull;
    v4i acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(bits_to_bytes(mrow, kg * 16),
                                               bits_to_bytes(fword, kg * 16), acc, 0, 0, 0);
    // decode: this lane owns rows 4 kg + i with frontier c
    uint64_t hits[4];
    int rowid[4];
    unsigned cnt = 0;
    const int counts[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      hits[i] = 0;
      rowid[i] = 0;
      const int er = e0 + 4 * kg + i;
      if (counts[i] > 0 && er < e_end && fc < a.nfront) {
        hits[i] = a.masks[er] & fword;
        rowid[i] = a.rows[er];
        if (__popcll(hits[i]) != counts[i]) atomicExch(&a.ctl->mismatch, 1);
        cnt += static_cast<unsigned>(__popcll(hits[i]));
      }
    }
    unsigned wtot = 0;
    const unsigned off = wave_excl_scan(cnt, &wtot);
    unsigned long long wbase = 0;
    if (lane == 0 && wtot) wbase = atomicAdd(&a.ctl->count, static_cast<unsigned long long>(wtot));
    wbase = __shfl(wbase, 0);
    if (lane == 0) atomicAdd(&a.ctl->tiles, 1ull);
    uint64_t slot = wbase + off;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      uint64_t h = hits[i];
      while (h) {
        const int bit = __ffsll(static_cast<long long>(h)) - 1;
        if (slot < a.cap) a.out[slot] = make_int4(fc, iv * 64 + bit, rowid[i], 0);
        ++slot;
        h &= h - 1;
      }
    }
  }
}

void launch_bvss_discover(const BvssArgs& a, const int* fptr, const int* fsrc, int mode,
                          int dense_min, int grid, hipStream_t s) {
  const int fb = std::max(1, std::min(64, grid / std::max(1, a.nfront)));
  hipLaunchKernelGGL(bvss_frontier_kernel, dim3(fb, a.nfront), dim3(kBvssBlock), 0, s, a, fptr,
                     fsrc);
  CPG_HIP(hipGetLastError());
  const int ig = std::max(1, std::min(grid, (a.n_int + kBvssBlock - 1) / kBvssBlock));
  hipLaunchKernelGGL(bvss_density_kernel, dim3(ig), dim3(kBvssBlock), 0, s, a, mode, dense_min);
  CPG_HIP(hipGetLastError());
  if (a.n_entries > 0) {
    hipLaunchKernelGGL(bvss_bitmap_kernel, dim3(grid), dim3(kBvssBlock), 0, s, a);
    CPG_HIP(hipGetLastError());
  }
  if (a.n_tiles > 0) {
    hipLaunchKernelGGL(bvss_mfma_kernel, dim3(grid), dim3(kBvssBlock), 0, s, a);
    CPG_HIP(hipGetLastError());
  }
}

}  // namespace cpg
