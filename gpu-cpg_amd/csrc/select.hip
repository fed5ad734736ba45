// Final selection: stable LSD radix sort of (cost key, level) -> node and
// assembly of the ranked report.
//
// The reference sorts its whole short pile with thrust::sort on slack
// (gpucpg/gpucpg.cu:21323-21331), erases zero-cost entries (:21338-21358)
// and returns the first K (get_slacks :21409-21420).  Here the sort input
// is every stored path with a nonzero cost at or below the K-th bound, and
// the sort key is (orderable cost bits, prefix-tree level) so that a
// parent always ranks before a child of equal cost.
#include "common.hpp"
#include "kernels.hpp"

namespace cpg {

namespace {

constexpr int kSortBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kSortBlock * kItems;
constexpr int kWaves = kSortBlock / 64;
constexpr int kScanBlock = 1024;

__global__ void __launch_bounds__(kSortBlock) sort_hist_kernel(const uint64_t* key, uint64_t n,
                                                               int shift, unsigned* ghist,
                                                               uint64_t ntiles) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t t = blockIdx.x;
  for (int s = 0; s < kItems; ++s) {
    const uint64_t i = t * kTile + static_cast<uint64_t>(s) * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(key[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  ghist[static_cast<uint64_t>(threadIdx.x) * ntiles + t] = h[threadIdx.x];
}

// Exclusive scan of ghist[0..len) in place, one block: chunks of
// kScanBlock * 4 counts, each thread taking four consecutive ones, a wave
// scan by shuffles, a scan of the wave totals, and a carry across chunks.
__global__ void __launch_bounds__(kScanBlock) sort_scan_kernel(unsigned* ghist, uint64_t len) {
  constexpr int kW = kScanBlock / 64;
  __shared__ unsigned s_wave[kW];
  __shared__ unsigned long long s_carry;
  const int t = threadIdx.x, lane = __lane_id(), wave = t >> 6;
  if (t == 0) s_carry = 0;
  __syncthreads();
  for (uint64_t c = 0; c < len; c += static_cast<uint64_t>(kScanBlock) * 4) {
    const uint64_t i0 = c + static_cast<uint64_t>(t) * 4;
    unsigned v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = i0 + q < len ? ghist[i0 + q] : 0u;
    const unsigned mine = v[0] + v[1] + v[2] + v[3];
    unsigned incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned x = __shfl_up(incl, o);
      if (lane >= o) incl += x;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
      unsigned w = lane < kW ? s_wave[lane] : 0u;
      for (int o = 1; o < kW; o <<= 1) {
        const unsigned x = __shfl_up(w, o);
        if (lane >= o) w += x;
      }
      if (lane < kW) s_wave[lane] = w;  // inclusive wave totals
    }
    __syncthreads();
    const unsigned long long carry = s_carry;
    unsigned long long run = carry + (wave ? s_wave[wave - 1] : 0u) + (incl - mine);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (i0 + q < len) ghist[i0 + q] = static_cast<unsigned>(run);
      run += v[q];
    }
    __syncthreads();
    if (t == kScanBlock - 1) s_carry = carry + s_wave[kW - 1];
    __syncthreads();
  }
}

constexpr int kRowBlock = 256;
static_assert(kRowBlock == kSortBlock, "the scatter kernel scans the digit totals with the row scan's helper");

// Block-wide exclusive scan of one value per thread (kRowBlock threads);
// *total receives the sum.  Every thread of the block must call it.
__device__ inline unsigned block_exclusive_scan(unsigned v, unsigned* total) {
  __shared__ unsigned s_w[kRowBlock / 64];
  const int lane = __lane_id(), wave = static_cast<int>(threadIdx.x) >> 6;
  unsigned incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned x = __shfl_up(incl, o);
    if (lane >= o) incl += x;
  }
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  unsigned before = 0, sum = 0;
  for (int w = 0; w < kRowBlock / 64; ++w) {
    const unsigned t = s_w[w];
    if (w < wave) before += t;
    sum += t;
  }
  __syncthreads();  // s_w is written again by the next call
  *total = sum;
  return before + incl - v;
}

// Exclusive scan of each digit's row of per-tile counts, ghist[d * ntiles
// + t], in place, one block per digit; the row's total goes to
// ghist[256 * ntiles + d].  The scatter kernel adds the digits' bases.
__global__ void __launch_bounds__(kRowBlock) sort_rowscan_kernel(unsigned* ghist, uint64_t ntiles) {
  unsigned* row = ghist + static_cast<uint64_t>(blockIdx.x) * ntiles;
  unsigned carry = 0;
  for (uint64_t c = 0; c < ntiles; c += static_cast<uint64_t>(kRowBlock) * 4) {
    const uint64_t i0 = c + static_cast<uint64_t>(threadIdx.x) * 4;
    unsigned v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = i0 + q < ntiles ? row[i0 + q] : 0u;
    unsigned tot = 0;
    unsigned run = carry + block_exclusive_scan(v[0] + v[1] + v[2] + v[3], &tot);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (i0 + q < ntiles) row[i0 + q] = run;
      run += v[q];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) ghist[256 * ntiles + blockIdx.x] = carry;
}

// goff: the row-scanned counts and, after them, the 256 digit totals.
__global__ void __launch_bounds__(kSortBlock) sort_scatter_kernel(const uint64_t* kin, const uint32_t* vin,
                                                                  uint64_t* kout, uint32_t* vout,
                                                                  uint64_t n, int shift,
                                                                  const unsigned* goff, uint64_t ntiles) {
  __shared__ unsigned wcnt[kWaves][256];
  __shared__ unsigned run[256];
  __shared__ unsigned base[256];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const uint64_t t = blockIdx.x;
  run[tid] = 0;
  unsigned all = 0;
  const unsigned digit_base = block_exclusive_scan(goff[256 * ntiles + static_cast<uint64_t>(tid)], &all);
  base[tid] = digit_base + goff[static_cast<uint64_t>(tid) * ntiles + t];
  for (int w = 0; w < kWaves; ++w) wcnt[w][tid] = 0;
  __syncthreads();
  for (int s = 0; s < kItems; ++s) {
    const uint64_t i = t * kTile + static_cast<uint64_t>(s) * kSortBlock + tid;
    const bool act = i < n;
    uint64_t k = 0;
    uint32_t v = 0;
    unsigned d = 0;
    if (act) {
      k = kin[i];
      v = vin[i];
      d = static_cast<unsigned>(k >> shift) & 255u;
    }
    uint64_t peers = __ballot(act);
    for (int bit = 0; bit < 8; ++bit) {
      const uint64_t bb = __ballot(act && ((d >> bit) & 1u));
      peers &= ((d >> bit) & 1u) ? bb : ~bb;
    }
    const unsigned rank = __popcll(peers & lanemask_lt());
    if (act && rank == 0) wcnt[wave][d] = __popcll(peers);
    __syncthreads();
    if (act) {
      unsigned off = run[d] + rank;
      for (int w = 0; w < wave; ++w) off += wcnt[w][d];
      const uint64_t dst = static_cast<uint64_t>(base[d]) + off;
      kout[dst] = k;
      vout[dst] = v;
    }
    __syncthreads();
    unsigned tot = 0;
    for (int w = 0; w < kWaves; ++w) {
      tot += wcnt[w][tid];
      wcnt[w][tid] = 0;
    }
    run[tid] += tot;
    __syncthreads();
  }
}

// Sort input: the stored paths with a nonzero key <= U as
// (key << 32 | level) -> node index; *count receives their number.  A
// thread takes eight consecutive slots (two 16-byte key loads).
__global__ void __launch_bounds__(256) build_keys_kernel(NodeStore nodes, const PfxtCtl* ctl,
                                                         uint64_t* key, uint32_t* val,
                                                         unsigned long long* count) {
  const uint64_t n = ctl->node_n < nodes.cap ? ctl->node_n : nodes.cap;
  const uint32_t uk = ctl->ukey;
  constexpr uint64_t kT = 256ull * 8;
  const uint4* key4 = reinterpret_cast<const uint4*>(nodes.key);
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kT; base < n;
       base += static_cast<uint64_t>(gridDim.x) * kT) {
    const uint64_t i0 = base + 8ull * threadIdx.x;
    unsigned mask = 0, cnt = 0;
    uint32_t ks[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i0 < n) {
      const uint4 ka = key4[i0 / 4], kb = key4[i0 / 4 + 1];
      ks[0] = ka.x; ks[1] = ka.y; ks[2] = ka.z; ks[3] = ka.w;
      ks[4] = kb.x; ks[5] = kb.y; ks[6] = kb.z; ks[7] = kb.w;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (i0 + q < n && ks[q] <= uk && !key_is_zero(ks[q])) {
          mask |= 1u << q;
          ++cnt;
        }
      }
    }
    uint64_t slot = block_reserve<256>(count, cnt);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (mask & (1u << q)) {
        key[slot] = (static_cast<uint64_t>(ks[q]) << 32) |
                    static_cast<uint32_t>(nodes.level[i0 + q]);
        val[slot] = static_cast<uint32_t>(i0 + q);
        ++slot;
      }
    }
  }
}

__global__ void emit_report_kernel(const uint32_t* val, uint64_t nout, NodeStore nodes, float* ocost,
                                   int* olevel, int* ofrom, int* oto, int* opidx, int* rank_of) {
  for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < nout;
       r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t id = val[r];
    ocost[r] = nodes.cost[id];
    olevel[r] = nodes.level[id];
    ofrom[r] = nodes.from[id];
    oto[r] = nodes.to[id];
    opidx[r] = nodes.parent[id];
    rank_of[id] = static_cast<int>(r);
  }
}

__global__ void link_parents_kernel(const int* opidx, const int* rank_of, uint64_t nout, int* oparent) {
  for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < nout;
       r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const int p = opidx[r];
    oparent[r] = p < 0 ? -1 : rank_of[p];
  }
}

__global__ void fill_int_kernel(int* p, int v, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

int grid_for(uint64_t n, int block) {
  const uint64_t g = (n + block - 1) / block;
  return static_cast<int>(g < 4096 ? (g ? g : 1) : 4096);
}

}  // namespace

uint64_t sort_tiles(uint64_t n) { return (n + kTile - 1) / kTile; }
uint64_t sort_ghist_len(uint64_t n) { return 256 * (std::max<uint64_t>(1, sort_tiles(n)) + 1); }

void launch_exclusive_scan_u32(unsigned* p, uint64_t len, hipStream_t s) {
  if (len == 0) return;
  hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(kScanBlock), 0, s, p, len);
  CPG_HIP(hipGetLastError());
}

int radix_sort_pairs(SortBufs& b, uint64_t n, const int* bytes, int nbytes, hipStream_t s) {
  int cur = 0;
  if (n <= 1) return cur;
  const uint64_t ntiles = sort_tiles(n);
  for (int p = 0; p < nbytes; ++p) {
    const int shift = 8 * bytes[p];
    hipLaunchKernelGGL(sort_hist_kernel, dim3(static_cast<unsigned>(ntiles)), dim3(kSortBlock), 0, s, b.key[cur], n, shift, b.ghist, ntiles);
    CPG_HIP(hipGetLastError());
    hipLaunchKernelGGL(sort_rowscan_kernel, dim3(256), dim3(kRowBlock), 0, s, b.ghist, ntiles);
    CPG_HIP(hipGetLastError());
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(static_cast<unsigned>(ntiles)), dim3(kSortBlock), 0, s, 
        b.key[cur], b.val[cur], b.key[cur ^ 1], b.val[cur ^ 1], n, shift, b.ghist, ntiles);
    CPG_HIP(hipGetLastError());
    cur ^= 1;
  }
  return cur;
}

void launch_build_sort_keys(const NodeStore& nodes, const PfxtCtl* ctl, uint64_t nmax,
                            uint64_t* key, uint32_t* val, unsigned long long* count, hipStream_t s) {
  if (nmax == 0) return;
  hipLaunchKernelGGL(build_keys_kernel, dim3(grid_for((nmax + 7) / 8, 256)), dim3(256), 0, s,
                     nodes, ctl, key, val, count);
  CPG_HIP(hipGetLastError());
}

void launch_emit_report(const uint32_t* val, uint64_t nout, NodeStore nodes, float* ocost,
                        int* olevel, int* ofrom, int* oto, int* opidx, int* rank_of,
                        hipStream_t s) {
  if (nout == 0) return;
  hipLaunchKernelGGL(emit_report_kernel, dim3(grid_for(nout, 256)), dim3(256), 0, s, val, nout,
                     nodes, ocost, olevel, ofrom, oto, opidx, rank_of);
  CPG_HIP(hipGetLastError());
}

void launch_link_parents(const int* opidx, const int* rank_of, uint64_t nout, int* oparent,
                         hipStream_t s) {
  if (nout == 0) return;
  hipLaunchKernelGGL(link_parents_kernel, dim3(grid_for(nout, 256)), dim3(256), 0, s, opidx, rank_of, nout, oparent);
  CPG_HIP(hipGetLastError());
}

void launch_fill_int(int* p, int v, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
  CPG_HIP(hipGetLastError());
}

}  // namespace cpg

namespace cpg {
namespace {
__global__ void gather_int_kernel(const int* src, const int* idx, int n, int* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = src[idx[i]];
}
}  // namespace

void launch_gather_int(const int* src, const int* idx, int n, int* out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gather_int_kernel, dim3(grid_for(static_cast<uint64_t>(n), 256)), dim3(256), 0, s, src, idx, n, out);
  CPG_HIP(hipGetLastError());
}
}  // namespace cpg
