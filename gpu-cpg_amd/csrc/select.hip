// Final selection: stable LSD radix sort of (cost key, level) -> node and
// assembly of the ranked report.
//
// The reference sorts its whole short pile with thrust::sort on slack
// (gpucpg/gpucpg.cu:21323-21331), erases zero-cost entries (:21338-21358)
// and returns the first K (get_slacks :21409-21420).  Here the live set
// already holds only paths at or below the K-th bound, the sort key is
// (orderable cost bits, prefix-tree level) so that a parent always ranks
// before a child of equal cost, and zero-cost paths are skipped while
// ranking.
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

// Exclusive scan of ghist[0..len) in place, one block.
__global__ void __launch_bounds__(kScanBlock) sort_scan_kernel(unsigned* ghist, uint64_t len) {
  __shared__ unsigned long long part[kScanBlock];
  const uint64_t per = (len + kScanBlock - 1) / kScanBlock;
  const uint64_t b = threadIdx.x * per;
  const uint64_t e = b + per < len ? b + per : len;
  unsigned long long sum = 0;
  for (uint64_t i = b; i < e; ++i) sum += ghist[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const unsigned long long v = threadIdx.x >= static_cast<unsigned>(off) ? part[threadIdx.x - off] : 0ull;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned long long run = part[threadIdx.x] - sum;
  for (uint64_t i = b; i < e; ++i) {
    const unsigned c = ghist[i];
    ghist[i] = static_cast<unsigned>(run);
    run += c;
  }
}

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
  base[tid] = goff[static_cast<uint64_t>(tid) * ntiles + t];
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

__global__ void build_keys_kernel(const uint32_t* live_key, const uint32_t* live_idx, uint64_t n,
                                  const int* node_level, uint64_t* key, uint32_t* val,
                                  unsigned long long* zc) {
  unsigned long long neg = 0, zero = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t k = live_key[i];
    const uint32_t id = live_idx[i] & ~kExpandedBit;
    const uint32_t lvl = static_cast<uint32_t>(node_level[id]);
    key[i] = (static_cast<uint64_t>(k) << 32) | lvl;
    val[i] = id;
    if (key_is_zero(k)) ++zero;
    else if (k < kKeyNegZero) ++neg;
  }
  for (int o = 32; o > 0; o >>= 1) {
    neg += __shfl_xor(neg, o);
    zero += __shfl_xor(zero, o);
  }
  if (__lane_id() == 0) {
    if (neg) atomicAdd(&zc[0], neg);
    if (zero) atomicAdd(&zc[1], zero);
  }
}

__global__ void emit_report_kernel(const uint32_t* val, uint64_t nout, uint64_t nneg, uint64_t nzero,
                                   NodeStore nodes, float* ocost, int* olevel, int* ofrom, int* oto,
                                   int* opidx, int* rank_of) {
  for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < nout;
       r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t pos = r < nneg ? r : r + nzero;
    const uint32_t id = val[pos];
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

int radix_sort_pairs(SortBufs& b, uint64_t n, const int* bytes, int nbytes, hipStream_t s) {
  int cur = 0;
  if (n <= 1) return cur;
  const uint64_t ntiles = sort_tiles(n);
  for (int p = 0; p < nbytes; ++p) {
    const int shift = 8 * bytes[p];
    hipLaunchKernelGGL(sort_hist_kernel, dim3(static_cast<unsigned>(ntiles)), dim3(kSortBlock), 0, s, b.key[cur], n, shift, b.ghist, ntiles);
    CPG_HIP(hipGetLastError());
    hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(kScanBlock), 0, s, b.ghist, 256 * ntiles);
    CPG_HIP(hipGetLastError());
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(static_cast<unsigned>(ntiles)), dim3(kSortBlock), 0, s, 
        b.key[cur], b.val[cur], b.key[cur ^ 1], b.val[cur ^ 1], n, shift, b.ghist, ntiles);
    CPG_HIP(hipGetLastError());
    cur ^= 1;
  }
  return cur;
}

void launch_build_sort_keys(const uint32_t* live_key, const uint32_t* live_idx, uint64_t n,
                            const int* node_level, uint64_t* key, uint32_t* val,
                            unsigned long long* zero_counts, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(build_keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, live_key, live_idx, n, node_level, key, val,
                                                     zero_counts);
  CPG_HIP(hipGetLastError());
}

void launch_emit_report(const uint64_t* /*key*/, const uint32_t* val, uint64_t nout, uint64_t nneg,
                        uint64_t nzero, NodeStore nodes, float* ocost, int* olevel, int* ofrom,
                        int* oto, int* opidx, int* rank_of, hipStream_t s) {
  if (nout == 0) return;
  hipLaunchKernelGGL(emit_report_kernel, dim3(grid_for(nout, 256)), dim3(256), 0, s, val, nout, nneg, nzero, nodes, ocost,
                                                         olevel, ofrom, oto, opidx, rank_of);
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
