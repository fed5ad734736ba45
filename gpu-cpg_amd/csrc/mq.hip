// Multi-query PFXT: Q independent top-K queries expanded together.
//
// Every query q owns a set of roots and its own bound U_q, histogram
// window and K_q; the node store, the batch and every kernel are shared,
// and a node carries the id of its query.  The queries run on the disjoint
// union of their graphs (batched designs), whose SFXT is each graph's own
// SFXT, so one level loop serves all of them.  Per round, for every query
// at once:
//   1. histograms of its stored paths at or below U_q (global atomics into
//      its own kMqBins bins: Q x kMqBins bins keep contention low),
//   2. the batch threshold from its unexpanded paths below U_q,
//   3. a count pass over the batch histogramming its children (< U_q),
//   4. U_q tightened to the top of the bin where the running count of its
//      stored paths and counted children reaches K_q (the bin's upper
//      edge: at least K_q paths cost no more; a per-bin largest key would
//      be tighter by under one of the kMqBins bins, at the price of a
//      second global atomic per counted child, and atomics are performed
//      at memory),
//   5. a fill pass storing its children <= U_q (inclusive: the counted
//      children at U_q are needed).
// The per-query semantics equal the single-query rounds' (see pfxt.hip);
// the final selection sorts (query, cost, level) so each query's K_q
// results are contiguous.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"
#include "walk.hpp"

namespace cpg {

namespace {

constexpr int kMqBlock = 256;
constexpr int kMqSelBlock = 256;

__device__ inline uint32_t mq_bin(uint32_t k, uint32_t klo, uint32_t shift) {
  if (k < klo) return 0;
  const uint32_t b = (k - klo) >> shift;
  return b < static_cast<uint32_t>(kMqBins) ? b : static_cast<uint32_t>(kMqBins - 1);
}

__device__ inline uint32_t mq_shift(uint32_t klo, uint32_t khi) {
  uint32_t s = 0;
  const uint64_t span = khi >= klo ? static_cast<uint64_t>(khi - klo) : 0;
  while ((span >> s) >= static_cast<uint64_t>(kMqBins)) ++s;
  return s;
}

// Block-wide first bin at which the running sum of h reaches target;
// kMqBins if the total is smaller.  *total receives the sum.
__device__ uint32_t mq_crossing(const unsigned* h, unsigned long long target,
                                unsigned long long* total) {
  constexpr int kPer = kMqBins / kMqSelBlock;
  __shared__ unsigned long long s_part[kMqSelBlock];
  __shared__ unsigned s_bin;
  const int t = threadIdx.x;
  unsigned long long mine = 0;
  for (int i = 0; i < kPer; ++i) mine += h[t * kPer + i];
  s_part[t] = mine;
  if (t == 0) s_bin = kMqBins;
  __syncthreads();
  for (int off = 1; off < kMqSelBlock; off <<= 1) {
    const unsigned long long v = t >= off ? s_part[t - off] : 0ull;
    __syncthreads();
    s_part[t] += v;
    __syncthreads();
  }
  const unsigned long long incl = s_part[t];
  const unsigned long long excl = incl - mine;
  *total = s_part[kMqSelBlock - 1];
  if (excl < target && incl >= target) {
    unsigned long long run = excl;
    for (int i = 0; i < kPer; ++i) {
      run += h[t * kPer + i];
      if (run >= target) { s_bin = static_cast<uint32_t>(t * kPer + i); break; }
    }
  }
  __syncthreads();
  const uint32_t b = s_bin;
  __syncthreads();
  return b;
}

}  // namespace

__global__ void __launch_bounds__(kMqBlock) mq_roots_kernel(MqArgs a, const int* root_v,
                                                            const int* root_q, int nroots) {
  for (int base = blockIdx.x * blockDim.x; base < nroots; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    const bool take = i < nroots;
    const uint64_t slot = wave_reserve(&a.ctl->node_n, take);
    if (take) {
      const int s = root_v[i];
      const float c = a.distf[s];
      a.nodes.cost[slot] = c;
      a.nodes.to[slot] = s;
      a.nodes.from[slot] = -1;
      a.nodes.parent[slot] = -1;
      a.nodes.level[slot] = 0;
      a.nodes.key[slot] = float_key(c);
      a.nodes.flag[slot] = 0;
      a.nq[slot] = root_q[i];
    }
  }
}

// Histograms of every stored path at or below its query's bound: nonzero
// keys, and the unexpanded paths below the bound.
__global__ void __launch_bounds__(kMqBlock) mq_hist_kernel(MqArgs a) {
  const uint64_t n = a.ctl->node_n < a.nodes.cap ? a.ctl->node_n : a.nodes.cap;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const int q = a.nq[i];
    const MqQuery qs = a.qs[q];
    const uint32_t k = a.nodes.key[i];
    if (qs.done || k > qs.ukey) continue;
    const uint32_t b = mq_bin(k, qs.klo, qs.shift);
    unsigned* h = a.hist + static_cast<size_t>(q) * kMqHistStride;
    if (!key_is_zero(k)) atomicAdd(&h[b], 1u);
    if (k < qs.ukey && a.nodes.flag[i] == 0) atomicAdd(&h[kMqBins + b], 1u);
  }
}

// One block per query: the batch threshold (target unexpanded paths below
// the bound); a query with none left is done.
__global__ void __launch_bounds__(kMqSelBlock) mq_select_batch_kernel(MqArgs a, int all) {
  const int q = blockIdx.x;
  MqQuery* qs = a.qs + q;
  if (qs->done) return;
  const unsigned* h = a.hist + static_cast<size_t>(q) * kMqHistStride;
  unsigned long long total = 0;
  const bool take_all = all || qs->all;
  const uint32_t b = mq_crossing(h + kMqBins, take_all ? ~0ull : qs->batch_target, &total);
  if (threadIdx.x == 0) {
    qs->all = 0;
    if (total == 0) {
      qs->done = 1;
    } else if (b >= static_cast<uint32_t>(kMqBins)) {
      qs->thkey = qs->ukey;
    } else {
      const uint64_t up = static_cast<uint64_t>(qs->klo) + ((static_cast<uint64_t>(b) + 1) << qs->shift) - 1;
      const uint32_t upk = up > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(up);
      qs->thkey = upk < qs->ukey ? upk : qs->ukey;
    }
    if (total != 0) atomicAdd(&a.ctl->active, 1);
  }
}

__global__ void __launch_bounds__(kMqBlock) mq_make_batch_kernel(MqArgs a) {
  const uint64_t n = a.ctl->node_n < a.nodes.cap ? a.ctl->node_n : a.nodes.cap;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x; base < n;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = base + threadIdx.x;
    bool take = false;
    if (i < n && a.nodes.flag[i] == 0) {
      const MqQuery qs = a.qs[a.nq[i]];
      const uint32_t k = a.nodes.key[i];
      take = !qs.done && k <= qs.thkey && k < qs.ukey;
    }
    const uint64_t slot = wave_reserve(&a.ctl->batch_n, take);
    if (take) {
      a.nodes.flag[i] = 1;
      a.batch[slot] = static_cast<uint32_t>(i);
    }
  }
}

// Count pass (kFill false): histograms each query's children < U_q.
// Fill pass: stores each query's children <= U_q.
// In the count pass a block counts the children of the query its first
// parent belongs to in an LDS histogram and adds its nonzero bins to the
// query's global histogram at the end of the step: a block's 256
// consecutive batch entries mostly belong to one query (a parent's
// children are stored together), and every global atomic is performed at
// memory.  Parents of other queries count with global atomics.
template <bool kFill>
__global__ void __launch_bounds__(kMqBlock) mq_walk_kernel(MqArgs a) {
  __shared__ unsigned s_h[kMqBins];
  __shared__ int s_q;
  MqCtl* ctl = a.ctl;
  const uint64_t nb = ctl->batch_n;
  unsigned long long hops = 0, devv = 0, nonmono = 0;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x; base < nb;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {  // block-uniform
    if (!kFill) {
      if (threadIdx.x == 0) s_q = a.nq[a.batch[base]];
      for (int j = threadIdx.x; j < kMqBins; j += blockDim.x) s_h[j] = 0u;
      __syncthreads();
    }
    const uint64_t i = base + threadIdx.x;
    int p = 0, lvl = 0, q = 0, to = -1;
    float c = 0.0f;
    MqQuery qs{};
    qs.ukey = kKeyInf;
    if (i < nb) {
      p = static_cast<int>(a.batch[i]);
      c = a.nodes.cost[p];
      to = a.nodes.to[p];
      lvl = a.nodes.level[p];
      q = a.nq[p];
      qs = a.qs[q];
    }
    const float U = key_float(qs.ukey);
    ChainWalk w;
    w.start(i < nb, c, to, U, qs.ukey != kKeyInf, qs.mag);
    unsigned* h = a.hist + static_cast<size_t>(q) * kMqHistStride;
    const bool local = !kFill && q == s_q;  // count in the block's LDS histogram
    unsigned stored = 0;  // fill: children this lane stored for its parent
    while (__ballot(w.mode != 2)) {
      float cc = 0.0f;
      int cw = 0, from = 0;
      const bool emit = w.step(qs.vrec, qs.devinfo, qs.devs, U, !kFill, qs.step_limit, &ctl->err,
                               &hops, &devv, &nonmono, &cc, &cw, &from);
      if (!kFill) {
        if (emit) {
          const uint32_t k = float_key(cc);
          if (!key_is_zero(k)) {
            const uint32_t b = mq_bin(k, qs.klo, qs.shift);
            if (local) atomicAdd(&s_h[b], 1u);
            else atomicAdd(&h[b], 1u);
          }
        }
        continue;
      }
      const uint64_t slot = wave_reserve(&ctl->node_n, emit);
      if (emit) {
        if (slot < a.nodes.cap) {
          a.nodes.cost[slot] = cc;
          a.nodes.to[slot] = cw;
          a.nodes.from[slot] = from;
          a.nodes.parent[slot] = p;
          a.nodes.level[slot] = lvl + 1;
          a.nodes.key[slot] = float_key(cc);
          a.nodes.flag[slot] = 0;
          a.nq[slot] = q;
          ++stored;
        } else {
          atomicExch(&ctl->overflow, 1);
        }
      }
    }
    if (kFill) {
      // per-query stored counts: one atomic per distinct query among the
      // wave's lanes that stored children
      uint64_t pend = __ballot(stored > 0);
      while (pend) {
        const int leader = __ffsll(static_cast<long long>(pend)) - 1;
        const int lq = __shfl(q, leader);
        const bool mine = stored > 0 && q == lq;
        unsigned long long sum = mine ? stored : 0;
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (__lane_id() == leader) atomicAdd(&a.qs[lq].stored, sum);
        pend &= ~__ballot(mine);
      }
    }
    if (!kFill) {  // the block's LDS counts into its query's histogram
      __syncthreads();
      unsigned* hq = a.hist + static_cast<size_t>(s_q) * kMqHistStride;
      for (int j = threadIdx.x; j < kMqBins; j += blockDim.x) {
        const unsigned v = s_h[j];
        if (v) atomicAdd(&hq[j], v);
      }
      __syncthreads();  // s_h and s_q are set again for the next step
    }
  }
  (void)hops;
  (void)devv;
  for (int o = 32; o > 0; o >>= 1) nonmono += __shfl_xor(nonmono, o);
  if (kFill && __lane_id() == 0 && nonmono) atomicAdd(&ctl->nonmono, nonmono);  // once per expansion
}

// One block per query: U_q = the upper edge of the bin where the running
// count of stored paths and counted children reaches K_q.
__global__ void __launch_bounds__(kMqSelBlock) mq_select_bound_kernel(MqArgs a) {
  const int q = blockIdx.x;
  MqQuery* qs = a.qs + q;
  if (qs->done) return;
  const unsigned* h = a.hist + static_cast<size_t>(q) * kMqHistStride;
  unsigned long long total = 0;
  const uint32_t b = mq_crossing(h, qs->k, &total);
  if (threadIdx.x == 0) {
    uint32_t uk = qs->ukey;
    if (b < static_cast<uint32_t>(kMqBins)) {
      const uint64_t up = static_cast<uint64_t>(qs->klo) + ((static_cast<uint64_t>(b) + 1) << qs->shift) - 1;
      if (up < uk) uk = static_cast<uint32_t>(up);
    }
    qs->ukey = uk;
    qs->shift = mq_shift(qs->klo, uk);
    qs->rounds += 1;
  }
}

// Sort input: each stored path with a nonzero key <= its query's bound as
// (query << 48 | key << 16 | level) -> node, for the queries sel marks
// (every query when sel is null).
__global__ void __launch_bounds__(kMqBlock) mq_build_keys_kernel(MqArgs a, const uint8_t* sel,
                                                                 uint64_t* key, uint32_t* val,
                                                                 unsigned long long* count,
                                                                 unsigned long long* per_q) {
  const uint64_t n = a.ctl->node_n < a.nodes.cap ? a.ctl->node_n : a.nodes.cap;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x; base < n;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = base + threadIdx.x;
    bool take = false;
    uint32_t k = 0;
    int q = 0;
    if (i < n) {
      q = a.nq[i];
      k = a.nodes.key[i];
      take = (!sel || sel[q]) && k <= a.qs[q].ukey && !key_is_zero(k);
    }
    const uint64_t slot = wave_reserve(count, take);
    if (take) {
      key[slot] = (static_cast<uint64_t>(q) << 48) | (static_cast<uint64_t>(k) << 16) |
                  static_cast<uint32_t>(a.nodes.level[i] & 0xffff);
      val[slot] = static_cast<uint32_t>(i);
    }
    // per-query counts: one atomic per distinct query among the wave's
    // taken lanes (consecutive nodes mostly share a query)
    uint64_t pend = __ballot(take);
    while (pend) {
      const int leader = __ffsll(static_cast<long long>(pend)) - 1;
      const int lq = __shfl(q, leader);
      const uint64_t peers = __ballot(take && q == lq);
      if (__lane_id() == leader) atomicAdd(&per_q[lq], static_cast<unsigned long long>(__popcll(peers)));
      pend &= ~peers;
    }
  }
}

// Compaction of the shared node store (served queries): a node is live
// while its query's slot is not done.  Tiles of kCompactTile nodes keep
// their order; parents are remapped through newidx afterwards.
constexpr int kCompactChunks = 16;
constexpr int kCompactTile = kMqBlock * kCompactChunks;

__global__ void __launch_bounds__(kMqBlock) mq_count_live_kernel(MqArgs a, uint64_t n,
                                                                 unsigned* tile_cnt) {
  __shared__ unsigned s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const uint64_t t = blockIdx.x;
  unsigned c = 0;
  for (int j = 0; j < kCompactChunks; ++j) {
    const uint64_t i = t * kCompactTile + static_cast<uint64_t>(j) * kMqBlock + threadIdx.x;
    if (i < n && !a.qs[a.nq[i]].done) ++c;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (__lane_id() == 0) atomicAdd(&s_cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) tile_cnt[t] = s_cnt;
}

__global__ void __launch_bounds__(kMqBlock) mq_compact_kernel(MqArgs a, uint64_t n, NodeStore dst,
                                                              int* dst_nq, const unsigned* tile_off,
                                                              int* newidx) {
  constexpr int kW = kMqBlock / 64;
  __shared__ unsigned s_wave[kW];
  __shared__ unsigned s_run;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_run = 0;
  __syncthreads();
  const uint64_t t = blockIdx.x;
  const uint64_t base = tile_off[t];
  for (int j = 0; j < kCompactChunks; ++j) {
    const uint64_t i = t * kCompactTile + static_cast<uint64_t>(j) * kMqBlock + threadIdx.x;
    int q = 0;
    bool live = false;
    if (i < n) {
      q = a.nq[i];
      live = !a.qs[q].done;
    }
    const uint64_t m = __ballot(live);
    const unsigned rank = static_cast<unsigned>(__popcll(m & lanemask_lt()));
    if (lane == 0) s_wave[wave] = static_cast<unsigned>(__popcll(m));
    __syncthreads();
    unsigned off = s_run;
    for (int w = 0; w < wave; ++w) off += s_wave[w];
    if (live) {
      const uint64_t d = base + off + rank;
      dst.cost[d] = a.nodes.cost[i];
      dst.to[d] = a.nodes.to[i];
      dst.from[d] = a.nodes.from[i];
      dst.parent[d] = a.nodes.parent[i];
      dst.level[d] = a.nodes.level[i];
      dst.key[d] = a.nodes.key[i];
      dst.flag[d] = a.nodes.flag[i];
      dst_nq[d] = q;
      newidx[i] = static_cast<int>(d);
    } else if (i < n) {
      newidx[i] = -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned tot = 0;
      for (int w = 0; w < kW; ++w) tot += s_wave[w];
      s_run += tot;
    }
    __syncthreads();
  }
}

// Appends a single-query node store's first cnt paths (a bound probe's
// roots and children) at dst[base..) for query slot q, parents shifted.
__global__ void mq_seed_kernel(NodeStore src, uint64_t cnt, NodeStore dst, int* dst_nq,
                               uint64_t base, int q) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < cnt;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t d = base + i;
    const int p = src.parent[i];
    dst.cost[d] = src.cost[i];
    dst.to[d] = src.to[i];
    dst.from[d] = src.from[i];
    dst.parent[d] = p < 0 ? -1 : static_cast<int>(base + static_cast<uint64_t>(p));
    dst.level[d] = src.level[i];
    dst.key[d] = src.key[i];
    dst.flag[d] = src.flag[i];
    dst_nq[d] = q;
  }
}

__global__ void mq_set_node_n_kernel(MqCtl* ctl, unsigned long long n) { ctl->node_n = n; }

__global__ void mq_remap_parents_kernel(int* parent, const int* newidx, uint64_t n) {
  for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < n;
       j += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const int p = parent[j];
    if (p >= 0) parent[j] = newidx[p];
  }
}

__global__ void mq_gather_costs_kernel(const uint32_t* val, const float* cost, uint64_t n,
                                       float* out) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    out[i] = cost[val[i]];
}

__global__ void zero_u32_kernel(unsigned* p, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = 0;
}

__global__ void zero_active_kernel(MqCtl* ctl) {
  ctl->active = 0;
  ctl->batch_n = 0;
}

void launch_zero_u32(unsigned* p, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  const uint64_t g = std::min<uint64_t>(1024, (n + 255) / 256);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(static_cast<unsigned>(g)), dim3(256), 0, s, p, n);
  CPG_HIP(hipGetLastError());
}

void launch_zero_active(MqCtl* ctl, hipStream_t s) {
  hipLaunchKernelGGL(zero_active_kernel, dim3(1), dim3(1), 0, s, ctl);
  CPG_HIP(hipGetLastError());
}

void launch_mq_roots(const MqArgs& a, const int* root_v, const int* root_q, int nroots, hipStream_t s) {
  if (nroots <= 0) return;
  hipLaunchKernelGGL(mq_roots_kernel, dim3((nroots + kMqBlock - 1) / kMqBlock), dim3(kMqBlock), 0, s,
                     a, root_v, root_q, nroots);
  CPG_HIP(hipGetLastError());
}

void launch_mq_round(const MqArgs& a, bool all, int grid, hipStream_t s) {
  // a kernel clears the histograms (the round may be captured in a graph)
  launch_zero_u32(a.hist, static_cast<uint64_t>(a.nq_count) * kMqHistStride, s);
  launch_zero_active(a.ctl, s);
  hipLaunchKernelGGL(mq_hist_kernel, dim3(grid), dim3(kMqBlock), 0, s, a);
  CPG_HIP(hipGetLastError());
  hipLaunchKernelGGL(mq_select_batch_kernel, dim3(a.nq_count), dim3(kMqSelBlock), 0, s, a, all ? 1 : 0);
  CPG_HIP(hipGetLastError());
  hipLaunchKernelGGL(mq_make_batch_kernel, dim3(grid), dim3(kMqBlock), 0, s, a);
  CPG_HIP(hipGetLastError());
  hipLaunchKernelGGL((mq_walk_kernel<false>), dim3(grid), dim3(kMqBlock), 0, s, a);
  CPG_HIP(hipGetLastError());
  hipLaunchKernelGGL(mq_select_bound_kernel, dim3(a.nq_count), dim3(kMqSelBlock), 0, s, a);
  CPG_HIP(hipGetLastError());
  hipLaunchKernelGGL((mq_walk_kernel<true>), dim3(grid), dim3(kMqBlock), 0, s, a);
  CPG_HIP(hipGetLastError());
}

void launch_mq_build_keys(const MqArgs& a, const uint8_t* sel, uint64_t nmax, uint64_t* key,
                          uint32_t* val, unsigned long long* count, unsigned long long* per_q,
                          int grid, hipStream_t s) {
  if (nmax == 0) return;
  hipLaunchKernelGGL(mq_build_keys_kernel, dim3(grid), dim3(kMqBlock), 0, s, a, sel, key, val,
                     count, per_q);
  CPG_HIP(hipGetLastError());
}

uint64_t mq_compact_tiles(uint64_t n) { return (n + kCompactTile - 1) / kCompactTile; }

void launch_mq_seed(const NodeStore& src, uint64_t cnt, const NodeStore& dst, int* dst_nq,
                    uint64_t base, int q, MqCtl* ctl, int grid, hipStream_t s) {
  if (cnt > 0) {
    hipLaunchKernelGGL(mq_seed_kernel, dim3(grid), dim3(kMqBlock), 0, s, src, cnt, dst, dst_nq, base, q);
    CPG_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(mq_set_node_n_kernel, dim3(1), dim3(1), 0, s, ctl,
                     static_cast<unsigned long long>(base + cnt));
  CPG_HIP(hipGetLastError());
}

void launch_mq_count_live(const MqArgs& a, uint64_t n, unsigned* tile_cnt, hipStream_t s) {
  const uint64_t nt = mq_compact_tiles(n);
  if (nt == 0) return;
  hipLaunchKernelGGL(mq_count_live_kernel, dim3(static_cast<unsigned>(nt)), dim3(kMqBlock), 0, s, a,
                     n, tile_cnt);
  CPG_HIP(hipGetLastError());
}

void launch_mq_compact(const MqArgs& a, uint64_t n, const NodeStore& dst, int* dst_nq,
                       const unsigned* tile_off, int* newidx, uint64_t n_live, int grid,
                       hipStream_t s) {
  const uint64_t nt = mq_compact_tiles(n);
  if (nt == 0) return;
  hipLaunchKernelGGL(mq_compact_kernel, dim3(static_cast<unsigned>(nt)), dim3(kMqBlock), 0, s, a, n,
                     dst, dst_nq, tile_off, newidx);
  CPG_HIP(hipGetLastError());
  if (n_live == 0) return;
  hipLaunchKernelGGL(mq_remap_parents_kernel, dim3(grid), dim3(kMqBlock), 0, s, dst.parent, newidx,
                     n_live);
  CPG_HIP(hipGetLastError());
}

void launch_mq_gather_costs(const uint32_t* val, const float* cost, uint64_t n, float* out,
                            hipStream_t s) {
  if (n == 0) return;
  const uint64_t g = std::min<uint64_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(mq_gather_costs_kernel, dim3(static_cast<unsigned>(g)), dim3(256), 0, s, val,
                     cost, n, out);
  CPG_HIP(hipGetLastError());
}

}  // namespace cpg
