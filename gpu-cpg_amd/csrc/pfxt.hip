// PFXT: exact top-K prefix-tree expansion in bounded best-first rounds.
//
// A path is a root (a source followed by its suffix chain,
// gpucpg/gpucpg.cu:18502-18513) or a parent plus one deviation edge u->w
// taken at a vertex u of the parent's suffix chain, with cost
//   ((parent + distf[w]) + w_e) - distf[u]
// evaluated left to right in binary32 exactly as
// compute_short_long_path_counts / expand_short_pile do
// (gpucpg/gpucpg.cu:2670-2676, :3271-3276).
//
// Instead of the reference's short/long piles and split schedule
// (gpucpg/gpucpg.cu:19134-21228), every generated path stays in the node
// store with its orderable cost key and an expanded flag, and each round
//   1. histograms the stored paths at or below the bound U (all nonzero
//      ones, and the unexpanded ones),
//   2. picks the cheapest unexpanded paths as the batch (histogram select),
//   3. walks their suffix chains and stores every child with cost <= U,
//   4. tightens U to the K-th smallest nonzero cost among the stored paths
//      (histogram of the old paths plus the new children).
// While U is unbounded a count pass histograms the children before the
// fill pass stores them, so that only paths at or below the tightened U
// are stored.  U is always the upper end of the histogram bin where the
// running count of K distinct stored paths is reached, so no path of the
// final top K lies above it while child cost >= parent cost.
// The walk skips chain segments via the deviation index built by the SFXT:
// a vertex whose smallest deviation delta exceeds X = (U - parent) + slack
// is skipped through its next-smaller pointer and its sorted deviation
// list is cut at the first delta > X.  slack bounds the rounding gap
// between delta and the binary32 cost recurrence, so the skip never hides
// a child whose computed cost is <= U; the exact comparison cost <= U
// decides emission.
#include "common.hpp"
#include "kernels.hpp"
#include "walk.hpp"

namespace cpg {

namespace {

constexpr int kBlock = 256;
constexpr int kSelBlock = 1024;

__device__ inline unsigned long long ld_u64(const unsigned long long* p) { return *p; }

// Block-wide search for the first bin at which the running sum of h
// reaches target (1-based count).  Returns kHistBins if the total is
// smaller; *total receives the sum.  Must be called by all threads of a
// kSelBlock block.
__device__ uint32_t find_crossing(const unsigned* h, unsigned long long target,
                                  unsigned long long* total) {
  constexpr int kPer = kHistBins / kSelBlock;
  __shared__ unsigned long long s_part[kSelBlock];
  __shared__ unsigned s_bin;
  const int t = threadIdx.x;
  unsigned long long mine = 0;
  for (int i = 0; i < kPer; ++i) mine += h[t * kPer + i];
  s_part[t] = mine;
  if (t == 0) s_bin = kHistBins;
  __syncthreads();
  // inclusive scan (Hillis-Steele) over kSelBlock partial sums
  for (int off = 1; off < kSelBlock; off <<= 1) {
    const unsigned long long v = t >= off ? s_part[t - off] : 0ull;
    __syncthreads();
    s_part[t] += v;
    __syncthreads();
  }
  const unsigned long long incl = s_part[t];
  const unsigned long long excl = incl - mine;
  *total = s_part[kSelBlock - 1];
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

// Clears a bound histogram: kHistBins counts followed by kHistBins per-bin
// maximum keys.
__device__ inline void clear_hist(unsigned* h) {
  for (int i = threadIdx.x; i < 2 * kHistBins; i += blockDim.x) h[i] = 0;
}

__global__ void __launch_bounds__(kBlock) roots_kernel(const int* srcs, const uint8_t* keep, int nsrc,
                                                       const int* dist, const float* distf,
                                                       NodeStore nodes, PfxtCtl* ctl) {
  for (int base = blockIdx.x * blockDim.x; base < nsrc; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    bool take = false;
    int s = 0;
    if (i < nsrc) {
      s = srcs[i];
      take = keep[i] != 0 && dist[s] != kIntMax;
    }
    const uint64_t slot = wave_reserve(&ctl->node_n, take);
    if (take) {
      const float c = distf[s];
      nodes.cost[slot] = c;
      nodes.to[slot] = s;
      nodes.from[slot] = -1;
      nodes.parent[slot] = -1;
      nodes.level[slot] = 0;
      nodes.key[slot] = float_key(c);
      nodes.flag[slot] = 0;
    }
  }
}

// Per-block histograms over the window (klo, shift), counted in LDS and
// written out as one partial per block: the nonzero keys <= U with the
// largest key of each bin, and (with_pend) the unexpanded paths with key
// < U (those still to be expanded).  mode 0 covers every stored path,
// mode 1 this round's children only, mode 2 the keys a count pass wrote
// (ckeys, no flags).  Each thread reads four consecutive keys and flags
// with one vector load each (the arrays are padded to a multiple of four
// slots).
__global__ void __launch_bounds__(kBlock) hist_nodes_kernel(NodeStore nodes, const uint32_t* ckeys,
                                                            uint64_t ckeys_cap, PfxtCtl* ctl,
                                                            int mode, int with_pend,
                                                            unsigned* part_all, unsigned* part_pend,
                                                            unsigned* part_max) {
  __shared__ unsigned h_all[kHistBins];
  __shared__ unsigned h_pend[kHistBins];
  __shared__ unsigned h_max[kHistBins];
  for (int j = threadIdx.x; j < kHistBins; j += blockDim.x) {
    h_all[j] = 0;
    h_pend[j] = 0;
    h_max[j] = 0;
  }
  __syncthreads();
  // [lo, nn) within the buffer: after a node-store overflow node_n and
  // fill_node_base exceed the capacity, and the vector loop below reads
  // the group of four holding lo, so lo must not pass nn
  uint64_t nn, lo = 0;
  if (mode == 2) {
    const uint64_t c = ctl->child_n;
    nn = c < ckeys_cap ? c : ckeys_cap;
  } else {
    const uint64_t c = ctl->node_n;
    nn = c < nodes.cap ? c : nodes.cap;
    if (mode == 1) {
      const uint64_t b = ctl->fill_node_base;
      lo = b < nn ? b : nn;
    }
  }
  if (ctl->done) lo = nn = 0;  // nothing left: write empty partials
  const uint32_t uk = ctl->ukey;
  const uint32_t klo = ctl->klo, shift = ctl->shift;
  const uint4* key4 = reinterpret_cast<const uint4*>(mode == 2 ? ckeys : nodes.key);
  const uint32_t* flag4 = reinterpret_cast<const uint32_t*>(nodes.flag);
  for (uint64_t g = lo / 4 + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       g < (nn + 3) / 4; g += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint4 kv = key4[g];
    const uint32_t fv = with_pend ? flag4[g] : 0u;
    const uint32_t ks[4] = {kv.x, kv.y, kv.z, kv.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint64_t i = 4 * g + q;
      const uint32_t k = ks[q];
      if (i < lo || i >= nn || k > uk) continue;
      const uint32_t b = key_bin(k, klo, shift);
      if (!key_is_zero(k)) {
        atomicAdd(&h_all[b], 1u);
        atomicMax(&h_max[b], k);
      }
      if (with_pend && k < uk && ((fv >> (8 * q)) & 0xffu) == 0) atomicAdd(&h_pend[b], 1u);
    }
  }
  __syncthreads();
  unsigned* pa = part_all + static_cast<size_t>(blockIdx.x) * kHistBins;
  unsigned* pp = part_pend + static_cast<size_t>(blockIdx.x) * kHistBins;
  unsigned* pm = part_max + static_cast<size_t>(blockIdx.x) * kHistBins;
  for (int j = threadIdx.x; j < kHistBins; j += blockDim.x) {
    pa[j] = h_all[j];
    pm[j] = h_max[j];
    if (with_pend) pp[j] = h_pend[j];
  }
}

// Sums (or, take_max, takes the maximum of) the kHistParts per-block
// partial histograms into out, combined with out when accumulate and
// written over it otherwise.  A block of kRedBlock
// threads covers 64 bins: each of its kRedBlock / 64 groups adds an
// interleaved share of the partials with all its loads in flight, and the
// groups' sums meet in LDS.
constexpr int kRedBlock = 1024;
constexpr int kRedGroups = kRedBlock / 64;
constexpr int kRedPer = kHistParts / kRedGroups;
static_assert(kHistParts % kRedGroups == 0, "partials split evenly over groups");

__global__ void __launch_bounds__(kRedBlock) reduce_hist_kernel(const unsigned* part, unsigned* out,
                                                                int accumulate, int take_max) {
  __shared__ unsigned s_sum[kRedGroups][64];
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  unsigned v[kRedPer];
#pragma unroll
  for (int q = 0; q < kRedPer; ++q)
    v[q] = part[static_cast<size_t>(grp + q * kRedGroups) * kHistBins + j];
  unsigned sum = 0;
#pragma unroll
  for (int q = 0; q < kRedPer; ++q) sum = take_max ? max(sum, v[q]) : sum + v[q];
  s_sum[grp][lane] = sum;
  __syncthreads();
  if (grp == 0) {
    unsigned tot = accumulate ? out[j] : 0u;
    for (int g2 = 0; g2 < kRedGroups; ++g2)
      tot = take_max ? max(tot, s_sum[g2][lane]) : tot + s_sum[g2][lane];
    out[j] = tot;
  }
}

__global__ void __launch_bounds__(kSelBlock) select_batch_kernel(PfxtCtl* ctl, unsigned* hist_pend,
                                                                 unsigned long long target) {
  if (ctl->done) return;
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl->batch_n = 0;
    ctl->child_n = 0;
    ctl->fill_node_base = ctl->node_n;
  }
  unsigned long long total = 0;
  const uint32_t b = find_crossing(hist_pend, target, &total);
  if (threadIdx.x == 0) {
    if (total == 0) {
      ctl->done = 1;
    } else if (b >= static_cast<uint32_t>(kHistBins)) {
      ctl->thkey = ctl->ukey;
    } else {
      const uint32_t up = bin_upper_key(b, ctl->klo, ctl->shift);
      ctl->thkey = up < ctl->ukey ? up : ctl->ukey;
    }
  }
}

// Marks the stored paths with key <= thkey and < U that are not expanded
// yet and appends them to the batch; one counter atomic per block and tile.
// A path at U itself is not expanded: at least K stored paths cost <= U,
// so its children (cost >= U) cannot change the K smallest costs.  A
// thread takes eight consecutive slots (two 16-byte key loads, one 8-byte
// flag load).
__global__ void __launch_bounds__(kBlock) make_batch_kernel(NodeStore nodes, PfxtCtl* ctl,
                                                            uint32_t* batch) {
  if (ctl->done) return;
  const uint64_t n = ctl->node_n < nodes.cap ? ctl->node_n : nodes.cap;
  const uint32_t th = ctl->thkey, uk = ctl->ukey;
  constexpr uint64_t kTile = static_cast<uint64_t>(kBlock) * 8;
  const uint4* key4 = reinterpret_cast<const uint4*>(nodes.key);
  const uint2* flag8 = reinterpret_cast<const uint2*>(nodes.flag);
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kTile; base < n;
       base += static_cast<uint64_t>(gridDim.x) * kTile) {
    const uint64_t i0 = base + 8ull * threadIdx.x;
    unsigned mask = 0, cnt = 0;
    if (i0 < n) {
      const uint4 ka = key4[i0 / 4], kb = key4[i0 / 4 + 1];
      const uint2 fv = flag8[i0 / 8];
      const uint32_t ks[8] = {ka.x, ka.y, ka.z, ka.w, kb.x, kb.y, kb.z, kb.w};
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint32_t f = ((q < 4 ? fv.x : fv.y) >> (8 * (q & 3))) & 0xffu;
        if (i0 + q < n && f == 0 && ks[q] <= th && ks[q] < uk) {
          mask |= 1u << q;
          ++cnt;
        }
      }
    }
    uint64_t slot = block_reserve<kBlock>(&ctl->batch_n, cnt);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (mask & (1u << q)) {
        nodes.flag[i0 + q] = 1;
        batch[slot++] = static_cast<uint32_t>(i0 + q);
      }
    }
  }
}

constexpr int kStage = 256;  // per-wave LDS slots for children awaiting a flush

// Walks the suffix chains of the batch's parents and finds every child
// with cost < U (cost <= U when a.strict == 0).  Count pass (kFill =
// false): write the children's keys to ckeys and count them in child_n
// (hist_nodes_kernel mode 2 histograms them).  Fill pass: append them to
// the node store, unexpanded.  Children are staged per wave in LDS and
// written with one counter atomic per flush.
template <bool kFill>
__global__ void __launch_bounds__(kBlock) walk_kernel(WalkArgs a) {
  constexpr int kWaves = kBlock / 64;
  constexpr int kS = kStage;
  constexpr int kSF = kFill ? kStage : 1;  // node fields are staged by the fill pass only
  __shared__ float s_cost[kWaves][kS];
  __shared__ int s_to[kWaves][kSF];
  __shared__ int s_from[kWaves][kSF];
  __shared__ int s_parent[kWaves][kSF];
  __shared__ int s_level[kWaves][kSF];
  PfxtCtl* ctl = a.ctl;
  if (ctl->done) return;
  const int wave = threadIdx.x >> 6;
  const int lane = __lane_id();
  const uint64_t nb = ctl->batch_n;
  const uint32_t ukey = ctl->ukey;
  const bool bounded = ukey != kKeyInf;
  const float U = key_float(ukey);
  const uint64_t cap = a.nodes.cap;
  const float mag = 2.0f * a.max_abs_distf + a.max_abs_wgt + 1.0f;
  unsigned long long hops = 0, devv = 0, nonmono = 0;
  int staged = 0;  // wave-uniform

  // Writes the wave's staged children: one atomic on node_n (fill) or
  // child_n (count: keys only, into ckeys), then each lane stores every
  // 64th child.  All lanes of the wave call it together.
  auto flush = [&]() {
    if (staged == 0) return;
    __builtin_amdgcn_wave_barrier();
    unsigned long long base = 0;
    if (lane == 0)
      base = atomicAdd(kFill ? &ctl->node_n : &ctl->child_n, static_cast<unsigned long long>(staged));
    base = __shfl(base, 0);
    if (!kFill) {
      for (int q = lane; q < staged; q += 64)
        if (base + q < a.ckeys_cap) a.ckeys[base + q] = float_key(s_cost[wave][q]);
      __builtin_amdgcn_wave_barrier();
      staged = 0;
      return;
    }
    for (int q = lane; q < staged; q += 64) {
      const uint64_t slot = base + q;
      const float cost = s_cost[wave][q];
      if (slot < cap) {
        a.nodes.cost[slot] = cost;
        a.nodes.to[slot] = s_to[wave][q];
        a.nodes.from[slot] = s_from[wave][q];
        a.nodes.parent[slot] = s_parent[wave][q];
        a.nodes.level[slot] = s_level[wave][q];
        a.nodes.key[slot] = float_key(cost);
        a.nodes.flag[slot] = 0;
      } else {
        atomicExch(&ctl->overflow, 1);
      }
    }
    __builtin_amdgcn_wave_barrier();
    staged = 0;
  };

  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x; base < nb;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = base + threadIdx.x;
    int p = 0, lvl = 0;
    float c = 0.0f;
    int to = -1;
    if (i < nb) {
      p = static_cast<int>(a.batch[i]);
      c = a.nodes.cost[p];
      to = a.nodes.to[p];
      lvl = a.nodes.level[p];
    }
    ChainWalk cw_state;
    cw_state.start(i < nb, c, to, U, bounded, mag);
    while (__ballot(cw_state.mode != 2)) {
      float cc = 0.0f;
      int cw = 0, from = 0;
      const bool emit = cw_state.step(a.vrec, a.devinfo, a.devs, U, a.strict != 0, a.step_limit,
                                      &ctl->err, &hops, &devv, &nonmono, &cc, &cw, &from);
      const uint64_t m = __ballot(emit);
      if (m) {
        const int add = static_cast<int>(__popcll(m));
        if (staged + add > kS) flush();
        if (emit) {
          const int q = staged + static_cast<int>(__popcll(m & lanemask_lt()));
          s_cost[wave][q] = cc;
          if (kFill) {
            s_to[wave][q] = cw;
            s_from[wave][q] = from;
            s_parent[wave][q] = p;
            s_level[wave][q] = lvl + 1;
          }
        }
        staged += add;
      }
    }
  }
  flush();
  for (int o = 32; o > 0; o >>= 1) nonmono += __shfl_xor(nonmono, o);
  if (kFill && lane == 0 && nonmono) atomicAdd(&ctl->nonmono, nonmono);  // once per expansion
  if (a.collect_stats) {
    for (int o = 32; o > 0; o >>= 1) {
      hops += __shfl_xor(hops, o);
      devv += __shfl_xor(devv, o);
    }
    if (lane == 0) {
      atomicAdd(&ctl->hops, hops);
      atomicAdd(&ctl->devvis, devv);
    }
  }
}

__global__ void __launch_bounds__(kSelBlock) select_bound_kernel(PfxtCtl* ctl, unsigned* hist_all,
                                                                 unsigned long long k) {
  if (ctl->done) return;
  unsigned long long total = 0;
  const uint32_t b = find_crossing(hist_all, k, &total);
  if (threadIdx.x == 0) {
    uint32_t uk = ctl->ukey;
    if (b < static_cast<uint32_t>(kHistBins)) {
      const uint32_t top = hist_all[kHistBins + b];  // largest key in the crossing bin
      if (top < uk) uk = top;
    }
    ctl->ukey = uk;
    ctl->shift = window_shift(ctl->klo, uk);
    ctl->expanded += ctl->batch_n;
    ctl->rounds += 1;
  }
  __syncthreads();
  clear_hist(hist_all);
}

// Bound probe: the walk's inputs for one trial bound ukey over the roots
// (node slots 0..nroots-1, listed in batch by iota_kernel); the node store
// is cut back to the roots, dropping an earlier trial's children.
__global__ void probe_setup_kernel(PfxtCtl* ctl, uint32_t ukey, uint32_t shift,
                                   unsigned long long nb, unsigned long long nroots) {
  ctl->ukey = ukey;
  ctl->shift = shift;
  ctl->batch_n = nb;
  ctl->child_n = 0;
  ctl->done = 0;
  ctl->node_n = nroots;
  ctl->fill_node_base = nroots;
  ctl->overflow = 0;
}

__global__ void iota_kernel(uint32_t* p, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = static_cast<uint32_t>(i);
}

// Bound probe succeeded (at least k nonzero paths with cost <= the trial
// bound are stored): the bound becomes the largest key of the bin where
// the running count reaches k, and the nroots roots count as expanded.
__global__ void __launch_bounds__(kSelBlock) probe_finish_kernel(PfxtCtl* ctl, unsigned* hist_all,
                                                                 unsigned long long k,
                                                                 unsigned long long nroots) {
  unsigned long long total = 0;
  const uint32_t b = find_crossing(hist_all, k, &total);
  if (threadIdx.x == 0) {
    uint32_t uk = ctl->ukey;
    if (b < static_cast<uint32_t>(kHistBins)) {
      const uint32_t top = hist_all[kHistBins + b];  // largest key in the crossing bin
      if (top < uk) uk = top;
    }
    ctl->ukey = uk;
    ctl->shift = window_shift(ctl->klo, uk);
    ctl->batch_n = 0;
    ctl->child_n = 0;
    ctl->expanded += nroots;
  }
  __syncthreads();
  clear_hist(hist_all);
}

}  // namespace

void launch_roots(const int* srcs, const uint8_t* keep, int nsrc, const int* dist,
                  const float* distf, NodeStore nodes, PfxtCtl* ctl, hipStream_t s) {
  if (nsrc <= 0) return;
  const int grid = (nsrc + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(roots_kernel, dim3(grid), dim3(kBlock), 0, s, srcs, keep, nsrc, dist, distf,
                     nodes, ctl);
  CPG_HIP(hipGetLastError());
}

void launch_probe_setup(PfxtCtl* ctl, uint32_t ukey, uint32_t shift, unsigned long long nb,
                        unsigned long long nroots, hipStream_t s) {
  hipLaunchKernelGGL(probe_setup_kernel, dim3(1), dim3(1), 0, s, ctl, ukey, shift, nb, nroots);
  CPG_HIP(hipGetLastError());
}

void launch_iota(uint32_t* p, uint64_t n, int grid, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(kBlock), 0, s, p, n);
  CPG_HIP(hipGetLastError());
}

void launch_probe_finish(PfxtCtl* ctl, unsigned* hist_all, unsigned long long k,
                         unsigned long long nroots, hipStream_t s) {
  hipLaunchKernelGGL(probe_finish_kernel, dim3(1), dim3(kSelBlock), 0, s, ctl, hist_all, k, nroots);
  CPG_HIP(hipGetLastError());
}

void launch_hist_nodes(const NodeStore& nodes, const uint32_t* ckeys, uint64_t ckeys_cap,
                       PfxtCtl* ctl, int mode, bool with_pend, unsigned* part, unsigned* hist_all,
                       unsigned* hist_pend, bool accumulate, hipStream_t s) {
  const size_t plane = static_cast<size_t>(kHistParts) * kHistBins;
  unsigned* part_all = part;
  unsigned* part_pend = part + plane;
  unsigned* part_max = part + 2 * plane;
  hipLaunchKernelGGL(hist_nodes_kernel, dim3(kHistParts), dim3(kBlock), 0, s, nodes, ckeys,
                     ckeys_cap, ctl, mode, with_pend ? 1 : 0, part_all, part_pend, part_max);
  CPG_HIP(hipGetLastError());
  const int rg = kHistBins / 64;
  const int acc = accumulate ? 1 : 0;
  hipLaunchKernelGGL(reduce_hist_kernel, dim3(rg), dim3(kRedBlock), 0, s, part_all, hist_all, acc, 0);
  CPG_HIP(hipGetLastError());
  hipLaunchKernelGGL(reduce_hist_kernel, dim3(rg), dim3(kRedBlock), 0, s, part_max,
                     hist_all + kHistBins, acc, 1);
  CPG_HIP(hipGetLastError());
  if (with_pend) {
    hipLaunchKernelGGL(reduce_hist_kernel, dim3(rg), dim3(kRedBlock), 0, s, part_pend, hist_pend,
                       acc, 0);
    CPG_HIP(hipGetLastError());
  }
}

void launch_select_batch(PfxtCtl* ctl, unsigned* hist_pend, unsigned long long target,
                         hipStream_t s) {
  hipLaunchKernelGGL(select_batch_kernel, dim3(1), dim3(kSelBlock), 0, s, ctl, hist_pend, target);
  CPG_HIP(hipGetLastError());
}

void launch_make_batch(const NodeStore& nodes, PfxtCtl* ctl, uint32_t* batch, int grid,
                       hipStream_t s) {
  hipLaunchKernelGGL(make_batch_kernel, dim3(grid), dim3(kBlock), 0, s, nodes, ctl, batch);
  CPG_HIP(hipGetLastError());
}

void launch_walk(const WalkArgs& a, bool fill, int grid, hipStream_t s) {
  if (fill) hipLaunchKernelGGL((walk_kernel<true>), dim3(grid), dim3(kBlock), 0, s, a);
  else hipLaunchKernelGGL((walk_kernel<false>), dim3(grid), dim3(kBlock), 0, s, a);
  CPG_HIP(hipGetLastError());
}

namespace {
// Stored paths when `rounds` rounds have completed.
__global__ void round_log_kernel(const PfxtCtl* ctl, unsigned long long* rlog, int cap) {
  const int r = ctl->rounds;
  if (r >= 0 && r < cap) rlog[r] = ctl->node_n < ctl->node_cap ? ctl->node_n : ctl->node_cap;
}
}  // namespace

void launch_round_log(const PfxtCtl* ctl, unsigned long long* rlog, int cap, hipStream_t s) {
  hipLaunchKernelGGL(round_log_kernel, dim3(1), dim3(1), 0, s, ctl, rlog, cap);
  CPG_HIP(hipGetLastError());
}

void launch_select_bound(PfxtCtl* ctl, unsigned* hist_all, unsigned long long k, hipStream_t s) {
  hipLaunchKernelGGL(select_bound_kernel, dim3(1), dim3(kSelBlock), 0, s, ctl, hist_all, k);
  CPG_HIP(hipGetLastError());
}

}  // namespace cpg
