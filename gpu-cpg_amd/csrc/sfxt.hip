// SFXT: levelize + relax + successor + distf, then the deviation index the
// PFXT walk reads.
//
// Semantics follow the reference path:
//   levelize    bfs_td_step_privatized   gpucpg/gpucpg.cu:2215-2283, loop :17642-17681
//               (a vertex joins level d+1 when its last fanout is consumed)
//   relax       relax_bu_step            gpucpg/gpucpg.cu:1167-1187, loop :17933-17961
//               dist[v] = min_e dist[w] + (int)(w_e * 10000.0f), sinks 0 (:17035-17039)
//   successor   update_successors_bu     gpucpg/gpucpg.cu:2606-2628 (largest id at the min)
//   distf       (float)dist / SCALE_UP, correctly rounded (used at :2672-2673)
// Every fanout of a level-d vertex sits on a lower level, so relax and
// successor for level d run in the same pass that peels level d.  That
// pass is the only part with a level-to-level dependency; it runs as one
// resident kernel with a grid barrier per level on a small grid (the
// barrier costs more with more blocks, and levels hold a few thousand
// vertices) and does no more than the relax, the successor and the fanin
// release.
//
// Endpoint subset (an extension; the reference has no such option): with a
// sink mask, only the chosen sinks start at distance 0.  The other sinks,
// and every vertex whose fanouts are all unreachable, keep dist INT_MAX
// and succ -1 and offer no candidates; levels are unchanged.  A path then
// ends at a chosen sink, as the reference's own unreachable test (dist[w]
// != INT_MAX, gpucpg/gpucpg.cu:2662-2668) already implies for deviations.
//
// The deviation index is this engine's own structure for the PFXT walk,
// built afterwards by fully parallel passes:
//   devs     per vertex u, its deviation edges (u->w, w != succ[u], dist[w]
//            finite: gpucpg/gpucpg.cu:2662-2668) as {w, w_e, distf[w], delta}
//            with delta = distf[w] + w_e - distf[u], sorted by delta (stable)
//   vrec     {smallest delta, nsv, succ, distf}; nsv[u] is the first vertex
//            after u on u's suffix chain whose smallest delta is strictly
//            smaller (next-smaller-value pointer), so a walk can skip chain
//            segments with no cheap deviation
// nsv comes from binary lifting over the successor forest: lift[k][u] =
// {2^k-th successor of u, smallest delta over the 2^k chain vertices after
// u}; a greedy descent from the top level finds the first chain vertex
// whose smallest delta is below u's with one dependent load per level.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace cpg {

namespace {

constexpr int kSfxtBlock = 256;
constexpr int kLdsQueue = 2048;
// Polls of a grid-barrier wait before the level loop reports an error; a
// build that runs blocks on CPU threads (the fiber emulator) raises it.
#ifndef CPG_SPIN_LIMIT
#define CPG_SPIN_LIMIT (1u << 22)
#endif
constexpr unsigned kSpinLimit = CPG_SPIN_LIMIT;
constexpr int kThreadSort = 16;    // deviation lists up to this length: one thread sorts
constexpr int kWaveSort = 64;      // up to this length: one wave ranks them in registers
constexpr int kLdsSort = kSfxtLdsSort;  // up to this length: one block sorts in LDS
constexpr int kFaninUnroll = 8;   // fanin records (and their atomics) in flight per step
constexpr int kMultiSteps = 4;    // multi-subset loop: record steps in flight per wave
constexpr int kSortBlock = 256;

__device__ inline unsigned long long atomic_load_rlx64(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Device-scope (relaxed) loads and stores of words other blocks write or
// read during the same kernel.
__device__ inline unsigned long long ld_dev64(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_dev64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_dev32(int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_dev32(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned ld_devu32(unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Generation barrier over the blocks of one level loop, which also hands
// every block the queue tail of the next level.  One release fence before
// arriving and one acquire fence after leaving; the wait polls with
// relaxed loads so it does not invalidate caches on every probe.
//
// Each block arrives with the number of vertices it pushed this level
// (its own queue writes are done by then): the arrival word counts blocks
// in its top 24 bits and sums their pushes in the low 40.  The last block
// to arrive therefore knows the level's tail, le + the sum, without
// reading qtail again, and releases the others with one 64-bit store of
// the next generation and that tail, which the waiters take from the word
// they poll.  (Reading qtail itself after leaving would race with blocks
// already pushing the next level.)  A spin that exceeds kSpinLimit raises
// ctl->err and every later barrier returns false at once, so the kernel
// always drains.  Returns false on error.
//
// kFenced = false: the fence-free variant for loops whose cross-block data
// (queue slots, best[] words) is itself written and read at device scope.
// A block then arrives once its own stores and atomics of the level have
// completed (s_waitcnt), and no block writes back or invalidates its L2.
template <bool kFenced = true>
__device__ bool grid_barrier(SfxtCtl* ctl, unsigned nblocks, unsigned pushed, unsigned le,
                             unsigned* next_tail) {
  if (!kFenced) __builtin_amdgcn_s_waitcnt(0);  // this thread's stores and atomics are done
  __syncthreads();
  __shared__ int s_ok;
  __shared__ unsigned s_tail;
  if (threadIdx.x == 0) {
    int ok = 1;
    unsigned tail = 0;
    const unsigned g = static_cast<unsigned>(atomic_load_rlx64(&ctl->bar_word) >> 32);
    if (kFenced) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned long long old = __hip_atomic_fetch_add(
        &ctl->bar_arrive, (1ull << 40) | pushed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (static_cast<unsigned>(old >> 40) == nblocks - 1) {
      if (kFenced) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      tail = le + static_cast<unsigned>(old & ((1ull << 40) - 1)) + pushed;
      __hip_atomic_store(&ctl->bar_arrive, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (kFenced) {
        __hip_atomic_store(&ctl->bar_word, (static_cast<unsigned long long>(g + 1u) << 32) | tail,
                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __builtin_amdgcn_s_waitcnt(0);  // the reset lands before any block can arrive again
        __hip_atomic_store(&ctl->bar_word, (static_cast<unsigned long long>(g + 1u) << 32) | tail,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else {
      unsigned spins = 0;
      for (;;) {
        const unsigned long long w = atomic_load_rlx64(&ctl->bar_word);
        if (static_cast<unsigned>(w >> 32) != g) {
          tail = static_cast<unsigned>(w);
          break;
        }
        if (__hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = 0; break; }
        if (++spins > kSpinLimit) { atomicExch(&ctl->err, 1); ok = 0; break; }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (kFenced) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (__hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) ok = 0;
    s_tail = tail;
    s_ok = ok;
  }
  __syncthreads();
  *next_tail = s_tail;
  return s_ok != 0;
}

__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Candidate key: high word = candidate distance with the sign bit flipped
// (unsigned order = signed order), low word = 0x7fffffff - fanout id (a
// larger id orders first among equal distances).
__device__ inline unsigned long long pack_cand(int dist, int w) {
  return (static_cast<unsigned long long>(static_cast<uint32_t>(dist) ^ 0x80000000u) << 32) |
         static_cast<uint32_t>(0x7fffffff - w);
}
__device__ inline int unpack_cand(unsigned long long k) {
  return static_cast<int>(static_cast<uint32_t>(k >> 32) ^ 0x80000000u);
}
__device__ inline int unpack_succ(unsigned long long k) {
  return 0x7fffffff - static_cast<int>(static_cast<uint32_t>(k));
}

__device__ inline int4 default_vrec() {
  return make_int4(__float_as_int(__builtin_huge_valf()), -1, -1, 0);
}

}  // namespace

// Per-vertex state before the level loop; the level-0 queue is the sink list.
__global__ void __launch_bounds__(256) sfxt_init_kernel(SfxtArgs a) {
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int gthreads = gridDim.x * blockDim.x;
  for (int v = gtid; v < a.n; v += gthreads) {
    a.deps[v] = a.oadjp[v + 1] - a.oadjp[v];
    a.level[v] = -1;
    a.dist[v] = kIntMax;
    a.succ[v] = -1;
    a.distf[v] = static_cast<float>(kIntMax) / static_cast<float>(kScaleUp);
    a.best[v] = ~0ull;
  }
  for (int i = gtid; i < a.nsinks; i += gthreads) a.queue[i] = a.sinks[i];
  if (gtid == 0) a.ctl->qtail = static_cast<unsigned>(a.nsinks);
}

// Level loop: peels level d (the vertices whose fanouts are all settled)
// and releases their fanins; a grid barrier separates levels.
//
// The relax runs push-style: a settled vertex w offers every fanin x the
// candidate dist[w] + (int)(w_e * 10000.0f) through a 64-bit atomic min on
// best[x], whose high word orders the candidate and whose low word orders
// larger w first.  When x is peeled, best[x] therefore holds exactly the
// minimum candidate and, among equal minima, the largest fanout id: the
// relax and successor rules above, read with one load instead of a gather
// over x's fanouts.  Per vertex the level loop makes four dependent memory
// round trips (queue entry; best and fanin range; fanin records; the
// atomics), with all loads of a round issued together (kFaninUnroll fanins
// per step).
//
// kFenced = false (the fence-free variant): queue slots are written with
// device-scope stores and read with device-scope loads (-1 until written),
// best[] is read with device-scope loads, and the grid barrier has no
// release/acquire fence; the queue must be initialised to the sinks and -1.
template <bool kFenced, int kU = kFaninUnroll>
__device__ void levels_loop(const SfxtArgs& a) {
  __shared__ int s_q[kLdsQueue];
  __shared__ int s_qn;
  __shared__ unsigned s_base;
  SfxtCtl* ctl = a.ctl;
  const unsigned nblocks = gridDim.x;
  const int gtid = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int gthreads = static_cast<int>(nblocks * blockDim.x);
  if (threadIdx.x == 0) s_qn = 0;
  __syncthreads();

  auto push = [&](int x) {
    const int p = atomicAdd(&s_qn, 1);
    if (p < kLdsQueue) s_q[p] = x;
    else if (kFenced) a.queue[atomicAdd(&ctl->qtail, 1u)] = x;
    else st_dev32(a.queue + atomicAdd(&ctl->qtail, 1u), x);
  };
  // Writes the block's LDS queue out; returns every vertex the block
  // pushed this level (those that spilled past the LDS queue included).
  auto flush = [&]() -> unsigned {
    __syncthreads();
    const int total = s_qn;
    const int cnt = min(total, kLdsQueue);
    if (threadIdx.x == 0) s_base = cnt ? atomicAdd(&ctl->qtail, static_cast<unsigned>(cnt)) : 0u;
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
      if (kFenced) a.queue[s_base + i] = s_q[i];
      else st_dev32(a.queue + s_base + i, s_q[i]);
    }
    __syncthreads();
    if (threadIdx.x == 0) s_qn = 0;
    return static_cast<unsigned>(total);
  };

  unsigned lb = 0, le = static_cast<unsigned>(a.nsinks);
  int d = 0;
  float local_max = 0.0f;
  while (true) {
    if (gtid == 0) a.lvlp[d] = static_cast<int>(lb);
    for (unsigned i = lb + gtid; i < le; i += gthreads) {
      int u;
      unsigned long long key;
      if (kFenced) {
        u = a.queue[i];
        key = a.best[u];
      } else {
        u = ld_dev32(a.queue + i);
        unsigned spins = 0;
        while (u < 0) {  // written before the barrier; a store still landing
          if (++spins > kSpinLimit) { atomicExch(&ctl->err, 1); break; }
          __builtin_amdgcn_s_sleep(1);
          u = ld_dev32(a.queue + i);
        }
        if (u < 0) continue;
        key = ld_dev64(a.best + u);
      }
      const int b = a.iadjp[u], e = a.iadjp[u + 1];
      // level 0: a sink outside the endpoint subset is unreachable; later
      // levels: no candidate arrived (every fanout unreachable)
      int du = 0, s = -1;
      if (d == 0) {
        if (a.sink_mask && !a.sink_mask[u]) du = kIntMax;
      } else if (key == ~0ull) {
        du = kIntMax;
      } else {
        du = unpack_cand(key);
        s = unpack_succ(key);
      }
      const bool reach = du != kIntMax;
      const float df = static_cast<float>(du) / static_cast<float>(kScaleUp);
      const int tie = a.orig ? a.orig[u] : u;  // successor tie-break: original id
      a.dist[u] = du;
      a.succ[u] = s;
      a.level[u] = d;
      a.distf[u] = df;
      if (reach) local_max = fmaxf(local_max, fabsf(df));
      for (int j = b; j < e; j += kU) {
        // kU fanin records; past the end the last one is loaded again, so
        // every load is unconditional
        int2 r[kU];
#pragma unroll
        for (int q = 0; q < kU; ++q) r[q] = a.iedge[min(j + q, e - 1)];
        // Branch-free atomics, all in flight together: with a branch per
        // record the compiler waits for each returning atomic before the
        // next.  A slot past the end issues no-op atomics (min with the
        // largest key, subtract 0) on a pad slot after the n vertices,
        // spread by thread and slot (atomics on one address serialise) in
        // a range small enough to stay in the L2.
        int old[kU];
#pragma unroll
        for (int q = 0; q < kU; ++q) {
          const bool in = j + q < e;
          const int wi = static_cast<int>(__int_as_float(r[q].y) * static_cast<float>(kScaleUp));
          const int x = in ? r[q].x
                           : a.n + static_cast<int>((static_cast<unsigned>(gtid) * kU + static_cast<unsigned>(q)) &
                                                    static_cast<unsigned>(kSfxtPad - 1));
          atomicMin(&a.best[x], in && reach ? pack_cand(du + wi, tie) : ~0ull);
          old[q] = atomicSub(&a.deps[x], in ? 1 : 0);
        }
#pragma unroll
        for (int q = 0; q < kU; ++q)
          if (j + q < e && old[q] == 1) push(r[q].x);
      }
    }
    const unsigned pushed = flush();
    unsigned ne = 0;
    if (!grid_barrier<kFenced>(ctl, nblocks, pushed, le, &ne)) return;
    if (ne == le) break;
    lb = le;
    le = ne;
    ++d;
  }
  const float m = wave_max(local_max);
  if (__lane_id() == 0) atomicMax(&ctl->max_abs_distf_bits, __float_as_uint(m));
  if (gtid == 0) {
    a.lvlp[d + 1] = static_cast<int>(le);
    ctl->num_levels = le > 0 ? d + 1 : 0;
    ctl->levelized = le;
  }
}

__global__ void __launch_bounds__(kSfxtBlock) sfxt_levels_kernel(SfxtArgs a) { levels_loop<true>(a); }
// The fence-free loop with 4 or 8 fanin records in flight per step, chosen
// by the graph's mean in-degree (each slot past a vertex's last record
// costs a no-op atomic pair; 16 per step measured slower on the 30M-edge
// DAG: 11.5 ms against 10.4 ms).
__global__ void __launch_bounds__(kSfxtBlock) sfxt_levels_nf_kernel(SfxtArgs a) { levels_loop<false, 8>(a); }
__global__ void __launch_bounds__(kSfxtBlock) sfxt_levels_nf4_kernel(SfxtArgs a) { levels_loop<false, 4>(a); }

// Counter barrier of the fence-free level loops (levels_edges_loop,
// levels_multi_part): two 64-bit counters, ctl->lvl_bar[d & 1], each holding
// the arrivals at its levels (top 24 bits) and the vertices pushed at them
// (low 40 bits), both running totals that are never reset.  A block arrives
// with one returning atomic that adds its LDS queue count; the value it
// returns is also where the block's entries go: the level's entries before
// it are the counter's pushes since this parity's previous level, so the
// block writes its entries at le + (old - those) after arriving, while the
// others are still on their way (the next level's readers wait on a slot
// that is still -1).  A vertex pushed past the LDS queue reserves its slot
// the same way, with a pushes-only add.  The waiters poll the counter itself
// until every block has arrived; the last to arrive sees that in its own
// atomic's result.  So a level ends after one returning atomic and, for the
// others, the polls that see it, where the generation barrier also needs a
// queue-tail atomic, a store-completion wait before arriving and the last
// block's release store.
//
// Two counters make the running totals safe to read: a block that has seen
// level d complete cannot be overtaken at level d + 2 (the same counter)
// before it has itself arrived at level d + 1, so the value it reads is level
// d's final one.  Arrival counts are compared modulo 2^24.
constexpr unsigned long long kLevelLow = (1ull << 40) - 1;

// Pushes counted on each parity's counter up to its previous level (two
// scalars, not an indexed array, so that they stay in registers).
struct LevelCounts {
  unsigned long long last0 = 0ull, last1 = 0ull;
  __device__ unsigned long long last(int d) const { return (d & 1) ? last1 : last0; }
};

// Queue slot for a vertex pushed at level d past the block's LDS queue.
__device__ __forceinline__ unsigned counter_reserve(SfxtCtl* ctl, int d, unsigned le, const LevelCounts& lc) {
  const unsigned long long old =
      __hip_atomic_fetch_add(&ctl->lvl_bar[d & 1], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return le + static_cast<unsigned>((old & kLevelLow) - lc.last(d));
}

// Arrives at level d's barrier with the block's LDS queue s_q[0, min(*s_qn,
// kLdsQueue)), writes those entries, waits for the other nblocks - 1 blocks
// and clears *s_qn.  Returns false on error, else the next level's end.
__device__ __forceinline__ bool counter_barrier(SfxtCtl* ctl, int* queue, const int* s_q, int* s_qn,
                                                unsigned nblocks, int d, unsigned le, LevelCounts& lc,
                                                unsigned* next_end) {
  __shared__ unsigned s_base;
  __shared__ unsigned long long s_low;
  __shared__ int s_ok;
  __builtin_amdgcn_s_waitcnt(0);  // this thread's atomics and stores are done
  __syncthreads();
  const int cnt = min(*s_qn, kLdsQueue);
  const int par = d & 1;
  const unsigned long long last = lc.last(d);
  unsigned long long mine = 0;
  if (threadIdx.x == 0) {
    const unsigned long long add = (1ull << 40) | static_cast<unsigned long long>(cnt);
    const unsigned long long old =
        __hip_atomic_fetch_add(&ctl->lvl_bar[par], add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_base = le + static_cast<unsigned>((old & kLevelLow) - last);
    mine = old + add;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) st_dev32(queue + s_base + i, s_q[i]);
  if (threadIdx.x == 0) {
    *s_qn = 0;  // every thread read the count before the barrier above
    const unsigned target = ((static_cast<unsigned>(d) >> 1) + 1u) * nblocks;
    unsigned long long w = mine;
    int ok = 1;
    unsigned spins = 0;
    // arrivals so far minus the target, as a signed 24-bit difference
    auto short_of = [&](unsigned long long v) {
      const unsigned diff = (static_cast<unsigned>(v >> 40) - target) & 0xffffffu;
      return (diff & 0x800000u) != 0;
    };
    while (short_of(w)) {
      if (__hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = 0; break; }
      if (++spins > kSpinLimit) { atomicExch(&ctl->err, 1); ok = 0; break; }
      __builtin_amdgcn_s_sleep(1);
      w = ld_dev64(&ctl->lvl_bar[par]);
    }
    s_low = w & kLevelLow;
    s_ok = ok;
  }
  __syncthreads();
  const unsigned long long low = s_low;
  *next_end = le + static_cast<unsigned>(low - last);
  if (par) lc.last1 = low;
  else lc.last0 = low;
  return s_ok != 0;
}

// Queue entry i of the fence-free loops: written by a block after it
// arrived at the previous level's barrier, so it may still be landing (-1).
__device__ __forceinline__ int queue_entry(const SfxtArgs& a, unsigned i) {
  int u = ld_dev32(a.queue + i);
  unsigned spins = 0;
  while (u < 0) {
    if (++spins > kSpinLimit) { atomicExch(&a.ctl->err, 1); break; }
    __builtin_amdgcn_s_sleep(1);
    u = ld_dev32(a.queue + i);
  }
  return u;
}

// Edge-parallel fence-free level loop: the same peel, relax and successor
// as levels_loop<false>, with the fanin release spread over the records of
// a wave's vertices instead of over one vertex per lane.
//
// A wave takes 64 queue entries at a time (one per lane; fewer, by
// wave_entries, where in-degrees are long), settles them as above, and
// prefix-sums their in-degrees over the wave; the wave's records
// then form one flat list, handed out kS per lane step (lane l takes list
// positions r0 + s*64 + l), and a lane finds the vertex owning its position
// by a binary search over the exclusive prefix sums held in the lanes
// (ds_bpermute).  Every device-scope atomic costs a memory-side write
// whether it changes anything or not, so the per-lane loop's no-op slots (a
// vertex's record count rounded up to its step) were 40% of that loop's
// atomics on the 1.5M-vertex DAG; here only the positions past the wave's
// last record in its final step issue no-ops (on the pad slots).  Records of
// one vertex sit in consecutive lanes, so their loads coalesce.  Each level
// ends at the counter barrier.
template <int kS>
__device__ void levels_edges_loop(const SfxtArgs& a) {
  __shared__ int s_q[kLdsQueue];
  __shared__ int s_qn;
  SfxtCtl* ctl = a.ctl;
  const unsigned stride = a.place_stride > 1 ? static_cast<unsigned>(a.place_stride) : 1u;
  if (blockIdx.x % stride != 0) return;  // block-uniform: not a working block
  const unsigned nblocks = gridDim.x / stride;
  const int gtid = static_cast<int>(blockIdx.x / stride * blockDim.x + threadIdx.x);
  const int lane = __lane_id();
  const unsigned wave = static_cast<unsigned>(gtid) >> 6;
  const unsigned nwaves = nblocks * blockDim.x / 64u;
  // entries per wave: fewer than 64 where in-degrees are long, so that a
  // level's records spread over more waves (lanes past them hold none)
  const unsigned we = a.wave_entries > 0 && a.wave_entries < 64 ? static_cast<unsigned>(a.wave_entries) : 64u;
  if (threadIdx.x == 0) s_qn = 0;
  __syncthreads();

  unsigned lb = 0, le = static_cast<unsigned>(a.nsinks);
  int d = 0;
  LevelCounts lc;
  auto push = [&](int x) {
    const int p = atomicAdd(&s_qn, 1);
    if (p < kLdsQueue) s_q[p] = x;
    else st_dev32(a.queue + counter_reserve(ctl, d, le, lc), x);  // past the LDS queue
  };

  float local_max = 0.0f;
  while (true) {
    if (gtid == 0) a.lvlp[d] = static_cast<int>(lb);
    // wave-uniform loop over we-entry chunks of the level
    for (unsigned i0 = lb + wave * we; i0 < le; i0 += nwaves * we) {
      const unsigned i = i0 + static_cast<unsigned>(lane);
      int b = 0, deg = 0, du = kIntMax, tie = 0;
      if (static_cast<unsigned>(lane) < we && i < le) {
        const int u = queue_entry(a, i);
        if (u >= 0) {
          const unsigned long long key = ld_dev64(a.best + u);
          b = a.iadjp[u];
          deg = a.iadjp[u + 1] - b;
          int s = -1;
          du = 0;
          if (d == 0) {
            if (a.sink_mask && !a.sink_mask[u]) du = kIntMax;
          } else if (key == ~0ull) {
            du = kIntMax;
          } else {
            du = unpack_cand(key);
            s = unpack_succ(key);
          }
          const float df = static_cast<float>(du) / static_cast<float>(kScaleUp);
          tie = a.orig ? a.orig[u] : u;  // successor tie-break: original id
          a.dist[u] = du;
          a.succ[u] = s;
          a.level[u] = d;
          a.distf[u] = df;
          if (du != kIntMax) local_max = fmaxf(local_max, fabsf(df));
        }
      }
      // inclusive prefix sum of the in-degrees over the wave
      int incl = deg;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
      }
      const int excl = incl - deg;
      const int total = __shfl(incl, 63);
      for (int r0 = 0; r0 < total; r0 += 64 * kS) {
        int2 rec[kS];
        int cand_du[kS], cand_tie[kS];
#pragma unroll
        for (int s = 0; s < kS; ++s) {
          // list position, clamped to the last record so that every load
          // is in range; positions past the end issue no-op atomics below
          const int r = min(r0 + s * 64 + lane, total - 1);
          int o = 0;  // owner: the last lane whose exclusive sum is <= r
#pragma unroll
          for (int step = 32; step > 0; step >>= 1)
            if (__shfl(excl, o + step) <= r) o += step;
          const int j = __shfl(b, o) + (r - __shfl(excl, o));
          rec[s] = a.iedge[j];
          cand_du[s] = __shfl(du, o);
          cand_tie[s] = __shfl(tie, o);
        }
        int old[kS];
#pragma unroll
        for (int s = 0; s < kS; ++s) {
          // branch-free: every slot issues both atomics, so they are all
          // in flight together (see levels_loop)
          const bool in = r0 + s * 64 + lane < total;
          const bool reach = cand_du[s] != kIntMax;
          const int wi = static_cast<int>(__int_as_float(rec[s].y) * static_cast<float>(kScaleUp));
          const int x = in ? rec[s].x
                           : a.n + static_cast<int>((static_cast<unsigned>(gtid) * kS + static_cast<unsigned>(s)) &
                                                    static_cast<unsigned>(kSfxtPad - 1));
          atomicMin(&a.best[x], in && reach ? pack_cand(cand_du[s] + wi, cand_tie[s]) : ~0ull);
          old[s] = atomicSub(&a.deps[x], in ? 1 : 0);
        }
#pragma unroll
        for (int s = 0; s < kS; ++s)
          if (r0 + s * 64 + lane < total && old[s] == 1) push(rec[s].x);
      }
    }
    unsigned ne = 0;
    if (!counter_barrier(ctl, a.queue, s_q, &s_qn, nblocks, d, le, lc, &ne)) return;
    if (ne == le) break;
    lb = le;
    le = ne;
    ++d;
  }
  const float m = wave_max(local_max);
  if (__lane_id() == 0) atomicMax(&ctl->max_abs_distf_bits, __float_as_uint(m));
  if (gtid == 0) {
    a.lvlp[d + 1] = static_cast<int>(le);
    ctl->num_levels = le > 0 ? d + 1 : 0;
    ctl->levelized = le;
  }
}

// 3 list positions per lane step where the mean in-degree is under 4 (a
// wave's 64 vertices then hold about 64 x 2.7 records: one step), 8 above.
__global__ void __launch_bounds__(kSfxtBlock) sfxt_levels_e3_kernel(SfxtArgs a) { levels_edges_loop<3>(a); }
__global__ void __launch_bounds__(kSfxtBlock) sfxt_levels_e8_kernel(SfxtArgs a) { levels_edges_loop<8>(a); }

// ---------------------------------------------------------------------------
// Dataflow loop: the same levelize / relax / successor results as the level
// loop, without a grid barrier per level.
//
// A vertex is ready once every fanout has settled (its deps counter reaches
// zero).  The block that settles a vertex's last fanout takes it into its
// LDS pool and settles it in its next iteration, so a chain of dependent
// levels costs one block iteration per level instead of one grid barrier.
// Settling is pull-style, as relax_bu_step (gpucpg/gpucpg.cu:1167-1187) and
// update_successors_bu (:2606-2628) read the fanouts: dist[v] = min over
// reachable fanouts w of dist[w] + (int)(w_e * 10000.0f), succ[v] = the
// largest such w at the minimum, level[v] = 1 + max level[w] (the peel depth
// of bfs_td_step_privatized, :2215-2283).
//
// Everything one block reads of another's work is a write-once 64-bit state
// word per vertex, state[w] = level << 32 | dist (all ones until settled),
// read and written with device-scope atomic loads and stores; a reader that
// still sees all ones (the store has not landed yet) polls again.  So no
// fence orders anything: each value carries its own validity, and the deps
// counters are device-scope atomics.
//
// Work distribution: per block an LDS pool of ready vertices with their CSR
// ranges, loaded when they are released (the releasing thread loads the
// ranges of every fanin it decrements, in the same round trip as the
// atomics).  A pool larger than kDfKeep after an iteration spills the rest
// to a global queue (slots written once, -1 until then); the sinks start
// there.  A block with an empty pool claims a window of kDfBlock queue
// slots.  Termination: each block publishes one word (kept releases << 32 |
// vertices settled); a block that finds no work reads every block's word
// twice (settled counts, then released counts and the queue tail) and
// leaves when the settled total equals the released total (the sinks and
// spilled vertices count in the queue tail).  Releases are published before
// the released vertices can be settled anywhere, so the check cannot pass
// while work is outstanding.  Vertices on a cycle are never released and
// keep their initial values, as the level loop leaves them.
// ---------------------------------------------------------------------------
namespace {
constexpr int kDfBlock = 256;
constexpr int kDfPool = 2048;   // LDS pool capacity (vertices)
constexpr int kDfKeep = 512;    // pool entries kept after an iteration
constexpr int kDfUnroll = 8;    // edge records (and their loads / atomics) in flight per step
constexpr unsigned long long kStateEmpty = ~0ull;

__device__ inline int wave_max_int(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}
__device__ inline unsigned long long wave_sum64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
}  // namespace

__global__ void __launch_bounds__(256) sfxt_dataflow_init_kernel(SfxtArgs a) {
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int gthreads = gridDim.x * blockDim.x;
  for (int v = gtid; v < a.n; v += gthreads) {
    a.deps[v] = a.oadjp[v + 1] - a.oadjp[v];
    a.level[v] = -1;
    a.dist[v] = kIntMax;
    a.succ[v] = -1;
    a.distf[v] = static_cast<float>(kIntMax) / static_cast<float>(kScaleUp);
    a.best[v] = kStateEmpty;
    a.queue[v] = v < a.nsinks ? a.sinks[v] : -1;
  }
  for (int b = gtid; b < a.dfw_n; b += gthreads) a.dfw[b] = 0ull;
  if (gtid == 0) a.ctl->qtail = static_cast<unsigned>(a.nsinks);
}

__global__ void __launch_bounds__(kDfBlock) sfxt_dataflow_kernel(SfxtArgs a) {
  __shared__ int s_v[kDfPool];
  __shared__ int4 s_r[kDfPool];   // {fanout begin, fanout end, fanin begin, fanin end}
  __shared__ int s_pn;            // pool entries (a push past kDfPool spills at once)
  __shared__ int s_pbase;         // entries left in the pool after this iteration's take
  __shared__ int s_take;
  __shared__ int s_win;           // a window of queue slots is held
  __shared__ unsigned s_clo;      // window: slots [s_clo, s_clo + kDfBlock)
  __shared__ unsigned s_tail;     // queue tail known to the block
  __shared__ unsigned long long s_wd[kDfBlock / 64];  // window positions taken
  __shared__ int s_proc;          // vertices settled this iteration
  __shared__ int s_spill;         // pool entries spilled this iteration
  __shared__ unsigned s_qbase;
  __shared__ int s_exit;
  __shared__ int s_check;         // this idle iteration checks for quiescence
  __shared__ unsigned long long s_red[kDfBlock / 64];
  static_assert(kDfBlock == 256, "window done bits are four 64-bit words");

  SfxtCtl* ctl = a.ctl;
  const int t = threadIdx.x;
  const int lane = __lane_id();
  const int wid = t >> 6;
  const unsigned n = static_cast<unsigned>(a.n);
  unsigned long long* state = a.best;
  unsigned my_r = 0, my_p = 0;  // thread 0: kept releases, vertices settled
  unsigned idle = 0;
  int maxlev = -1;
  float local_max = 0.0f;
  if (t == 0) {
    s_pn = 0;
    s_win = 0;
    s_clo = 0;
    s_tail = 0;
    s_exit = 0;
  }
  __syncthreads();

  auto push = [&](int z, int4 rz) {
    const int p = atomicAdd(&s_pn, 1);
    if (p < kDfPool) {
      s_v[p] = z;
      s_r[p] = rz;
    } else {
      // pool full: straight to the global queue (counted in the tail first)
      const unsigned q = atomicAdd(&ctl->qtail, 1u);
      st_dev32(a.queue + q, z);
    }
  };

  for (;;) {
    // -- take: the pool's top entries, then window slots for idle threads
    if (t == 0) {
      s_proc = 0;
      if (!s_win && s_pn == 0) {
        const unsigned tail = ld_devu32(&ctl->qtail);
        const unsigned head = ld_devu32(&ctl->qhead);
        if (tail > head) {
          s_clo = atomicAdd(&ctl->qhead, static_cast<unsigned>(kDfBlock));
          s_win = 1;
          s_tail = tail;
          for (int i = 0; i < kDfBlock / 64; ++i) s_wd[i] = 0ull;
        }
      } else if (s_win && s_clo + kDfBlock > s_tail) {
        s_tail = ld_devu32(&ctl->qtail);
      }
      const int pn = s_pn;
      s_take = min(pn, kDfBlock);
      s_pbase = pn - s_take;
    }
    __syncthreads();
    int v = -1;
    int4 r = make_int4(0, 0, 0, 0);
    bool have_r = false;
    bool taken = false;
    if (t < s_take) {
      v = s_v[s_pbase + t];
      r = s_r[s_pbase + t];
      have_r = true;
    } else if (s_win && !((s_wd[wid] >> lane) & 1ull)) {
      const unsigned slot = s_clo + static_cast<unsigned>(t);
      if (slot >= n) {
        taken = true;  // never written
      } else if (slot < s_tail) {
        const int x = ld_dev32(a.queue + slot);
        if (x >= 0) {
          v = x;
          taken = true;
        }
      }
    }
    const unsigned long long tb = __ballot(taken);
    if (lane == 0 && tb) s_wd[wid] |= tb;
    __syncthreads();
    if (t == 0) s_pn = s_pbase;
    __syncthreads();

    // -- settle v
    if (v >= 0) {
      if (!have_r) r = make_int4(a.oadjp[v], a.oadjp[v + 1], a.iadjp[v], a.iadjp[v + 1]);
      int bd = kIntMax, bw = -1, lev = 0;
      if (r.x == r.y && !(a.sink_mask && !a.sink_mask[v])) bd = 0;  // a chosen sink
      for (int j = r.x; j < r.y; j += kDfUnroll) {
        int w[kDfUnroll];
        float wf[kDfUnroll];
        unsigned long long st[kDfUnroll];
#pragma unroll
        for (int q = 0; q < kDfUnroll; ++q) {
          const int e = min(j + q, r.y - 1);
          w[q] = a.oadj[e];
          wf[q] = a.owgt[e];
        }
#pragma unroll
        for (int q = 0; q < kDfUnroll; ++q) st[q] = ld_dev64(state + w[q]);
#pragma unroll
        for (int q = 0; q < kDfUnroll; ++q) {
          if (j + q >= r.y) continue;
          unsigned spins = 0;
          while (st[q] == kStateEmpty) {  // settled, store not landed yet
            if (++spins > kSpinLimit) {
              atomicExch(&ctl->err, 1);
              st[q] = static_cast<uint32_t>(kIntMax);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            st[q] = ld_dev64(state + w[q]);
          }
          const int d = static_cast<int>(static_cast<uint32_t>(st[q]));
          lev = max(lev, static_cast<int>(st[q] >> 32) + 1);
          if (d != kIntMax) {
            const int wi = static_cast<int>(wf[q] * static_cast<float>(kScaleUp));
            const int c = static_cast<int>(static_cast<unsigned>(d) + static_cast<unsigned>(wi));
            if (c < bd || (c == bd && w[q] > bw)) {
              bd = c;
              bw = w[q];
            }
          }
        }
      }
      const float df = static_cast<float>(bd) / static_cast<float>(kScaleUp);
      a.dist[v] = bd;
      a.succ[v] = bw;
      a.level[v] = lev;
      a.distf[v] = df;
      if (bd != kIntMax) local_max = fmaxf(local_max, fabsf(df));
      maxlev = max(maxlev, lev);
      st_dev64(state + v, (static_cast<unsigned long long>(static_cast<uint32_t>(lev)) << 32) |
                              static_cast<uint32_t>(bd));
      // release the fanins; load each one's CSR ranges with its atomic
      for (int j = r.z; j < r.w; j += kDfUnroll) {
        int z[kDfUnroll];
#pragma unroll
        for (int q = 0; q < kDfUnroll; ++q) z[q] = a.iedge[min(j + q, r.w - 1)].x;
        int old[kDfUnroll];
        int4 rz[kDfUnroll];
#pragma unroll
        for (int q = 0; q < kDfUnroll; ++q) {
          old[q] = 0;
          if (j + q < r.w) {
            old[q] = atomicSub(&a.deps[z[q]], 1);
            rz[q] = make_int4(a.oadjp[z[q]], a.oadjp[z[q] + 1], a.iadjp[z[q]], a.iadjp[z[q] + 1]);
          }
        }
#pragma unroll
        for (int q = 0; q < kDfUnroll; ++q)
          if (j + q < r.w && old[q] == 1) push(z[q], rz[q]);
      }
    }
    const unsigned long long pb = __ballot(v >= 0);
    if (lane == 0 && pb) atomicAdd(&s_proc, __popcll(pb));
    __syncthreads();

    // -- publish, spill the pool beyond kDfKeep, retire a finished window
    if (t == 0) {
      const int pool_now = min(s_pn, kDfPool);
      const int keep = min(pool_now, kDfKeep);  // >= s_pbase (<= kDfKeep)
      const int spill = pool_now - keep;
      s_spill = spill;
      if (spill > 0) s_qbase = atomicAdd(&ctl->qtail, static_cast<unsigned>(spill));
      s_pn = keep;
      my_r += static_cast<unsigned>(keep - s_pbase);
      my_p += static_cast<unsigned>(s_proc);
      st_dev64(a.dfw + blockIdx.x, (static_cast<unsigned long long>(my_r) << 32) | my_p);
      if (s_win && (s_wd[0] & s_wd[1] & s_wd[2] & s_wd[3]) == ~0ull) s_win = 0;
      if (s_proc == 0 && keep == 0) {
        ++idle;
      } else {
        idle = 0;
      }
      s_exit = 0;
      s_check = (idle & 3u) == 3u;
      if (idle > kSpinLimit) {
        atomicExch(&ctl->err, 1);
        s_exit = 1;
      } else if (idle > 0 && __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        s_exit = 1;
      }
    }
    __syncthreads();
    for (int i = t; i < s_spill; i += kDfBlock) st_dev32(a.queue + s_qbase + i, s_v[kDfKeep + i]);
    if (s_exit) break;
    if (s_proc == 0 && s_pn == 0) {
      // no work here: check for quiescence every fourth idle iteration
      // (thread 0 keeps the idle count; s_check is set before the barrier
      // that the spill loop above is followed by)
      __builtin_amdgcn_s_sleep(8);
      if (s_check) {
        unsigned long long sp = 0;
        for (int b = t; b < a.dfw_n; b += kDfBlock) sp += ld_dev64(a.dfw + b) & 0xffffffffull;
        sp = wave_sum64(sp);
        if (lane == 0) s_red[wid] = sp;
        __syncthreads();
        const unsigned long long settled = s_red[0] + s_red[1] + s_red[2] + s_red[3];
        __syncthreads();
        unsigned long long sr = 0;
        for (int b = t; b < a.dfw_n; b += kDfBlock) sr += ld_dev64(a.dfw + b) >> 32;
        if (t == 0) sr += ld_devu32(&ctl->qtail);
        sr = wave_sum64(sr);
        if (lane == 0) s_red[wid] = sr;
        __syncthreads();
        const unsigned long long released = s_red[0] + s_red[1] + s_red[2] + s_red[3];
        if (settled == released) break;
        __syncthreads();
      }
    }
  }
  const int ml = wave_max_int(maxlev);
  const float mx = wave_max(local_max);
  if (lane == 0) {
    if (ml >= 0) atomicMax(&ctl->num_levels, ml + 1);
    atomicMax(&ctl->max_abs_distf_bits, __float_as_uint(mx));
  }
  if (t == 0) atomicAdd(&ctl->levelized, my_p);
}

// ---------------------------------------------------------------------------
// Local-queue level loop: the level loop's levels, dist and succ with no
// queue flush, no global queue read and no cache maintenance per level.
//
// A block settles the vertices it released itself during the previous
// level.  They wait in an LDS queue together with their fanin ranges, which
// the releasing thread loads beside its release atomics, so settling a
// vertex starts with its best[] word and its fanin records in one round
// trip.  Releases past the LDS queue spill to a global queue that every
// block shares out at the next level (slots -1 until written).  Level 0 is
// the sink list, shared out the same way.
//
// Everything one block reads of another block's work is read at device
// scope: best[] with device-scope atomic loads (the candidates arrive by
// atomicMin, a read-modify-write performed at the device's coherence
// point), spill slots with device-scope loads of device-scope stores.  A
// block arrives at the level's barrier once its own stores and atomics of
// the level have completed (s_waitcnt), and every read of the next level
// comes after the barrier.  So the barrier needs no release or acquire
// fence: the L2s keep the read-only CSR lines from level to level instead
// of writing back and invalidating at every level.
//
// The barrier's arrival word counts blocks in bits 56-63, sums the blocks'
// releases in bits 28-55 and their spills in bits 0-27 (graphs of fewer
// than 2^28 vertices).  The last block to arrive publishes the next
// generation with the level's spill count, bit 31 set when nothing was
// released (the loop ends).  Spills of level d+1 are exactly the slots
// [end of level d's spills, + count): they are all appended during level d.
// ---------------------------------------------------------------------------
namespace {
constexpr int kLocBlock = 256;
constexpr int kLocCap = 1536;   // LDS queue entries per level buffer
constexpr int kLocUnroll = 8;   // fanin records (loads and atomics) in flight per step
constexpr unsigned kLocMask = (1u << 28) - 1;

// Returns false when the spin limit was exceeded or another block raised
// the error flag; *spills = the next level's spill count, *more = whether
// any block released a vertex.
__device__ bool local_barrier(SfxtCtl* ctl, unsigned nblocks, unsigned pushed, unsigned spilled,
                              unsigned* spills, bool* more) {
  __builtin_amdgcn_s_waitcnt(0);  // this thread's stores and atomics of the level are done
  __syncthreads();
  __shared__ int s_ok;
  __shared__ unsigned s_low;
  if (threadIdx.x == 0) {
    int ok = 1;
    unsigned low = 0;
    const unsigned g = static_cast<unsigned>(ld_dev64(&ctl->bar_word) >> 32);
    const unsigned long long add = (1ull << 56) | (static_cast<unsigned long long>(pushed) << 28) | spilled;
    const unsigned long long old = __hip_atomic_fetch_add(&ctl->bar_arrive, add, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT);
    if (static_cast<unsigned>(old >> 56) == nblocks - 1) {
      const unsigned long long all = old + add;
      const unsigned tp = static_cast<unsigned>(all >> 28) & kLocMask;
      low = (static_cast<unsigned>(all) & kLocMask) | (tp == 0 ? 0x80000000u : 0u);
      st_dev64(&ctl->bar_arrive, 0ull);
      __builtin_amdgcn_s_waitcnt(0);  // the reset lands before any block can arrive again
      st_dev64(&ctl->bar_word, (static_cast<unsigned long long>(g + 1u) << 32) | low);
    } else {
      unsigned spins = 0;
      for (;;) {
        const unsigned long long w = ld_dev64(&ctl->bar_word);
        if (static_cast<unsigned>(w >> 32) != g) {
          low = static_cast<unsigned>(w);
          break;
        }
        if (ld_dev32(&ctl->err) != 0) { ok = 0; break; }
        if (++spins > kSpinLimit) { atomicExch(&ctl->err, 1); ok = 0; break; }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (ld_dev32(&ctl->err) != 0) ok = 0;
    s_low = low;
    s_ok = ok;
  }
  __syncthreads();
  *spills = s_low & kLocMask;
  *more = (s_low & 0x80000000u) == 0;
  return s_ok != 0;
}
}  // namespace

// Per-vertex state before the local-queue loop; spill slots start at -1.
__global__ void __launch_bounds__(256) sfxt_local_init_kernel(SfxtArgs a) {
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int gthreads = gridDim.x * blockDim.x;
  for (int v = gtid; v < a.n; v += gthreads) {
    a.deps[v] = a.oadjp[v + 1] - a.oadjp[v];
    a.level[v] = -1;
    a.dist[v] = kIntMax;
    a.succ[v] = -1;
    a.distf[v] = static_cast<float>(kIntMax) / static_cast<float>(kScaleUp);
    a.best[v] = ~0ull;
    a.queue[v] = -1;
  }
}

__global__ void __launch_bounds__(kLocBlock) sfxt_local_kernel(SfxtArgs a) {
  __shared__ int s_x[2][kLocCap];
  __shared__ int s_b[2][kLocCap];
  __shared__ int s_e[2][kLocCap];
  __shared__ int s_n[2];
  __shared__ unsigned s_done;
  SfxtCtl* ctl = a.ctl;
  const unsigned G = gridDim.x;
  const unsigned bid = blockIdx.x;
  const int t = threadIdx.x;
  if (t == 0) {
    s_n[0] = 0;
    s_n[1] = 0;
    s_done = 0;
  }
  __syncthreads();
  unsigned sp_lo = 0, sp_n = 0;  // this level's spill slots [sp_lo, sp_lo + sp_n)
  int d = 0;
  unsigned settled = 0;
  float local_max = 0.0f;
  bool ok = true;
  while (true) {
    const int cur = d & 1, nxt = cur ^ 1;
    if (t == 0) s_n[nxt] = 0;
    __syncthreads();
    unsigned nloc, q0, q1;
    if (d == 0) {
      nloc = 0;
      q0 = static_cast<unsigned>(static_cast<unsigned long long>(a.nsinks) * bid / G);
      q1 = static_cast<unsigned>(static_cast<unsigned long long>(a.nsinks) * (bid + 1) / G);
    } else {
      nloc = static_cast<unsigned>(min(s_n[cur], kLocCap));
      q0 = sp_lo + static_cast<unsigned>(static_cast<unsigned long long>(sp_n) * bid / G);
      q1 = sp_lo + static_cast<unsigned>(static_cast<unsigned long long>(sp_n) * (bid + 1) / G);
    }
    const unsigned items = nloc + (q1 - q0);

    auto push = [&](int x, int xb, int xe) {
      const int p = atomicAdd(&s_n[nxt], 1);
      if (p < kLocCap) {
        s_x[nxt][p] = x;
        s_b[nxt][p] = xb;
        s_e[nxt][p] = xe;
      } else {
        const unsigned q = atomicAdd(&ctl->qtail, 1u);
        st_dev32(a.queue + q, x);
      }
    };

    for (unsigned i = static_cast<unsigned>(t); i < items; i += kLocBlock) {
      int u, b, e;
      if (i < nloc) {
        u = s_x[cur][i];
        b = s_b[cur][i];
        e = s_e[cur][i];
      } else {
        const unsigned qi = q0 + (i - nloc);
        if (d == 0) {
          u = a.sinks[qi];
        } else {
          u = ld_dev32(a.queue + qi);
          unsigned spins = 0;
          while (u < 0) {  // written before the barrier; a store still landing
            if (++spins > kSpinLimit) { atomicExch(&ctl->err, 1); break; }
            __builtin_amdgcn_s_sleep(1);
            u = ld_dev32(a.queue + qi);
          }
          if (u < 0) continue;
        }
        b = a.iadjp[u];
        e = a.iadjp[u + 1];
      }
      int du = 0, s = -1;
      if (d == 0) {
        if (a.sink_mask && !a.sink_mask[u]) du = kIntMax;
      } else {
        const unsigned long long key = ld_dev64(a.best + u);
        if (key == ~0ull) {
          du = kIntMax;
        } else {
          du = unpack_cand(key);
          s = unpack_succ(key);
        }
      }
      const bool reach = du != kIntMax;
      const float df = static_cast<float>(du) / static_cast<float>(kScaleUp);
      const int tie = a.orig ? a.orig[u] : u;  // successor tie-break: original id
      a.dist[u] = du;
      a.succ[u] = s;
      a.level[u] = d;
      a.distf[u] = df;
      ++settled;
      if (reach) local_max = fmaxf(local_max, fabsf(df));
      for (int j = b; j < e; j += kLocUnroll) {
        // kLocUnroll fanin records; past the end the last one is loaded
        // again (every load unconditional) and its atomics are skipped.  The
        // fanin's CSR range is loaded beside the atomics, for the LDS queue.
        int2 r[kLocUnroll];
#pragma unroll
        for (int q = 0; q < kLocUnroll; ++q) r[q] = a.iedge[min(j + q, e - 1)];
        int xb[kLocUnroll], xe[kLocUnroll], old[kLocUnroll];
#pragma unroll
        for (int q = 0; q < kLocUnroll; ++q) {
          xb[q] = a.iadjp[r[q].x];
          xe[q] = a.iadjp[r[q].x + 1];
          // branch-free: a slot past the end issues no-op atomics on a pad
          // slot after the n vertices, spread by thread and slot
          const bool in = j + q < e;
          const int wi = static_cast<int>(__int_as_float(r[q].y) * static_cast<float>(kScaleUp));
          const int x = in ? r[q].x
                           : a.n + static_cast<int>(((bid * kLocBlock + static_cast<unsigned>(t)) * kLocUnroll +
                                                     static_cast<unsigned>(q)) &
                                                    static_cast<unsigned>(kSfxtPad - 1));
          atomicMin(&a.best[x], in && reach ? pack_cand(du + wi, tie) : ~0ull);
          old[q] = atomicSub(&a.deps[x], in ? 1 : 0);
        }
#pragma unroll
        for (int q = 0; q < kLocUnroll; ++q)
          if (j + q < e && old[q] == 1) push(r[q].x, xb[q], xe[q]);
      }
    }
    __syncthreads();
    const int pushed = s_n[nxt];
    const unsigned spilled = pushed > kLocCap ? static_cast<unsigned>(pushed - kLocCap) : 0u;
    unsigned next_spills = 0;
    bool more = false;
    if (!local_barrier(ctl, G, static_cast<unsigned>(pushed), spilled, &next_spills, &more)) {
      ok = false;
      break;
    }
    if (!more) break;
    sp_lo += sp_n;
    sp_n = next_spills;
    ++d;
  }
  const float m = wave_max(local_max);
  if (__lane_id() == 0) atomicMax(&ctl->max_abs_distf_bits, __float_as_uint(m));
  atomicAdd(&s_done, settled);
  __syncthreads();
  if (t == 0) atomicAdd(&ctl->levelized, s_done);
  if (ok && bid == 0 && t == 0) ctl->num_levels = a.nsinks > 0 ? d + 1 : 0;
}

// Relabelled graph: the level loops store the successor's original id;
// this maps it to the relabelled id before the deviation index is built.
__global__ void __launch_bounds__(256) sfxt_succ_relabel_kernel(int* succ, const int* perm, int n) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const int s = succ[v];
    if (s >= 0) succ[v] = perm[s];
  }
}

// Per-vertex state before the multi-subset level loop.
__global__ void __launch_bounds__(256) sfxt_init_multi_kernel(SfxtArgs a, SfxtMulti m) {
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int gthreads = gridDim.x * blockDim.x;
  const size_t n = static_cast<size_t>(a.n);
  for (int v = gtid; v < a.n; v += gthreads) {
    a.deps[v] = a.oadjp[v + 1] - a.oadjp[v];
    a.level[v] = -1;
    for (int q = 0; q < m.q; ++q) {
      m.dv[static_cast<size_t>(v) * m.q + q] = kIntMax;
      m.dist[q * n + v] = kIntMax;
      m.succ[q * n + v] = -1;
      m.distf[q * n + v] = static_cast<float>(kIntMax) / static_cast<float>(kScaleUp);
    }
  }
  // the queue: the sinks, then -1 until the level loop writes a slot
  for (int v = gtid; v < a.n; v += gthreads) a.queue[v] = v < a.nsinks ? a.sinks[v] : -1;
  for (int q = gtid; q < m.q; q += gthreads) m.maxd[q] = 0u;
  if (gtid == 0) a.ctl->qtail = static_cast<unsigned>(a.nsinks);
}

// The level loop for Q endpoint subsets at once.  Levels, the queue and
// the fanin release are the single loop's; one wave takes a vertex and
// settles it in all Q subsets pull-style, as relax_bu_step and
// update_successors_bu read the fanouts (gpucpg/gpucpg.cu:1167-1187,
// :2606-2628): lane l reads fanout record l / Q in subset l % Q from the
// vertex-major distances dv[w*Q + q] of the settled fanouts, and a
// shuffle reduction over each subset's lanes gives the minimum and the
// largest fanout id at it.  No per-subset atomics: the only atomics are
// the fanin releases, shared by the subsets.  The per-level barrier is
// shared by all Q subsets, so Q subsets cost about one loop.  Fence-free:
// queue entries and dv[] are written and read at device scope, and each
// level ends at the counter barrier (the queue must hold the sinks and -1
// beyond, as sfxt_init_multi leaves it).
// It runs on the blocks [b0, b0 + nblk) of a launch that holds the level
// loops of several graphs (one partition each, with its own barrier).
__device__ void levels_multi_part(const SfxtArgs& a, const SfxtMulti& m, unsigned bid, unsigned nblk) {
  __shared__ int s_q[kLdsQueue];
  __shared__ int s_qn;
  SfxtCtl* ctl = a.ctl;
  const unsigned nblocks = nblk;
  const int lane = __lane_id();
  const int wave = static_cast<int>((bid * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = static_cast<int>((nblk * blockDim.x) >> 6);
  const int Q = m.q;
  const int R = 64 / Q;            // fanin records per wave step
  const int sub = lane & (Q - 1);  // this lane's subset in the record loop
  const int rec = lane / Q;        // this lane's record in the step
  const size_t n = static_cast<size_t>(a.n);
  if (threadIdx.x == 0) s_qn = 0;
  __syncthreads();

  unsigned lb = 0, le = static_cast<unsigned>(a.nsinks);
  int d = 0;
  LevelCounts lc;
  auto push = [&](int x) {
    const int p = atomicAdd(&s_qn, 1);
    if (p < kLdsQueue) s_q[p] = x;
    else st_dev32(a.queue + counter_reserve(ctl, d, le, lc), x);  // past the LDS queue
  };

  float local_max = 0.0f;  // lanes < Q: subset lane's largest finite |distf|
  while (true) {
    if (bid == 0 && threadIdx.x == 0) a.lvlp[d] = static_cast<int>(lb);
    for (unsigned i = lb + static_cast<unsigned>(wave); i < le; i += static_cast<unsigned>(nwaves)) {
      const int u = queue_entry(a, i);
      if (u < 0) continue;  // (only after a spin-limit error; the barrier ends the loop)
      const int ob = a.oadjp[u], oe = a.oadjp[u + 1];
      const int ib = a.iadjp[u], ie = a.iadjp[u + 1];
      int du = kIntMax, s = -1;
      if (d == 0) {  // level 0 is the sink list in order: i is the sink's position
        if (lane < Q && (!m.masks || m.masks[static_cast<size_t>(lane) * a.nsinks + i])) du = 0;
      } else {
        // pull: lane l reads fanout record l / Q (64/Q at a time,
        // kMultiSteps steps in flight) in subset l % Q; every fanout sits on
        // a lower level, so its distance in every subset is settled
        int bd = kIntMax, bw = -1;
        for (int j0 = ob; j0 < oe; j0 += kMultiSteps * R) {
          int w[kMultiSteps], dw[kMultiSteps];
          float wf[kMultiSteps];
#pragma unroll
          for (int t = 0; t < kMultiSteps; ++t) {
            const int j = j0 + t * R + rec;
            w[t] = j < oe ? a.oadj[j] : -1;
            wf[t] = j < oe ? a.owgt[j] : 0.0f;
          }
#pragma unroll
          for (int t = 0; t < kMultiSteps; ++t)
            dw[t] = w[t] >= 0 ? ld_dev32(m.dv + static_cast<size_t>(w[t]) * Q + sub) : kIntMax;
#pragma unroll
          for (int t = 0; t < kMultiSteps; ++t) {
            if (dw[t] == kIntMax) continue;
            const int c = dw[t] + static_cast<int>(wf[t] * static_cast<float>(kScaleUp));
            if (c < bd || (c == bd && w[t] > bw)) {
              bd = c;
              bw = w[t];
            }
          }
        }
        // minimum over the lanes of one subset (lanes sub, sub + Q, ...),
        // the largest fanout id among equal minima
        for (int o = Q; o < 64; o <<= 1) {
          const int od = __shfl_xor(bd, o);
          const int ow = __shfl_xor(bw, o);
          if (od < bd || (od == bd && ow > bw)) {
            bd = od;
            bw = ow;
          }
        }
        du = bd;
        s = bw;
      }
      if (lane < Q) {
        const float df = static_cast<float>(du) / static_cast<float>(kScaleUp);
        m.dist[lane * n + u] = du;
        m.succ[lane * n + u] = s;
        m.distf[lane * n + u] = df;
        st_dev32(m.dv + static_cast<size_t>(u) * Q + lane, du);
        if (du != kIntMax) local_max = fmaxf(local_max, fabsf(df));
      }
      if (lane == 0) a.level[u] = d;
      // release: every lane takes fanin records (the subsets share the
      // levels, so each record is released once)
      for (int j0 = ib; j0 < ie; j0 += 64 * kMultiSteps) {
        if (ie - j0 <= 64) {
          // at most one record per lane (u, and so the branch, is uniform
          // over the wave): one atomic per lane that has a record, no no-ops
          const int j = j0 + lane;
          if (j < ie) {
            const int x = a.iedge[j].x;
            if (atomicSub(&a.deps[x], 1) == 1) push(x);
          }
          continue;
        }
        int x[kMultiSteps], old[kMultiSteps];
        // branch-free, so the atomics are in flight together: a slot past
        // the end subtracts 0 (a no-op) from a pad slot after the n
        // vertices, spread by lane and step (atomics on one address
        // serialise)
#pragma unroll
        for (int t = 0; t < kMultiSteps; ++t) {
          const int j = j0 + t * 64 + lane;
          x[t] = j < ie ? a.iedge[j].x
                        : a.n + static_cast<int>((static_cast<unsigned>(wave) * 64u * kMultiSteps +
                                                  static_cast<unsigned>(t * 64 + lane)) &
                                                 static_cast<unsigned>(kSfxtPad - 1));
        }
#pragma unroll
        for (int t = 0; t < kMultiSteps; ++t) old[t] = atomicSub(&a.deps[x[t]], j0 + t * 64 + lane < ie ? 1 : 0);
#pragma unroll
        for (int t = 0; t < kMultiSteps; ++t)
          if (j0 + t * 64 + lane < ie && old[t] == 1) push(x[t]);
      }
    }
    unsigned ne = 0;
    if (!counter_barrier(ctl, a.queue, s_q, &s_qn, nblocks, d, le, lc, &ne)) return;
    if (ne == le) break;
    lb = le;
    le = ne;
    ++d;
  }
  if (lane < Q) atomicMax(&m.maxd[lane], __float_as_uint(local_max));
  if (bid == 0 && threadIdx.x == 0) {
    a.lvlp[d + 1] = static_cast<int>(le);
    ctl->num_levels = le > 0 ? d + 1 : 0;
    ctl->levelized = le;
  }
}

// Level loops of several graphs in one resident launch: each block
// runs the loop of the partition its index falls in, and a partition's
// barrier counts only its own blocks, so every graph advances at its own
// pace while all of them are resident together.
__global__ void __launch_bounds__(kSfxtBlock) sfxt_levels_parts_kernel(SfxtParts parts) {
  __shared__ int s_part;
  if (threadIdx.x == 0) {
    int p = 0;
    while (p + 1 < parts.n && static_cast<int>(blockIdx.x) >= parts.p[p + 1].b0) ++p;
    s_part = p;
  }
  __syncthreads();
  const SfxtPart P = parts.p[s_part];
  levels_multi_part(P.a, P.m, blockIdx.x - static_cast<unsigned>(P.b0), static_cast<unsigned>(P.nb));
}

// Logical index of a block such that the blocks the hardware places on one
// XCD (round-robin: block b on XCD b % 8) take consecutive indices, so that
// a kernel giving each block one contiguous chunk gives each XCD one
// contiguous range (identity when the grid is not a multiple of 8).
__device__ inline unsigned xcd_block(unsigned b, unsigned nb) {
  constexpr unsigned kX = 8;
  return nb % kX == 0 ? (b % kX) * (nb / kX) + b / kX : b;
}

// Deviation records of every settled vertex in fanout order, their smallest
// delta, and the vertex record without its nsv pointer.  Lists of up to
// kThreadSort records are sorted here (stable insertion sort by delta);
// longer ones are queued for sfxt_sort_mid_kernel (up to kWaveSort, from
// the front of big[]) or sfxt_sort_big_kernel (from the back).  The loop
// is wave-uniform so the queue reservations are wave-aggregated.
// Each block takes one contiguous chunk of vertices and each XCD one range
// of chunks (xcd_block): a vertex's fanouts sit on the levels just below
// it, and with level-ordered ids their dist/distf words then lie in the
// range the same XCD's L2 already holds.
__global__ void __launch_bounds__(256) sfxt_devs_kernel(SfxtArgs a) {
  const int chunk = static_cast<int>(((static_cast<int64_t>(a.n) + gridDim.x - 1) / gridDim.x + 255) / 256 * 256);
  const int64_t lo64 = static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * chunk;
  const int lo = static_cast<int>(lo64 < a.n ? lo64 : a.n);
  const int hi = static_cast<int>(lo64 + chunk < a.n ? lo64 + chunk : a.n);
  for (int base = lo; base < hi; base += blockDim.x) {
    const int u = base + threadIdx.x;
    int cnt = 0;
    bool sorted = true;
    if (u < hi) {
      const int b = a.oadjp[u], e = a.oadjp[u + 1];
      if (a.level[u] < 0 || a.dist[u] == kIntMax) {  // unsettled or reaches no chosen sink
        a.vrec[u] = default_vrec();
        a.devinfo[u] = make_int2(b, 0);
      } else {
        const int s = a.succ[u];
        const float df = a.distf[u];
        float mind = __builtin_huge_valf();
        for (int j = b; j < e; ++j) {
          const int w = a.oadj[j];
          if (w == s) continue;
          if (a.dist[w] == kIntMax) continue;
          const float fw = a.distf[w];
          const float we = a.owgt[j];
          const float delta = static_cast<float>(static_cast<double>(fw) +
                                                 static_cast<double>(we) -
                                                 static_cast<double>(df));
          a.devs[b + cnt] =
              make_int4(w, __float_as_int(we), __float_as_int(fw), __float_as_int(delta));
          mind = fminf(mind, delta);
          ++cnt;
        }
        if (cnt <= kThreadSort) {
          for (int i = 1; i < cnt; ++i) {
            const int4 rec = a.devs[b + i];
            const float key = __int_as_float(rec.w);
            int k = i;
            while (k > 0) {
              const int4 prev = a.devs[b + k - 1];
              if (__int_as_float(prev.w) <= key) break;
              a.devs[b + k] = prev;
              --k;
            }
            if (k != i) a.devs[b + k] = rec;
          }
        } else {
          sorted = false;
        }
        a.vrec[u] = make_int4(__float_as_int(mind), -1, s, __float_as_int(df));
        a.devinfo[u] = make_int2(b, cnt);
      }
    }
    const bool mid = !sorted && cnt <= kWaveSort;
    const bool large = !sorted && cnt > kWaveSort;
    const uint64_t sm = wave_reserve(reinterpret_cast<unsigned long long*>(&a.ctl->nbig), mid);
    const uint64_t sl = wave_reserve(reinterpret_cast<unsigned long long*>(&a.ctl->nbig_l), large);
    if (mid) a.big[sm] = u;
    if (large) a.big[a.n - 1 - static_cast<int>(sl)] = u;
  }
}

// Total order used to rank deviation records: float order with -0 == +0
// (ties keep fanout order, as the insertion sort does); NaN ranks last.
__device__ inline uint32_t delta_rank_key(int bits) {
  const float x = __int_as_float(bits);
  return x == 0.0f ? kKeyPosZero : float_key(x);
}

// Deviation records with one wave per vertex, for graphs whose fanouts
// are long on average: lanes take the fanout edges 64 at a time and a
// ballot compacts the deviation records in fanout order.  A fanout of at
// most 128 edges keeps its records in registers (two per lane), ranks them
// there (ties in fanout order, as sfxt_sort_mid_kernel) and stores each
// once, sorted: each valid record's key is broadcast once and compared
// with both of a lane's keys, so ranking costs one shuffle per record.
// Longer fanouts store their records in fanout order and are queued for
// sfxt_sort_big_kernel.  vrec and devinfo as in sfxt_devs_kernel.
__device__ inline bool devs_record(const SfxtArgs& a, int j, int s, float df, int4& rec) {
  const int w = a.oadj[j];
  if (w == s || a.dist[w] == kIntMax) return false;
  const float fw = a.distf[w];
  const float we = a.owgt[j];
  const float delta = static_cast<float>(static_cast<double>(fw) + static_cast<double>(we) -
                                         static_cast<double>(df));
  rec = make_int4(w, __float_as_int(we), __float_as_int(fw), __float_as_int(delta));
  return true;
}

__global__ void __launch_bounds__(256) sfxt_devs_wave_kernel(SfxtArgs a) {
  const int lane = __lane_id();
  const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = static_cast<int>((gridDim.x * blockDim.x) >> 6);
  for (int u = wave; u < a.n; u += nwaves) {
    const int b = a.oadjp[u], e = a.oadjp[u + 1];
    if (a.level[u] < 0 || a.dist[u] == kIntMax) {  // unsettled or reaches no chosen sink
      if (lane == 0) {
        a.vrec[u] = default_vrec();
        a.devinfo[u] = make_int2(b, 0);
      }
      continue;
    }
    const int s = a.succ[u];
    const float df = a.distf[u];
    int cnt = 0;
    float mind = __builtin_huge_valf();
    if (e - b <= 128) {
      int4 r0 = make_int4(0, 0, 0, 0), r1 = make_int4(0, 0, 0, 0);
      const bool ok0 = b + lane < e && devs_record(a, b + lane, s, df, r0);
      const bool ok1 = b + 64 + lane < e && devs_record(a, b + 64 + lane, s, df, r1);
      if (ok0) mind = fminf(mind, __int_as_float(r0.w));
      if (ok1) mind = fminf(mind, __int_as_float(r1.w));
      const uint64_t m0 = __ballot(ok0), m1 = __ballot(ok1);
      cnt = static_cast<int>(__popcll(m0) + __popcll(m1));
      const uint32_t k0 = ok0 ? delta_rank_key(r0.w) : 0u;
      const uint32_t k1 = ok1 ? delta_rank_key(r1.w) : 0u;
      // rank = records ordering before this one: a smaller key, or an
      // equal key earlier in fanout order (every first-half record is
      // earlier than every second-half one)
      int rank0 = 0, rank1 = 0;
      for (uint64_t mm = m0; mm; mm &= mm - 1) {
        const int jj = static_cast<int>(__builtin_ctzll(mm));
        const uint32_t kj = __shfl(k0, jj);
        rank0 += (kj < k0 || (kj == k0 && jj < lane)) ? 1 : 0;
        rank1 += kj <= k1 ? 1 : 0;
      }
      for (uint64_t mm = m1; mm; mm &= mm - 1) {
        const int jj = static_cast<int>(__builtin_ctzll(mm));
        const uint32_t kj = __shfl(k1, jj);
        rank0 += kj < k0 ? 1 : 0;
        rank1 += (kj < k1 || (kj == k1 && jj < lane)) ? 1 : 0;
      }
      if (ok0) a.devs[b + rank0] = r0;
      if (ok1) a.devs[b + rank1] = r1;
    } else {
      for (int j0 = b; j0 < e; j0 += 64) {
        int4 rec = make_int4(0, 0, 0, 0);
        const bool ok = j0 + lane < e && devs_record(a, j0 + lane, s, df, rec);
        if (ok) mind = fminf(mind, __int_as_float(rec.w));
        const uint64_t bm = __ballot(ok);
        if (ok) a.devs[b + cnt + static_cast<int>(__popcll(bm & lanemask_lt()))] = rec;  // sorted later
        cnt += static_cast<int>(__popcll(bm));
      }
      if (lane == 0) {
        const unsigned long long sl = atomicAdd(&a.ctl->nbig_l, 1ull);
        a.big[a.n - 1 - static_cast<int>(sl)] = u;
      }
    }
    for (int o = 32; o > 0; o >>= 1) mind = fminf(mind, __shfl_xor(mind, o));
    if (lane == 0) {
      a.vrec[u] = make_int4(__float_as_int(mind), -1, s, __float_as_int(df));
      a.devinfo[u] = make_int2(b, cnt);
    }
  }
}

// Stable sort by delta of the lists of kThreadSort+1 .. kWaveSort records,
// one wave per list: lane i holds record i and its rank is the number of
// records that order before it (shuffled keys), so the records land in
// place with one store each.
__global__ void __launch_bounds__(256) sfxt_sort_mid_kernel(SfxtArgs a) {
  const unsigned nmid = static_cast<unsigned>(a.ctl->nbig);
  const int lane = __lane_id();
  const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const unsigned nwaves = (gridDim.x * blockDim.x) >> 6;
  for (unsigned t = wave; t < nmid; t += nwaves) {
    const int u = a.big[t];
    const int2 di = a.devinfo[u];
    const int b = di.x, cnt = di.y;
    int4 rec = make_int4(0, 0, 0, 0);
    uint32_t key = 0xffffffffu;
    if (lane < cnt) {
      rec = a.devs[b + lane];
      key = delta_rank_key(rec.w);
    }
    int rank = 0;
    for (int j = 0; j < cnt; ++j) {
      const uint32_t kj = __shfl(key, j);
      rank += (kj < key || (kj == key && j < lane)) ? 1 : 0;
    }
    if (lane < cnt) a.devs[b + rank] = rec;
  }
}

// Stable sort by delta of the lists longer than kWaveSort, one block per
// list (ranks as in sfxt_sort_mid_kernel).  Lists of up to kLdsSort records
// are ranked in LDS; longer ones (only when the graph has such fanouts)
// rank from global memory and permute through devs_tmp, the same slot
// range of a second m-record buffer.
__global__ void __launch_bounds__(kSortBlock) sfxt_sort_big_kernel(SfxtArgs a) {
  __shared__ int4 s_rec[kLdsSort];
  const unsigned nl = static_cast<unsigned>(a.ctl->nbig_l);
  for (unsigned t = blockIdx.x; t < nl; t += gridDim.x) {
    const int u = a.big[a.n - 1 - static_cast<int>(t)];
    const int2 di = a.devinfo[u];
    const int b = di.x, cnt = di.y;
    if (cnt <= kLdsSort) {
      for (int i = threadIdx.x; i < cnt; i += blockDim.x) s_rec[i] = a.devs[b + i];
      __syncthreads();
      for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const uint32_t key = delta_rank_key(s_rec[i].w);
        int rank = 0;
        for (int j = 0; j < cnt; ++j) {
          const uint32_t kj = delta_rank_key(s_rec[j].w);
          rank += (kj < key || (kj == key && j < i)) ? 1 : 0;
        }
        a.devs[b + rank] = s_rec[i];
      }
      __syncthreads();
    } else if (a.devs_tmp) {
      for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const int4 rec = a.devs[b + i];
        const uint32_t key = delta_rank_key(rec.w);
        int rank = 0;
        for (int j = 0; j < cnt; ++j) {
          const uint32_t kj = delta_rank_key(a.devs[b + j].w);
          rank += (kj < key || (kj == key && j < i)) ? 1 : 0;
        }
        a.devs_tmp[b + rank] = rec;
      }
      __syncthreads();
      for (int i = threadIdx.x; i < cnt; i += blockDim.x) a.devs[b + i] = a.devs_tmp[b + i];
      __syncthreads();
    } else if (threadIdx.x == 0) {
      atomicExch(&a.ctl->nsv_err, 2);  // host did not provide devs_tmp
    }
  }
}

// lift[0][u] = {succ[u], smallest delta of succ[u]} ({-1, +inf} at a chain end).
__global__ void __launch_bounds__(256) sfxt_lift0_kernel(SfxtArgs a) {
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < a.n; u += gridDim.x * blockDim.x) {
    const int s = a.succ[u];
    a.lift[u] = s < 0 ? make_int2(-1, __float_as_int(__builtin_huge_valf()))
                      : make_int2(s, a.vrec[s].x);
  }
}

// lift[k][u] from lift[k-1]: jump twice, minimum over both segments.
__global__ void __launch_bounds__(256) sfxt_liftk_kernel(SfxtArgs a, int k) {
  const size_t n = static_cast<size_t>(a.n);
  const int2* prev = a.lift + static_cast<size_t>(k - 1) * n;
  int2* cur = a.lift + static_cast<size_t>(k) * n;
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < a.n; u += gridDim.x * blockDim.x) {
    const int2 p = prev[u];
    if (p.x < 0) {
      cur[u] = p;
    } else {
      const int2 q = prev[p.x];
      cur[u] = make_int2(q.x, __float_as_int(fminf(__int_as_float(p.y), __int_as_float(q.y))));
    }
  }
}

// nsv[u]: greedy descent over the lift levels.  At level k the walk jumps
// 2^k chain vertices when none of them has a smallest delta strictly below
// t = u's; the chain vertex after the final position is the answer.  With
// 2^levels above the longest chain, the descent lands on the vertex just
// before the first qualifying one, or runs off the chain end (no answer).
__global__ void __launch_bounds__(256) sfxt_nsv_kernel(SfxtArgs a, int levels) {
  const size_t n = static_cast<size_t>(a.n);
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < a.n; u += gridDim.x * blockDim.x) {
    int4 r = a.vrec[u];
    if (r.z < 0) continue;  // chain end or unsettled: nsv stays -1
    const float t = __int_as_float(r.x);
    int x = u;
    for (int k = levels - 1; k >= 0 && x >= 0; --k) {
      const int2 e = a.lift[static_cast<size_t>(k) * n + x];
      if (!(__int_as_float(e.y) < t)) x = e.x;  // skip 2^k vertices (or fall off the end)
    }
    int nsv = -1;
    if (x >= 0) {
      const int2 e0 = a.lift[x];
      if (e0.x >= 0 && __int_as_float(e0.y) < t) nsv = e0.x;
      else atomicExch(&a.ctl->nsv_err, 1);  // cannot happen with enough levels
    }
    r.y = nsv;
    a.vrec[u] = r;
  }
}

namespace {

int device_cus() {
  int dev = 0, cus = 0;
  CPG_HIP(hipGetDevice(&dev));
  CPG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus;
}

int flat_grid(int n, int cus) {
  return std::max(1, std::min((n + 255) / 256, cus * 8));
}

// Kernels with a grid barrier launch as ordinary kernels on a grid of at
// most half the blocks the device keeps resident at once, so every block
// of the grid is dispatched at once even beside another kernel of this
// process.  (A cooperative launch would promise the same, but on this
// platform it fails with "cannot be performed in the present state" while
// another process on the GPU holds the cooperative-launch resources, e.g. a
// driver started from a process that has made cooperative launches.)  A
// barrier wait that never completes ends at its spin limit with an error,
// never in a hang.
template <typename K, typename A>
void launch_resident(K kernel, int grid, int block, const A& args, hipStream_t stream) {
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, args);
  CPG_HIP(hipGetLastError());
}

template <typename K>
int resident_blocks(K kernel, int block) {
  int per_cu = 0;
  CPG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
  if (per_cu < 1) throw std::runtime_error("grid-barrier kernel cannot be resident");
  return std::max(1, device_cus() * per_cu / 2);
}

}  // namespace

int sfxt_lift_levels(int num_levels) {
  int k = 1;
  while ((1 << k) < num_levels + 1 && k < 30) ++k;
  return k;
}

void launch_sfxt_init(const SfxtArgs& args, hipStream_t stream) {
  hipLaunchKernelGGL(sfxt_init_kernel, dim3(flat_grid(args.n, device_cus())), dim3(256), 0, stream,
                     args);
  CPG_HIP(hipGetLastError());
}

// Level-loop grid: the blocks asked for, within the resident capacity.
static int loop_grid(int blocks, int resident) { return std::max(1, std::min(blocks, resident)); }

int launch_sfxt_levels(const SfxtArgs& args, int blocks, hipStream_t stream) {
  const int grid = loop_grid(blocks, resident_blocks(sfxt_levels_kernel, kSfxtBlock));
  launch_resident(sfxt_levels_kernel, grid, kSfxtBlock, args, stream);
  return grid;
}

int launch_sfxt_levels_nf(const SfxtArgs& args, int blocks, hipStream_t stream) {
  // mean in-degree m / n: under 4 -> 4 records per step, else 8
  const int64_t m = args.m, n = std::max(1, args.n);
  auto k = m < 4 * n ? sfxt_levels_nf4_kernel : sfxt_levels_nf_kernel;
  const int grid = loop_grid(blocks, resident_blocks(k, kSfxtBlock));
  launch_resident(k, grid, kSfxtBlock, args, stream);
  return grid;
}

int launch_sfxt_levels_edges(const SfxtArgs& args, int blocks, hipStream_t stream) {
  // mean in-degree m / n: under 4 -> 3 list positions per lane step, else 8
  const int64_t m = args.m, n = std::max(1, args.n);
  auto k = m < 4 * n ? sfxt_levels_e3_kernel : sfxt_levels_e8_kernel;
  // place_stride launch blocks per working block (the others return at once)
  const int stride = std::max(1, args.place_stride);
  const int working = loop_grid(blocks, resident_blocks(k, kSfxtBlock) / stride);
  launch_resident(k, working * stride, kSfxtBlock, args, stream);
  return working;
}

int sfxt_dataflow_grid(int blocks) {
  int per_cu = 0;
  CPG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sfxt_dataflow_kernel, kDfBlock, 0));
  if (per_cu < 1) throw std::runtime_error("sfxt dataflow kernel cannot be resident");
  const int cus = device_cus();
  const int want = blocks > 0 ? blocks : cus;
  return std::max(1, std::min(want, cus * per_cu));
}

void launch_sfxt_dataflow_init(const SfxtArgs& args, hipStream_t stream) {
  const int items = std::max(args.n, args.dfw_n);
  hipLaunchKernelGGL(sfxt_dataflow_init_kernel, dim3(flat_grid(items, device_cus())), dim3(256), 0,
                     stream, args);
  CPG_HIP(hipGetLastError());
}

// An ordinary launch: no block waits for a block that has not started (a
// wait on a fanout's state word ends once that vertex's store lands, and
// an idle block leaves at quiescence), so the loop needs no co-residency
// guarantee.  A cooperative launch would also tie up the device's
// cooperative-launch resources, which another process on the same GPU may
// hold.
int launch_sfxt_dataflow(const SfxtArgs& args, int blocks, hipStream_t stream) {
  const int grid = sfxt_dataflow_grid(blocks);
  if (args.dfw_n < grid || !args.dfw) throw std::runtime_error("dataflow progress words too few");
  hipLaunchKernelGGL(sfxt_dataflow_kernel, dim3(grid), dim3(kDfBlock), 0, stream, args);
  CPG_HIP(hipGetLastError());
  return grid;
}

void launch_sfxt_succ_relabel(int* succ, const int* perm, int n, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(sfxt_succ_relabel_kernel, dim3(flat_grid(n, device_cus())), dim3(256), 0, stream,
                     succ, perm, n);
  CPG_HIP(hipGetLastError());
}

void launch_sfxt_local_init(const SfxtArgs& args, hipStream_t stream) {
  hipLaunchKernelGGL(sfxt_local_init_kernel, dim3(flat_grid(args.n, device_cus())), dim3(256), 0,
                     stream, args);
  CPG_HIP(hipGetLastError());
}

int launch_sfxt_local(const SfxtArgs& args, int blocks, hipStream_t stream) {
  if (args.n >= (1 << 28)) throw std::runtime_error("local-queue level loop: graph too large");
  const int grid = std::max(1, std::min({blocks, resident_blocks(sfxt_local_kernel, kLocBlock), 255}));
  launch_resident(sfxt_local_kernel, grid, kLocBlock, args, stream);
  return grid;
}

void launch_sfxt_init_multi(const SfxtArgs& args, const SfxtMulti& m, hipStream_t stream) {
  hipLaunchKernelGGL(sfxt_init_multi_kernel, dim3(flat_grid(args.n, device_cus())), dim3(256), 0,
                     stream, args, m);
  CPG_HIP(hipGetLastError());
}

int sfxt_parts_capacity() { return resident_blocks(sfxt_levels_parts_kernel, kSfxtBlock); }

void launch_sfxt_levels_parts(const SfxtParts& parts, int total_blocks, hipStream_t stream) {
  if (parts.n < 1 || parts.n > kMaxParts || total_blocks > sfxt_parts_capacity())
    throw std::runtime_error("level-loop partitions do not fit one resident launch");
  for (int i = 0; i < parts.n; ++i) {
    const int end = i + 1 < parts.n ? parts.p[i + 1].b0 : total_blocks;
    if (parts.p[i].nb < 1 || parts.p[i].b0 + parts.p[i].nb != end || (i == 0 && parts.p[0].b0 != 0))
      throw std::runtime_error("level-loop partitions must tile the launch");
  }
  launch_resident(sfxt_levels_parts_kernel, total_blocks, kSfxtBlock, parts, stream);
}

void launch_sfxt_index(const SfxtArgs& args, int lift_levels, int64_t m, hipStream_t stream) {
  const int cus = device_cus();
  const int g = flat_grid(args.n, cus);
  if (args.n > 0 && m >= 8 * static_cast<int64_t>(args.n)) {
    // mean fanout of 8 or more: one wave per vertex
    const int64_t wb = (static_cast<int64_t>(args.n) + 3) / 4;
    const int gw = static_cast<int>(std::min<int64_t>(wb, static_cast<int64_t>(cus) * 64));
    hipLaunchKernelGGL(sfxt_devs_wave_kernel, dim3(gw), dim3(256), 0, stream, args);
  } else {
    hipLaunchKernelGGL(sfxt_devs_kernel, dim3(g), dim3(256), 0, stream, args);
  }
  CPG_HIP(hipGetLastError());
  hipLaunchKernelGGL(sfxt_sort_mid_kernel, dim3(cus * 8), dim3(256), 0, stream, args);
  CPG_HIP(hipGetLastError());
  hipLaunchKernelGGL(sfxt_sort_big_kernel, dim3(cus * 2), dim3(kSortBlock), 0, stream, args);
  CPG_HIP(hipGetLastError());
  hipLaunchKernelGGL(sfxt_lift0_kernel, dim3(g), dim3(256), 0, stream, args);
  CPG_HIP(hipGetLastError());
  for (int k = 1; k < lift_levels; ++k) {
    hipLaunchKernelGGL(sfxt_liftk_kernel, dim3(g), dim3(256), 0, stream, args, k);
    CPG_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(sfxt_nsv_kernel, dim3(g), dim3(256), 0, stream, args, lift_levels);
  CPG_HIP(hipGetLastError());
}

}  // namespace cpg
