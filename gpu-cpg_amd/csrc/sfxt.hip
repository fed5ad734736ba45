// SFXT: levelize + relax + successor + distf + deviation index, one
// cooperative persistent kernel with a grid barrier between levels.
//
// Semantics follow the reference path:
//   levelize    bfs_td_step_privatized   gpucpg/gpucpg.cu:2215-2283, loop :17642-17681
//               (a vertex joins level d+1 when its last fanout is consumed)
//   relax       relax_bu_step            gpucpg/gpucpg.cu:1167-1187, loop :17933-17961
//               dist[v] = min_e dist[w] + (int)(w_e * 10000.0f), sinks 0 (:17035-17039)
//   successor   update_successors_bu     gpucpg/gpucpg.cu:2606-2628 (largest id at the min)
//   distf       (float)dist / SCALE_UP, correctly rounded (used at :2672-2673)
// Because every fanout of a level-d vertex sits on a lower level, relax and
// successor for level d run in the same pass that peels level d.
//
// The deviation index is this engine's own structure for the PFXT walk:
// per vertex u, its deviation edges (u->w, w != succ[u], dist[w] finite:
// gpucpg/gpucpg.cu:2662-2668) sorted by delta = distf[w] + w_e - distf[u],
// the smallest delta, and nsv[u] = the first vertex after u on u's suffix
// chain whose smallest delta is strictly smaller (next-smaller-value
// pointer), so a walk can skip chain segments with no cheap deviation.
#include "common.hpp"
#include "kernels.hpp"

namespace cpg {

namespace {

constexpr int kSfxtBlock = 256;
constexpr int kLdsQueue = 2048;
constexpr unsigned kSpinLimit = 1u << 22;

__device__ inline unsigned atomic_load_rlx(unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Generation barrier over all co-resident blocks of a cooperative launch.
// One release fence before arriving and one acquire fence after leaving;
// the wait polls with relaxed loads so it does not invalidate caches on
// every probe.  A spin that exceeds kSpinLimit raises ctl->err and every
// later barrier returns false at once, so the kernel always drains.
// Returns false on error; *qtail receives the queue tail read after the
// barrier by one thread and broadcast through LDS.
__device__ bool grid_barrier(SfxtCtl* ctl, unsigned nblocks, unsigned* qtail) {
  __syncthreads();
  __shared__ int s_ok;
  __shared__ unsigned s_tail;
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned g = atomic_load_rlx(&ctl->bar_gen);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned arrived = __hip_atomic_fetch_add(&ctl->bar_count, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nblocks - 1) {
      __hip_atomic_store(&ctl->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&ctl->bar_gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned spins = 0;
      while (atomic_load_rlx(&ctl->bar_gen) == g) {
        if (__hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = 0; break; }
        if (++spins > kSpinLimit) { atomicExch(&ctl->err, 1); ok = 0; break; }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (__hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) ok = 0;
    s_tail = atomic_load_rlx(&ctl->qtail);
    s_ok = ok;
  }
  __syncthreads();
  *qtail = s_tail;
  return s_ok != 0;
}

__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Relax, successor, distf and deviation index of one vertex of level d.
__device__ void settle_vertex(const SfxtArgs& a, int u, int d, float* local_max) {
  const int b = a.oadjp[u], e = a.oadjp[u + 1];
  int best = d == 0 ? 0 : kIntMax;
  int s = -1;
  if (d > 0) {
    for (int j = b; j < e; ++j) {
      const int w = a.oadj[j];
      const int wi = static_cast<int>(a.owgt[j] * static_cast<float>(kScaleUp));
      const int cand = a.dist[w] + wi;
      if (cand < best) { best = cand; s = w; }
      else if (cand == best && w > s) s = w;
    }
  }
  const float df = static_cast<float>(best) / static_cast<float>(kScaleUp);
  a.dist[u] = best;
  a.succ[u] = s;
  a.level[u] = d;
  a.distf[u] = df;
  *local_max = fmaxf(*local_max, fabsf(df));
  // deviation records {w, w_e bits, distf[w] bits, delta bits}, sorted by delta
  int cnt = 0;
  for (int j = b; j < e; ++j) {
    const int w = a.oadj[j];
    if (w == s) continue;
    const int dw = a.dist[w];
    if (dw == kIntMax) continue;
    const float fw = a.distf[w];
    const float we = a.owgt[j];
    const float delta = static_cast<float>(static_cast<double>(fw) + static_cast<double>(we) -
                                           static_cast<double>(df));
    int4 rec = make_int4(w, __float_as_int(we), __float_as_int(fw), __float_as_int(delta));
    int k = cnt;
    while (k > 0) {
      const int4 prev = a.devs[b + k - 1];
      if (__int_as_float(prev.w) <= delta) break;
      a.devs[b + k] = prev;
      --k;
    }
    a.devs[b + k] = rec;
    ++cnt;
  }
  const float mind = cnt ? __int_as_float(a.devs[b].w) : __builtin_huge_valf();
  int x = s;
  for (int guard = 0; x != -1; ++guard) {
    if (guard > a.n) {  // pointers only move toward the sinks; never taken
      atomicExch(&a.ctl->nsv_err, 1);
      x = -1;
      break;
    }
    const int4 r = a.vrec[x];
    if (__int_as_float(r.x) < mind) break;
    x = r.y;
  }
  a.vrec[u] = make_int4(__float_as_int(mind), x, s, __float_as_int(df));
  a.devinfo[u] = make_int2(b, cnt);
}

}  // namespace

__global__ void __launch_bounds__(kSfxtBlock) sfxt_kernel(SfxtArgs a) {
  __shared__ int s_q[kLdsQueue];
  __shared__ int s_qn;
  __shared__ unsigned s_base;
  SfxtCtl* ctl = a.ctl;
  const unsigned nblocks = gridDim.x;
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int gthreads = gridDim.x * blockDim.x;
  if (threadIdx.x == 0) s_qn = 0;
  __syncthreads();

  auto push = [&](int x) {
    const int p = atomicAdd(&s_qn, 1);
    if (p < kLdsQueue) s_q[p] = x;
    else a.queue[atomicAdd(&ctl->qtail, 1u)] = x;
  };
  auto flush = [&]() {
    __syncthreads();
    const int cnt = min(s_qn, kLdsQueue);
    if (threadIdx.x == 0) s_base = cnt ? atomicAdd(&ctl->qtail, static_cast<unsigned>(cnt)) : 0u;
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) a.queue[s_base + i] = s_q[i];
    __syncthreads();
    if (threadIdx.x == 0) s_qn = 0;
  };

  // phase 0: per-vertex init; sinks (out-degree 0) form level 0
  for (int v = gtid; v < a.n; v += gthreads) {
    const int deg = a.oadjp[v + 1] - a.oadjp[v];
    a.deps[v] = deg;
    a.level[v] = -1;
    a.dist[v] = kIntMax;
    a.succ[v] = -1;
    a.distf[v] = static_cast<float>(kIntMax) / static_cast<float>(kScaleUp);
    a.vrec[v] = make_int4(__float_as_int(__builtin_huge_valf()), -1, -1, 0);
    a.devinfo[v] = make_int2(a.oadjp[v], 0);
    if (deg == 0) push(v);
  }
  flush();
  unsigned le = 0;
  if (!grid_barrier(ctl, nblocks, &le)) return;

  unsigned lb = 0;
  int d = 0;
  float local_max = 0.0f;
  while (true) {
    if (gtid == 0) a.lvlp[d] = static_cast<int>(lb);
    for (unsigned i = lb + gtid; i < le; i += gthreads) {
      const int u = a.queue[i];
      settle_vertex(a, u, d, &local_max);
      for (int j = a.iadjp[u]; j < a.iadjp[u + 1]; ++j) {
        const int x = a.iadj[j];
        if (atomicSub(&a.deps[x], 1) == 1) push(x);
      }
    }
    flush();
    unsigned ne = 0;
    if (!grid_barrier(ctl, nblocks, &ne)) return;
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

void launch_sfxt(const SfxtArgs& args, hipStream_t stream, int* grid_out) {
  int dev = 0;
  CPG_HIP(hipGetDevice(&dev));
  int cus = 0, coop = 0;
  CPG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  CPG_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop) throw std::runtime_error("device does not support cooperative launch");
  int per_cu = 0;
  CPG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sfxt_kernel, kSfxtBlock, 0));
  if (per_cu < 1) throw std::runtime_error("sfxt kernel cannot be resident");
  const int grid = cus;  // one block per CU: fewer barrier arrivals and fences
  if (grid_out) *grid_out = grid;
  SfxtArgs a = args;
  void* params[] = {&a};
  CPG_HIP(hipLaunchCooperativeKernel(sfxt_kernel, dim3(grid),
                                     dim3(kSfxtBlock), params, 0, stream));
}

}  // namespace cpg
