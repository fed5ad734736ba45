// Host orchestration of the CPG query and the C ABI (include/cpg.h).
//
// The handle mirrors gpucpg::CpGen (gpucpg/gpucpg.cuh:119-431): it owns the
// host CSR, keeps the graph and its SFXT resident on the device between
// queries, and runs one query at a time on its own HIP stream.
#include "../../include/cpg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <string>
#include <vector>

#include "common.hpp"
#include "graph.hpp"
#include "kernels.hpp"

namespace cpg {
void launch_gather_int(const int* src, const int* idx, int n, int* out, hipStream_t s);
}

using namespace cpg;

namespace {

thread_local std::string g_last_error;

// Wall-clock limit of one query's PFXT rounds; a build that emulates the
// kernels on the CPU raises it.
#ifndef CPG_QUERY_LIMIT_MS
#define CPG_QUERY_LIMIT_MS 120000.0
#endif

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void reserve(size_t want) {
    if (want <= n && p) return;
    release();
    CPG_HIP(hipMalloc(&p, sizeof(T) * (want ? want : 1)));
    n = want;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
};

// Pinned host memory (hipHostMalloc) for small transfers on the query path:
// copies from and to it go straight to the DMA engine, with no staging.
template <class T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  void reserve(size_t want) {
    if (want <= n) return;
    release();
    void* q = nullptr;
    CPG_HIP(hipHostMalloc(&q, sizeof(T) * want, hipHostMallocDefault));
    p = static_cast<T*>(q);
    n = want;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
};

template <class T>
void h2d(T* d, const T* h, size_t n, hipStream_t s) {
  if (n) CPG_HIP(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
}
template <class T>
void d2h(T* h, const T* d, size_t n, hipStream_t s) {
  if (n) CPG_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
}

// Roots of a query: the sources (ascending id) with a finite distance
// (gpucpg/gpucpg.cu:18502-18513).  With the tie quirk, the initial split of
// gpucpg/gpucpg.cu:19136-19180: the short pile keeps the first (int)(p*n+1)
// roots with cost <= split = sorted[(size_t)(p*n)]; roots tied at <= split
// beyond that count are lost.  sd holds the sources' int distances; kp
// receives one keep flag per source, minc the smallest kept cost.
void pick_roots(const std::vector<int>& sd, bool quirk, float perc, std::vector<uint8_t>& kp,
                int64_t& nroots, int64_t& dropped, float& minc) {
  const int nsrc = static_cast<int>(sd.size());
  kp.assign(std::max(1, nsrc), 0);
  nroots = 0;
  for (int i = 0; i < nsrc; ++i)
    if (sd[i] != kIntMax) {
      kp[i] = 1;
      ++nroots;
    }
  dropped = 0;
  if (quirk && nroots > 0) {
    // the (size_t)(p*n)-th smallest root cost (0-based): the largest of the
    // at + 1 smallest distances, kept in a max-heap (at is 0.5% of the roots
    // by default, so most distances are rejected against its top at once).
    // The cost (float)d / SCALE_UP never reverses the order of distances.
    const float pn = perc * static_cast<float>(nroots);
    const size_t at = std::min(static_cast<size_t>(pn), static_cast<size_t>(nroots) - 1);
    std::vector<int> heap;
    heap.reserve(at + 1);
    for (int i = 0; i < nsrc; ++i) {
      const int x = sd[i];
      if (x == kIntMax) continue;
      if (heap.size() < at + 1) {
        heap.push_back(x);
        std::push_heap(heap.begin(), heap.end());
      } else if (x < heap.front()) {
        std::pop_heap(heap.begin(), heap.end());
        heap.back() = x;
        std::push_heap(heap.begin(), heap.end());
      }
    }
    const float split = static_cast<float>(heap.front()) / static_cast<float>(kScaleUp);
    const int num_short = static_cast<int>(pn + 1);
    int kept_le = 0;
    for (int i = 0; i < nsrc; ++i) {
      if (!kp[i]) continue;
      const float c = static_cast<float>(sd[i]) / static_cast<float>(kScaleUp);
      if (c > split) continue;
      if (kept_le < num_short) ++kept_le;
      else { kp[i] = 0; ++dropped; }
    }
  }
  minc = __builtin_huge_valf();
  for (int i = 0; i < nsrc; ++i)
    if (kp[i]) minc = std::min(minc, static_cast<float>(sd[i]) / static_cast<float>(kScaleUp));
}

// A pair of HIP events bracketing work on one stream.
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  void create() {
    if (!a) CPG_HIP(hipEventCreate(&a));
    if (!b) CPG_HIP(hipEventCreate(&b));
  }
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  float ms() const {
    float t = 0.0f;
    CPG_HIP(hipEventElapsedTime(&t, a, b));
    return t;
  }
};

}  // namespace

struct cpg_handle {
  HostGraph g;
  bool loaded = false;
  hipStream_t stream = nullptr;
  int grid = 0;  // grid-stride kernels

  // device graph
  DevBuf<int> oadjp, oadj, iadjp, srcs, sinks;
  DevBuf<int2> iedge;
  DevBuf<float> owgt;
  bool graph_on_device = false;

  // SFXT
  DevBuf<int> deps, queue, lvlp, level, dist, succ, src_dist, big;
  PinnedBuf<int> h_src_dist;      // the sources' distances, copied back per query
  PinnedBuf<uint8_t> h_keep;      // the roots' keep flags, copied up per query
  PinnedBuf<SfxtCtl> h_sctl;      // the SFXT control block, copied back per query
  PinnedBuf<PfxtCtl> h_pctl;      // the PFXT control block, copied back per round
  PinnedBuf<unsigned long long> h_nsel;  // the final selection's path count

  // The PFXT control block after the work on the stream so far.
  PfxtCtl read_pctl() {
    h_pctl.reserve(1);
    d2h(h_pctl.p, pctl.p, 1, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    return *h_pctl.p;
  }
  DevBuf<unsigned long long> best;
  DevBuf<float> distf;
  DevBuf<int4> vrec, devs, devs_tmp;
  DevBuf<int2> devinfo, lift;
  DevBuf<SfxtCtl> sctl;
  DevBuf<unsigned long long> dfw;  // dataflow loop: one progress word per block
  bool sfxt_valid = false;
  std::vector<uint8_t> sfxt_sinks;  // endpoint subset of the current SFXT, per sink (empty = all)
  DevBuf<uint8_t> emask;
  int diameter = 0;
  int sfxt_blocks = 0;       // level-loop grid requested (0 = automatic)
  int sfxt_method = 0;       // 0 automatic, 1 level loop, 2 dataflow, 3 local queues, 4 fence-free, 5 edge-parallel
  int last_diameter = 0;     // levels of the last SFXT on this graph (automatic grid)
  float max_abs_distf = 0.0f;
  bool cache_enabled = false;
  int cache_hits = 0, cache_misses = 0;
  std::vector<int> h_succ;  // lazily copied for path reconstruction

  // CSR relabelling (cpg_set_csr_reorder; 0 none, 1 level order as the
  // reference's E_ORIENTED reorder): the device graph is held relabelled,
  // h_perm maps original -> relabelled ids and h_orig back.  Every result
  // leaves the engine in original ids, and ties break on original ids, so
  // results equal those of the unrelabelled graph.
  int csr_reorder = 0;
  bool relabelled = false;
  std::vector<int> h_perm, h_orig;
  DevBuf<int> perm_d, orig_d;
  DevBuf<uint8_t> mask_tmp;
  double reorder_ms = 0.0;

  // PFXT
  DevBuf<float> n_cost;
  DevBuf<int> n_to, n_from, n_parent, n_level;
  DevBuf<uint32_t> n_key, batch, ckeys;
  DevBuf<uint8_t> n_flag;
  DevBuf<unsigned> hist_all, hist_pend, hist_part;
  DevBuf<PfxtCtl> pctl;
  DevBuf<uint8_t> keep;
  uint64_t node_cap = 0;

  // final selection
  DevBuf<uint64_t> skey[2];
  DevBuf<uint32_t> sval[2];
  DevBuf<unsigned> ghist;
  DevBuf<unsigned long long> nsel;
  DevBuf<float> r_cost;
  DevBuf<int> r_level, r_from, r_to, r_pidx, r_parent, rank_of;
  DevBuf<uint32_t> r_node;   // node-store slot of each reported path
  int64_t nres = 0;
  uint64_t nodes_stored = 0; // node-store slots the last query filled
  bool have_result = false;
  std::vector<float> h_cost;
  std::vector<int> h_level, h_from, h_to, h_parent;
  bool host_result = false;
  // host copies of the prefix tree for vertex reports (filled on first use)
  std::vector<uint32_t> h_node;
  std::vector<int> hn_from, hn_to, hn_parent;
  bool host_tree = false;

  // BVSS deviation index (host copy for decoding, device copy for discovery)
  struct HostBvss {
    bool valid = false;
    int n_int = 0;
    std::vector<int> int_ptr, rows, entry_int;
    std::vector<unsigned long long> masks;
    std::vector<int2> tiles;
    uint64_t set_bits = 0;
  } hb;
  DevBuf<int> bv_int_ptr, bv_rows, bv_entry_int, bv_fptr, bv_fsrc;
  DevBuf<unsigned long long> bv_masks, bv_front;
  DevBuf<int2> bv_tiles;
  DevBuf<uint8_t> bv_dense;
  DevBuf<int4> bv_out;
  DevBuf<BvssCtl> bv_ctl;
  EventPair ev_bvss;

  cpg_timing timing{};
  bool kernel_timing = false;
  bool probe_overflowed = false;  // the last bound probe stopped at a full node store
  EventPair ev_sfxt, ev_levels, ev_count, ev_fill;
  // per-round record (step_times option): an event after every round and
  // the stored-path count the round ended with
  static constexpr int kRoundLog = 4096;
  std::vector<hipEvent_t> rev;
  DevBuf<unsigned long long> rlog;
  std::vector<double> step_ms;
  std::vector<int64_t> step_paths;

  cpg_handle() {}
  ~cpg_handle() {
    for (hipEvent_t e : rev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  hipEvent_t round_event(size_t i) {
    while (rev.size() <= i) {
      hipEvent_t e = nullptr;
      CPG_HIP(hipEventCreate(&e));
      rev.push_back(e);
    }
    return rev[i];
  }

  void ensure_stream() {
    if (stream) return;
    CPG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    int dev = 0, cus = 0;
    CPG_HIP(hipGetDevice(&dev));
    CPG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid = std::max(1, cus) * 8;
  }

  void set_graph(const EdgeList& el) {
    HostGraph ng;
    build_csr(el, ng);
    g = std::move(ng);
    loaded = true;
    graph_on_device = false;
    relabelled = false;
    last_diameter = 0;
    invalidate_cache_payload();
    cache_hits = cache_misses = 0;  // read_input clears the static cache (gpucpg.cu:454)
    reset();
  }

  void invalidate_cache_payload() {
    sfxt_valid = false;
    hb.valid = false;
    h_succ.clear();
  }

  void reset() {  // CpGen::reset, gpucpg/gpucpg.cuh:211-250
    have_result = false;
    host_result = false;
    host_tree = false;
    nres = 0;
    timing = cpg_timing{};
    if (!cache_enabled) invalidate_cache_payload();
  }

  void upload_graph() {
    if (graph_on_device) return;
    ensure_stream();
    const size_t n = static_cast<size_t>(g.n), m = static_cast<size_t>(g.m);
    oadjp.reserve(n + 1); oadj.reserve(m); owgt.reserve(m);
    iadjp.reserve(n + 1); iedge.reserve(m); srcs.reserve(g.srcs.size());
    sinks.reserve(g.sinks.size());
    h2d(oadjp.p, g.oadjp.data(), n + 1, stream);
    h2d(oadj.p, g.oadj.data(), m, stream);
    h2d(owgt.p, g.owgt.data(), m, stream);
    h2d(iadjp.p, g.iadjp.data(), n + 1, stream);
    {
      std::vector<int2> ie(m);
      for (size_t j = 0; j < m; ++j) {
        int wb;
        std::memcpy(&wb, &g.iwgt[j], 4);
        ie[j] = make_int2(g.iadj[j], wb);
      }
      h2d(iedge.p, ie.data(), m, stream);
      CPG_HIP(hipStreamSynchronize(stream));  // ie goes out of scope
    }
    h2d(srcs.p, g.srcs.data(), g.srcs.size(), stream);
    h2d(sinks.p, g.sinks.data(), g.sinks.size(), stream);
    CPG_HIP(hipStreamSynchronize(stream));
    graph_on_device = true;
    sfxt_valid = false;
    relabelled = false;
    if (csr_reorder == 1) relabel_graph();
  }

  // Level-order relabelling of the device graph (E_ORIENTED): levels from
  // one SFXT of the graph as loaded, the permutation by a counting sort of
  // the levels (ties and unlevelled vertices in ascending id), then the
  // edge-parallel scatter of both CSRs on the device.
  void relabel_graph() {
    const double t0 = now_ms();
    relabelled = false;
    const size_t n = static_cast<size_t>(g.n), m = static_cast<size_t>(g.m);
    if (n == 0) return;
    run_sfxt(false, nullptr);
    std::vector<int> lv(n);
    d2h(lv.data(), level.p, n, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    const int L = std::max(diameter, 0);
    std::vector<int> start(static_cast<size_t>(L) + 2, 0);
    for (size_t v = 0; v < n; ++v) start[static_cast<size_t>(lv[v] < 0 || lv[v] >= L ? L : lv[v]) + 1]++;
    for (int b = 0; b <= L; ++b) start[b + 1] += start[b];
    h_orig.assign(n, 0);
    h_perm.assign(n, 0);
    for (size_t v = 0; v < n; ++v) {
      const int b = lv[v] < 0 || lv[v] >= L ? L : lv[v];
      const int p = start[b]++;
      h_orig[p] = static_cast<int>(v);
      h_perm[v] = p;
    }
    std::vector<int> noadjp(n + 1, 0), niadjp(n + 1, 0);
    for (size_t p = 0; p < n; ++p) {
      const int v = h_orig[p];
      noadjp[p + 1] = noadjp[p] + (g.oadjp[v + 1] - g.oadjp[v]);
      niadjp[p + 1] = niadjp[p] + (g.iadjp[v + 1] - g.iadjp[v]);
    }
    perm_d.reserve(n); orig_d.reserve(n);
    h2d(perm_d.p, h_perm.data(), n, stream);
    h2d(orig_d.p, h_orig.data(), n, stream);
    DevBuf<int> owner, nadjp, nadj, niadjp_d;
    DevBuf<float> nw;
    DevBuf<int2> nie;
    owner.reserve(std::max<size_t>(m, 1));
    nadjp.reserve(n + 1); niadjp_d.reserve(n + 1);
    nadj.reserve(std::max<size_t>(m, 1)); nw.reserve(std::max<size_t>(m, 1)); nie.reserve(std::max<size_t>(m, 1));
    h2d(nadjp.p, noadjp.data(), n + 1, stream);
    h2d(niadjp_d.p, niadjp.data(), n + 1, stream);
    launch_edge_owner(oadjp.p, g.n, owner.p, stream);
    launch_reorder_fanout(oadjp.p, oadj.p, owgt.p, owner.p, perm_d.p, nadjp.p, static_cast<int64_t>(m),
                          nadj.p, nw.p, stream);
    launch_edge_owner(iadjp.p, g.n, owner.p, stream);
    launch_reorder_fanin(iadjp.p, iedge.p, owner.p, perm_d.p, niadjp_d.p, static_cast<int64_t>(m), nie.p,
                         stream);
    std::vector<int> ns(g.srcs.size()), nk(g.sinks.size());
    for (size_t i = 0; i < ns.size(); ++i) ns[i] = h_perm[static_cast<size_t>(g.srcs[i])];
    for (size_t i = 0; i < nk.size(); ++i) nk[i] = h_perm[static_cast<size_t>(g.sinks[i])];
    h2d(srcs.p, ns.data(), ns.size(), stream);
    h2d(sinks.p, nk.data(), nk.size(), stream);
    CPG_HIP(hipStreamSynchronize(stream));
    oadjp.swap(nadjp); oadj.swap(nadj); owgt.swap(nw); iadjp.swap(niadjp_d); iedge.swap(nie);
    relabelled = true;
    sfxt_valid = false;
    h_succ.clear();
    hb.valid = false;
    reorder_ms = now_ms() - t0;
  }

  // Original id of a relabelled vertex id (-1 stays -1).
  int to_orig(int v) const { return relabelled && v >= 0 ? h_orig[static_cast<size_t>(v)] : v; }

  // Level-loop grid: about one thread per vertex of an average level (known
  // from the last levelization of this graph), within [16, 128] blocks; a
  // grid barrier per level costs more with more blocks.
  static constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin
  static constexpr int64_t kOneXcdRecords = 50000;  // fanin records per level kept on one XCD
  int level_loop_blocks() const {
    if (sfxt_blocks > 0) return sfxt_blocks;
    if (last_diameter <= 0) return 64;
    const int64_t per_level = g.n / last_diameter;
    return static_cast<int>(std::min<int64_t>(128, std::max<int64_t>(16, (per_level + 255) / 256)));
  }

  // The endpoint subset of an n-byte mask as one flag per sink (in sink
  // order); empty when the mask is null or chooses every sink.
  std::vector<uint8_t> sink_subset(const uint8_t* mask) const {
    std::vector<uint8_t> f;
    if (!mask) return f;
    bool all = true;
    f.resize(g.sinks.size());
    for (size_t i = 0; i < g.sinks.size(); ++i) {
      f[i] = mask[g.sinks[i]] ? 1 : 0;
      all = all && f[i];
    }
    if (all) f.clear();
    return f;
  }

  // Runs the SFXT for the endpoint subset of mask (n bytes, null = every
  // sink) unless a cached SFXT for that subset is valid.  Returns elapsed ms.
  double run_sfxt(bool count_cache, const uint8_t* mask = nullptr) {
    const double t0 = now_ms();
    if (!sfxt_enqueue(count_cache, mask)) return 0.0;
    SfxtCtl c{};
    d2h(&c, sctl.p, 1, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    sfxt_finish(c, count_cache, mask);
    return now_ms() - t0;
  }

  // The SFXT in two halves, so that a caller can put its own work on the
  // stream before the one synchronisation: sfxt_enqueue puts the kernels on
  // the stream (false: the cached SFXT stands, nothing was queued), and
  // sfxt_finish checks the control block the caller copied back after them.
  std::vector<uint8_t> pending_sinks;  // endpoint subset of the queued SFXT
  int pending_lift = 0;                // lifting levels its index was built with

  bool sfxt_enqueue(bool count_cache, const uint8_t* mask) {
    upload_graph();
    std::vector<uint8_t> subset = sink_subset(mask);
    if (sfxt_valid && cache_enabled && subset == sfxt_sinks) {
      if (count_cache) ++cache_hits;
      return false;
    }
    hb.valid = false;
    sfxt_valid = false;
    if (cache_enabled && count_cache) ++cache_misses;
    const size_t n = static_cast<size_t>(g.n), m = static_cast<size_t>(g.m);
    // best[] and deps[] carry the level loops' pad slots for no-op atomics
    deps.reserve(n + kSfxtPad); queue.reserve(n); lvlp.reserve(n + 2); level.reserve(n);
    dist.reserve(n); succ.reserve(n); distf.reserve(n); vrec.reserve(n);
    devinfo.reserve(n); devs.reserve(m); big.reserve(n); best.reserve(n + kSfxtPad); sctl.reserve(1);
    if (g.max_out_degree > kSfxtLdsSort) devs_tmp.reserve(m);
    CPG_HIP(hipMemsetAsync(sctl.p, 0, sizeof(SfxtCtl), stream));
    SfxtCtl c{};
    if (n > 0) {
      SfxtArgs a{};
      a.n = g.n;
      a.m = g.m;
      a.nsinks = static_cast<int>(g.sinks.size());
      a.oadjp = oadjp.p; a.oadj = oadj.p; a.owgt = owgt.p;
      a.iadjp = iadjp.p; a.iedge = iedge.p; a.sinks = sinks.p; a.best = best.p;
      a.deps = deps.p; a.queue = queue.p; a.lvlp = lvlp.p; a.level = level.p;
      a.dist = dist.p; a.succ = succ.p; a.distf = distf.p;
      a.vrec = vrec.p; a.devinfo = devinfo.p; a.devs = devs.p; a.big = big.p;
      a.devs_tmp = g.max_out_degree > kSfxtLdsSort ? devs_tmp.p : nullptr;
      a.ctl = sctl.p;
      if (!subset.empty()) {
        emask.reserve(n);
        if (relabelled) {  // the mask is indexed by original id
          mask_tmp.reserve(n);
          h2d(mask_tmp.p, mask, n, stream);
          launch_gather_u8(mask_tmp.p, orig_d.p, g.n, emask.p, stream);
        } else {
          h2d(emask.p, mask, n, stream);
        }
        a.sink_mask = emask.p;
      }
      a.orig = relabelled ? orig_d.p : nullptr;
      // 0 (automatic) is the edge-parallel fence-free level loop
      const int method = sfxt_method == 0 ? 5 : sfxt_method;
      // Its blocks on one XCD (round-robin dispatch puts every 8th block on
      // the same XCD) while a level holds few fanin records, over all XCDs
      // above that: on the 1.5M-vertex DAG (13.6K records per level) one
      // XCD moves 459 MB per launch against 782 MB over all XCDs in the
      // same time; the 30M-edge DAG (100K per level) runs 8.26 ms over all
      // against 8.67 on one, the 256-design union graph 25 against 38 ms
      const int64_t records_per_level = g.m / std::max(1, last_diameter);
      a.place_stride = method == 5 && records_per_level <= kOneXcdRecords ? kXcds : 1;
      // Queue entries per wave in that loop, on one XCD: as many as fill
      // about one lane step of fanin records (64 x 3 records per step below
      // a mean in-degree of 4, 64 x 8 above), at least 8, so that a level of
      // few vertices with long in-degrees spreads over more waves.  The
      // 2.6M-vertex graph of mean in-degree 27 (1,609 levels) takes 35.4 ms
      // at 16 entries against 55.0 at 64, the 385K-vertex one 7.1 against
      // 20.2; over all XCDs the 30M-edge DAG ran slower (11.2 against 8.4),
      // so there every wave keeps 64.
      int we = 64;
      if (a.place_stride > 1) {
        const double mean_in = static_cast<double>(g.m) / static_cast<double>(std::max(1, g.n));
        const double per_step = 64.0 * (g.m < 4 * static_cast<int64_t>(g.n) ? 3 : 8);
        while (we > 8 && we * mean_in > per_step) we >>= 1;
      }
      a.wave_entries = we;
      const bool dataflow = method == 2;
      if (dataflow && relabelled)
        throw std::runtime_error("the dataflow SFXT loop does not run on a relabelled graph");
      int df_grid = 0;
      if (dataflow) {
        df_grid = sfxt_dataflow_grid(sfxt_blocks);
        dfw.reserve(static_cast<size_t>(df_grid));
        a.dfw = dfw.p;
        a.dfw_n = df_grid;
      }
      if (kernel_timing) {
        ev_sfxt.create();
        ev_levels.create();
        CPG_HIP(hipEventRecord(ev_sfxt.a, stream));
      }
      if (dataflow || method >= 4) launch_sfxt_dataflow_init(a, stream);
      else if (method == 3) launch_sfxt_local_init(a, stream);
      else launch_sfxt_init(a, stream);
      if (kernel_timing) CPG_HIP(hipEventRecord(ev_levels.a, stream));
      timing.sfxt_blocks = dataflow      ? launch_sfxt_dataflow(a, df_grid, stream)
                           : method == 3 ? launch_sfxt_local(a, level_loop_blocks(), stream)
                           : method == 4 ? launch_sfxt_levels_nf(a, level_loop_blocks(), stream)
                           : method == 5 ? launch_sfxt_levels_edges(a, level_loop_blocks(), stream)
                                         : launch_sfxt_levels(a, level_loop_blocks(), stream);
      if (kernel_timing) CPG_HIP(hipEventRecord(ev_levels.b, stream));
      if (relabelled) launch_sfxt_succ_relabel(succ.p, perm_d.p, g.n, stream);
      // The lifting table needs the level count.  Levels depend on the
      // graph alone (not on the endpoint subset), so after this graph's
      // first levelization the count is known without waiting for the loop.
      int lv = 0;
      if (last_diameter > 0) {
        lv = sfxt_lift_levels(last_diameter);
      } else {
        d2h(&c, sctl.p, 1, stream);
        CPG_HIP(hipStreamSynchronize(stream));
        if (c.err) throw std::runtime_error("SFXT grid barrier did not complete");
        lv = sfxt_lift_levels(c.num_levels);
      }
      lift.reserve(static_cast<size_t>(lv) * n);
      a.lift = lift.p;
      launch_sfxt_index(a, lv, g.m, stream);
      pending_lift = lv;
      if (kernel_timing) CPG_HIP(hipEventRecord(ev_sfxt.b, stream));
    }
    pending_sinks = std::move(subset);
    return true;
  }

  void sfxt_finish(const SfxtCtl& c, bool count_cache, const uint8_t* mask) {
    if (kernel_timing && g.n > 0) {
      timing.sfxt_kernel_ms = ev_sfxt.ms();
      timing.sfxt_levels_ms = ev_levels.ms();
    }
    if (c.err) throw std::runtime_error("SFXT grid barrier did not complete");
    if (g.n > 0 && sfxt_lift_levels(c.num_levels) > pending_lift) {
      // more levels than the table was sized for (not seen on one graph,
      // whose levels do not change): run it again with the count known
      if (last_diameter == c.num_levels) throw std::runtime_error("SFXT lifting table sizing failed");
      last_diameter = c.num_levels;
      run_sfxt(count_cache, mask);
      return;
    }
    if (c.nsv_err == 2) throw std::runtime_error("SFXT long deviation list without sort buffer");
    if (c.nsv_err) throw std::runtime_error("SFXT next-smaller descent found no answer");
    diameter = c.num_levels;
    last_diameter = diameter;
    uint32_t bits = c.max_abs_distf_bits;
    std::memcpy(&max_abs_distf, &bits, 4);
    sfxt_valid = true;
    sfxt_sinks = std::move(pending_sinks);
    h_succ.clear();
  }

  // Device bytes per node-store slot: cost, to, from, parent, level, key (4
  // each), flag (1), count-pass key and batch entry (4 each), and the final
  // sort's two key/value buffer pairs (24) and rank (4).
  static constexpr uint64_t kNodeBytes = 61;

  // Device bytes the graph and the SFXT hold.
  uint64_t sfxt_device_bytes() const {
    auto b = [](const auto& d) { return static_cast<uint64_t>(d.n) * sizeof(*d.p); };
    return b(oadjp) + b(oadj) + b(owgt) + b(iadjp) + b(iedge) + b(srcs) + b(sinks) + b(deps) +
           b(queue) + b(lvlp) + b(level) + b(dist) + b(succ) + b(src_dist) + b(big) + b(best) +
           b(distf) + b(vrec) + b(devs) + b(devs_tmp) + b(devinfo) + b(lift) + b(sctl) + b(dfw) +
           b(emask) + b(keep);
  }

  void release_pfxt_buffers() {
    n_cost.release(); n_to.release(); n_from.release(); n_parent.release(); n_level.release();
    n_key.release(); n_flag.release(); ckeys.release(); batch.release();
    skey[0].release(); skey[1].release(); sval[0].release(); sval[1].release(); rank_of.release();
    node_cap = 0;
  }

  void ensure_pfxt_buffers(uint64_t cap) {
    if (cap > 0x7fffffffull) throw std::runtime_error("node capacity exceeds 2^31");
    n_cost.reserve(cap); n_to.reserve(cap); n_from.reserve(cap);
    n_parent.reserve(cap); n_level.reserve(cap);
    n_key.reserve(cap + 16); n_flag.reserve(cap + 16);  // vector loads read whole groups of 8
    ckeys.reserve(cap + 16);
    batch.reserve(cap);
    hist_all.reserve(2 * kHistBins);
    hist_pend.reserve(kHistBins);
    hist_part.reserve(3 * static_cast<size_t>(kHistParts) * kHistBins);
    pctl.reserve(1);
    node_cap = cap;
  }

  // Bound on one chain walk's steps: chain vertices plus their deviations.
  int step_limit() const {
    return static_cast<int>(std::min<int64_t>(
        0x3fffffff, (static_cast<int64_t>(diameter) + 2) * (g.max_out_degree + 2) + 16));
  }

  NodeStore node_store() {
    return NodeStore{n_cost.p, n_to.p, n_from.p, n_parent.p, n_level.p, n_key.p, n_flag.p, node_cap};
  }

  // Histograms of the stored paths at or below the bound (see launch_hist_nodes).
  void hist_nodes(int mode, bool with_pend, bool accumulate) {
    launch_hist_nodes(node_store(), ckeys.p, node_cap, pctl.p, mode, with_pend, hist_part.p,
                      hist_all.p, hist_pend.p, accumulate, stream);
  }

  // Finds a finite first bound without walking the roots' whole chains:
  // expand the roots with a pruned fill pass that stores every child up to
  // a trial bound c_min + g, and widen g (dropping the stored children)
  // until at least k nonzero paths fall under it.  The bound is then the
  // largest key of the histogram bin where the running count of the stored
  // paths reaches k (k distinct stored paths cost no more, so no path above
  // it can be among the k cheapest), and the roots count as expanded.
  // Returns the bound's key, or kKeyInf when eight trials fall short, the
  // roots and their children are fewer than k, or a trial overflows the
  // node store (the rounds then start unbounded from the roots alone).
  uint32_t probe_bound(WalkArgs& wa, int64_t k, const std::vector<uint8_t>& kp,
                       const std::vector<int>& sd, float cmin, uint32_t klo, bool trace) {
    std::vector<float> rc;
    for (size_t i = 0; i < sd.size(); ++i)
      if (kp[i]) rc.push_back(static_cast<float>(sd[i]) / static_cast<float>(kScaleUp));
    probe_overflowed = false;
    const double tb0 = now_ms();
    const uint64_t nb = rc.size();
    launch_iota(batch.p, nb, grid, stream);
    // the roots' 10th-percentile cost, which only sets the first trial's
    // width: from an evenly strided sample of about 2048 roots
    const size_t step = std::max<size_t>(1, rc.size() / 2048);
    std::vector<float> sample;
    sample.reserve(rc.size() / step + 1);
    for (size_t i = 0; i < rc.size(); i += step) sample.push_back(rc[i]);
    std::nth_element(sample.begin(), sample.begin() + static_cast<std::ptrdiff_t>(sample.size() / 10), sample.end());
    const float p10 = sample[sample.size() / 10];
    if (trace) std::fprintf(stderr, "probe: host preparation %.3f ms\n", now_ms() - tb0);
    double gap = static_cast<double>(p10) - static_cast<double>(cmin);
    gap = std::max({gap, std::fabs(static_cast<double>(cmin)) * 0x1p-8,
                    static_cast<double>(max_abs_distf) * 0x1p-8, 1e-6});
    uint64_t prev_total = 0;
    double prev_gap = gap;
    wa.strict = 0;  // a trial keeps the children at the trial bound itself
    int t = 0;
    for (; t < 8; ++t) {
      const float u0 = static_cast<float>(static_cast<double>(cmin) + gap);
      const uint32_t uk = float_key(u0);
      launch_probe_setup(pctl.p, uk, window_shift(klo, uk), nb, nb, stream);
      launch_walk(wa, true, grid, stream);
      PfxtCtl pc = read_pctl();
      if (pc.err) throw std::runtime_error("PFXT walk exceeded its step bound");
      if (pc.overflow) {  // the trial bound's children do not fit the store
        probe_overflowed = true;
        break;
      }
      uint64_t roots_le = 0;
      for (float c : rc)
        if (c <= u0 && c != 0.0f) ++roots_le;
      const uint64_t children = pc.node_n - nb;
      const uint64_t total = roots_le + children;
      if (trace)
        std::fprintf(stderr, "probe %d bound %.9g roots %llu children %llu\n", t,
                     static_cast<double>(u0), static_cast<unsigned long long>(roots_le),
                     static_cast<unsigned long long>(children));
      if (total >= static_cast<uint64_t>(k)) {
        // the stored roots and children give the crossing; the roots are
        // expanded (every child of theirs up to the trial bound is stored)
        CPG_HIP(hipMemsetAsync(n_flag.p, 1, nb, stream));
        hist_nodes(0, false, false);
        launch_probe_finish(pctl.p, hist_all.p, static_cast<unsigned long long>(k), nb, stream);
        pc = read_pctl();
        timing.probes = t + 1;
        return pc.ukey;
      }
      // a wider bound that finds nothing new means every child of every
      // root is already stored: k needs deeper paths, stop probing
      if (t > 0 && total == prev_total) break;
      // Counts grow steeply with the bound (children cluster in cost):
      // model total ~ gap^alpha, alpha fitted to the last two trials (4 at
      // first), aim 10% past k, and widen by at most 2x so that a
      // successful trial stores at most the children of twice the needed
      // width.
      double alpha = 4.0;
      if (t > 0 && prev_total > 0 && total > prev_total)
        alpha = std::min(8.0, std::max(1.0, std::log(static_cast<double>(total) / prev_total) /
                                                std::log(gap / prev_gap)));
      prev_total = total;
      prev_gap = gap;
      const double f = std::pow(static_cast<double>(k) / static_cast<double>(std::max<uint64_t>(total, 1)),
                                1.0 / alpha) * 1.1;
      gap *= std::min(2.0, std::max(1.1, f));
    }
    timing.probes = std::min(t + 1, 8);
    launch_probe_setup(pctl.p, kKeyInf, window_shift(klo, kKeyInf), 0, nb, stream);
    CPG_HIP(hipMemsetAsync(hist_all.p, 0, sizeof(unsigned) * 2 * kHistBins, stream));
    return kKeyInf;
  }

  void query(int64_t k, const cpg_query_opts& o) {
    if (!loaded) throw std::runtime_error("no graph loaded");
    have_result = false;
    host_result = false;
    host_tree = false;
    nres = 0;
    nodes_stored = 0;
    timing = cpg_timing{};
    kernel_timing = o.kernel_timing != 0;
    const bool step_times = o.step_times != 0;
    step_ms.clear();
    step_paths.clear();
    if (o.sfxt_blocks < 0) throw std::runtime_error("sfxt_blocks must be >= 0");
    if (o.sfxt_method < 0 || o.sfxt_method > 5) throw std::runtime_error("sfxt_method must be 0-5");
    sfxt_blocks = o.sfxt_blocks;
    sfxt_method = o.sfxt_method;
    const double tq0 = now_ms();
    const bool trace = std::getenv("CPG_TRACE") != nullptr;
    // The SFXT and the sources' distances (roots: sources with finite dist,
    // gpucpg/gpucpg.cu:18502-18513) come back in one synchronisation;
    // sfxt_ms covers both.
    const int nsrc = static_cast<int>(g.srcs.size());
    src_dist.reserve(std::max(1, nsrc));
    h_src_dist.reserve(std::max(1, nsrc));
    h_sctl.reserve(1);
    auto copy_source_dists = [&]() {
      if (g.n > 0 && nsrc > 0) {
        launch_gather_int(dist.p, srcs.p, nsrc, src_dist.p, stream);
        d2h(h_src_dist.p, src_dist.p, nsrc, stream);
      }
    };
    const bool queued = sfxt_enqueue(true, o.endpoint_mask);
    copy_source_dists();
    if (queued) d2h(h_sctl.p, sctl.p, 1, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    if (queued) {
      const int lift_before = pending_lift;
      sfxt_finish(*h_sctl.p, true, o.endpoint_mask);
      if (pending_lift != lift_before) {  // it ran the SFXT again (a short lifting table)
        copy_source_dists();
        CPG_HIP(hipStreamSynchronize(stream));
      }
    }
    timing.sfxt_ms = now_ms() - tq0;
    if (k <= 0 || g.n == 0) {
      timing.query_ms = now_ms() - tq0;
      have_result = true;
      return;
    }
    const double tp0 = now_ms();
    std::vector<int> sd(h_src_dist.p, h_src_dist.p + nsrc);
    const double tr_copy = now_ms();
    std::vector<uint8_t> kp;
    int64_t nroots = 0, dropped = 0;
    float minc = 0.0f;
    pick_roots(sd, o.root_tie_quirk != 0, o.init_split_perc, kp, nroots, dropped, minc);
    if (trace)
      std::fprintf(stderr, "query roots: source distances to host %.3f ms, pick %.3f ms (%d sources)\n",
                   tr_copy - tp0, now_ms() - tr_copy, nsrc);
    timing.roots = nroots - dropped;
    timing.roots_dropped = dropped;
    keep.reserve(std::max(1, nsrc));
    h_keep.reserve(std::max(1, nsrc));
    std::copy(kp.begin(), kp.begin() + nsrc, h_keep.p);  // the previous query's copy has completed
    h2d(keep.p, h_keep.p, nsrc, stream);

    const uint64_t batch_target = o.batch > 0 ? static_cast<uint64_t>(o.batch) : 131072ull;
    uint64_t cap = o.max_nodes > 0 ? static_cast<uint64_t>(o.max_nodes)
                                   : std::max<uint64_t>(uint64_t(1) << 22,
                                                        16ull * static_cast<uint64_t>(k) +
                                                            static_cast<uint64_t>(nroots) + (1u << 20));
    cap = std::min<uint64_t>(cap, 0x7fffffffull);
    // Device-byte budget: the graph and its SFXT are held already; the node
    // store and the final sort take kNodeBytes per slot, so the store gets
    // what is left, and a store the budget cannot grow ends the query.
    uint64_t budget_cap = 0x7fffffffull;
    if (o.mem_budget_bytes > 0) {
      const uint64_t held = sfxt_device_bytes();
      const uint64_t budget = static_cast<uint64_t>(o.mem_budget_bytes);
      budget_cap = budget > held ? std::min<uint64_t>((budget - held) / kNodeBytes, 0x7fffffffull) : 0;
      if (budget_cap < static_cast<uint64_t>(nroots) + 1024)
        throw std::runtime_error("mem_budget_bytes leaves no room for the prefix-tree node store (graph "
                                 "and SFXT hold " + std::to_string(held) + " bytes)");
      cap = std::min(cap, budget_cap);
      if (node_cap > cap) release_pfxt_buffers();
    }
    PfxtCtl c{};
    // Rounds with a finite bound walk once: the fill pass stores children at
    // the bound, their histogram joins the stored paths', and the bound is
    // tightened.  A node-store overflow restarts the query with a larger
    // store and a count pass every round, which stores only children at or
    // below the tightened bound.
    bool two_pass = false;
    for (int attempt = 0;; ++attempt) {
      ensure_pfxt_buffers(cap);
      PfxtCtl init{};
      init.ukey = kKeyInf;
      init.klo = nroots - dropped > 0 ? float_key(minc) : 0u;
      init.shift = window_shift(init.klo, kKeyInf);
      init.node_cap = cap;
      h2d(pctl.p, &init, 1, stream);
      const NodeStore ns = node_store();
      launch_roots(srcs.p, keep.p, nsrc, dist.p, distf.p, ns, pctl.p, stream);
      WalkArgs wa{};
      wa.vrec = vrec.p; wa.devinfo = devinfo.p; wa.devs = devs.p; wa.batch = batch.p;
      wa.nodes = ns; wa.ckeys = ckeys.p; wa.ckeys_cap = cap; wa.ctl = pctl.p;
      wa.max_abs_distf = max_abs_distf; wa.max_abs_wgt = g.max_abs_wgt;
      wa.collect_stats = o.collect_stats;
      wa.step_limit = step_limit();
      const int64_t nkept = nroots - dropped;
      const bool probe = nkept > 0 && (o.bound_probe > 0 ||
                                       (o.bound_probe == 0 && nkept * (diameter + 1) > (1 << 20)));
      uint32_t ukey_host = kKeyInf;
      const double tpr0 = now_ms();
      if (trace) std::fprintf(stderr, "query roots: buffers, flags and root launch %.3f ms\n", tpr0 - tp0);
      if (probe) ukey_host = probe_bound(wa, k, kp, sd, minc, init.klo, trace);
      if (probe && probe_overflowed && attempt < 4 && cap < budget_cap && cap < 0x7fffffffull) {
        // a trial bound's children filled the store: probe again in a larger
        // store rather than start the rounds unbounded
        cap = std::min<uint64_t>({cap * 2, 0x7fffffffull, budget_cap});
        if (o.mem_budget_bytes > 0) release_pfxt_buffers();
        timing.retries++;
        continue;
      }
      if (trace)
        std::fprintf(stderr, "query phases: sfxt %.3f ms, roots %.3f ms, probe %.3f ms\n", timing.sfxt_ms,
                     tpr0 - tp0, now_ms() - tpr0);
      // A bound set from stored paths lets a fill pass drop children at the
      // bound (k stored paths already cost no more).  A bound set from
      // counted children that are not stored yet (a count pass) needs an
      // inclusive fill, or the paths at the bound would be lost.  The probe
      // sets its bound from stored paths.
      const bool probed = ukey_host != kKeyInf;
      bool bound_from_store = probed;
      bool overflow = false;
      int r = 0;
      if (step_times) {
        rlog.reserve(kRoundLog + 1);
        launch_round_log(pctl.p, rlog.p, kRoundLog + 1, stream);
        CPG_HIP(hipEventRecord(round_event(0), stream));
      }
      while (true) {
        if (r > 100000) throw std::runtime_error("PFXT round limit exceeded");
        if (now_ms() - tq0 > CPG_QUERY_LIMIT_MS) throw std::runtime_error("PFXT query exceeded its time limit");
        // Once the bound is finite and set from stored paths, every round is
        // the same fill-only sequence and every kernel does nothing after
        // the done flag is set: launch a group of rounds per host sync.
        const bool count_pass = two_pass || ukey_host == kKeyInf;
        const int group = (!kernel_timing && !count_pass && bound_from_store) ? 4 : 1;
        for (int gi = 0; gi < group; ++gi, ++r) {
          hist_nodes(0, true, false);
          // the first round expands every root, unless the probe already did
          const bool all_roots = r == 0 && !probed;
          launch_select_batch(pctl.p, hist_pend.p, all_roots ? ~0ull : batch_target, stream);
          launch_make_batch(ns, pctl.p, batch.p, grid, stream);
          if (kernel_timing) {
            ev_count.create();
            ev_fill.create();
            CPG_HIP(hipEventRecord(ev_count.a, stream));
          }
          if (count_pass) {
            wa.strict = ukey_host == kKeyInf || bound_from_store ? 1 : 0;
            launch_walk(wa, false, grid, stream);
            hist_nodes(2, false, true);
            if (kernel_timing) CPG_HIP(hipEventRecord(ev_count.b, stream));
            launch_select_bound(pctl.p, hist_all.p, static_cast<unsigned long long>(k), stream);
            if (kernel_timing) CPG_HIP(hipEventRecord(ev_fill.a, stream));
            wa.strict = 0;
            launch_walk(wa, true, grid, stream);
            if (kernel_timing) CPG_HIP(hipEventRecord(ev_fill.b, stream));
          } else {
            wa.strict = bound_from_store ? 1 : 0;
            if (kernel_timing) {
              CPG_HIP(hipEventRecord(ev_count.b, stream));
              CPG_HIP(hipEventRecord(ev_fill.a, stream));
            }
            launch_walk(wa, true, grid, stream);
            if (kernel_timing) CPG_HIP(hipEventRecord(ev_fill.b, stream));
            hist_nodes(1, false, true);
            launch_select_bound(pctl.p, hist_all.p, static_cast<unsigned long long>(k), stream);
          }
          if (step_times && r < kRoundLog) {
            launch_round_log(pctl.p, rlog.p, kRoundLog + 1, stream);
            CPG_HIP(hipEventRecord(round_event(static_cast<size_t>(r) + 1), stream));
          }
        }
        c = read_pctl();
        if (kernel_timing) {
          const double cm = ev_count.ms(), fm = ev_fill.ms();
          timing.walk_count_ms += cm;
          timing.walk_fill_ms += fm;
          timing.walk_launches += count_pass ? 2 : 1;
          if (trace)
            std::fprintf(stderr,
                         "round %d %s batch %llu expanded %llu nodes %llu U %.9g "
                         "count %.3f ms fill %.3f ms\n",
                         r - 1, count_pass ? "count+fill" : "fill", c.batch_n, c.expanded, c.node_n,
                         static_cast<double>(key_float(c.ukey)), cm, fm);
        }
        ukey_host = c.ukey;
        bound_from_store = true;  // every path counted so far is stored now
        if (c.err) throw std::runtime_error("PFXT walk exceeded its step bound");
        if (c.overflow) { overflow = true; break; }
        if (c.done || c.batch_n == 0) break;  // no unexpanded path below the bound
      }
      if (!overflow) break;
      if (cap >= budget_cap)
        throw std::runtime_error("prefix-tree node store would exceed mem_budget_bytes (" +
                                 std::to_string(cap) + " slots fit)");
      if (attempt >= 4 || cap >= 0x7fffffffull) throw std::runtime_error("prefix-tree node capacity exceeded");
      cap = std::min<uint64_t>({cap * 2, 0x7fffffffull, budget_cap});
      if (o.mem_budget_bytes > 0) release_pfxt_buffers();  // reallocated at the new size
      two_pass = true;
      timing.retries++;
    }
    timing.rounds = c.rounds;
    timing.nonmonotone = static_cast<int64_t>(c.nonmono);
    if (o.monotone_check && c.nonmono)
      throw std::runtime_error("a child path cost less than its parent (" + std::to_string(c.nonmono) +
                               " times): the top K may miss a path just below the bound");
    if (step_times) {
      const int nr = std::min(c.rounds, kRoundLog);
      std::vector<unsigned long long> hl(static_cast<size_t>(nr) + 1);
      d2h(hl.data(), rlog.p, static_cast<size_t>(nr) + 1, stream);
      CPG_HIP(hipStreamSynchronize(stream));
      for (int i = 0; i < nr; ++i) {
        float ms = 0.0f;
        CPG_HIP(hipEventElapsedTime(&ms, rev[static_cast<size_t>(i)], rev[static_cast<size_t>(i) + 1]));
        step_ms.push_back(ms);
        step_paths.push_back(static_cast<int64_t>(hl[i + 1] > hl[i] ? hl[i + 1] - hl[i] : 0));
      }
    }
    timing.expanded = static_cast<int64_t>(c.expanded);
    timing.generated = static_cast<int64_t>(c.node_n);
    timing.hops = static_cast<int64_t>(c.hops);
    timing.dev_visits = static_cast<int64_t>(c.devvis);
    const double ts0 = now_ms();
    timing.pfxt_ms = ts0 - tp0;

    // final selection: every stored path with a nonzero cost <= bound
    const uint64_t nn = std::min<uint64_t>(c.node_n, node_cap);
    skey[0].reserve(std::max<uint64_t>(1, nn)); skey[1].reserve(std::max<uint64_t>(1, nn));
    sval[0].reserve(std::max<uint64_t>(1, nn)); sval[1].reserve(std::max<uint64_t>(1, nn));
    nsel.reserve(1);
    CPG_HIP(hipMemsetAsync(nsel.p, 0, sizeof(unsigned long long), stream));
    const NodeStore ns = node_store();
    launch_build_sort_keys(ns, pctl.p, nn, skey[0].p, sval[0].p, nsel.p, stream);
    h_nsel.reserve(1);
    d2h(h_nsel.p, nsel.p, 1, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    const unsigned long long nl = *h_nsel.p;
    ghist.reserve(sort_ghist_len(nl));
    SortBufs sb{{skey[0].p, skey[1].p}, {sval[0].p, sval[1].p}, ghist.p, nl};
    const int bytes[6] = {0, 1, 4, 5, 6, 7};
    const int res = radix_sort_pairs(sb, nl, bytes, 6, stream);
    const uint64_t nout = std::min<uint64_t>(static_cast<uint64_t>(k), nl);
    r_cost.reserve(std::max<uint64_t>(1, nout));
    r_level.reserve(std::max<uint64_t>(1, nout)); r_from.reserve(std::max<uint64_t>(1, nout));
    r_to.reserve(std::max<uint64_t>(1, nout)); r_pidx.reserve(std::max<uint64_t>(1, nout));
    r_parent.reserve(std::max<uint64_t>(1, nout));
    rank_of.reserve(std::max<uint64_t>(1, nn));
    launch_fill_int(rank_of.p, -1, nn, stream);
    launch_emit_report(sval[res].p, nout, ns, r_cost.p, r_level.p, r_from.p, r_to.p, r_pidx.p,
                       rank_of.p, stream);
    launch_link_parents(r_pidx.p, rank_of.p, nout, r_parent.p, stream);
    r_node.reserve(std::max<uint64_t>(1, nout));
    if (nout)
      CPG_HIP(hipMemcpyAsync(r_node.p, sval[res].p, sizeof(uint32_t) * nout, hipMemcpyDeviceToDevice,
                             stream));
    CPG_HIP(hipStreamSynchronize(stream));
    nodes_stored = nn;
    timing.device_bytes = static_cast<int64_t>(
        sfxt_device_bytes() + node_cap * kNodeBytes);
    timing.node_slots = static_cast<int64_t>(node_cap);
    const double te = now_ms();
    timing.select_ms = te - ts0;
    timing.query_ms = te - tq0;
    if (std::getenv("CPG_TRACE"))
      std::fprintf(stderr, "query phases: pfxt %.3f ms (%d rounds), select %.3f ms, total %.3f ms\n",
                   timing.pfxt_ms, c.rounds, timing.select_ms, timing.query_ms);
    nres = static_cast<int64_t>(nout);
    have_result = true;
  }

  void fetch_result() {
    if (host_result) return;
    if (!have_result) throw std::runtime_error("no query result");
    const size_t n = static_cast<size_t>(nres);
    h_cost.resize(n); h_level.resize(n); h_from.resize(n); h_to.resize(n); h_parent.resize(n);
    if (n) {
      d2h(h_cost.data(), r_cost.p, n, stream);
      d2h(h_level.data(), r_level.p, n, stream);
      d2h(h_from.data(), r_from.p, n, stream);
      d2h(h_to.data(), r_to.p, n, stream);
      d2h(h_parent.data(), r_parent.p, n, stream);
      CPG_HIP(hipStreamSynchronize(stream));
      if (relabelled)
        for (size_t i = 0; i < n; ++i) {
          h_from[i] = to_orig(h_from[i]);
          h_to[i] = to_orig(h_to[i]);
        }
    }
    host_result = true;
  }

  // The prefix tree on the host: the report's node slots and every stored
  // node's deviation and parent (a path's ancestors may be unreported, e.g.
  // zero-cost paths, so the walk follows node slots, not report ranks).
  void fetch_tree() {
    if (host_tree) return;
    fetch_result();
    const size_t nr = static_cast<size_t>(nres), nn = static_cast<size_t>(nodes_stored);
    h_node.resize(nr);
    hn_from.resize(nn); hn_to.resize(nn); hn_parent.resize(nn);
    d2h(h_node.data(), r_node.p, nr, stream);
    d2h(hn_from.data(), n_from.p, nn, stream);
    d2h(hn_to.data(), n_to.p, nn, stream);
    d2h(hn_parent.data(), n_parent.p, nn, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    host_tree = true;
  }

  // Vertices of reported path i: its root's suffix chain with every
  // deviation (from -> to) spliced in at its from-vertex.
  std::vector<int> path_vertices(int64_t i) {
    fetch_tree();
    if (i < 0 || i >= nres) throw std::runtime_error("path index out of range");
    fetch_succ();
    std::vector<std::pair<int, int>> devs;  // (from, to), leaf to root
    int64_t x = h_node[static_cast<size_t>(i)];
    for (int guard = 0;; ++guard) {
      if (x < 0 || x >= static_cast<int64_t>(nodes_stored) || guard > g.n + 1)
        throw std::runtime_error("prefix-tree parent chain is broken");
      if (hn_from[x] < 0) break;
      devs.emplace_back(hn_from[x], hn_to[x]);
      x = hn_parent[x];
    }
    std::vector<int> path;
    int v = hn_to[x];
    size_t di = devs.size();
    while (v != -1) {
      path.push_back(v);
      if (di > 0 && devs[di - 1].first == v) {
        v = devs[di - 1].second;
        --di;
      } else {
        v = h_succ[v];
      }
      if (path.size() > static_cast<size_t>(g.n) + 1) throw std::runtime_error("path does not end");
    }
    if (di != 0) throw std::runtime_error("deviation vertex not on the suffix chain");
    if (relabelled)
      for (int& v2 : path) v2 = to_orig(v2);
    return path;
  }

  // A_dev index: one entry per (64-source interval, destination row) with
  // a deviation edge, entries ordered by interval then row.
  void bvss_build() {
    if (csr_reorder != 0)
      throw std::runtime_error("the BVSS index is built on the graph in its original vertex order "
                               "(csr reorder off)");
    if (!sfxt_valid) run_sfxt(false);
    fetch_succ();
    const int n = g.n;
    HostBvss b;
    b.n_int = (n + 63) / 64;
    std::vector<std::pair<uint64_t, int>> kb;  // {interval << 32 | row, source bit}
    kb.reserve(static_cast<size_t>(g.m));
    for (int u = 0; u < n; ++u)
      for (int e = g.oadjp[u]; e < g.oadjp[u + 1]; ++e) {
        const int v = g.oadj[e];
        if (v == h_succ[u]) continue;
        kb.emplace_back((static_cast<uint64_t>(u >> 6) << 32) | static_cast<uint32_t>(v), u & 63);
      }
    std::sort(kb.begin(), kb.end());
    b.int_ptr.assign(static_cast<size_t>(b.n_int) + 1, 0);
    for (size_t i = 0; i < kb.size();) {
      const uint64_t key = kb[i].first;
      unsigned long long m = 0;
      for (; i < kb.size() && kb[i].first == key; ++i) m |= 1ull << kb[i].second;
      const int iv = static_cast<int>(key >> 32);
      b.rows.push_back(static_cast<int>(key & 0xffffffffu));
      b.entry_int.push_back(iv);
      b.masks.push_back(m);
      b.set_bits += static_cast<uint64_t>(__builtin_popcountll(m));
      b.int_ptr[static_cast<size_t>(iv) + 1]++;
    }
    for (int i = 0; i < b.n_int; ++i) b.int_ptr[i + 1] += b.int_ptr[i];
    for (int i = 0; i < b.n_int; ++i)
      for (int e = b.int_ptr[i]; e < b.int_ptr[i + 1]; e += 16) b.tiles.push_back(make_int2(i, e));
    b.valid = true;
    const size_t ne = b.rows.size();
    bv_int_ptr.reserve(b.int_ptr.size()); bv_rows.reserve(std::max<size_t>(1, ne));
    bv_entry_int.reserve(std::max<size_t>(1, ne)); bv_masks.reserve(std::max<size_t>(1, ne));
    bv_tiles.reserve(std::max<size_t>(1, b.tiles.size())); bv_dense.reserve(std::max(1, b.n_int));
    h2d(bv_int_ptr.p, b.int_ptr.data(), b.int_ptr.size(), stream);
    h2d(bv_rows.p, b.rows.data(), ne, stream);
    h2d(bv_entry_int.p, b.entry_int.data(), ne, stream);
    h2d(bv_masks.p, b.masks.data(), ne, stream);
    h2d(bv_tiles.p, b.tiles.data(), b.tiles.size(), stream);
    CPG_HIP(hipStreamSynchronize(stream));
    hb = std::move(b);
  }

  int64_t bvss_discover(int nfront, const int* fptr, const int* fsrc, int mode, int dense_min,
                        int* of, int* ou, int* ov, int64_t cap) {
    if (nfront < 0 || (nfront > 0 && (!fptr || (fptr[nfront] > 0 && !fsrc))))
      throw std::runtime_error("bad frontier arguments");
    if (mode < 0 || mode > 2) throw std::runtime_error("mode must be 0, 1 or 2");
    if (cap < 0) throw std::runtime_error("cap must be >= 0");
    if (!hb.valid) bvss_build();
    timing.bvss_ms = 0.0;
    timing.bvss_tiles = 0;
    if (nfront == 0) return 0;
    for (int f = 0; f < nfront; ++f)
      if (fptr[f] > fptr[f + 1] || fptr[f] < 0) throw std::runtime_error("fptr must be ascending");
    const int nsrc = fptr[nfront];
    bv_fptr.reserve(static_cast<size_t>(nfront) + 1);
    bv_fsrc.reserve(std::max(1, nsrc));
    h2d(bv_fptr.p, fptr, static_cast<size_t>(nfront) + 1, stream);
    h2d(bv_fsrc.p, fsrc, static_cast<size_t>(nsrc), stream);
    const size_t fw = static_cast<size_t>(nfront) * std::max(1, hb.n_int);
    bv_front.reserve(fw);
    CPG_HIP(hipMemsetAsync(bv_front.p, 0, fw * sizeof(unsigned long long), stream));
    bv_out.reserve(std::max<int64_t>(1, cap));
    bv_ctl.reserve(1);
    CPG_HIP(hipMemsetAsync(bv_ctl.p, 0, sizeof(BvssCtl), stream));
    BvssArgs a{};
    a.n = g.n; a.n_int = hb.n_int; a.n_entries = static_cast<int>(hb.rows.size());
    a.n_tiles = static_cast<int>(hb.tiles.size()); a.nfront = nfront;
    a.int_ptr = bv_int_ptr.p; a.rows = bv_rows.p; a.entry_int = bv_entry_int.p;
    a.masks = bv_masks.p; a.tiles = bv_tiles.p; a.front = bv_front.p; a.dense = bv_dense.p;
    a.out = bv_out.p; a.cap = static_cast<uint64_t>(cap); a.ctl = bv_ctl.p;
    ev_bvss.create();
    CPG_HIP(hipEventRecord(ev_bvss.a, stream));
    launch_bvss_discover(a, bv_fptr.p, bv_fsrc.p, mode, dense_min, grid, stream);
    CPG_HIP(hipEventRecord(ev_bvss.b, stream));
    BvssCtl c{};
    d2h(&c, bv_ctl.p, 1, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    timing.bvss_ms = ev_bvss.ms();
    timing.bvss_tiles = static_cast<int64_t>(c.tiles);
    if (c.mismatch) throw std::runtime_error("BVSS MFMA count disagrees with the decoded words");
    const uint64_t cnt = c.count;
    if (cnt > static_cast<uint64_t>(cap)) return -static_cast<int64_t>(cnt) - 1000;
    std::vector<int4> t(cnt);
    d2h(t.data(), bv_out.p, cnt, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    std::sort(t.begin(), t.end(), [](const int4& x, const int4& y) {
      if (x.x != y.x) return x.x < y.x;
      if (x.y != y.y) return x.y < y.y;
      return x.z < y.z;
    });
    for (uint64_t i = 0; i < cnt; ++i) {
      if (of) of[i] = t[i].x;
      if (ou) ou[i] = t[i].y;
      if (ov) ov[i] = t[i].z;
    }
    return static_cast<int64_t>(cnt);
  }

  void fetch_succ() {
    if (!h_succ.empty() || g.n == 0) return;
    if (!sfxt_valid) run_sfxt(false);
    h_succ.resize(g.n);
    d2h(h_succ.data(), succ.p, g.n, stream);
    CPG_HIP(hipStreamSynchronize(stream));
  }
};

// Batched queries: the disjoint union of several designs, one top-K query
// each, expanded together by the multi-query PFXT (mq.hip).  The union's
// SFXT is every design's own SFXT, so one level loop serves all of them;
// the rounds after the first are captured once in a HIP graph and
// replayed.
struct cpg_batch {
  struct Design {
    int n = 0;
    std::vector<int> u, v;
    std::vector<float> w;
    int64_t k = 0;
  };
  std::vector<Design> designs;
  std::vector<int> voff;              // first union vertex of each design
  cpg_handle uh;                      // the union graph (CSR + SFXT)
  bool built = false;

  DevBuf<int> nq, root_v, root_q;
  DevBuf<MqQuery> qs;
  DevBuf<MqCtl> mctl;
  DevBuf<unsigned> hist;
  DevBuf<float> n_cost;
  DevBuf<int> n_to, n_from, n_parent, n_level;
  DevBuf<uint32_t> n_key, batch;
  DevBuf<uint8_t> n_flag;
  DevBuf<uint64_t> skey[2];
  DevBuf<uint32_t> sval[2];
  DevBuf<unsigned> ghist;
  DevBuf<unsigned long long> nsel, per_q;
  DevBuf<float> out_cost;
  uint64_t cap = 0;

  std::vector<std::vector<float>> results;
  cpg_timing timing{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  MqArgs graph_args{};   // the arguments the captured round was recorded with
  int graph_grid = 0;
  // Round shapes (cpg_batch_autotune picks them by measurement):
  //   group       replayed rounds per host synchronisation
  //   grid_mult   round kernels' grid = CUs x grid_mult blocks
  //   target_div  per K class (K < 1000, < 10000, >= 10000): a query
  //               expands about K / div parents per round
  cpg_batch_tuning tune = default_tuning();

  static cpg_batch_tuning default_tuning() {
    cpg_batch_tuning t{};
    t.group = 4;
    t.grid_mult = 8;
    t.target_div[0] = t.target_div[1] = t.target_div[2] = 8;
    return t;
  }
  static int k_class(uint64_t k) { return k < 1000 ? 0 : k < 10000 ? 1 : 2; }

  ~cpg_batch() { drop_graph(); }

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }

  void build_union() {
    EdgeList el;
    voff.assign(designs.size() + 1, 0);
    size_t m = 0;
    for (size_t q = 0; q < designs.size(); ++q) {
      voff[q + 1] = voff[q] + designs[q].n;
      m += designs[q].u.size();
    }
    el.n = voff.back();
    el.u.reserve(m); el.v.reserve(m); el.w.reserve(m);
    for (size_t q = 0; q < designs.size(); ++q) {
      const Design& d = designs[q];
      for (size_t e = 0; e < d.u.size(); ++e) {
        el.u.push_back(d.u[e] + voff[q]);
        el.v.push_back(d.v[e] + voff[q]);
        el.w.push_back(d.w[e]);
      }
    }
    uh.set_graph(el);
    built = true;
  }

  void ensure(uint64_t c, int nqc) {
    n_cost.reserve(c); n_to.reserve(c); n_from.reserve(c); n_parent.reserve(c);
    n_level.reserve(c); n_key.reserve(c); n_flag.reserve(c); nq.reserve(c); batch.reserve(c);
    qs.reserve(nqc); mctl.reserve(1);
    hist.reserve(static_cast<size_t>(nqc) * kMqHistStride);
    cap = c;
  }

  // Measurement-driven choice of the round shapes: coordinate descent
  // from the current shapes, one whole batch run (SFXT, rounds, selection)
  // per trial, a candidate kept when it is at least 2% faster.  Results do
  // not depend on the shapes (every shape gives the exact top K).
  void autotune(const cpg_query_opts& o, int max_trials) {
    if (max_trials < 1) throw std::runtime_error("max_trials must be >= 1");
    auto timed = [&]() {
      const double t0 = now_ms();
      run(o);
      return now_ms() - t0;
    };
    timed();  // sizes the buffers and captures the round
    cpg_batch_tuning best = tune;
    double best_ms = timed();
    const double start_ms = best_ms;
    int trials = 1;
    static const int kGroups[] = {2, 4, 8, 16};
    static const int kGrids[] = {4, 8, 16};
    static const int kDivs[] = {2, 4, 8, 16, 32};
    auto try_value = [&](int* field, const int* vals, int nv) {
      const int keep = *field;
      int pick = keep;
      for (int i = 0; i < nv && trials < max_trials; ++i) {
        if (vals[i] == keep) continue;
        *field = vals[i];
        ++trials;
        const double ms = timed();
        if (ms < best_ms * 0.98) {
          best_ms = ms;
          pick = vals[i];
        }
      }
      *field = pick;
      best = tune;
    };
    try_value(&tune.group, kGroups, 4);
    try_value(&tune.grid_mult, kGrids, 3);
    for (int c = 2; c >= 0; --c) try_value(&tune.target_div[c], kDivs, 5);
    tune = best;
    tune.trials = trials;
    tune.best_ms = best_ms;
    tune.start_ms = start_ms;
    run(o);  // leave the results of the chosen shapes
  }

  void run(const cpg_query_opts& o) {
    if (designs.empty()) throw std::runtime_error("no designs in the batch");
    if (!built) build_union();
    timing = cpg_timing{};
    const double t0 = now_ms();
    uh.kernel_timing = false;
    if (o.sfxt_blocks < 0) throw std::runtime_error("sfxt_blocks must be >= 0");
    if (o.sfxt_method < 0 || o.sfxt_method > 5) throw std::runtime_error("sfxt_method must be 0-5");
    uh.sfxt_blocks = o.sfxt_blocks;
    uh.sfxt_method = o.sfxt_method;
    uh.reset();
    timing.sfxt_ms = uh.run_sfxt(false);
    hipStream_t stream = uh.stream;
    const double tp0 = now_ms();
    const int nqc = static_cast<int>(designs.size());
    // roots per design (ascending id), with each design's own initial-split quirk
    const int nsrc = static_cast<int>(uh.g.srcs.size());
    std::vector<int> sd(nsrc);
    uh.src_dist.reserve(std::max(1, nsrc));
    launch_gather_int(uh.dist.p, uh.srcs.p, nsrc, uh.src_dist.p, stream);
    d2h(sd.data(), uh.src_dist.p, nsrc, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    std::vector<std::vector<int>> rv(nqc);
    std::vector<std::vector<float>> rcost(nqc);
    for (int i = 0; i < nsrc; ++i) {
      if (sd[i] == kIntMax) continue;
      const int s = uh.g.srcs[i];
      const int q = static_cast<int>(std::upper_bound(voff.begin(), voff.end(), s) - voff.begin()) - 1;
      rv[q].push_back(s);
      rcost[q].push_back(static_cast<float>(sd[i]) / static_cast<float>(kScaleUp));
    }
    std::vector<int> root_vh, root_qh;
    std::vector<MqQuery> qh(nqc);
    uint64_t need = 0;
    for (int q = 0; q < nqc; ++q) {
      std::vector<uint8_t> keep(rv[q].size(), 1);
      const size_t nr = rv[q].size();
      if (o.root_tie_quirk && nr > 0) {
        std::vector<float> sel = rcost[q];
        const float pn = o.init_split_perc * static_cast<float>(nr);
        const size_t at = std::min(static_cast<size_t>(pn), nr - 1);
        std::nth_element(sel.begin(), sel.begin() + static_cast<std::ptrdiff_t>(at), sel.end());
        const float split = sel[at];
        const int num_short = static_cast<int>(pn + 1);
        int kept = 0;
        for (size_t i = 0; i < nr; ++i) {
          if (rcost[q][i] > split) continue;
          if (kept < num_short) ++kept;
          else keep[i] = 0;
        }
      }
      float minc = __builtin_huge_valf();
      int cnt = 0;
      for (size_t i = 0; i < nr; ++i)
        if (keep[i]) {
          root_vh.push_back(rv[q][i]);
          root_qh.push_back(q);
          minc = std::min(minc, rcost[q][i]);
          ++cnt;
        }
      MqQuery& m = qh[q];
      m = MqQuery{};
      m.ukey = kKeyInf;
      m.klo = cnt ? float_key(minc) : 0u;
      m.shift = window_shift_bins(m.klo, kKeyInf, kMqBins);
      m.k = static_cast<unsigned long long>(std::max<int64_t>(0, designs[q].k));
      m.batch_target = std::max<unsigned long long>(
          256, m.k / static_cast<unsigned long long>(std::max(1, tune.target_div[k_class(m.k)])));
      m.done = (cnt == 0 || m.k == 0) ? 1 : 0;
      m.vrec = uh.vrec.p;
      m.devinfo = uh.devinfo.p;
      m.devs = uh.devs.p;
      m.mag = 2.0f * uh.max_abs_distf + uh.g.max_abs_wgt + 1.0f;
      m.step_limit = uh.step_limit();
      need += 16ull * m.k + static_cast<uint64_t>(cnt) + 4096;
      timing.roots += cnt;
      timing.roots_dropped += static_cast<int64_t>(nr) - cnt;
    }
    const int nroots = static_cast<int>(root_vh.size());
    root_v.reserve(std::max(1, nroots)); root_q.reserve(std::max(1, nroots));
    h2d(root_v.p, root_vh.data(), nroots, stream);
    h2d(root_q.p, root_qh.data(), nroots, stream);
    uint64_t c = std::min<uint64_t>(std::max<uint64_t>(need, 1u << 20), 0x7fffffffull);
    MqCtl mc{};
    for (int attempt = 0;; ++attempt) {
      ensure(c, nqc);
      h2d(qs.p, qh.data(), nqc, stream);
      CPG_HIP(hipMemsetAsync(mctl.p, 0, sizeof(MqCtl), stream));
      MqArgs a = args(nqc);
      launch_mq_roots(a, root_v.p, root_q.p, nroots, stream);
      bool overflow = false;
      for (int r = 0;; ++r) {
        if (r > 100000) throw std::runtime_error("batched PFXT round limit exceeded");
        if (now_ms() - t0 > CPG_QUERY_LIMIT_MS) throw std::runtime_error("batched query exceeded its time limit");
        // round 0 expands every root; later rounds replay the captured round
        const int group = r == 0 ? 1 : std::max(1, tune.group);
        const int rgrid = std::max(1, uh.grid / 8 * std::max(1, tune.grid_mult));
        if (r == 0) {
          launch_mq_round(a, true, rgrid, stream);
        } else {
          // the captured round holds buffer addresses and sizes: capture
          // again whenever any of them changed (node store growth, designs
          // added, a rebuilt union graph, another grid)
          if (graph_exec && (std::memcmp(&graph_args, &a, sizeof(MqArgs)) != 0 || graph_grid != rgrid))
            drop_graph();
          if (!graph_exec) {
            CPG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            launch_mq_round(a, false, rgrid, stream);
            CPG_HIP(hipStreamEndCapture(stream, &graph));
            CPG_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
            graph_args = a;
            graph_grid = rgrid;
          }
          for (int gi = 0; gi < group; ++gi) CPG_HIP(hipGraphLaunch(graph_exec, stream));
        }
        d2h(&mc, mctl.p, 1, stream);
        CPG_HIP(hipStreamSynchronize(stream));
        timing.rounds += group;
        if (mc.err) throw std::runtime_error("batched PFXT walk exceeded its step bound");
        if (mc.overflow) { overflow = true; break; }
        if (mc.active == 0) break;
      }
      if (!overflow) break;
      if (attempt >= 4 || c >= 0x7fffffffull) throw std::runtime_error("prefix-tree node capacity exceeded");
      c = std::min<uint64_t>(c * 2, 0x7fffffffull);
      timing.retries++;
    }
    timing.generated = static_cast<int64_t>(mc.node_n);
    timing.nonmonotone = static_cast<int64_t>(mc.nonmono);
    if (o.monotone_check && mc.nonmono)
      throw std::runtime_error("a child path cost less than its parent (" + std::to_string(mc.nonmono) +
                               " times): a query's top K may miss a path just below its bound");
    const double ts0 = now_ms();
    timing.pfxt_ms = ts0 - tp0;
    // final selection: (query, cost, level) sort of the paths at or below
    // each query's bound; each query's first k are contiguous
    const uint64_t nn = std::min<uint64_t>(mc.node_n, cap);
    skey[0].reserve(std::max<uint64_t>(1, nn)); skey[1].reserve(std::max<uint64_t>(1, nn));
    sval[0].reserve(std::max<uint64_t>(1, nn)); sval[1].reserve(std::max<uint64_t>(1, nn));
    nsel.reserve(1); per_q.reserve(nqc);
    CPG_HIP(hipMemsetAsync(nsel.p, 0, sizeof(unsigned long long), stream));
    CPG_HIP(hipMemsetAsync(per_q.p, 0, sizeof(unsigned long long) * nqc, stream));
    MqArgs a = args(nqc);
    launch_mq_build_keys(a, nullptr, nn, skey[0].p, sval[0].p, nsel.p, per_q.p, uh.grid, stream);
    unsigned long long nl = 0;
    std::vector<unsigned long long> pq(nqc);
    d2h(&nl, nsel.p, 1, stream);
    d2h(pq.data(), per_q.p, nqc, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    ghist.reserve(sort_ghist_len(nl));
    SortBufs sb{{skey[0].p, skey[1].p}, {sval[0].p, sval[1].p}, ghist.p, nl};
    std::vector<int> bytes = {0, 1, 2, 3, 4, 5};
    if (nqc > 1) bytes.push_back(6);
    if (nqc > 256) bytes.push_back(7);
    const int res = radix_sort_pairs(sb, nl, bytes.data(), static_cast<int>(bytes.size()), stream);
    out_cost.reserve(std::max<uint64_t>(1, nl));
    launch_mq_gather_costs(sval[res].p, n_cost.p, nl, out_cost.p, stream);
    std::vector<float> all(nl);
    d2h(all.data(), out_cost.p, nl, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    results.assign(nqc, {});
    uint64_t off = 0;
    for (int q = 0; q < nqc; ++q) {
      const uint64_t take = std::min<uint64_t>(pq[q], qh[q].k);
      results[q].assign(all.begin() + off, all.begin() + off + take);
      off += pq[q];
    }
    const double te = now_ms();
    timing.select_ms = te - ts0;
    timing.query_ms = te - t0;
  }

  MqArgs args(int nqc) {
    MqArgs a{};
    a.nodes = NodeStore{n_cost.p, n_to.p, n_from.p, n_parent.p, n_level.p, n_key.p, n_flag.p, cap};
    a.nq = nq.p;
    a.distf = uh.distf.p;
    a.batch = batch.p; a.hist = hist.p; a.qs = qs.p; a.ctl = mctl.p; a.nq_count = nqc;
    return a;
  }
};

// Served queries: resident graphs, a stream of top-K queries each with its
// own endpoint subset, continuous batching of their PFXT rounds.
//
// A query is admitted between rounds when a slot is free and the device
// bytes of the in-flight queries' deviation indexes stay within the
// budget.  Admission runs the query's SFXT for its endpoint subset on its
// graph's handle and takes the index buffers over (a buffer swap, no
// copy), then appends its roots to the shared node store; every round then
// expands every in-flight query at once (the multi-query kernels of
// mq.hip, each query reading its own index).  A query whose expansion is
// done is extracted at the end of that round (its paths sorted by cost),
// and its slot and index are released.  Nodes of finished queries stay in
// the store until a compaction copies the live ones to the second buffer
// set; a round that overflows the store restarts the in-flight queries
// from their roots in a store twice as large.
// Node-store headroom reserved per requested path at admission (a test
// build sets 0 to force the overflow restart).
#ifndef CPG_SERVER_NODES_PER_PATH
#define CPG_SERVER_NODES_PER_PATH 16ull
#endif
#ifndef CPG_SERVER_HEADROOM_NODES
#define CPG_SERVER_HEADROOM_NODES 65536ull
#endif

struct cpg_server {
  struct IndexSet {
    DevBuf<int4> vrec, devs;
    DevBuf<int2> devinfo;
  };
  struct Graph {
    std::unique_ptr<cpg_handle> h;
    std::vector<std::unique_ptr<IndexSet>> pool;  // index buffers of finished queries
    int64_t index_bytes = 0;
    // stored paths per asked path of this graph's finished queries (an
    // exponential average; admission's node estimate)
    double nodes_per_k = static_cast<double>(kNodesPerK);
    struct Multi {  // the multi-subset SFXT of an admission group
      DevBuf<int> dv;  // settled distances, vertex-major (n x Q)
      DevBuf<int> dist, succ;
      DevBuf<float> distf;
      DevBuf<uint8_t> masks;
      DevBuf<unsigned> maxd;
      DevBuf<SfxtCtl> ctls;
    } mm;
    // the admission group in progress on this graph's stream
    bool busy = false;
    std::vector<int64_t> ids;
    std::vector<SfxtArgs> args;   // per query: its distance columns and index
    SfxtArgs base{};              // the group's shared level-loop arguments
    int lv = 0;                   // lift depth
    PinnedBuf<uint8_t> hmask;     // host copy of the subsets' sink flags (pinned: the copy is asynchronous)
    PinnedBuf<SfxtCtl> hctl;      // the group's SFXT controls, copied out before its event: graph's, then each query's
    PinnedBuf<unsigned> hmaxd;    // each query's largest finite |distf| (float bits), likewise
    hipEvent_t ev = nullptr;
    ~Graph() {
      if (ev) (void)hipEventDestroy(ev);
    }
  };
  struct Query {
    int64_t id = 0;
    int graph = 0;
    int64_t k = 0;
    std::vector<uint8_t> mask;  // n bytes, or empty for every sink
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
    int slot = -1;
    int state = 0;              // 0 queued, 1 in flight, 2 done
    std::unique_ptr<IndexSet> index;
    std::vector<int> roots;     // kept roots (vertex ids, ascending)
    uint64_t seed_n = 0;        // paths of the first-bound probe left in the graph handle's store
    int64_t node_est = 0;       // node-store estimate held while admitted (0 = not held)
    MqQuery init{};             // slot state at admission
    std::vector<float> result;
  };
  struct Store {
    DevBuf<float> cost;
    DevBuf<int> to, from, parent, level, nq;
    DevBuf<uint32_t> key;
    DevBuf<uint8_t> flag;
    uint64_t cap = 0;
    void reserve(uint64_t c) {
      cost.reserve(c); to.reserve(c); from.reserve(c); parent.reserve(c); level.reserve(c);
      nq.reserve(c); key.reserve(c + 16); flag.reserve(c + 16);
      cap = c;
    }
    NodeStore view() { return NodeStore{cost.p, to.p, from.p, parent.p, level.p, key.p, flag.p, cap}; }
  };

  std::vector<std::unique_ptr<Graph>> graphs;
  std::deque<int64_t> queue;  // submitted, not yet admitted
  std::unordered_map<int64_t, std::unique_ptr<Query>> queries;
  std::vector<int64_t> slot_query;  // query id per slot (-1 free)
  std::vector<uint8_t> slot_reserved;  // held by a query of a group in progress
  std::vector<int> free_slots, dead_slots;
  std::vector<int64_t> completed;   // finished since the last poll
  int64_t next_id = 1;
  int64_t budget = 0;
  int64_t index_bytes = 0;          // indexes of the in-flight queries
  // Node-store admission: every round scans the whole store (histograms,
  // batch selection), so its cost grows with the nodes the in-flight
  // queries hold.  A query is expected to store about kNodesPerK x K
  // paths (7.04 x K on the 1.5M-vertex DAG at K = 1M); admission keeps the
  // in-flight estimate within node_budget, a lone query always admitted.
  static constexpr int64_t kNodesPerK = 8;
  int64_t node_budget = int64_t(128) << 20;
  uint64_t dead_nodes = 0;          // nodes of finished queries still in the store
  int64_t nodes_held = 0;           // node estimates of the admitted queries
  static bool trace_on() {
    static const bool on = std::getenv("CPG_TRACE") != nullptr;
    return on;
  }
  static int trace_level() {
    static const int lv = std::getenv("CPG_TRACE") ? std::atoi(std::getenv("CPG_TRACE")) : 0;
    return lv;
  }
  // a query's node estimate, fixed when it is admitted (released with it)
  int64_t node_estimate(const Query& q) const {
    if (q.node_est > 0) return q.node_est;
    const double r = graphs[q.graph]->nodes_per_k;
    return static_cast<int64_t>(r * static_cast<double>(std::max<int64_t>(q.k, 0))) + 4096;
  }
  int nslots = 0;
  int in_flight = 0;
  hipStream_t stream = nullptr;
  int grid = 0;
  Store st[2];
  int cur = 0;
  uint64_t node_n = 0;
  std::vector<MqQuery> hq;
  DevBuf<uint32_t> batch;
  DevBuf<unsigned> hist, tiles, ghist;
  DevBuf<MqQuery> qs;
  DevBuf<MqCtl> mctl;
  DevBuf<int> newidx;
  DevBuf<uint8_t> sel;
  DevBuf<uint64_t> skey[2];
  DevBuf<uint32_t> sval[2];
  DevBuf<unsigned long long> nsel, per_q;
  DevBuf<float> out_cost;
  cpg_server_stats stats{};
  int64_t nonmono_base = 0;       // non-monotone children counted before the last ctl reset
  bool ready = false;
  hipStream_t sstream = nullptr;  // level-loop launches of admission groups
  hipEvent_t ev_levels = nullptr;
  std::mutex stats_mu;  // statistics the probes of several graphs update
  int parts_cap = 1;              // blocks one partitioned level-loop launch keeps resident

  uint64_t first_cap = uint64_t(1) << 24;

  cpg_server(int64_t budget_bytes, int max_in_flight, int64_t initial_nodes)
      : budget(budget_bytes > 0 ? budget_bytes : (int64_t(64) << 30)),
        nslots(max_in_flight > 0 ? std::min(max_in_flight, 256) : 64) {
    if (initial_nodes > 0) first_cap = std::min<uint64_t>(static_cast<uint64_t>(initial_nodes), 0x7fffffffull);
  }
  ~cpg_server() {
    if (ev_levels) (void)hipEventDestroy(ev_levels);
    if (sstream) (void)hipStreamDestroy(sstream);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void init() {
    if (ready) return;
    CPG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    CPG_HIP(hipStreamCreateWithFlags(&sstream, hipStreamNonBlocking));
    parts_cap = sfxt_parts_capacity();
    int dev = 0, cus = 0;
    CPG_HIP(hipGetDevice(&dev));
    CPG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid = std::max(1, cus) * 8;
    slot_query.assign(nslots, -1);
    slot_reserved.assign(nslots, 0);
    for (int i = nslots - 1; i >= 0; --i) free_slots.push_back(i);
    hq.assign(nslots, MqQuery{});
    for (auto& m : hq) m.done = 1;
    qs.reserve(nslots);
    h2d(qs.p, hq.data(), nslots, stream);
    hist.reserve(static_cast<size_t>(nslots) * kMqHistStride);
    mctl.reserve(1);
    const MqCtl z{};
    h2d(mctl.p, &z, 1, stream);
    sel.reserve(nslots);
    per_q.reserve(nslots);
    st[0].reserve(first_cap);
    batch.reserve(st[0].cap);
    CPG_HIP(hipStreamSynchronize(stream));
    ready = true;
  }

  int add_graph(const EdgeList& el) {
    auto G = std::make_unique<Graph>();
    G->h = std::make_unique<cpg_handle>();
    G->h->set_graph(el);
    G->h->upload_graph();
    G->h->run_sfxt(false);  // the level count (the same for every endpoint subset)
    const int64_t n = G->h->g.n, m = G->h->g.m;
    G->index_bytes = 24 * n + 16 * m;  // vrec 16, devinfo 8 per vertex; devs 16 per edge
    graphs.push_back(std::move(G));
    return static_cast<int>(graphs.size()) - 1;
  }

  int64_t submit(int graph, int64_t k, const uint8_t* mask) {
    if (graph < 0 || graph >= static_cast<int>(graphs.size())) throw std::runtime_error("no such graph");
    if (k < 0) throw std::runtime_error("k must be >= 0");
    auto q = std::make_unique<Query>();
    q->id = next_id++;
    q->graph = graph;
    q->k = k;
    if (mask) q->mask.assign(mask, mask + graphs[graph]->h->g.n);
    q->t_submit = now_ms();
    const int64_t id = q->id;
    queries[id] = std::move(q);
    queue.push_back(id);
    stats.submitted++;
    return id;
  }

  MqArgs args(const float* distf = nullptr) {
    MqArgs a{};
    a.nodes = st[cur].view();
    a.nq = st[cur].nq.p;
    a.distf = distf;
    a.batch = batch.p;
    a.hist = hist.p;
    a.qs = qs.p;
    a.ctl = mctl.p;
    a.nq_count = nslots;
    return a;
  }

  // The first bound of query q, whose SFXT columns and index are in place:
  // its roots, then the single query's probe on the graph handle's own node
  // store, which stores the roots (expanded) and their children up to a
  // bound u0 under which at least k paths lie (admission copies them into
  // the shared store, so the query's rounds start where a single query's
  // would).  A probe that finds no bound leaves the roots alone (u0 = no
  // bound, the first round expands them all).
  void probe_query(Query& q, cpg_handle& h, const SfxtArgs& aq, float max_abs_distf) {
    const int nsrc = static_cast<int>(h.g.srcs.size());
    std::vector<int> sd(nsrc);
    h.src_dist.reserve(std::max(1, nsrc));
    launch_gather_int(aq.dist, h.srcs.p, nsrc, h.src_dist.p, h.stream);
    d2h(sd.data(), h.src_dist.p, nsrc, h.stream);
    CPG_HIP(hipStreamSynchronize(h.stream));
    std::vector<uint8_t> kp;
    int64_t nroots = 0, dropped = 0;
    float minc = 0.0f;
    pick_roots(sd, true, 0.005f, kp, nroots, dropped, minc);
    const int64_t nkept = nroots - dropped;
    const uint32_t klo = nkept > 0 ? float_key(minc) : 0u;
    uint32_t u0 = kKeyInf;
    q.seed_n = 0;
    // At least k nonzero roots: the k-th smallest root cost bounds the
    // answer already (roots are paths), and the rounds start from the
    // roots alone, unexpanded, with no probe walks.
    std::vector<float> rc;
    for (int i = 0; i < nsrc; ++i)
      if (kp[i]) {
        const float c = static_cast<float>(sd[i]) / static_cast<float>(kScaleUp);
        if (c != 0.0f) rc.push_back(c);
      }
    // The K-th root cost bounds the answer, but loosely: the children of
    // the cheapest roots are cheaper than the K-th root, and a first round
    // under that bound stores every child up to it.  It is taken only
    // where the roots' chains are short (the single query's rule for
    // skipping its probe); otherwise the probe finds a tight bound.
    const bool short_chains = nkept * (static_cast<int64_t>(h.diameter) + 1) <= (int64_t(1) << 20);
    const bool roots_bound = q.k > 0 && static_cast<int64_t>(rc.size()) >= q.k && short_chains;
    if (roots_bound) {
      std::nth_element(rc.begin(), rc.begin() + (q.k - 1), rc.end());
      const uint64_t cap = std::max<uint64_t>(uint64_t(1) << 16, static_cast<uint64_t>(nkept));
      h.ensure_pfxt_buffers(std::max<uint64_t>(cap, h.node_cap));
      PfxtCtl init{};
      init.ukey = kKeyInf;
      init.node_cap = h.node_cap;
      h2d(h.pctl.p, &init, 1, h.stream);
      h.keep.reserve(std::max(1, nsrc));
      h2d(h.keep.p, kp.data(), nsrc, h.stream);
      launch_roots(h.srcs.p, h.keep.p, nsrc, aq.dist, aq.distf, h.node_store(), h.pctl.p, h.stream);
      PfxtCtl pc = h.read_pctl();
      u0 = float_key(rc[q.k - 1]);
      q.seed_n = pc.node_n;
      std::lock_guard<std::mutex> lk(stats_mu);
      stats.root_bounds++;
    } else if (nkept > 0 && q.k > 0) {
      uint64_t cap = std::min<uint64_t>(
          std::max<uint64_t>(uint64_t(1) << 22, 16ull * static_cast<uint64_t>(q.k) + nkept + (1u << 20)),
          0x7fffffffull);
      // a probe whose trial bound's children fill the handle's store runs
      // again in a larger one: without a bound the first round would expand
      // every root under a coarse bound and store far more
      for (int attempt = 0;; ++attempt) {
        h.ensure_pfxt_buffers(cap);
        PfxtCtl init{};
        init.ukey = kKeyInf;
        init.klo = klo;
        init.shift = window_shift(klo, kKeyInf);
        init.node_cap = cap;
        h2d(h.pctl.p, &init, 1, h.stream);
        h.keep.reserve(std::max(1, nsrc));
        h2d(h.keep.p, kp.data(), nsrc, h.stream);
        launch_roots(h.srcs.p, h.keep.p, nsrc, aq.dist, aq.distf, h.node_store(), h.pctl.p, h.stream);
        WalkArgs wa{};
        wa.vrec = aq.vrec; wa.devinfo = aq.devinfo; wa.devs = aq.devs; wa.batch = h.batch.p;
        wa.nodes = h.node_store(); wa.ckeys = h.ckeys.p; wa.ckeys_cap = cap; wa.ctl = h.pctl.p;
        wa.max_abs_distf = max_abs_distf; wa.max_abs_wgt = h.g.max_abs_wgt;
        wa.step_limit = h.step_limit();
        const double tp = now_ms();
        u0 = h.probe_bound(wa, q.k, kp, sd, minc, klo, false);
        {
          std::lock_guard<std::mutex> lk(stats_mu);
          stats.probe_ms += now_ms() - tp;
          if (trace_on())
            std::fprintf(stderr, "served t %.2f probe query %lld graph %d k %lld trials %d ms %.2f\n", now_ms(),
                         static_cast<long long>(q.id), q.graph, static_cast<long long>(q.k), h.timing.probes,
                         now_ms() - tp);
        }
        PfxtCtl pc = h.read_pctl();
        if (pc.err) throw std::runtime_error("PFXT walk exceeded its step bound");
        q.seed_n = std::min<uint64_t>(pc.node_n, cap);
        if (!(u0 == kKeyInf && h.probe_overflowed) || attempt >= 4 || cap >= 0x7fffffffull) break;
        cap = std::min<uint64_t>(cap * 2, 0x7fffffffull);
      }
    }
    MqQuery m{};
    m.ukey = u0;
    m.klo = klo;
    m.shift = window_shift_bins(m.klo, u0, kMqBins);
    m.k = static_cast<unsigned long long>(q.k);
    m.batch_target = std::max<unsigned long long>(256, m.k / 8);
    m.done = (nkept == 0 || q.k == 0) ? 1 : 0;
    m.vrec = aq.vrec;
    m.devinfo = aq.devinfo;
    m.devs = aq.devs;
    m.mag = 2.0f * max_abs_distf + h.g.max_abs_wgt + 1.0f;
    m.step_limit = h.step_limit();
    m.all = u0 == kKeyInf ? 1 : 0;
    q.init = m;
  }

  // Puts q in its reserved slot and copies its first-bound paths (still in
  // the graph handle's node store) into the shared store.  The server
  // stream must be idle (no round running).
  void admit_one(Query& q) {
    Graph& G = *graphs[q.graph];
    const uint64_t need = CPG_SERVER_NODES_PER_PATH * static_cast<uint64_t>(q.k) + q.seed_n +
                          CPG_SERVER_HEADROOM_NODES;
    if (node_n + need > st[cur].cap) {
      compact(std::max(st[cur].cap, node_n + 2 * need));
      if (node_n + need > st[cur].cap) compact(2 * (node_n + need));
    }
    const int slot = q.slot;  // reserved when its group started
    slot_reserved[slot] = 0;
    q.state = 1;
    q.t_admit = now_ms();
    slot_query[slot] = q.id;
    ++in_flight;
    h2d(qs.p + slot, &q.init, 1, stream);
    if (q.seed_n > 0)
      launch_mq_seed(G.h->node_store(), q.seed_n, st[cur].view(), st[cur].nq.p, node_n, slot,
                     mctl.p, grid, stream);
    node_n += q.seed_n;
    CPG_HIP(hipStreamSynchronize(stream));
    stats.admitted++;
  }

  // Starts admission groups on several idle graphs (at most kMaxParts):
  // reserves their queries' slots and budget, then on the server's SFXT
  // stream one cooperative launch runs every group's multi-subset level
  // loop, each graph on its own partition of blocks with its own barrier
  // (the launch takes as long as the slowest graph, not the sum).  Each
  // graph's stream then builds its queries' deviation indexes after the
  // launch and records the group's event; finish_groups completes the
  // admission once it has fired, while rounds of the queries in flight run
  // on the server stream meanwhile.  The level count depends on the graph
  // only, so the lift depth is known from the graph's first SFXT.
  void start_groups(const std::vector<std::pair<int, std::vector<int64_t>>>& list) {
    const double t0 = now_ms();
    SfxtParts parts{};
    std::vector<int> part_of(list.size(), -1);
    for (size_t li = 0; li < list.size(); ++li) {
      const int gi = list[li].first;
      const std::vector<int64_t>& ids = list[li].second;
      Graph& G = *graphs[gi];
      cpg_handle& h = *G.h;
      const size_t n = static_cast<size_t>(h.g.n), m = static_cast<size_t>(h.g.m);
      int Q = 1;
      while (Q < static_cast<int>(ids.size())) Q <<= 1;
      for (int64_t id : ids) {
        Query& q = *queries[id];
        q.slot = free_slots.back();
        free_slots.pop_back();
        slot_reserved[q.slot] = 1;
        q.state = 3;
        index_bytes += G.index_bytes;
        q.node_est = node_estimate(q);
        nodes_held += q.node_est;
      }
      h.deps.reserve(n + kSfxtPad); h.queue.reserve(n); h.lvlp.reserve(n + 2); h.level.reserve(n);
      h.big.reserve(n); h.sctl.reserve(1);
      if (h.g.max_out_degree > kSfxtLdsSort) h.devs_tmp.reserve(m);
      Graph::Multi& M = G.mm;
      M.dv.reserve(n * Q); M.dist.reserve(n * Q); M.succ.reserve(n * Q); M.distf.reserve(n * Q);
      const size_t ns = h.g.sinks.size();
      M.masks.reserve(std::max<size_t>(1, ns * Q)); M.maxd.reserve(Q); M.ctls.reserve(Q);
      // one flag per sink (in sink order) and subset; padding subsets choose none
      G.hmask.reserve(std::max<size_t>(1, ns * Q));  // kept until the group finishes (asynchronous copy)
      std::memset(G.hmask.p, 0, ns * Q);
      for (size_t q = 0; q < ids.size(); ++q) {
        const Query& qq = *queries[ids[q]];
        for (size_t i = 0; i < ns; ++i)
          G.hmask.p[q * ns + i] = qq.mask.empty() ? 1 : (qq.mask[h.g.sinks[i]] ? 1 : 0);
      }
      h2d(M.masks.p, G.hmask.p, ns * Q, sstream);
      CPG_HIP(hipMemsetAsync(h.sctl.p, 0, sizeof(SfxtCtl), sstream));
      CPG_HIP(hipMemsetAsync(M.ctls.p, 0, sizeof(SfxtCtl) * Q, sstream));
      SfxtArgs a{};
      a.n = h.g.n;
      a.nsinks = static_cast<int>(ns);
      a.oadjp = h.oadjp.p; a.oadj = h.oadj.p; a.owgt = h.owgt.p;
      a.iadjp = h.iadjp.p; a.iedge = h.iedge.p; a.sinks = h.sinks.p;
      a.deps = h.deps.p; a.queue = h.queue.p; a.lvlp = h.lvlp.p; a.level = h.level.p;
      a.big = h.big.p;
      a.devs_tmp = h.g.max_out_degree > kSfxtLdsSort ? h.devs_tmp.p : nullptr;
      a.ctl = h.sctl.p;
      SfxtMulti mm{};
      mm.q = Q;
      mm.masks = M.masks.p; mm.dv = M.dv.p; mm.dist = M.dist.p; mm.succ = M.succ.p;
      mm.distf = M.distf.p; mm.maxd = M.maxd.p;
      const int lv = sfxt_lift_levels(h.diameter);
      h.lift.reserve(static_cast<size_t>(lv) * std::max<size_t>(n, 1));
      G.args.clear();
      G.base = a;
      G.lv = lv;
      if (n > 0) {
        launch_sfxt_init_multi(a, mm, sstream);
        // one wave per vertex: about one block per 8 vertices of an average level
        const int64_t per_level = h.diameter > 0 ? static_cast<int64_t>(n) / h.diameter : 4096;
        SfxtPart& P = parts.p[parts.n];
        P.a = a;
        P.m = mm;
        P.nb = static_cast<int>(std::min<int64_t>(128, std::max<int64_t>(16, per_level / 8)));
        part_of[li] = parts.n++;
      }
      G.ids = ids;
    }
    stats.peak_index_bytes = std::max(stats.peak_index_bytes, index_bytes);
    if (parts.n > 0) {
      // fit every partition into one resident launch (admission forms at
      // most parts_cap groups, so one block each always fits)
      const int cap = parts_cap;
      int total = 0;
      for (int i = 0; i < parts.n; ++i) total += parts.p[i].nb;
      if (total > cap) {
        for (int i = 0; i < parts.n; ++i)
          parts.p[i].nb = std::max(1, static_cast<int>(static_cast<int64_t>(parts.p[i].nb) * cap / total));
        for (;;) {
          int sum = 0, big = 0;
          for (int i = 0; i < parts.n; ++i) {
            sum += parts.p[i].nb;
            if (parts.p[i].nb > parts.p[big].nb) big = i;
          }
          if (sum <= cap || parts.p[big].nb <= 1) break;
          --parts.p[big].nb;
        }
      }
      total = 0;
      for (int i = 0; i < parts.n; ++i) {
        parts.p[i].b0 = total;
        total += parts.p[i].nb;
      }
      launch_sfxt_levels_parts(parts, total, sstream);
      if (trace_on())
        for (size_t li = 0; li < list.size(); ++li)
          std::fprintf(stderr, "served t %.2f start graph %d queries %zu blocks %d\n", now_ms(), list[li].first,
                       list[li].second.size(), part_of[li] >= 0 ? parts.p[part_of[li]].nb : 0);
    }
    const double t_mid = now_ms();
    if (!ev_levels) CPG_HIP(hipEventCreateWithFlags(&ev_levels, hipEventDisableTiming));
    CPG_HIP(hipEventRecord(ev_levels, sstream));
    for (size_t li = 0; li < list.size(); ++li) {
      Graph& G = *graphs[list[li].first];
      cpg_handle& h = *G.h;
      const size_t n = static_cast<size_t>(h.g.n), m = static_cast<size_t>(h.g.m);
      CPG_HIP(hipStreamWaitEvent(h.stream, ev_levels, 0));
      for (size_t q = 0; q < G.ids.size(); ++q) {
        Query& qq = *queries[G.ids[q]];
        std::unique_ptr<IndexSet> ix;
        if (!G.pool.empty()) {
          ix = std::move(G.pool.back());
          G.pool.pop_back();
        } else {
          ix = std::make_unique<IndexSet>();
        }
        ix->vrec.reserve(std::max<size_t>(n, 1));
        ix->devinfo.reserve(std::max<size_t>(n, 1));
        ix->devs.reserve(std::max<size_t>(m, 1));
        SfxtArgs aq = G.base;
        aq.dist = G.mm.dist.p + q * n;
        aq.succ = G.mm.succ.p + q * n;
        aq.distf = G.mm.distf.p + q * n;
        aq.vrec = ix->vrec.p; aq.devinfo = ix->devinfo.p; aq.devs = ix->devs.p;
        aq.lift = h.lift.p;
        aq.ctl = G.mm.ctls.p + q;
        if (n > 0) launch_sfxt_index(aq, G.lv, h.g.m, h.stream);
        qq.index = std::move(ix);
        G.args.push_back(aq);
      }
      // the values check_group reads, copied out ahead of the group's event
      // (so that finishing the group waits for no copy queued behind other
      // streams' work)
      const size_t qa = G.args.size();
      G.hctl.reserve(qa + 1);
      G.hmaxd.reserve(std::max<size_t>(qa, 1));
      if (n > 0) {
        d2h(G.hctl.p, h.sctl.p, 1, h.stream);
        d2h(G.hctl.p + 1, G.mm.ctls.p, qa, h.stream);
        d2h(G.hmaxd.p, G.mm.maxd.p, qa, h.stream);
      }
      if (!G.ev) CPG_HIP(hipEventCreateWithFlags(&G.ev, hipEventDisableTiming));
      CPG_HIP(hipEventRecord(G.ev, h.stream));
      G.busy = true;
      stats.sfxt_groups++;
    }
    if (trace_on())
      std::fprintf(stderr, "served t %.2f groups started: %zu graphs, loops %.2f ms, indexes %.2f ms\n", now_ms(),
                   list.size(), t_mid - t0, now_ms() - t_mid);
    stats.sfxt_ms += now_ms() - t0;
  }

  // Completes started groups whose stream work has finished: checks each
  // group's level loop and indexes, then gives every query its first bound
  // and admits it.  The server stream must be idle.  A graph's queries are
  // probed one after another (a probe leaves its paths in the graph
  // handle's node store until admission has copied them), but the graphs'
  // probes run on host threads of their own, each on its graph's stream;
  // admissions, which use the shared store and the server stream, take a
  // lock one at a time.
  void finish_groups(const std::vector<int>& gis) {
    if (gis.empty()) return;
    if (trace_on())
      for (int gi : gis) std::fprintf(stderr, "served t %.2f ready graph %d\n", now_ms(), gi);
    std::vector<std::vector<unsigned>> maxd(gis.size());
    std::vector<std::vector<Query*>> qp(gis.size());  // looked up here, not in the threads
    // every group is checked before any is committed: a failed check puts
    // the queries of all these groups back on the queue with their slots
    // and index budget released, then reports the error
    try {
      for (size_t i = 0; i < gis.size(); ++i) maxd[i] = check_group(gis[i]);
    } catch (...) {
      for (auto it = gis.rbegin(); it != gis.rend(); ++it) unwind_group(*it);
      throw;
    }
    for (size_t i = 0; i < gis.size(); ++i) {
      graphs[gis[i]]->busy = false;
      for (int64_t id : graphs[gis[i]]->ids) qp[i].push_back(queries.at(id).get());
    }
    auto run = [&](size_t i, std::mutex* mu) {
      Graph& G = *graphs[gis[i]];
      for (size_t q = 0; q < G.ids.size(); ++q) {
        Query& qq = *qp[i][q];
        float mad = 0.0f;
        std::memcpy(&mad, &maxd[i][q], 4);
        probe_query(qq, *G.h, G.args[q], mad);
        if (mu) {
          std::lock_guard<std::mutex> lk(*mu);
          admit_one(qq);
        } else {
          admit_one(qq);
        }
      }
    };
    if (gis.size() == 1) {
      run(0, nullptr);
      return;
    }
    int dev = 0;
    CPG_HIP(hipGetDevice(&dev));
    std::mutex admit_mu;
    std::vector<std::exception_ptr> errs(gis.size());
    std::vector<std::thread> th;
    th.reserve(gis.size());
    for (size_t i = 0; i < gis.size(); ++i)
      th.emplace_back([&, i]() {
        try {
          CPG_HIP(hipSetDevice(dev));
          run(i, &admit_mu);
        } catch (...) {
          errs[i] = std::current_exception();
        }
      });
    for (auto& t : th) t.join();
    for (auto& e : errs)
      if (e) std::rethrow_exception(e);
  }

  // Checks a started group whose stream work has finished (its level loop
  // and its queries' indexes) and returns each query's largest finite
  // |distf| as float bits.
  std::vector<unsigned> check_group(int gi) {
    Graph& G = *graphs[gi];
    cpg_handle& h = *G.h;
    const double t0 = now_ms();
    const int Q = static_cast<int>(G.args.size());
    SfxtCtl c{};
    std::vector<SfxtCtl> cq(std::max(Q, 1));
    std::vector<unsigned> maxd(std::max(Q, 1), 0u);
    if (h.g.n > 0) {  // copied out before the group's event, which has fired
      c = G.hctl.p[0];
      for (int q = 0; q < Q; ++q) {
        cq[q] = G.hctl.p[1 + q];
        maxd[q] = G.hmaxd.p[q];
      }
    }
    if (h.g.n > 0) {
      if (c.err) throw std::runtime_error("SFXT grid barrier did not complete");
      if (c.num_levels != h.diameter) throw std::runtime_error("SFXT levels differ from the graph's");
      for (int q = 0; q < Q; ++q) {
        if (cq[q].nsv_err == 2) throw std::runtime_error("SFXT long deviation list without sort buffer");
        if (cq[q].nsv_err) throw std::runtime_error("SFXT next-smaller descent found no answer");
      }
    }
    stats.sfxt_ms += now_ms() - t0;
    return maxd;
  }

  // Returns a started group's queries to the front of the queue (in their
  // order), with their reserved slots, index buffers and budget released.
  void unwind_group(int gi) {
    Graph& G = *graphs[gi];
    for (auto it = G.ids.rbegin(); it != G.ids.rend(); ++it) {
      Query& q = *queries.at(*it);
      if (q.slot >= 0) {
        slot_reserved[q.slot] = 0;
        free_slots.push_back(q.slot);
      }
      q.slot = -1;
      q.state = 0;
      if (q.index) G.pool.push_back(std::move(q.index));
      index_bytes -= G.index_bytes;
      nodes_held -= q.node_est;
      q.node_est = 0;
      queue.push_front(q.id);
    }
    G.ids.clear();
    G.args.clear();
    G.busy = false;
  }

  // Copies the live nodes (slots not done) into the other buffer set, of
  // capacity at least want; the finished queries' slots become free.
  void compact(uint64_t want) {
    const double t0 = now_ms();
    const uint64_t n = node_n;
    const uint64_t nt = mq_compact_tiles(n);
    tiles.reserve(nt + 1);
    launch_zero_u32(tiles.p + nt, 1, stream);
    MqArgs a = args();
    launch_mq_count_live(a, n, tiles.p, stream);
    launch_exclusive_scan_u32(tiles.p, nt + 1, stream);
    unsigned live = 0;
    d2h(&live, tiles.p + nt, 1, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    Store& d = st[cur ^ 1];
    d.reserve(std::max<uint64_t>({want, st[cur].cap, 1}));
    newidx.reserve(std::max<uint64_t>(n, 1));
    launch_mq_compact(a, n, d.view(), d.nq.p, tiles.p, newidx.p, live, grid, stream);
    cur ^= 1;
    node_n = live;
    MqCtl c{};
    c.node_n = live;
    nonmono_base = stats.nonmonotone;
    h2d(mctl.p, &c, 1, stream);
    batch.reserve(st[cur].cap);
    CPG_HIP(hipStreamSynchronize(stream));
    for (int sl : dead_slots) free_slots.push_back(sl);
    dead_slots.clear();
    dead_nodes = 0;
    stats.compactions++;
    stats.compact_ms += now_ms() - t0;
  }

  // Admission between rounds (the server stream is idle).  Groups whose
  // stream work has finished are admitted.  Then, in FIFO order, each graph
  // without a group in progress forms one with the first queued query on
  // it and the next queued queries on the same graph (at most 16), while
  // slots are free and the in-flight and starting indexes stay within the
  // budget (a lone query is taken whatever its size); the groups formed
  // start in at most two level-loop launches (short loops first).  With nothing in flight,
  // admission waits for the first group in progress.
  void admit() {
    finish_groups(ready_groups());
    std::vector<std::pair<int, std::vector<int64_t>>> batch;
    std::vector<uint8_t> in_batch(graphs.size(), 0);
    int held = 0;             // slots promised to the groups formed so far
    int64_t held_bytes = 0;
    int64_t held_nodes = 0;
    while (!queue.empty() && static_cast<int>(batch.size()) < std::min(kMaxParts, parts_cap)) {
      if (static_cast<int>(free_slots.size()) - held <= 0) {
        if (dead_slots.empty()) break;
        compact(st[cur].cap);
        continue;
      }
      int gi = -1;
      for (int64_t id : queue) {
        const int g = queries[id]->graph;
        if (!graphs[g]->busy && !in_batch[g]) {
          gi = g;
          break;
        }
      }
      if (gi < 0) break;
      const int64_t gb = graphs[gi]->index_bytes;
      const bool idle = in_flight == 0 && !any_busy() && batch.empty();
      const int avail = static_cast<int>(free_slots.size()) - held;
      std::vector<int64_t> ids;
      int64_t bytes = index_bytes + held_bytes;
      int64_t nodes = nodes_held + held_nodes;
      for (int64_t id : queue) {
        if (static_cast<int>(ids.size()) >= std::min(16, avail)) break;
        if (queries[id]->graph != gi) continue;
        const int64_t ne = node_estimate(*queries[id]);
        // the store as it stands (live and not yet compacted) counts too
        const int64_t store = static_cast<int64_t>(std::min<uint64_t>(node_n, st[cur].cap));
        if ((!idle || !ids.empty()) &&
            (bytes + gb > budget || nodes + ne > node_budget || store > node_budget))
          break;
        ids.push_back(id);
        bytes += gb;
        nodes += ne;
      }
      if (ids.empty()) break;
      for (int64_t id : ids) queue.erase(std::find(queue.begin(), queue.end(), id));
      held += static_cast<int>(ids.size());
      held_bytes += gb * static_cast<int64_t>(ids.size());
      for (int64_t id : ids) held_nodes += node_estimate(*queries[id]);
      in_batch[gi] = 1;
      batch.emplace_back(gi, std::move(ids));
    }
    if (!batch.empty()) {
      // A launch lasts as long as its longest level loop, so graphs of at
      // most half the batch's largest level count start in a launch of
      // their own first: their queries reach the rounds without waiting
      // for the long loops, which follow on the same stream.
      int dmax = 0;
      for (const auto& b : batch) dmax = std::max(dmax, static_cast<int>(graphs[b.first]->h->diameter));
      std::vector<std::pair<int, std::vector<int64_t>>> shorter, longer;
      for (auto& b : batch)
        (2 * static_cast<int>(graphs[b.first]->h->diameter) <= dmax ? shorter : longer).push_back(std::move(b));
      if (!shorter.empty()) start_groups(shorter);
      start_groups(longer);
    }
    if (in_flight == 0 && any_busy()) {
      // nothing to run meanwhile: wait for whichever groups finish first
      // (the graphs' index builds end at different times)
      const double tw = now_ms();
      std::vector<int> r = ready_groups();
      while (r.empty()) {
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        r = ready_groups();
      }
      stats.group_wait_ms += now_ms() - tw;
      finish_groups(r);
    }
  }

  // The graphs whose started group has finished its stream work.
  std::vector<int> ready_groups() {
    std::vector<int> r;
    for (size_t gi = 0; gi < graphs.size(); ++gi) {
      Graph& G = *graphs[gi];
      if (!G.busy) continue;
      const hipError_t e = hipEventQuery(G.ev);
      if (e == hipSuccess) r.push_back(static_cast<int>(gi));
      else if (e != hipErrorNotReady) CPG_HIP(e);
    }
    return r;
  }

  bool any_busy() const {
    for (const auto& G : graphs)
      if (G->busy) return true;
    return false;
  }

  int64_t preparing() const {
    int64_t k = 0;
    for (const auto& G : graphs)
      if (G->busy) k += static_cast<int64_t>(G->ids.size());
    return k;
  }

  // Sorted costs of the finished slots' paths (at or below each bound).
  void extract(const std::vector<int>& fin) {
    const double t0 = now_ms();
    std::vector<uint8_t> hs(nslots, 0);
    for (int sl : fin) hs[sl] = 1;
    h2d(sel.p, hs.data(), nslots, stream);
    const uint64_t nn = std::min<uint64_t>(node_n, st[cur].cap);
    skey[0].reserve(std::max<uint64_t>(1, nn)); skey[1].reserve(std::max<uint64_t>(1, nn));
    sval[0].reserve(std::max<uint64_t>(1, nn)); sval[1].reserve(std::max<uint64_t>(1, nn));
    nsel.reserve(1);
    CPG_HIP(hipMemsetAsync(nsel.p, 0, sizeof(unsigned long long), stream));
    CPG_HIP(hipMemsetAsync(per_q.p, 0, sizeof(unsigned long long) * nslots, stream));
    launch_mq_build_keys(args(), sel.p, nn, skey[0].p, sval[0].p, nsel.p, per_q.p, grid, stream);
    unsigned long long nl = 0;
    std::vector<unsigned long long> pq(nslots);
    d2h(&nl, nsel.p, 1, stream);
    d2h(pq.data(), per_q.p, nslots, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    ghist.reserve(sort_ghist_len(nl));
    SortBufs sb{{skey[0].p, skey[1].p}, {sval[0].p, sval[1].p}, ghist.p, nl};
    const int bytes[7] = {0, 1, 2, 3, 4, 5, 6};  // slot < 256: bits 48..55
    const int res = radix_sort_pairs(sb, nl, bytes, 7, stream);
    out_cost.reserve(std::max<uint64_t>(1, nl));
    launch_mq_gather_costs(sval[res].p, st[cur].cost.p, nl, out_cost.p, stream);
    std::vector<float> all(nl);
    d2h(all.data(), out_cost.p, nl, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    const double t1 = now_ms();
    uint64_t off = 0;
    for (int sl = 0; sl < nslots; ++sl) {
      if (!hs[sl]) continue;
      Query& q = *queries[slot_query[sl]];
      const uint64_t take = std::min<uint64_t>(pq[sl], static_cast<uint64_t>(q.k));
      {
        Graph& Gq = *graphs[q.graph];
        const double used = static_cast<double>(hq[sl].stored + q.seed_n + q.roots.size());
        const double r = used / static_cast<double>(std::max<int64_t>(q.k, 1));
        Gq.nodes_per_k = 0.5 * Gq.nodes_per_k + 0.5 * std::max(1.0, r);
        dead_nodes += static_cast<uint64_t>(used);
        if (trace_on())
          std::fprintf(stderr, "served t %.2f query %lld graph %d k %lld stored %llu seed %llu roots %zu ratio %.1f\n",
                       now_ms(), static_cast<long long>(q.id), q.graph, static_cast<long long>(q.k),
                       static_cast<unsigned long long>(hq[sl].stored),
                       static_cast<unsigned long long>(q.seed_n), q.roots.size(), r);
      }
      q.result.assign(all.begin() + off, all.begin() + off + take);
      off += pq[sl];
      q.state = 2;
      q.t_done = t1;
      completed.push_back(q.id);
      Graph& G = *graphs[q.graph];
      G.pool.push_back(std::move(q.index));
      index_bytes -= G.index_bytes;
      nodes_held -= q.node_est;
      q.node_est = 0;
      --in_flight;
      slot_query[sl] = -1;
      dead_slots.push_back(sl);
      stats.completed++;
    }
    stats.extract_ms += now_ms() - t0;
  }

  // A round overflowed the store: every in-flight query starts again from
  // its roots (its index is kept) in an empty store of twice the capacity.
  void restart_all() {
    std::vector<int64_t> back;
    for (int sl = 0; sl < nslots; ++sl) {
      if (slot_query[sl] < 0) continue;
      Query& q = *queries[slot_query[sl]];
      q.state = 0;
      q.slot = -1;
      back.push_back(q.id);
      index_bytes -= graphs[q.graph]->index_bytes;
      nodes_held -= q.node_est;
      q.node_est = 0;
      graphs[q.graph]->pool.push_back(std::move(q.index));  // prepared again on re-admission
      slot_query[sl] = -1;
    }
    for (auto it = back.rbegin(); it != back.rend(); ++it) queue.push_front(*it);
    in_flight = 0;
    for (auto& m : hq) m = MqQuery{}, m.done = 1;
    h2d(qs.p, hq.data(), nslots, stream);
    free_slots.clear();
    dead_slots.clear();
    dead_nodes = 0;
    for (int i = nslots - 1; i >= 0; --i)
      if (!slot_reserved[i]) free_slots.push_back(i);
    const uint64_t c = std::min<uint64_t>(2 * st[cur].cap, 0x7fffffffull);
    st[cur].reserve(c);
    batch.reserve(c);
    node_n = 0;
    const MqCtl z{};
    nonmono_base = stats.nonmonotone;
    h2d(mctl.p, &z, 1, stream);
    CPG_HIP(hipStreamSynchronize(stream));
    stats.restarts++;
  }

  // Admission, one round for every in-flight query, extraction of the
  // finished ones.  Returns the queries still queued or in flight.
  int64_t step() {
    init();
    admit();
    if (in_flight > 0) {
      const double t0 = now_ms();
      MqArgs a = args();
      launch_mq_round(a, false, grid, stream);
      MqCtl mc{};
      d2h(&mc, mctl.p, 1, stream);
      d2h(hq.data(), qs.p, nslots, stream);
      CPG_HIP(hipStreamSynchronize(stream));
      stats.rounds++;
      stats.round_ms += now_ms() - t0;
      if (mc.err) throw std::runtime_error("served PFXT walk exceeded its step bound");
      stats.nonmonotone = nonmono_base + static_cast<int64_t>(mc.nonmono);
      node_n = mc.node_n;
      stats.peak_nodes = std::max<int64_t>(stats.peak_nodes, static_cast<int64_t>(std::min(node_n, st[cur].cap)));
      if (trace_level() >= 2)
        for (int sl = 0; sl < nslots; ++sl)
          if (slot_query[sl] >= 0)
            std::fprintf(stderr, "  slot %d query %lld U %.9g th %.9g stored %llu done %d rounds %d batch %llu\n", sl,
                         static_cast<long long>(slot_query[sl]), static_cast<double>(key_float(hq[sl].ukey)),
                         static_cast<double>(key_float(hq[sl].thkey)),
                         static_cast<unsigned long long>(hq[sl].stored), hq[sl].done, hq[sl].rounds,
                         static_cast<unsigned long long>(mc.batch_n));
      if (trace_on() && (stats.rounds % 16 == 0 || mc.overflow || trace_level() >= 2))
        std::fprintf(stderr, "served round %lld node_n %llu cap %llu in_flight %d queued %zu dead %zu held %lld ms %.2f%s\n",
                     static_cast<long long>(stats.rounds), static_cast<unsigned long long>(node_n),
                     static_cast<unsigned long long>(st[cur].cap), in_flight, queue.size(), dead_slots.size(),
                     static_cast<long long>(nodes_held), now_ms() - t0, mc.overflow ? " OVERFLOW" : "");
      if (mc.overflow) {
        restart_all();
      } else {
        std::vector<int> fin;
        for (int sl = 0; sl < nslots; ++sl)
          if (slot_query[sl] >= 0 && hq[sl].done) fin.push_back(sl);
        if (!fin.empty()) extract(fin);
        // every round scans the whole store: drop finished queries' nodes
        // once they are a quarter of it (or the store is half full, or as
        // many queries finished as are in flight)
        if (!dead_slots.empty() && node_n > (uint64_t(1) << 22) &&
            (node_n > st[cur].cap / 2 || dead_nodes * 4 > node_n ||
             static_cast<int>(dead_slots.size()) >= std::max(1, in_flight)))
          compact(st[cur].cap);
      }
    }
    stats.in_flight = in_flight;
    stats.queued = static_cast<int64_t>(queue.size()) + preparing();
    stats.node_cap = static_cast<int64_t>(st[cur].cap);
    return stats.in_flight + stats.queued;
  }

  Query& get(int64_t id) {
    auto it = queries.find(id);
    if (it == queries.end()) throw std::runtime_error("no such query");
    return *it->second;
  }
};

// ------------------------------------------------------------------ ABI

// Runs f, turning any exception into fail_value and a thread-local message.
template <class R, class F>
R guarded(R fail_value, F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return fail_value;
  } catch (...) {
    g_last_error = "unknown error";
    return fail_value;
  }
}

#define CPG_CHECK_HANDLE(h, fail_value)                 \
  if (!(h)) {                                           \
    g_last_error = "null handle";                       \
    return fail_value;                                  \
  }

extern "C" {

cpg_handle* cpg_create(void) {
  return guarded<cpg_handle*>(nullptr, [&]() { return new cpg_handle(); });
}

void cpg_destroy(cpg_handle* h) { delete h; }

const char* cpg_last_error(void) { return g_last_error.c_str(); }

int cpg_load_text(cpg_handle* h, const char* path, int ignore_wgts) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    EdgeList el;
    parse_text(path ? path : "", ignore_wgts != 0, el);
    h->set_graph(el);
    return 0;
  });
}

int cpg_load_bin(cpg_handle* h, const char* path) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    EdgeList el;
    read_bin(path ? path : "", el);
    h->set_graph(el);
    return 0;
  });
}

int cpg_write_bin(cpg_handle* h, const char* path) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    if (!h->loaded) throw std::runtime_error("no graph loaded");
    write_bin(h->g, path ? path : "");
    return 0;
  });
}

int cpg_load_edges(cpg_handle* h, int n, int64_t m, const int* eu, const int* ev, const float* ew) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    if (n < 0 || m < 0 || (m > 0 && (!eu || !ev || !ew))) throw std::runtime_error("bad edge list");
    EdgeList el;
    el.n = n;
    el.u.assign(eu, eu + m);
    el.v.assign(ev, ev + m);
    el.w.assign(ew, ew + m);
    h->set_graph(el);
    return 0;
  });
}

int cpg_load_csr(cpg_handle* h, int n, int64_t m, const int* adjp, const int* adjncy, const float* wgts) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    if (n < 0 || m < 0 || !adjp || (m > 0 && (!adjncy || !wgts))) throw std::runtime_error("bad CSR");
    if (adjp[0] != 0 || adjp[n] != m) throw std::runtime_error("CSR offsets do not span the edges");
    EdgeList el;
    el.n = n;
    el.u.resize(m);
    el.v.assign(adjncy, adjncy + m);
    el.w.assign(wgts, wgts + m);
    for (int u = 0; u < n; ++u) {
      if (adjp[u + 1] < adjp[u]) throw std::runtime_error("CSR offsets decrease");
      for (int e = adjp[u]; e < adjp[u + 1]; ++e) el.u[e] = u;
    }
    h->set_graph(el);
    return 0;
  });
}

int cpg_get_csr(cpg_handle* h, int* adjp, int* adjncy, float* wgts) {
  CPG_CHECK_HANDLE(h, -1)
  if (adjp) std::memcpy(adjp, h->g.oadjp.data(), sizeof(int) * h->g.oadjp.size());
  if (adjncy) std::memcpy(adjncy, h->g.oadj.data(), sizeof(int) * h->g.oadj.size());
  if (wgts) std::memcpy(wgts, h->g.owgt.data(), sizeof(float) * h->g.owgt.size());
  return 0;
}

int64_t cpg_num_verts(cpg_handle* h) { return h ? h->g.n : -1; }
int64_t cpg_num_edges(cpg_handle* h) { return h ? h->g.m : -1; }
int cpg_graph_diameter(cpg_handle* h) { return h ? h->diameter : -1; }

int cpg_reset(cpg_handle* h) {
  CPG_CHECK_HANDLE(h, -1)
  h->reset();
  return 0;
}

void cpg_query_opts_default(cpg_query_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->struct_size = sizeof(cpg_query_opts);
  o->init_split_perc = 0.005f;
  o->root_tie_quirk = 1;
}

int cpg_query(cpg_handle* h, int64_t k, const cpg_query_opts* o) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    cpg_query_opts d;
    cpg_query_opts_default(&d);
    if (o) {
      if (o->struct_size != static_cast<int>(sizeof(cpg_query_opts)))
        throw std::runtime_error("cpg_query_opts.struct_size mismatch");
      d = *o;
    }
    h->query(k, d);
    return 0;
  });
}

int64_t cpg_num_paths(cpg_handle* h) { return h && h->have_result ? h->nres : -1; }

int64_t cpg_get_costs(cpg_handle* h, int64_t k, float* out) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int64_t>(-1, [&]() {
    if (!h->have_result) throw std::runtime_error("no query result");
    const int64_t n = std::min<int64_t>(std::max<int64_t>(k, 0), h->nres);
    if (n && out) {
      if (h->host_result) {
        std::memcpy(out, h->h_cost.data(), sizeof(float) * n);
      } else {  // only the costs cross PCIe (the other report arrays stay on the device)
        d2h(out, h->r_cost.p, static_cast<size_t>(n), h->stream);
        CPG_HIP(hipStreamSynchronize(h->stream));
      }
    }
    return n;
  });
}

int64_t cpg_get_nodes(cpg_handle* h, int64_t k, cpg_node* out) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int64_t>(-1, [&]() {
    h->fetch_result();
    const int64_t n = std::min<int64_t>(std::max<int64_t>(k, 0), h->nres);
    for (int64_t i = 0; i < n && out; ++i) {
      out[i].level = h->h_level[i];
      out[i].from = h->h_from[i];
      out[i].to = h->h_to[i];
      out[i].parent = h->h_parent[i];
      out[i].num_children = 0;
      out[i].cost = h->h_cost[i];
    }
    return n;
  });
}

int64_t cpg_get_path_vertices(cpg_handle* h, int64_t i, int* out, int64_t cap) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int64_t>(-1, [&]() {
    const std::vector<int> path = h->path_vertices(i);
    const int64_t len = static_cast<int64_t>(path.size());
    if (len > cap) return -len - 1000;
    if (out) std::memcpy(out, path.data(), sizeof(int) * len);
    return len;
  });
}

int cpg_timings(cpg_handle* h, cpg_timing* out) {
  CPG_CHECK_HANDLE(h, -1)
  if (out) *out = h->timing;
  return 0;
}

int64_t cpg_step_times(cpg_handle* h, double* ms, int64_t* paths, int64_t cap) {
  CPG_CHECK_HANDLE(h, -1)
  const int64_t n = static_cast<int64_t>(h->step_ms.size());
  const int64_t c = std::min(std::max<int64_t>(cap, 0), n);
  for (int64_t i = 0; i < c; ++i) {
    if (ms) ms[i] = h->step_ms[static_cast<size_t>(i)];
    if (paths) paths[i] = h->step_paths[static_cast<size_t>(i)];
  }
  return n;
}

int cpg_set_csr_reorder(cpg_handle* h, int method) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    if (method < 0 || method > 1) throw std::runtime_error("csr reorder method must be 0 (none) or 1 (level order)");
    if (method != h->csr_reorder) {
      h->csr_reorder = method;
      h->graph_on_device = false;  // re-uploaded (and relabelled) on next use
      h->relabelled = false;
      h->invalidate_cache_payload();
    }
    return 0;
  });
}

double cpg_reorder_ms(cpg_handle* h) { return h ? h->reorder_ms : -1.0; }

int cpg_sfxt_noalloc(int n, const int* d_oadjp, const int* d_iadjp, const int* d_iedge,
                     const int* d_sinks, int nsinks, int* d_dist, int* d_succ, int* d_level,
                     void* stream) {
  return guarded<int>(-1, [&]() {
    return sfxt_noalloc(n, d_oadjp, d_iadjp, reinterpret_cast<const int2*>(d_iedge), d_sinks, nsinks,
                        d_dist, d_succ, d_level, static_cast<hipStream_t>(stream));
  });
}

int cpg_sfxt_noalloc_max_verts(void) { return sfxt_noalloc_max_verts(); }

int cpg_sfxt(cpg_handle* h) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    if (!h->loaded) throw std::runtime_error("no graph loaded");
    h->timing.sfxt_ms = h->run_sfxt(false);
    return 0;
  });
}

int cpg_get_sfxt(cpg_handle* h, int* dist_out, int* succ_out, int* level_out) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    if (!h->loaded) throw std::runtime_error("no graph loaded");
    if (!h->sfxt_valid) h->run_sfxt(false);
    const size_t n = static_cast<size_t>(h->g.n);
    if (!h->relabelled) {
      if (dist_out) d2h(dist_out, h->dist.p, n, h->stream);
      if (succ_out) d2h(succ_out, h->succ.p, n, h->stream);
      if (level_out) d2h(level_out, h->level.p, n, h->stream);
      CPG_HIP(hipStreamSynchronize(h->stream));
      return 0;
    }
    // relabelled graph: entry v of each array is original vertex v
    std::vector<int> t(n);
    int* outs[3] = {dist_out, succ_out, level_out};
    int* srcs[3] = {h->dist.p, h->succ.p, h->level.p};
    for (int k = 0; k < 3; ++k) {
      if (!outs[k]) continue;
      d2h(t.data(), srcs[k], n, h->stream);
      CPG_HIP(hipStreamSynchronize(h->stream));
      for (size_t v = 0; v < n; ++v) {
        const int x = t[static_cast<size_t>(h->h_perm[v])];
        outs[k][v] = k == 1 ? h->to_orig(x) : x;
      }
    }
    return 0;
  });
}

int cpg_bvss_build(cpg_handle* h) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    if (!h->loaded) throw std::runtime_error("no graph loaded");
    h->bvss_build();
    return 0;
  });
}

int cpg_bvss_info(cpg_handle* h, int64_t* n_intervals, int64_t* n_entries, int64_t* set_bits) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int>(-1, [&]() {
    if (!h->hb.valid) throw std::runtime_error("BVSS index not built");
    if (n_intervals) *n_intervals = h->hb.n_int;
    if (n_entries) *n_entries = static_cast<int64_t>(h->hb.rows.size());
    if (set_bits) *set_bits = static_cast<int64_t>(h->hb.set_bits);
    return 0;
  });
}

int64_t cpg_bvss_row_sources(cpg_handle* h, int row, int* out, int64_t cap) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int64_t>(-1, [&]() -> int64_t {
    if (!h->hb.valid) throw std::runtime_error("BVSS index not built");
    if (row < 0 || row >= h->g.n) throw std::runtime_error("row out of range");
    std::vector<int> srcs;
    for (size_t e = 0; e < h->hb.rows.size(); ++e) {
      if (h->hb.rows[e] != row) continue;
      unsigned long long m = h->hb.masks[e];
      while (m) {
        const int bit = __builtin_ctzll(m);
        srcs.push_back(h->hb.entry_int[e] * 64 + bit);
        m &= m - 1;
      }
    }
    std::sort(srcs.begin(), srcs.end());
    const int64_t len = static_cast<int64_t>(srcs.size());
    if (len > cap) return -len - 1000;
    if (out) std::copy(srcs.begin(), srcs.end(), out);
    return len;
  });
}

int64_t cpg_bvss_discover(cpg_handle* h, int nfront, const int* fptr, const int* fsrc, int mode,
                          int dense_min, int* out_f, int* out_u, int* out_v, int64_t cap) {
  CPG_CHECK_HANDLE(h, -1)
  return guarded<int64_t>(-1, [&]() -> int64_t {
    if (!h->loaded) throw std::runtime_error("no graph loaded");
    return h->bvss_discover(nfront, fptr, fsrc, mode, dense_min, out_f, out_u, out_v, cap);
  });
}

cpg_batch* cpg_batch_create(void) {
  return guarded<cpg_batch*>(nullptr, []() { return new cpg_batch(); });
}

void cpg_batch_destroy(cpg_batch* b) { delete b; }

int cpg_batch_add_edges(cpg_batch* b, int n, int64_t m, const int* eu, const int* ev,
                        const float* ew, int64_t k) {
  if (!b) { g_last_error = "null batch"; return -1; }
  return guarded<int>(-1, [&]() {
    if (n < 0 || m < 0 || (m > 0 && (!eu || !ev || !ew))) throw std::runtime_error("bad edge list");
    if (k < 0) throw std::runtime_error("k must be >= 0");
    for (int64_t e = 0; e < m; ++e)
      if (eu[e] < 0 || eu[e] >= n || ev[e] < 0 || ev[e] >= n) throw std::runtime_error("vertex id out of range");
    cpg_batch::Design d;
    d.n = n;
    d.u.assign(eu, eu + m);
    d.v.assign(ev, ev + m);
    d.w.assign(ew, ew + m);
    d.k = k;
    b->designs.push_back(std::move(d));
    b->built = false;
    return static_cast<int>(b->designs.size()) - 1;
  });
}

int cpg_batch_num_queries(cpg_batch* b) { return b ? static_cast<int>(b->designs.size()) : -1; }

int cpg_batch_autotune(cpg_batch* b, const cpg_query_opts* o, int max_trials) {
  if (!b) { g_last_error = "null batch"; return -1; }
  return guarded<int>(-1, [&]() {
    cpg_query_opts d;
    cpg_query_opts_default(&d);
    if (o) {
      if (o->struct_size != static_cast<int>(sizeof(cpg_query_opts)))
        throw std::runtime_error("cpg_query_opts.struct_size mismatch");
      d = *o;
    }
    if (d.endpoint_mask) throw std::runtime_error("endpoint_mask is not supported in a batch");
    b->autotune(d, max_trials);
    return b->tune.trials;
  });
}

int cpg_batch_get_tuning(cpg_batch* b, cpg_batch_tuning* out) {
  if (!b || !out) { g_last_error = "null argument"; return -1; }
  *out = b->tune;
  return 0;
}

int cpg_batch_set_tuning(cpg_batch* b, const cpg_batch_tuning* t) {
  if (!b || !t) { g_last_error = "null argument"; return -1; }
  if (t->group < 1 || t->group > 64 || t->grid_mult < 1 || t->grid_mult > 64 || t->target_div[0] < 1 ||
      t->target_div[1] < 1 || t->target_div[2] < 1) {
    g_last_error = "tuning values out of range";
    return -1;
  }
  b->tune = *t;
  return 0;
}

int cpg_batch_run(cpg_batch* b, const cpg_query_opts* o) {
  if (!b) { g_last_error = "null batch"; return -1; }
  return guarded<int>(-1, [&]() {
    cpg_query_opts d;
    cpg_query_opts_default(&d);
    if (o) {
      if (o->struct_size != static_cast<int>(sizeof(cpg_query_opts)))
        throw std::runtime_error("cpg_query_opts.struct_size mismatch");
      d = *o;
    }
    if (d.endpoint_mask) throw std::runtime_error("endpoint_mask is not supported in a batch");
    b->run(d);
    return 0;
  });
}

int64_t cpg_batch_get_costs(cpg_batch* b, int q, int64_t k, float* out) {
  if (!b) { g_last_error = "null batch"; return -1; }
  return guarded<int64_t>(-1, [&]() -> int64_t {
    if (q < 0 || q >= static_cast<int>(b->results.size())) throw std::runtime_error("no result for that query");
    const auto& r = b->results[q];
    const int64_t n = std::min<int64_t>(std::max<int64_t>(k, 0), static_cast<int64_t>(r.size()));
    if (n > 0 && !out) throw std::runtime_error("null output");
    std::copy(r.begin(), r.begin() + n, out);
    return n;
  });
}

int cpg_batch_timings(cpg_batch* b, cpg_timing* out) {
  if (!b || !out) { g_last_error = "null argument"; return -1; }
  *out = b->timing;
  return 0;
}

cpg_server* cpg_server_create(int64_t budget_bytes, int max_in_flight, int64_t initial_nodes) {
  return guarded<cpg_server*>(nullptr, [&]() {
    return new cpg_server(budget_bytes, max_in_flight, initial_nodes);
  });
}

void cpg_server_destroy(cpg_server* s) { delete s; }

#define CPG_CHECK_SERVER(s, fail_value)                 \
  if (!(s)) {                                           \
    g_last_error = "null server";                       \
    return fail_value;                                  \
  }

int cpg_server_add_graph(cpg_server* s, int n, int64_t m, const int* eu, const int* ev,
                         const float* ew) {
  CPG_CHECK_SERVER(s, -1)
  return guarded<int>(-1, [&]() {
    if (n < 0 || m < 0 || (m > 0 && (!eu || !ev || !ew))) throw std::runtime_error("bad edge list");
    EdgeList el;
    el.n = n;
    el.u.assign(eu, eu + m);
    el.v.assign(ev, ev + m);
    el.w.assign(ew, ew + m);
    return s->add_graph(el);
  });
}

int64_t cpg_server_submit(cpg_server* s, int graph, int64_t k, const uint8_t* endpoint_mask) {
  CPG_CHECK_SERVER(s, -1)
  return guarded<int64_t>(-1, [&]() { return s->submit(graph, k, endpoint_mask); });
}

int64_t cpg_server_step(cpg_server* s) {
  CPG_CHECK_SERVER(s, -1)
  return guarded<int64_t>(-1, [&]() { return s->step(); });
}

int64_t cpg_server_poll(cpg_server* s, int64_t* ids, int64_t cap) {
  CPG_CHECK_SERVER(s, -1)
  return guarded<int64_t>(-1, [&]() {
    const int64_t n = std::min<int64_t>(std::max<int64_t>(cap, 0), static_cast<int64_t>(s->completed.size()));
    if (n > 0 && !ids) throw std::runtime_error("null ids");
    for (int64_t i = 0; i < n; ++i) ids[i] = s->completed[i];
    s->completed.erase(s->completed.begin(), s->completed.begin() + n);
    return n;
  });
}

int64_t cpg_server_get_costs(cpg_server* s, int64_t id, int64_t k, float* out) {
  CPG_CHECK_SERVER(s, -1)
  return guarded<int64_t>(-1, [&]() {
    auto& q = s->get(id);
    if (q.state != 2) throw std::runtime_error("query not finished");
    const int64_t n = std::min<int64_t>(std::max<int64_t>(k, 0), static_cast<int64_t>(q.result.size()));
    if (n && out) std::memcpy(out, q.result.data(), sizeof(float) * n);
    return n;
  });
}

int cpg_server_query_times(cpg_server* s, int64_t id, double* submit_ms, double* admit_ms,
                           double* done_ms) {
  CPG_CHECK_SERVER(s, -1)
  return guarded<int>(-1, [&]() {
    auto& q = s->get(id);
    if (submit_ms) *submit_ms = q.t_submit;
    if (admit_ms) *admit_ms = q.t_admit;
    if (done_ms) *done_ms = q.t_done;
    return 0;
  });
}

int cpg_server_release(cpg_server* s, int64_t id) {
  CPG_CHECK_SERVER(s, -1)
  return guarded<int>(-1, [&]() {
    auto& q = s->get(id);
    if (q.state != 2) throw std::runtime_error("query not finished");
    s->queries.erase(id);
    return 0;
  });
}

int cpg_server_get_stats(cpg_server* s, cpg_server_stats* out) {
  CPG_CHECK_SERVER(s, -1)
  if (!out) { g_last_error = "null argument"; return -1; }
  *out = s->stats;
  return 0;
}

double cpg_now_ms(void) { return now_ms(); }

void* cpg_host_alloc(int64_t bytes) {
  return guarded<void*>(nullptr, [&]() -> void* {
    if (bytes <= 0) throw std::runtime_error("cpg_host_alloc needs a positive size");
    void* p = nullptr;
    CPG_HIP(hipHostMalloc(&p, static_cast<size_t>(bytes), hipHostMallocDefault));
    return p;
  });
}

int cpg_host_free(void* p) {
  return guarded<int>(-1, [&]() {
    if (p) CPG_HIP(hipHostFree(p));
    return 0;
  });
}

int cpg_cache_enable(cpg_handle* h, int enabled) {
  CPG_CHECK_HANDLE(h, -1)
  h->cache_enabled = enabled != 0;
  return 0;
}
int cpg_cache_clear(cpg_handle* h) {
  CPG_CHECK_HANDLE(h, -1)
  h->invalidate_cache_payload();
  h->cache_hits = h->cache_misses = 0;
  return 0;
}
int cpg_cache_hits(cpg_handle* h) { return h ? h->cache_hits : -1; }
int cpg_cache_misses(cpg_handle* h) { return h ? h->cache_misses : -1; }

int cpg_gen_preset(int which, int index, int unit_wgts, cpg_edge_list* out) {
  if (!out) { g_last_error = "null output"; return -1; }
  return guarded<int>(-1, [&]() {
    EdgeList el;
    generate(preset(which, index, unit_wgts), el);
    const size_t m = el.u.size();
    out->n = el.n;
    out->m = static_cast<int64_t>(m);
    out->u = static_cast<int*>(std::malloc(sizeof(int) * (m ? m : 1)));
    out->v = static_cast<int*>(std::malloc(sizeof(int) * (m ? m : 1)));
    out->w = static_cast<float*>(std::malloc(sizeof(float) * (m ? m : 1)));
    if (!out->u || !out->v || !out->w) throw std::runtime_error("out of memory");
    std::memcpy(out->u, el.u.data(), sizeof(int) * m);
    std::memcpy(out->v, el.v.data(), sizeof(int) * m);
    std::memcpy(out->w, el.w.data(), sizeof(float) * m);
    return 0;
  });
}

void cpg_edge_list_free(cpg_edge_list* el) {
  if (!el) return;
  std::free(el->u);
  std::free(el->v);
  std::free(el->w);
  el->u = el->v = nullptr;
  el->w = nullptr;
  el->n = 0;
  el->m = 0;
}

int64_t cpg_preset_k(int which, int index) {
  return guarded<int64_t>(-1, [&]() { return preset_k(which, index); });
}

int cpg_write_text(const cpg_edge_list* el, const char* path) {
  if (!el || !path) { g_last_error = "null argument"; return -1; }
  return guarded<int>(-1, [&]() {
    EdgeList e;
    e.n = el->n;
    e.u.assign(el->u, el->u + el->m);
    e.v.assign(el->v, el->v + el->m);
    e.w.assign(el->w, el->w + el->m);
    write_text(e, path);
    return 0;
  });
}

}  // extern "C"
