// gpucpg_amd.hpp -- header-only C++ facade over the C ABI in cpg.h.
//
// Class gpucpg::CpGen keeps the method names, argument order and error
// behaviour of the reference class (gpucpg/gpucpg.cuh:119-431, paths
// relative to the gpu-cpg repository root), so a driver written against
// the reference compiles against this header and links libcpgamd.so:
//   read_input          gpucpg/gpucpg.cuh:123, gpucpg/gpucpg.cu:453-548
//   reset               gpucpg/gpucpg.cuh:211-250
//   report_paths        gpucpg/gpucpg.cuh:164-181 (GPG mode: the engine always
//                       runs its exact GPU path; method and tuning arguments
//                       are accepted and ignored, init_split_perc is used)
//   get_slacks          gpucpg/gpucpg.cuh:183, gpucpg/gpucpg.cu:21409-21420
//   get_pfxt_nodes      gpucpg/gpucpg.cuh:184, gpucpg/gpucpg.cu:21422-21440
//   num_verts/num_edges gpucpg/gpucpg.cuh:208-209
//   static cache        gpucpg/gpucpg.cuh:186-189
//   timing members      gpucpg/gpucpg.cuh:327-356 (short_long_step_times and
//                       paths_gen_per_step: one entry per expansion round)
//   write_to_csr_bin    gpucpg/gpucpg.cuh:310-323, binary layout gpucpg/graph.h:24-55
// Errors throw std::runtime_error with the library's message, as the
// reference throws on I/O and capacity failures (device errors there end
// the process; here they are reported the same way as other errors).
#ifndef GPUCPG_AMD_HPP
#define GPUCPG_AMD_HPP

#include <chrono>
#include <cmath>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "cpg.h"

namespace gpucpg {

// Enumerations of gpucpg/gpucpg.cuh:41-72, kept for call compatibility.
enum class PropDistMethod {
  BASIC = 0,
  CUDA_GRAPH,
  BFS_TOP_DOWN_PRIVATIZED,
  BFS_TOP_DOWN,
  BFS_PRIVATIZED_MERGED,
  BFS_HYBRID,
  BFS_HYBRID_PRIVATIZED,
  TEST_COUNT_MF,
  LEVELIZE_THEN_RELAX,
  LEVELIZE_HYBRID_THEN_RELAX,
  BFS_TD_RELAX_BU_PRIVATIZED
};
enum class PfxtExpMethod { BASIC = 0, PRECOMP_SPURS, ATOMIC_ENQ, SHORT_LONG, SEQUENTIAL };
enum class CsrReorderMethod {
  NONE = -1,
  V_ORIENTED = 0,
  V_ORIENTED_TILE_SCAN,
  E_ORIENTED,
  E_ORIENTED_VEC2,
  RABBIT,
  GORDER,
  CORDER
};

// PfxtNode, gpucpg/gpucpg.cuh:74-107.  parent is the rank of the parent
// path in the same report (-1 for a root or a parent not reported).
struct PfxtNode {
  int level = -1;
  int from = -1;
  int to = -1;
  int parent = -1;
  int num_children = 0;
  float slack = 0.0f;
};
static_assert(sizeof(PfxtNode) == sizeof(cpg_node), "PfxtNode layout");

class CpGen {
 public:
  using usec = std::chrono::duration<double, std::micro>;

  CpGen() : h_(cpg_create()) {
    if (!h_) throw std::runtime_error(cpg_last_error());
  }
  ~CpGen() { cpg_destroy(h_); }
  CpGen(const CpGen&) = delete;
  CpGen& operator=(const CpGen&) = delete;

  void read_input(const std::string& filename, bool ignore_wgts = false) {
    check(cpg_load_text(h_, filename.c_str(), ignore_wgts ? 1 : 0));
    benchmark_path = filename;
  }

  void reset() { check(cpg_reset(h_)); }

  void report_paths(const int k, const int /*max_dev_lvls*/, const bool /*enable_compress*/,
                    const PropDistMethod /*pd_method*/, const PfxtExpMethod /*pe_method*/,
                    const bool /*enable_runtime_log_file*/ = false,
                    const float init_split_perc = 0.005f, const float /*alpha*/ = 5.0f,
                    const int /*per_thread_work_items*/ = 8, bool /*enable_reindex_cpu*/ = false,
                    bool enable_reindex_gpu = false, bool /*enable_fuse_steps*/ = false,
                    bool enable_interm_perf_log = false,
                    const CsrReorderMethod csr_reorder_method = CsrReorderMethod::E_ORIENTED,
                    bool /*enable_tile_spur*/ = false,
                    std::optional<float> /*fixed_split_inc_amount*/ = std::nullopt,
                    std::optional<std::string> /*rabbit_vmap_file*/ = std::nullopt) {
    // The reference relabels the CSR only with enable_reindex_gpu, and its
    // V_/E_ORIENTED kernels all build the same level-order CSR
    // (gpucpg/gpucpg.cu:17690-17781); RABBIT/GORDER/CORDER orders are not
    // built here and run unrelabelled.  Results are the same either way.
    const bool level_order = enable_reindex_gpu && (csr_reorder_method == CsrReorderMethod::V_ORIENTED ||
                                                    csr_reorder_method == CsrReorderMethod::V_ORIENTED_TILE_SCAN ||
                                                    csr_reorder_method == CsrReorderMethod::E_ORIENTED ||
                                                    csr_reorder_method == CsrReorderMethod::E_ORIENTED_VEC2);
    check(cpg_set_csr_reorder(h_, level_order ? 1 : 0));
    cpg_query_opts o;
    cpg_query_opts_default(&o);
    o.init_split_perc = init_split_perc;
    o.step_times = 1;
    o.kernel_timing = enable_interm_perf_log ? 1 : 0;
    check(cpg_query(h_, k, &o));
    cpg_timing t{};
    check(cpg_timings(h_, &t));
    prop_time = usec(t.sfxt_ms * 1e3);
    expand_time = usec(t.pfxt_ms * 1e3);
    // levelize, relax and successor run as one pass over the levels: its
    // device time is lvlize_time (with enable_interm_perf_log) and relax_time
    // stays 0, so lvlize_time + relax_time is the time of both
    lvlize_time = usec(t.sfxt_levels_ms * 1e3);
    relax_time = usec(0.0);
    short_long_expansion_steps = static_cast<int>(t.rounds);
    total_gen_paths = static_cast<int>(t.generated);
    graph_diameter = cpg_graph_diameter(h_);
    const int64_t nr = cpg_step_times(h_, nullptr, nullptr, 0);
    check(nr < 0 ? -1 : 0);
    std::vector<double> ms(static_cast<size_t>(nr));
    std::vector<int64_t> paths(static_cast<size_t>(nr));
    cpg_step_times(h_, ms.data(), paths.data(), nr);
    short_long_step_times.clear();
    paths_gen_per_step.clear();
    for (int64_t i = 0; i < nr; ++i) {
      short_long_step_times.push_back(usec(ms[static_cast<size_t>(i)] * 1e3));
      paths_gen_per_step.push_back(static_cast<int>(paths[static_cast<size_t>(i)]));
    }
  }

  // gpucpg/gpucpg.cuh:292-299 (initial split increment of the reference's
  // schedule; the engine's exact rounds do not use it)
  float compute_split_inc_amount(float avg_deg) {
    const float lo = 0.1f, hi = 100.0f, d0 = 3.0f, kk = 0.8f;
    return lo + (hi - lo) / (1 + std::exp(kk * (avg_deg - d0)));
  }

  // gpucpg/gpucpg.cuh:310-323 with graph.h:24-38: the fanout CSR as
  // N, M, offsets[N+1], adjncy[M], wgts[M] (native-endian 4-byte words)
  void write_to_csr_bin(std::string& filename) { check(cpg_write_bin(h_, filename.c_str())); }
  // The same format back in (graph.h:40-55), as read_input does for text.
  void read_input_bin(const std::string& filename) {
    check(cpg_load_bin(h_, filename.c_str()));
    benchmark_path = filename;
  }

  std::vector<float> get_slacks(int k) {
    std::vector<float> out(k > 0 ? static_cast<size_t>(k) : 0);
    const int64_t n = cpg_get_costs(h_, k, out.data());
    check(n < 0 ? -1 : 0);
    out.resize(static_cast<size_t>(n));
    return out;
  }

  std::vector<PfxtNode> get_pfxt_nodes(int k) {
    std::vector<PfxtNode> out(k > 0 ? static_cast<size_t>(k) : 0);
    const int64_t n = cpg_get_nodes(h_, k, reinterpret_cast<cpg_node*>(out.data()));
    check(n < 0 ? -1 : 0);
    out.resize(static_cast<size_t>(n));
    return out;
  }

  // Vertex sequence (source ... sink) of the i-th reported path.
  std::vector<int> get_path_vertices(int i) {
    std::vector<int> out(64);
    for (;;) {
      const int64_t n = cpg_get_path_vertices(h_, i, out.data(), static_cast<int64_t>(out.size()));
      if (n >= 0) {
        out.resize(static_cast<size_t>(n));
        return out;
      }
      if (n > -1000) check(-1);
      out.resize(static_cast<size_t>(-(n + 1000)));
    }
  }

  void enable_tc_pfxt_static_cache(bool enabled) { check(cpg_cache_enable(h_, enabled ? 1 : 0)); }
  void clear_tc_pfxt_static_cache() { check(cpg_cache_clear(h_)); }
  int tc_pfxt_static_cache_hits() const { return cpg_cache_hits(h_); }
  int tc_pfxt_static_cache_misses() const { return cpg_cache_misses(h_); }

  size_t num_verts() const { return static_cast<size_t>(cpg_num_verts(h_)); }
  size_t num_edges() const { return static_cast<size_t>(cpg_num_edges(h_)); }

  // The library's own timing record of the last query.
  cpg_timing timings() const {
    cpg_timing t{};
    check(cpg_timings(h_, &t));
    return t;
  }
  cpg_handle* handle() const { return h_; }

  // timing members, gpucpg/gpucpg.cuh:327-356
  usec prop_time{0};
  usec expand_time{0};
  usec lvlize_time{0};
  usec prefix_scan_time{0};
  usec csr_reorder_time{0};
  usec relax_time{0};
  usec rabbit_copy_to_gpu_time{0};
  usec rabbit_update_csr_time{0};
  usec gorder_copy_to_gpu_time{0};
  usec gorder_update_csr_time{0};
  usec corder_copy_to_gpu_time{0};
  usec corder_update_csr_time{0};
  usec gpba_transfer_time{0};
  usec gpba_pfxt_expand_time{0};
  usec gpba_sort_and_prune_time{0};
  int short_long_expansion_steps{0};
  std::string benchmark_path;
  int graph_diameter{0};
  std::vector<int> paths_gen_per_step;
  std::vector<usec> short_long_step_times;
  int total_gen_paths{0};

 private:
  static void check(int rc) {
    if (rc < 0) throw std::runtime_error(cpg_last_error());
  }
  cpg_handle* h_;
};

}  // namespace gpucpg

#endif
