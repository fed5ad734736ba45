// Host-side timing-graph storage for the CPG engine: the benchmark text
// loader, CSR construction and the seeded synthetic layered-DAG generator.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cpg {

// Fanout + fanin CSR.  Within every vertex the edges keep input order,
// which is how CpGen::read_input builds its CSR
// (gpucpg/gpucpg.cu:507-546): per-vertex vectors appended in input order,
// concatenated in vertex-id order.
struct HostGraph {
  int n = 0;
  int64_t m = 0;
  std::vector<int> oadjp, oadj;   // fanout
  std::vector<float> owgt;
  std::vector<int> iadjp, iadj;   // fanin
  std::vector<float> iwgt;
  std::vector<int> srcs, sinks;   // in-degree 0 / out-degree 0, ascending id
  float max_abs_wgt = 0.0f;
  int max_out_degree = 0;
};

// Edge list in input order.
struct EdgeList {
  int n = 0;
  std::vector<int> u, v;
  std::vector<float> w;
};

// Builds the CSR pair from an edge list; throws std::runtime_error on an
// out-of-range vertex id.
void build_csr(const EdgeList& el, HostGraph& g);

// Benchmark text format (gpucpg/gpucpg.cu:453-548):
//   line 1: N; then N vertex-name lines (skipped);
//   then one edge per line:  "u" -> "v"[, w];   (default weight 1.0f,
//   ignore_wgts forces 1.0f).  Throws std::runtime_error on a malformed
//   line, matching the reference's std::stoi / std::stof failures.
//   The edge lines are parsed on up to 16 host threads (threads = 0: one per
//   hardware thread), joined in file order.
void parse_text(const std::string& path, bool ignore_wgts, EdgeList& el, int threads = 0);
void write_text(const EdgeList& el, const std::string& path);
// Binary fanout CSR, gpucpg/graph.h:24-55: int32 N, int32 M, offsets[N+1],
// adjncy[M], float wgts[M] (native-endian).  read_bin checks the size and
// the offsets and returns the edges in CSR order.
void write_bin(const HostGraph& g, const std::string& path);
void read_bin(const std::string& path, EdgeList& el);

// Seeded layered DAG generator.
struct GenParams {
  uint64_t seed = 1;
  int n = 50000;
  int levels = 120;
  double src_frac = 0.01;     // share of vertices on level 0 (sources)
  double sink_frac = 0.01;    // share of vertices on the last level (sinks)
  int deg_kind = 0;           // 0: uniform {deg_lo..deg_hi}; 1: geometric(mean)
  int deg_lo = 1, deg_hi = 4;
  double deg_mean = 2.7;
  int deg_cap = 32;
  int span_kind = 0;          // 0: uniform {1..span_hi}; 1: 1 w.p. p1 else {2..span_hi}
  int span_hi = 3;
  double span_p1 = 0.9;
  int sink_edge = 1;          // every non-sink vertex gets one edge to a random sink
  int wgt_kind = 0;           // 0: uniform [w_lo,w_hi]; 1: lognormal(mu,sigma)
  int unit_wgts = 0;          // replace every drawn weight by 1.0f
  double w_lo = 0.01, w_hi = 1.0;
  double mu = -2.0, sigma = 0.7;
  int permute_ids = 1;        // seeded random vertex ids instead of level order
  // densification: extra forward edges until mean out-degree reaches
  // dense_mean (0 = off), targets dense_span_hi levels ahead at most.
  double dense_mean = 0.0;
  uint64_t dense_seed = 3;
  int dense_span_hi = 5;
};

void generate(const GenParams& p, EdgeList& el);

// Named workloads.  which: 1 = 50K-vertex / 120-level DAG,
// 2 = 1.5M-vertex / 300-level DAG, 3 = the same densified to mean
// out-degree 20, 4 = the i-th graph of the 256-graph batch,
// 5 = the i-th of the 8 resident served graphs.  unit_wgts forces 1.0f.
GenParams preset(int which, int index, int unit_wgts);
// Per-graph K for the batch (which = 4) and the served stream (which = 5).
int64_t preset_k(int which, int index);

}  // namespace cpg
