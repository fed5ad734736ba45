/*
 * oracle.h -- CPU restatement of gpu-cpg's critical-path-generation (CPG)
 * query, used ONLY as test infrastructure.
 *
 * This library is the checker.  It is loaded by the test-suite, by the
 * smoke check and by the CPU-baseline leg of the benchmark, and never by
 * the product library (gpu-cpg_amd/).  Every function names the reference
 * file:line it restates (paths relative to the gpu-cpg repository root).
 *
 * Pinning (what each function is checked against, in tests/):
 *   orc_parse_text      benchmarks/{simple,dac21_paper_example,tv80_unit_wgts}.txt
 *                       and the inline 5-vertex fixture of
 *                       unittests/tc_pfxt_inprocess.cu:56-67 (hand-derived CSR).
 *   orc_levelize        closed form: level(v) = longest hop count to a sink
 *                       (memoised DFS in the tests) and the known answer
 *                       {3,3,2,1,2,1,0,0} of unittests/tc_sfxt_bvss.cu:153-172.
 *   orc_relax           independent Dijkstra over int-scaled weights.
 *   orc_relax_masked    the same Dijkstra started from the chosen sinks only;
 *                       a mask of every sink equals orc_relax.
 *   orc_successors      brute-force max-id arg-min per vertex.
 *   orc_distf           IEEE float division (numpy float32 divide).
 *   orc_candidate_cost  known answer -0.75 of unittests/tc_pfxt_candidates.cu:85-90.
 *   orc_classify        known answers of unittests/tc_pfxt_candidates.cu:69-83.
 *   orc_deviations      known answer of unittests/tc_pfxt_candidates.cu:386-408.
 *   orc_topk            brute-force enumeration of every source->sink path
 *                       (orc_all_paths) on tiny DAGs; the faithful SEQUENTIAL
 *                       and SHORT_LONG transliterations below on small DAGs.
 *   orc_topk_paths      costs: orc_topk; vertex sequences: per cost class
 *                       the same multiset as orc_all_paths_vertices (tiny
 *                       DAGs and the reference benchmarks' small cases);
 *                       each sequence a fanout-CSR path whose recurrence
 *                       cost equals its reported cost.
 *   orc_all_paths_vertices  every sequence is a source->sink walk of the CSR,
 *                       their number equals a path count by DP over the DAG.
 *   orc_sequential      orc_all_paths (tiny DAGs).
 *   orc_short_long      orc_all_paths (tiny DAGs), orc_topk (small DAGs).
 *   orc_compare_prefix  unittests/tc_pfxt_inprocess.cu:26-51 (known answers).
 */
#ifndef CPG_ORACLE_H
#define CPG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SCALE_UP 10000   /* gpucpg/gpucpg.cuh:35 */
#define ORC_INT_MAX 2147483647

/* Fanout + fanin CSR in the reference's edge order
 * (gpucpg/gpucpg.cu:507-546): per vertex, edges in input order. */
typedef struct {
  int n;
  long long m;
  int* oadjp;   /* n+1 */
  int* oadj;    /* m   */
  float* owgt;  /* m   */
  int* iadjp;   /* n+1 */
  int* iadj;    /* m   */
  float* iwgt;  /* m   */
} orc_graph;

/* PfxtNode, gpucpg/gpucpg.cuh:74-107 (24 bytes). */
typedef struct {
  int level;
  int from;
  int to;
  int parent;
  int num_children;
  float slack;
} orc_node;

typedef struct {
  long long pops;            /* paths popped (expanded)                   */
  long long pushes;          /* children pushed                           */
  long long hops;            /* suffix-chain hops walked (H_K)            */
  long long edges;           /* fanout edges visited (E_K)                */
  long long children;        /* children emitted before pruning (C_K)     */
  long long nonmonotone;     /* children whose cost < parent cost         */
  long long zero_cost;       /* popped paths with cost == 0.0f (erased)   */
  long long roots;           /* roots seeded                              */
  long long roots_dropped;   /* roots dropped by the initial-split quirk  */
  long long long_dummies;    /* default slack-0 nodes left in the long pile */
  long long steps;           /* SHORT_LONG steps                          */
  long long short_pile;      /* SHORT_LONG final short pile size          */
  long long final_window;    /* SHORT_LONG prefetched final window size (0 = unused) */
  int split_index_mismatch;  /* (int)(p*n+1) != (size_t)(p*n)+1           */
  int aborted;               /* 0 ok; 1 node budget; 2 final-window index past the
                                long pile; 3 initial split out of range; 4 split
                                update made no progress (reference UB or hang) */
} orc_stats;

/* ---- input (a1) ---- */
int orc_parse_text(const char* path, int ignore_wgts, orc_graph* g);
int orc_build_csr(int n, long long m, const int* eu, const int* ev,
                  const float* ew, orc_graph* g);
void orc_graph_free(orc_graph* g);
const char* orc_last_error(void);

/* ---- SFXT (a3-a6) ---- */
int orc_levelize(const orc_graph* g, int* level, int* queue, int* lvlp,
                 int* num_levels);
void orc_relax(const orc_graph* g, const int* queue, const int* lvlp,
               int num_levels, int* dist);
void orc_successors(const orc_graph* g, const int* dist, int* succ);
void orc_distf(int n, const int* dist, float* distf);
int orc_sfxt(const orc_graph* g, int* dist, int* succ, int* level,
             int* num_levels);
/* endpoint subset: sink_mask n bytes, nonzero = chosen sink (NULL = all) */
void orc_relax_masked(const orc_graph* g, const int* queue, const int* lvlp,
                      int num_levels, const unsigned char* sink_mask, int* dist);
int orc_sfxt_masked(const orc_graph* g, const unsigned char* sink_mask, int* dist,
                    int* succ, int* level, int* num_levels);

/* ---- candidate cost (a10/a17) ---- */
float orc_candidate_cost(float parent, int src_dist, int dst_dist, float wgt);
float orc_split_inc_amount(float avg_deg);
/* classify_candidate, gpucpg/tc_pfxt_candidates.cuh:905-918:
 * 0 = short (<= split), 1 = long, 2 = skip. */
int orc_classify(float slack, float split, float final_split, int use_final_split,
                 int skip_long_paths);
/* Deviation edges of every vertex (u->w with w != succ[u] and dist[w]
 * finite, gpucpg/gpucpg.cu:2662-2668, tc_pfxt_candidates.cuh:264-276) in
 * fanout order, with delta = distf[w] + w_e - distf[u] evaluated in that
 * order in binary32 (tc_pfxt_candidates.cuh:662-669).  offsets has n+1
 * entries; returns the number of deviations written. */
long long orc_deviations(int n, const int* adjp, const int* adj, const float* wgt,
                         const int* succ, const int* dist, int* offsets, int* dsts,
                         float* deltas);

/* ---- PFXT (a7-a15) ---- */
long long orc_topk(const orc_graph* g, const int* dist, const int* succ,
                   long long k, int root_tie_quirk, float* out_costs,
                   orc_stats* st);
long long orc_sequential(const orc_graph* g, const int* dist,
                         const int* succ, long long k, float* out_costs,
                         orc_stats* st);
long long orc_short_long(const orc_graph* g, const int* dist,
                         const int* succ, long long k, float init_split_perc,
                         int use_fixed_inc, float fixed_inc,
                         double long_pile_limit_bytes,
                         long long node_budget, float* out_costs,
                         orc_stats* st);
long long orc_all_paths(const orc_graph* g, const int* dist,
                        const int* succ, long long cap, float* out_costs);
/* orc_topk that also reports each path's vertices (source ... sink),
 * out_vptr[0..got] CSR offsets into out_verts (vcap entries). */
long long orc_topk_paths(const orc_graph* g, const int* dist, const int* succ, long long k,
                         int root_tie_quirk, float* out_costs, long long* out_vptr,
                         int* out_verts, long long vcap, orc_stats* st);
/* Brute force with vertices: every path (DFS order, zero costs kept). */
long long orc_all_paths_vertices(const orc_graph* g, const int* dist, const int* succ,
                                 long long cap, float* out_costs, long long* out_vptr,
                                 int* out_verts, long long vcap);

/* ---- compare (gate semantics) ---- */
typedef struct {
  long long baseline_count;
  long long tc_count;
  long long compared;
  float max_diff;
  int max_diff_rank;
  int first_mismatch_rank;
  int pass;
} orc_compare;
void orc_compare_prefix(const float* baseline, long long nb, const float* tc,
                        long long nt, int k, float tol, orc_compare* out);
long long orc_first_bit_mismatch(const float* a, const float* b, long long n);

/* ---- timing helper for the CPU baseline ---- */
double orc_query_ms(const orc_graph* g, long long k, float* out_costs,
                    long long* out_count, double* sfxt_ms, double* pfxt_ms);

#ifdef __cplusplus
}
#endif
#endif
