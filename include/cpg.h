/*
 * cpg.h -- C ABI of the MI355X critical-path-generation (CPG) engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (0 = ok, negative = error, message in cpg_last_error()).  No exception
 * crosses this boundary.  One handle per thread; calls on one handle are
 * serialised by the caller (the reference CpGen is not thread-safe either).
 *
 * Each entry point names the gpucpg::CpGen member it replaces
 * (gpucpg/gpucpg.cuh, gpucpg/gpucpg.cu, paths relative to the reference
 * repository root).  The C++ facade include/gpucpg_amd.hpp wraps this ABI
 * in a class with the reference's method names.
 */
#ifndef CPG_AMD_H
#define CPG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cpg_handle cpg_handle;

/* PfxtNode, gpucpg/gpucpg.cuh:74-107 (24 bytes, same field order).
 * parent is the 0-based rank of the parent path in the same report (the
 * reference's parent index goes stale after its final sort,
 * gpucpg/gpucpg.cu:3285 vs :21326); -1 for a root, or when the parent is
 * not among the reported paths.  num_children is 0 as in the GPG mode's
 * expand kernels (gpucpg/gpucpg.cu:3284-3288). */
typedef struct {
  int level;          /* number of deviation edges from the root path */
  int from;           /* deviation vertex u (-1 for a root)           */
  int to;             /* head of the path's suffix chain (source for roots) */
  int parent;
  int num_children;
  float cost;         /* the reference's "slack"                      */
} cpg_node;

/* Query options.  Defaults (cpg_query_opts_default) are the GPG golden
 * configuration of examples/tc-pfxt-inprocess-common.cuh:112-143. */
typedef struct {
  int struct_size;          /* sizeof(cpg_query_opts)                         */
  float init_split_perc;    /* 0.005f: initial split percentile (gpucpg.cu:19140) */
  int root_tie_quirk;       /* 1: drop roots tied at the initial split beyond
                               (int)(p*n+1) as gpucpg.cu:19141-19180 does     */
  int64_t batch;            /* parents expanded per round (0 = automatic)     */
  int64_t max_nodes;        /* prefix-tree node capacity (0 = automatic)      */
  int64_t mem_budget_bytes; /* device bytes a query may hold (0 = no limit): the
                               prefix-tree node store and the final sort get
                               what the graph and its SFXT leave; a query whose
                               store would have to grow past it fails          */
  const uint8_t* endpoint_mask; /* optional n-byte sink subset (NULL = all sinks) */
  int collect_stats;        /* 1: count hops / deviations / emitted children  */
  int kernel_timing;        /* 1: time the SFXT and walk kernels with HIP events
                               on the engine's stream (cpg_timing.*_kernel_ms) */
  int sfxt_blocks;          /* SFXT level-loop grid in blocks (0 = automatic) */
  int bound_probe;          /* first bound from pruned count passes: 0 = when the
                               roots' chains are long (automatic), 1 = always,
                               -1 = never (the first round walks unbounded)   */
  int sfxt_method;          /* SFXT loop: 0 automatic (5), 1 level-synchronous
                               loop with a global
                               queue and a fenced grid barrier, 2 barrier-free
                               dataflow loop, 3 level-synchronous loop with
                               block-local queues, 4 loop 1 with device-scope
                               queue access and a fence-free barrier, 5 loop 4
                               with each wave's fanin records spread over its
                               lanes and a counter barrier; all give
                               identical results                              */
  int step_times;           /* 1: record every expansion round's device time and
                               the paths it stored (cpg_step_times)           */
  int monotone_check;       /* 1: fail the query (error) when a child's binary32
                               cost came out below its parent's; 0: only count
                               them (cpg_timing.nonmonotone)                  */
  int pad;
} cpg_query_opts;

/* Timings of the last query, milliseconds (gpucpg/gpucpg.cuh:327-356). */
typedef struct {
  double sfxt_ms;     /* levelize + relax + successor + deviation index (0 when cached) */
  double pfxt_ms;     /* prefix-tree expansion rounds (reference-comparable pfxt_ms)     */
  double select_ms;   /* final top-K sort and report assembly                            */
  double query_ms;    /* whole device query, results resident in HBM                     */
  int64_t rounds;     /* expansion rounds (the reference's "steps")                      */
  int64_t expanded;   /* prefix-tree nodes expanded                                      */
  int64_t generated;  /* prefix-tree nodes emitted (roots included)                      */
  int64_t hops;       /* suffix-chain vertices visited (collect_stats)                   */
  int64_t dev_visits; /* deviation records examined (collect_stats)                      */
  int64_t roots;      /* roots seeded                                                    */
  int64_t roots_dropped; /* roots lost to the initial-split tie truncation               */
  int64_t retries;    /* query restarts after a node-capacity overflow                   */
  double sfxt_kernel_ms;  /* kernel_timing: SFXT kernel, HIP events                      */
  double walk_count_ms;   /* kernel_timing: count-pass walk kernels, summed              */
  double walk_fill_ms;    /* kernel_timing: fill-pass walk kernels, summed               */
  int64_t walk_launches;  /* kernel_timing: walk kernel launches, both passes            */
  double sfxt_levels_ms;  /* kernel_timing: SFXT level-loop kernel (part of sfxt_kernel_ms)  */
  int64_t sfxt_blocks;    /* SFXT level-loop grid used                                    */
  int64_t probes;         /* bound-probe fill passes run before the first round            */
  double bvss_ms;         /* last cpg_bvss_discover: device time (frontier build + both paths) */
  int64_t bvss_tiles;     /* last cpg_bvss_discover: 16x16x64 int8 MFMA tiles issued        */
  int64_t device_bytes;   /* device memory the handle holds after the query (graph, SFXT,
                             node store, final sort)                                          */
  int64_t node_slots;     /* prefix-tree node store capacity (61 device bytes per slot with
                             the final sort's buffers)                                        */
  int64_t nonmonotone;    /* children whose binary32 cost came out below their parent's: the
                             top K is the exact K cheapest whenever none of them sits at the
                             bound; the reference's own best-first assumes there are none */
} cpg_timing;

/* ---- lifetime ---- */
cpg_handle* cpg_create(void);                      /* CpGen::CpGen   */
void cpg_destroy(cpg_handle* h);                   /* CpGen::~CpGen  */
const char* cpg_last_error(void);                  /* thread-local message */

/* ---- graph input ---- */
/* CpGen::read_input, gpucpg/gpucpg.cu:453-548. */
int cpg_load_text(cpg_handle* h, const char* path, int ignore_wgts);
/* Binary fanout CSR (gpucpg/graph.h:24-55: int32 N, int32 M, offsets[N+1],
 * adjncy[M], float wgts[M], native-endian), the format
 * CpGen::write_to_csr_bin (gpucpg/gpucpg.cuh:310-323) writes and
 * Graph::read_from_bin reads; the fanin CSR is derived as for text input. */
int cpg_load_bin(cpg_handle* h, const char* path);
/* Writes the loaded graph's fanout CSR in that format. */
int cpg_write_bin(cpg_handle* h, const char* path);
/* Edge list in input order: same CSR semantics as the text loader. */
int cpg_load_edges(cpg_handle* h, int n, int64_t m, const int* eu,
                   const int* ev, const float* ew);
/* Fanout CSR (adjp[n+1], adjncy[m], wgts[m]); the fanin CSR is derived in
 * the same order as gpucpg/gpucpg.cu:531-546 would build it. */
int cpg_load_csr(cpg_handle* h, int n, int64_t m, const int* adjp,
                 const int* adjncy, const float* wgts);
/* Copies of the host CSR the handle holds (NULL pointers are skipped). */
int cpg_get_csr(cpg_handle* h, int* adjp, int* adjncy, float* wgts);

int64_t cpg_num_verts(cpg_handle* h);              /* CpGen::num_verts, gpucpg.cuh:208 */
int64_t cpg_num_edges(cpg_handle* h);              /* CpGen::num_edges, gpucpg.cuh:209 */
int cpg_graph_diameter(cpg_handle* h);             /* CpGen::graph_diameter, gpucpg.cuh:348 */

/* ---- query ---- */
int cpg_reset(cpg_handle* h);                      /* CpGen::reset, gpucpg.cuh:211-250 */
void cpg_query_opts_default(cpg_query_opts* o);
/* CpGen::report_paths(k, 10, true, LEVELIZE_THEN_RELAX, SHORT_LONG, ...),
 * gpucpg/gpucpg.cu:16910-21366: the k lowest-cost source->sink paths, exact,
 * ascending cost, zero-cost paths excluded (gpucpg.cu:21338-21358). */
int cpg_query(cpg_handle* h, int64_t k, const cpg_query_opts* o);
int64_t cpg_num_paths(cpg_handle* h);              /* paths held by the last query */
/* CpGen::get_slacks, gpucpg/gpucpg.cu:21409-21420: first k costs. */
int64_t cpg_get_costs(cpg_handle* h, int64_t k, float* out);
/* CpGen::get_pfxt_nodes, gpucpg/gpucpg.cu:21422-21440. */
int64_t cpg_get_nodes(cpg_handle* h, int64_t k, cpg_node* out);
/* Vertex sequence of the i-th reported path (source ... sink); returns the
 * length, or a negative status (cap too small: -(length) - 1000).  No
 * counterpart in the reference (its parents are stale after the sort). */
int64_t cpg_get_path_vertices(cpg_handle* h, int64_t i, int* out, int64_t cap);
int cpg_timings(cpg_handle* h, cpg_timing* out);
/* Per-round record of the last query run with step_times: ms[i] the device
 * time of expansion round i (HIP events on the engine's stream), paths[i]
 * the prefix-tree paths it stored (NULL arrays are skipped).  Returns the
 * number of rounds.  CpGen::short_long_step_times and paths_gen_per_step,
 * gpucpg/gpucpg.cuh:351-352 (filled at gpucpg/gpucpg.cu:20404-20405). */
int64_t cpg_step_times(cpg_handle* h, double* ms, int64_t* paths, int64_t cap);

/* ---- SFXT results (dist/succ/level of the last SFXT) ---- */
/* dist: int-scaled suffix distance (relax_bu_step, gpucpg.cu:1167-1187),
 * succ: update_successors_bu (gpucpg.cu:2606-2628), level: peel depth
 * (bfs_td_step_privatized, gpucpg.cu:2215-2283).  Runs the SFXT if needed. */
/* CSR reorder of the resident graph (CsrReorderMethod, gpucpg.cuh:41-72):
 * 0 NONE, 1 E_ORIENTED -- vertices relabelled in level order (sinks first)
 * on the device by the edge-parallel scatter of reorder_csr_e_oriented
 * (gpucpg/gpucpg.cu:2048-2079, run at :17753-17765).  A layout change only:
 * ties still break on original ids and every id the engine returns is an
 * original id, so results equal those with 0.  Applied at the next upload
 * (one SFXT for the levels, then the scatter); cpg_reorder_ms reports its
 * time.  The BVSS entry points need 0. */
int cpg_set_csr_reorder(cpg_handle* h, int method);
double cpg_reorder_ms(cpg_handle* h);
int cpg_sfxt(cpg_handle* h);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops or harnesses that forbid hipMalloc): the SFXT level
 * loop (levelize, int-scaled relax, successor; the same dist/succ/level as
 * cpg_sfxt + cpg_get_sfxt) of a graph already on the device, with every
 * intermediate buffer in one static device array (about 25 MiB per device,
 * held until the process ends), so the call allocates nothing.  Inputs are
 * device pointers: the fanout offsets (n+1), the fanin offsets (n+1), the
 * fanin records (2 ints per edge: fanin vertex, weight as float bits) and
 * the sink list (out-degree-0 vertices, ascending); outputs are n ints each.
 * At most cpg_sfxt_noalloc_max_verts() vertices (1,048,576).  Synchronous on
 * `stream` (a hipStream_t, null = the null stream); calls that overlap on one
 * device race on the shared array.  Returns the number of levels. */
int cpg_sfxt_noalloc(int n, const int* d_oadjp, const int* d_iadjp, const int* d_iedge,
                     const int* d_sinks, int nsinks, int* d_dist, int* d_succ, int* d_level,
                     void* stream);
int cpg_sfxt_noalloc_max_verts(void);
int cpg_get_sfxt(cpg_handle* h, int* dist, int* succ, int* level);

/* ---- BVSS deviation discovery, gpucpg/tc_pfxt_bvss.cuh ---- */
/* Builds the A_dev index from the fanout CSR and the SFXT successors
 * (build_adev_bvss_from_fanout_csr, tc_pfxt_bvss.cuh:117-176): for every
 * 64-source interval, the destination rows v with a deviation edge u->v
 * (v != succ[u]) from it and a 64-bit mask of those sources.  Runs the
 * SFXT if needed; kept until the graph changes. */
int cpg_bvss_build(cpg_handle* h);
/* Index shape: intervals, (interval, row) entries, set source bits. */
int cpg_bvss_info(cpg_handle* h, int64_t* n_intervals, int64_t* n_entries, int64_t* set_bits);
/* Sources u with a deviation edge u->row, ascending (decode_row_neighbors,
 * tc_pfxt_bvss.cuh:228-251); same return convention as
 * cpg_get_path_vertices. */
int64_t cpg_bvss_row_sources(cpg_handle* h, int row, int* out, int64_t cap);
/* The deviation edges leaving each of nfront source frontiers
 * (discover_pairs_for_sources, tc_pfxt_bvss.cuh:289-439, :572-640):
 * frontier f is fsrc[fptr[f] .. fptr[f+1]).  Writes (f, u, v) triples
 * sorted by (f, u, v) and returns their number (-(number) - 1000 when cap
 * is too small).  mode 0: each interval takes the int8 MFMA path when its
 * frontier block holds at least dense_min source bits per frontier on
 * average, else the wave64 bitmap path; 1: bitmap only; 2: MFMA only. */
int64_t cpg_bvss_discover(cpg_handle* h, int nfront, const int* fptr, const int* fsrc, int mode,
                          int dense_min, int* out_f, int* out_u, int* out_v, int64_t cap);

/* ---- batched designs ----
 * Many independent designs, one K-path query each, solved together: the
 * designs' disjoint union gets one SFXT and one multi-query PFXT whose
 * rounds run for every query at once (after the first round they replay
 * a captured HIP graph).  Each query's result equals cpg_query's on its
 * design alone (same options). */
typedef struct cpg_batch cpg_batch;
cpg_batch* cpg_batch_create(void);
void cpg_batch_destroy(cpg_batch* b);
/* Adds a design (edge list, same semantics as cpg_load_edges) and its K;
 * returns the query index. */
int cpg_batch_add_edges(cpg_batch* b, int n, int64_t m, const int* eu, const int* ev,
                        const float* ew, int64_t k);
int cpg_batch_num_queries(cpg_batch* b);
int cpg_batch_run(cpg_batch* b, const cpg_query_opts* o);
/* Round shapes of a batch, chosen per workload by cpg_batch_autotune (no
 * counterpart in the reference, which runs one query at a time).  None of
 * them changes a result. */
typedef struct cpg_batch_tuning {
  int group;          /* replayed rounds per host synchronisation (default 4) */
  int grid_mult;      /* round kernels' grid = CUs x grid_mult blocks (8)     */
  int target_div[3];  /* per K class (K < 1000, < 10000, >= 10000): a query
                         expands about K / div parents per round (8, 8, 8)    */
  int trials;         /* batch runs the last autotune made                    */
  double start_ms;    /* last autotune: batch ms with the shapes it began from */
  double best_ms;     /* last autotune: batch ms with the shapes it kept       */
} cpg_batch_tuning;
/* Coordinate descent over the shapes: whole batch runs (at most max_trials),
 * a shape kept when it is >= 2% faster; leaves the chosen shapes set and
 * the batch's results computed with them.  Returns the trials made. */
int cpg_batch_autotune(cpg_batch* b, const cpg_query_opts* o, int max_trials);
int cpg_batch_get_tuning(cpg_batch* b, cpg_batch_tuning* out);
int cpg_batch_set_tuning(cpg_batch* b, const cpg_batch_tuning* t);
/* First k costs of query q (ascending, exact), as cpg_get_costs. */
int64_t cpg_batch_get_costs(cpg_batch* b, int q, int64_t k, float* out);
/* Whole-batch timings (sfxt/pfxt/select/query ms, rounds, generated). */
int cpg_batch_timings(cpg_batch* b, cpg_timing* out);

/* ---- served queries ----
 * No counterpart in the reference (its CpGen answers one query at a time).
 * Resident graphs and a stream of top-K queries, each with an optional
 * endpoint subset; queries are admitted between rounds while the device
 * bytes of the in-flight queries' deviation indexes stay within the budget
 * and a slot is free, and the PFXT rounds of every in-flight query run
 * together.  Each result equals cpg_query's for the same graph, K and
 * endpoint mask (default options).  Not thread-safe: one thread drives a
 * server. */
typedef struct cpg_server cpg_server;
typedef struct {
  int64_t submitted, admitted, completed, rounds, compactions, restarts;
  int64_t peak_nodes;        /* nodes stored at once (shared store)          */
  int64_t node_cap;          /* current store capacity                       */
  int64_t peak_index_bytes;  /* in-flight deviation indexes, device bytes    */
  int64_t in_flight, queued;
  int64_t sfxt_groups;       /* level loops run (each for a group of queries)  */
  int64_t nonmonotone;       /* children whose cost came out below their parent's */
  int64_t root_bounds;       /* queries whose first bound came from their roots */
  double sfxt_ms;            /* host time starting and checking groups' SFXTs  */
  double group_wait_ms;      /* host time waiting for a group, none in flight  */
  double probe_ms, round_ms, extract_ms, compact_ms;  /* host wall time      */
} cpg_server_stats;
/* budget_bytes: device bytes for in-flight indexes (0 = 64 GiB);
 * max_in_flight: query slots, at most 256 (0 = 64); initial_nodes: first
 * capacity of the shared node store (0 = 2^24; it grows as needed). */
cpg_server* cpg_server_create(int64_t budget_bytes, int max_in_flight, int64_t initial_nodes);
void cpg_server_destroy(cpg_server* s);
/* Makes a graph resident (edge list as cpg_load_edges); returns its id. */
int cpg_server_add_graph(cpg_server* s, int n, int64_t m, const int* eu, const int* ev,
                         const float* ew);
/* Queues a query; endpoint_mask: NULL or n bytes (nonzero = chosen sink),
 * copied.  Returns the query id (> 0). */
int64_t cpg_server_submit(cpg_server* s, int graph, int64_t k, const uint8_t* endpoint_mask);
/* Admission, one round of every in-flight query, extraction of the
 * finished ones; returns the number of queries still queued or in flight. */
int64_t cpg_server_step(cpg_server* s);
/* Ids of the queries finished since the last poll (up to cap). */
int64_t cpg_server_poll(cpg_server* s, int64_t* ids, int64_t cap);
/* First k costs of a finished query (ascending, exact); -1 if not done. */
int64_t cpg_server_get_costs(cpg_server* s, int64_t id, int64_t k, float* out);
/* Submit/admit/done times of a query in ms on the engine's steady clock. */
int cpg_server_query_times(cpg_server* s, int64_t id, double* submit_ms, double* admit_ms,
                           double* done_ms);
/* Forgets a finished query. */
int cpg_server_release(cpg_server* s, int64_t id);
int cpg_server_get_stats(cpg_server* s, cpg_server_stats* out);
/* The engine's steady clock in ms (for arrival schedules). */
double cpg_now_ms(void);
/* Pinned (page-locked) host memory for result buffers: cpg_get_costs and
 * the like copy into it straight from the device, without the staging a
 * pageable buffer needs.  No counterpart in the reference (its get_slacks
 * fills a std::vector).  cpg_host_alloc returns null on failure (see
 * cpg_last_error); cpg_host_free takes what it returned. */
void* cpg_host_alloc(int64_t bytes);
int cpg_host_free(void* p);

/* ---- static cache, gpucpg/gpucpg.cuh:186-189 ---- */
int cpg_cache_enable(cpg_handle* h, int enabled);
int cpg_cache_clear(cpg_handle* h);
int cpg_cache_hits(cpg_handle* h);
int cpg_cache_misses(cpg_handle* h);

/* ---- synthetic workloads (host only; no device call) ---- */
typedef struct {
  int n;
  int64_t m;
  int* u;
  int* v;
  float* w;
} cpg_edge_list;
/* which: 1 = 50K-vertex/120-level DAG, 2 = 1.5M-vertex/300-level DAG,
 * 3 = that DAG densified to mean out-degree 20, 4 = graph `index` of the
 * 256-graph batch, 5 = graph `index` of the 8 served graphs. */
int cpg_gen_preset(int which, int index, int unit_wgts, cpg_edge_list* out);
void cpg_edge_list_free(cpg_edge_list* el);
int64_t cpg_preset_k(int which, int index);
/* Writes the benchmark text format read by cpg_load_text. */
int cpg_write_text(const cpg_edge_list* el, const char* path);

#ifdef __cplusplus
}
#endif
#endif
