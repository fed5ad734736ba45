// Device-side argument blocks and host launch wrappers of the CPG kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cpg {

// ---------------- SFXT ----------------
struct SfxtCtl {
  unsigned qtail;               // level-order queue tail
  int err;                      // barrier spin limit exceeded
  int num_levels;               // graph diameter (levels peeled)
  unsigned levelized;           // vertices peeled
  unsigned max_abs_distf_bits;  // max |distf| as float bits
  int nsv_err;                  // 1: nsv descent found no answer, 2: long list without devs_tmp
  unsigned long long bar_arrive;  // grid barrier: arrivals << 40 | vertices the arrived blocks pushed
  unsigned long long bar_word;    // grid barrier: generation << 32 | queue tail when it completed
  unsigned long long nbig;      // deviation lists one wave sorts (front of big[])
  unsigned long long nbig_l;    // deviation lists one block sorts (back of big[])
  unsigned qhead;               // dataflow loop: queue slots claimed
  int pad;
  // edge-parallel level loop: one counter per level parity, arrivals << 40 |
  // vertices pushed (both running totals, never reset)
  unsigned long long lvl_bar[2];
};

struct SfxtArgs {
  int n;
  int nsinks;
  const int* oadjp;
  const int* oadj;
  const float* owgt;
  const int* iadjp;
  const int2* iedge;  // fanin records {x, w_e bits}, in-CSR order
  const int* sinks;   // out-degree-0 vertices: level 0
  // Relabelled graph (level-order CSR): original id of every vertex, the
  // successor tie-break's order (null: ids are the original ones)
  const int* orig;
  const uint8_t* sink_mask;  // optional endpoint subset (n bytes, nonzero = chosen sink)
  unsigned long long* best;  // packed relax candidate per vertex (see sfxt.hip)
  int* deps;
  int* queue;
  int* lvlp;
  int* level;
  int* dist;
  int* succ;
  float* distf;
  int4* vrec;      // {min delta bits, nsv, succ, distf bits}
  int2* devinfo;   // {first deviation slot, deviation count}
  int4* devs;      // {w, w_e bits, distf[w] bits, delta bits}, m slots
  int4* devs_tmp;  // m slots, only when some fanout exceeds the LDS sort (else null)
  int* big;        // vertices whose deviation lists are sorted block-wide, n slots
  int2* lift;      // lift levels x n: {2^k-th successor, min delta over those 2^k}
  SfxtCtl* ctl;
  unsigned long long* dfw;  // dataflow loop: one progress word per block (dfw_n slots)
  int dfw_n;
  // edge-parallel level loop: only blocks whose index is a multiple of
  // place_stride take part (the others return at once); with a stride of 8
  // (the XCD count) every working block lands on one XCD under round-robin
  // dispatch
  int place_stride;
  int wave_entries;  // edge-parallel loop: queue entries a wave takes at a time (1-64; 0 = 64)
  int64_t m;  // edges (the fence-free loop picks its records per step by m / n)
};

// Q endpoint subsets relaxed in one level loop (served queries): the
// levels, queue and deps of SfxtArgs are shared; per subset q its chosen
// sinks masks[q*nsinks + i] (i = position in the sink list), distances
// dist/succ/distf[q*n..] and max |distf| bits maxd[q]; candidates
// best[v*Q + q].  Q is a power of two <= 64.
struct SfxtMulti {
  int q;
  int pad;
  const uint8_t* masks;       // Q x nsinks bytes, nonzero = chosen sink (null = all)
  int* dv;                    // n x Q settled distances, vertex-major (the pull reads)
  int* dist;                  // Q x n
  int* succ;                  // Q x n
  float* distf;               // Q x n
  unsigned* maxd;             // Q
};

// Number of lift levels for a graph of num_levels levels (2^levels > chain length + 1).
int sfxt_lift_levels(int num_levels);
void launch_sfxt_init_multi(const SfxtArgs& args, const SfxtMulti& m, hipStream_t stream);
// One graph's multi-subset level loop on the blocks [b0, b0 + nb) of a
// launch holding several graphs' loops (partitions in ascending b0).
struct SfxtPart {
  SfxtArgs a;
  SfxtMulti m;
  int b0;
  int nb;
};
constexpr int kMaxParts = 8;  // partitions per launch (passed by value: ~2 KB of arguments)
struct SfxtParts {
  SfxtPart p[kMaxParts];
  int n;
  int pad;
};
// Blocks the partitioned level loop can keep resident at once.
int sfxt_parts_capacity();
// parts.p[0..n) cover the blocks [0, total_blocks) in ascending b0.
void launch_sfxt_levels_parts(const SfxtParts& parts, int total_blocks, hipStream_t stream);
// Per-vertex init, then the cooperative level loop (returns the grid used).
void launch_sfxt_init(const SfxtArgs& args, hipStream_t stream);
int launch_sfxt_levels(const SfxtArgs& args, int blocks, hipStream_t stream);
// The same loop without cache fences (device-scope queue and best[] access);
// the queue must hold the sinks and -1 beyond: launch_sfxt_dataflow_init
// with args.dfw_n = 0 prepares it.
int launch_sfxt_levels_nf(const SfxtArgs& args, int blocks, hipStream_t stream);
// The fence-free loop with the fanin release spread over a wave's records
// (same preparation as launch_sfxt_levels_nf).
int launch_sfxt_levels_edges(const SfxtArgs& args, int blocks, hipStream_t stream);
// The barrier-free dataflow loop (same results as the level loop); args.dfw
// must hold at least sfxt_dataflow_grid(blocks) words.  Returns the grid used.
int sfxt_dataflow_grid(int blocks);
void launch_sfxt_dataflow_init(const SfxtArgs& args, hipStream_t stream);
int launch_sfxt_dataflow(const SfxtArgs& args, int blocks, hipStream_t stream);
// The local-queue level loop (block-local LDS queues, fence-free grid
// barrier; same results as the level loop).  Returns the grid used.
void launch_sfxt_local_init(const SfxtArgs& args, hipStream_t stream);
int launch_sfxt_local(const SfxtArgs& args, int blocks, hipStream_t stream);
// Relabelled graph: succ[] holds original ids after a level loop; map
// them to the relabelled ids (perm: original -> relabelled).
void launch_sfxt_succ_relabel(int* succ, const int* perm, int n, hipStream_t stream);
// Deviation records, their sort, the lift tables and the nsv pointers.
// m: the graph's edge count (chooses thread- or wave-per-vertex records).
void launch_sfxt_index(const SfxtArgs& args, int lift_levels, int64_t m, hipStream_t stream);
constexpr int kSfxtLdsSort = 2048;  // longest deviation list sorted in LDS
// best[] and deps[] hold kSfxtPad slots after their n vertices: the level
// loops aim their no-op atomics there (a slot past a vertex's last fanin
// record), so those stay in a small range that the L2 keeps
constexpr int kSfxtPad = 1 << 16;

// ---------------- PFXT ----------------
constexpr int kHistParts = 256;  // per-block partial histograms per pass

struct PfxtCtl {
  unsigned long long batch_n;      // parents selected this round
  unsigned long long node_n;       // prefix-tree nodes stored (may pass node_cap on overflow)
  unsigned long long child_n;      // count pass: children <= bound
  unsigned long long expanded;     // parents expanded, all rounds
  unsigned long long hops;         // chain vertices visited (stats)
  unsigned long long devvis;       // deviation records examined (stats)
  unsigned long long fill_node_base;  // node_n when the round began
  unsigned long long node_cap;     // node store capacity
  unsigned ukey;                   // bound: K-th smallest nonzero cost key (kKeyInf = none)
  unsigned thkey;                  // batch threshold key
  unsigned klo;                    // histogram window base key
  unsigned shift;                  // histogram window shift
  int done;                        // no parent left to expand
  int overflow;                    // node capacity exceeded in a fill pass
  int rounds;
  int err;                         // a walk exceeded its step bound (corrupt index)
  unsigned long long nonmono;      // children whose cost came out below their parent's
};

// Prefix-tree node store: one slot per generated path, roots first.
struct NodeStore {
  float* cost;
  int* to;        // head of the path's suffix chain
  int* from;      // deviation vertex (-1 for a root)
  int* parent;
  int* level;
  uint32_t* key;  // orderable cost key
  uint8_t* flag;  // 1 once expanded
  uint64_t cap;
};

struct WalkArgs {
  const int4* vrec;
  const int2* devinfo;
  const int4* devs;
  const uint32_t* batch;
  NodeStore nodes;
  uint32_t* ckeys;        // count pass: children's keys (ckeys_cap slots, padded by 16)
  uint64_t ckeys_cap;
  PfxtCtl* ctl;
  float max_abs_distf;
  float max_abs_wgt;
  int collect_stats;
  int strict;             // emit children with cost < U (rounds) or <= U (bound probe)
  int step_limit;         // bound on one walk's steps: (levels + 2) * (max out-degree + 2)
};

void launch_roots(const int* srcs, const uint8_t* keep, int nsrc, const int* dist,
                  const float* distf, NodeStore nodes, PfxtCtl* ctl, hipStream_t s);
void launch_probe_setup(PfxtCtl* ctl, uint32_t ukey, uint32_t shift, unsigned long long nb,
                        unsigned long long nroots, hipStream_t s);
void launch_iota(uint32_t* p, uint64_t n, int grid, hipStream_t s);
void launch_probe_finish(PfxtCtl* ctl, unsigned* hist_all, unsigned long long k,
                         unsigned long long nroots, hipStream_t s);
// Histograms of stored paths (mode 0: all, 1: this round's children, 2: the
// keys of the last count pass) into
// hist_all (2 x kHistBins: counts of nonzero keys <= U, then each bin's
// largest key) and, with_pend, hist_pend (unexpanded paths with key < U).
// part holds 3 x kHistParts x kHistBins per-block partials.
void launch_hist_nodes(const NodeStore& nodes, const uint32_t* ckeys, uint64_t ckeys_cap,
                       PfxtCtl* ctl, int mode, bool with_pend, unsigned* part, unsigned* hist_all,
                       unsigned* hist_pend, bool accumulate, hipStream_t s);
void launch_select_batch(PfxtCtl* ctl, unsigned* hist_pend, unsigned long long batch_target,
                         hipStream_t s);
void launch_make_batch(const NodeStore& nodes, PfxtCtl* ctl, uint32_t* batch, int grid,
                       hipStream_t s);
void launch_walk(const WalkArgs& a, bool fill, int grid, hipStream_t s);
void launch_select_bound(PfxtCtl* ctl, unsigned* hist_all, unsigned long long k,
                         hipStream_t s);
// rlog[rounds] = paths stored once `rounds` rounds have completed (rounds < cap).
void launch_round_log(const PfxtCtl* ctl, unsigned long long* rlog, int cap, hipStream_t s);

// ---------------- multi-query PFXT (batched designs) ----------------
constexpr int kMqBins = 1024;                 // per-query histogram bins
constexpr int kMqHistStride = 2 * kMqBins;    // counts, then unexpanded counts

struct MqQuery {
  uint32_t ukey;                  // bound (kKeyInf = none)
  uint32_t thkey;                 // batch threshold
  uint32_t klo;                   // histogram window base key
  uint32_t shift;                 // histogram window shift
  unsigned long long k;           // paths asked for
  unsigned long long batch_target;
  int done;                       // nothing left to expand (or the slot is free)
  int rounds;
  // the query's deviation index: the batched designs share the union's,
  // a served query has its own (its endpoint subset's SFXT)
  const int4* vrec;
  const int2* devinfo;
  const int4* devs;
  float mag;                      // 2 max|distf| + max|w| + 1 (walk slack)
  int step_limit;                 // bound on one walk's steps
  int all;                        // the next round expands every unexpanded path
  int pad;
  unsigned long long stored;      // children the fill passes stored for this query
};

struct MqCtl {
  unsigned long long node_n;
  unsigned long long batch_n;
  int overflow;
  int err;
  int active;                     // queries with unexpanded paths this round
  int pad;
  unsigned long long nonmono;     // children whose cost came out below their parent's
};

struct MqArgs {
  NodeStore nodes;
  int* nq;                        // query of each node
  const float* distf;             // roots' costs (mq_roots)
  uint32_t* batch;
  unsigned* hist;                 // nq_count x kMqHistStride
  MqQuery* qs;
  MqCtl* ctl;
  int nq_count;
  int pad;
};

void launch_mq_roots(const MqArgs& a, const int* root_v, const int* root_q, int nroots, hipStream_t s);
// One round for every query (all: the batch takes every unexpanded path).
void launch_mq_round(const MqArgs& a, bool all, int grid, hipStream_t s);
// sel: one byte per query slot choosing the queries whose paths are taken
// (null = every query).
void launch_mq_build_keys(const MqArgs& a, const uint8_t* sel, uint64_t nmax, uint64_t* key,
                          uint32_t* val, unsigned long long* count, unsigned long long* per_q,
                          int grid, hipStream_t s);
void launch_mq_gather_costs(const uint32_t* val, const float* cost, uint64_t n, float* out,
                            hipStream_t s);
void launch_zero_u32(unsigned* p, uint64_t n, hipStream_t s);
// Compaction of the shared store: live nodes (query slot not done) per
// tile, then (after an exclusive scan of the tile counts) the ordered copy
// into dst with newidx[old] = new slot (-1 for a dead node) and the
// parents remapped.
uint64_t mq_compact_tiles(uint64_t n);
// Appends src[0..cnt) at dst[base..) as query slot q and sets node_n to
// base + cnt (no fill pass may run concurrently).
void launch_mq_seed(const NodeStore& src, uint64_t cnt, const NodeStore& dst, int* dst_nq,
                    uint64_t base, int q, MqCtl* ctl, int grid, hipStream_t s);
void launch_mq_count_live(const MqArgs& a, uint64_t n, unsigned* tile_cnt, hipStream_t s);
void launch_mq_compact(const MqArgs& a, uint64_t n, const NodeStore& dst, int* dst_nq,
                       const unsigned* tile_off, int* newidx, uint64_t n_live, int grid,
                       hipStream_t s);
void launch_exclusive_scan_u32(unsigned* p, uint64_t len, hipStream_t s);
void launch_zero_active(MqCtl* ctl, hipStream_t s);

// ---------------- BVSS deviation discovery ----------------
struct BvssCtl {
  unsigned long long count;   // triples found (written while < cap)
  unsigned long long tiles;   // MFMA tiles issued
  int mismatch;               // an MFMA count disagreed with the decoded words
  int pad;
};

struct BvssArgs {
  int n;                      // vertices
  int n_int;                  // 64-source intervals
  int n_entries;              // (interval, destination row) entries
  int n_tiles;                // 16-entry tiles (per interval, rounded up)
  int nfront;                 // frontiers in this call
  const int* int_ptr;         // n_int + 1: first entry of each interval
  const int* rows;            // destination row of each entry
  const int* entry_int;       // interval of each entry
  const unsigned long long* masks;  // 64 source bits of each entry
  const int2* tiles;          // {interval, first entry}
  unsigned long long* front;  // nfront x n_int frontier words (zeroed by the caller)
  uint8_t* dense;             // n_int: interval takes the MFMA path
  int4* out;                  // {frontier, u, v, 0} triples
  uint64_t cap;
  BvssCtl* ctl;
};

// Frontier words, interval classification (mode 0 auto, 1 bitmap, 2 MFMA)
// and both discovery paths.
void launch_bvss_discover(const BvssArgs& a, const int* fptr, const int* fsrc, int mode,
                          int dense_min, int grid, hipStream_t s);

// ---------------- final selection ----------------
// Stable LSD radix sort of (key, value) pairs over the given key bytes.
struct SortBufs {
  uint64_t* key[2];
  uint32_t* val[2];
  unsigned* ghist;        // sort_ghist_len(n) counts: 256 digits x tiles, then 256 digit totals
  uint64_t cap;
};
int radix_sort_pairs(SortBufs& b, uint64_t n, const int* bytes, int nbytes, hipStream_t s);
uint64_t sort_tiles(uint64_t n);
uint64_t sort_ghist_len(uint64_t n);

// Sort input: stored paths with a nonzero key <= U (nmax = stored count).
void launch_build_sort_keys(const NodeStore& nodes, const PfxtCtl* ctl, uint64_t nmax,
                            uint64_t* key, uint32_t* val, unsigned long long* count, hipStream_t s);
void launch_emit_report(const uint32_t* val, uint64_t nout, NodeStore nodes, float* ocost,
                        int* olevel, int* ofrom, int* oto, int* opidx, int* rank_of,
                        hipStream_t s);
void launch_link_parents(const int* opidx, const int* rank_of, uint64_t nout, int* oparent,
                         hipStream_t s);
void launch_fill_int(int* p, int v, uint64_t n, hipStream_t s);

// ---------------- CSR relabelling (reorder.hip) ----------------
// E_ORIENTED (gpucpg/gpucpg.cu:2048-2079): one thread per edge copies edge
// e of u to slot new_adjp[perm[u]] + (e - adjp[u]) of the relabelled CSR,
// its head relabelled; each vertex keeps its edge order.  owner[e] = u is
// filled first by launch_edge_owner.
void launch_edge_owner(const int* adjp, int n, int* owner, hipStream_t stream);
void launch_reorder_fanout(const int* adjp, const int* adj, const float* w, const int* owner,
                           const int* perm, const int* new_adjp, int64_t m, int* new_adj,
                           float* new_w, hipStream_t stream);
void launch_reorder_fanin(const int* iadjp, const int2* iedge, const int* owner, const int* perm,
                          const int* new_iadjp, int64_t m, int2* new_iedge, hipStream_t stream);
// out[i] = in[idx[i]] for bytes (endpoint masks into relabelled order)
void launch_gather_u8(const uint8_t* in, const int* idx, int n, uint8_t* out, hipStream_t stream);

// ---------------- SFXT without allocation (noalloc.hip) ----------------
// Levelize + relax + successor of a device-resident graph with every
// intermediate buffer in a static device array; synchronous on `stream`.
// Returns the number of levels; throws on a bad argument, a graph larger
// than sfxt_noalloc_max_verts() or an incomplete loop.
int sfxt_noalloc_max_verts();
int sfxt_noalloc(int n, const int* oadjp, const int* iadjp, const int2* iedge, const int* sinks,
                 int nsinks, int* dist, int* succ, int* level, hipStream_t stream);

}  // namespace cpg
