/*
 * oracle.c -- CPU restatement of gpu-cpg's GPG critical-path query.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * check and the benchmark's cpu_baseline leg load it; the product library
 * never does.  See oracle.h for what pins each function.
 *
 * Floating point: built with -O2 -ffp-contract=off and without -ffast-math,
 * so every float operation below is a single IEEE-754 binary32 operation in
 * source order (x86-64 SSE, FLT_EVAL_METHOD == 0).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <errno.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static char g_err[512];
const char* orc_last_error(void) { return g_err; }
static int fail(const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return -1;
}

/* ------------------------------------------------------------------ */
/* a1: benchmark text loader + CSR build, gpucpg/gpucpg.cu:453-548      */
/* ------------------------------------------------------------------ */

void orc_graph_free(orc_graph* g) {
  if (!g) return;
  free(g->oadjp); free(g->oadj); free(g->owgt);
  free(g->iadjp); free(g->iadj); free(g->iwgt);
  memset(g, 0, sizeof(*g));
}

/* Counting-sort an edge list into fanout/fanin CSR, keeping input order
 * within every vertex (gpucpg/gpucpg.cu:507-546 appends per-vertex vectors
 * in input order and then concatenates them in vertex order). */
int orc_build_csr(int n, long long m, const int* eu, const int* ev,
                  const float* ew, orc_graph* g) {
  memset(g, 0, sizeof(*g));
  if (n < 0 || m < 0) return fail("negative size");
  g->n = n;
  g->m = m;
  g->oadjp = calloc((size_t)n + 1, sizeof(int));
  g->iadjp = calloc((size_t)n + 1, sizeof(int));
  g->oadj = malloc(sizeof(int) * (size_t)(m ? m : 1));
  g->owgt = malloc(sizeof(float) * (size_t)(m ? m : 1));
  g->iadj = malloc(sizeof(int) * (size_t)(m ? m : 1));
  g->iwgt = malloc(sizeof(float) * (size_t)(m ? m : 1));
  if (!g->oadjp || !g->iadjp || !g->oadj || !g->owgt || !g->iadj || !g->iwgt) {
    orc_graph_free(g);
    return fail("out of memory");
  }
  for (long long e = 0; e < m; ++e) {
    if (eu[e] < 0 || eu[e] >= n || ev[e] < 0 || ev[e] >= n) {
      orc_graph_free(g);
      return fail("vertex id out of range");
    }
    g->oadjp[eu[e] + 1]++;
    g->iadjp[ev[e] + 1]++;
  }
  for (int i = 0; i < n; ++i) {
    g->oadjp[i + 1] += g->oadjp[i];
    g->iadjp[i + 1] += g->iadjp[i];
  }
  int* ocur = malloc(sizeof(int) * (size_t)(n ? n : 1));
  int* icur = malloc(sizeof(int) * (size_t)(n ? n : 1));
  if (!ocur || !icur) { free(ocur); free(icur); orc_graph_free(g); return fail("out of memory"); }
  memcpy(ocur, g->oadjp, sizeof(int) * (size_t)n);
  memcpy(icur, g->iadjp, sizeof(int) * (size_t)n);
  for (long long e = 0; e < m; ++e) {
    int p = ocur[eu[e]]++;
    g->oadj[p] = ev[e];
    g->owgt[p] = ew[e];
    int q = icur[ev[e]]++;
    g->iadj[q] = eu[e];
    g->iwgt[q] = ew[e];
  }
  free(ocur);
  free(icur);
  return 0;
}

/* std::stoi: skip leading blanks, optional sign, decimal digits. */
static int parse_int(const char* s, size_t len, int* out) {
  char buf[64];
  if (len >= sizeof(buf)) return -1;
  memcpy(buf, s, len);
  buf[len] = 0;
  char* end = NULL;
  errno = 0;
  long v = strtol(buf, &end, 10);
  if (end == buf || errno == ERANGE || v > 2147483647L || v < -2147483647L - 1) return -1;
  *out = (int)v;
  return 0;
}

int orc_parse_text(const char* path, int ignore_wgts, orc_graph* g) {
  memset(g, 0, sizeof(*g));
  FILE* f = fopen(path, "r");
  if (!f) return fail("Unable to open file");
  char* line = NULL;
  size_t cap = 0;
  ssize_t len = getline(&line, &cap, f);
  int n = 0;
  if (len < 0 || parse_int(line, (size_t)len, &n) != 0 || n < 0) {
    free(line); fclose(f);
    return fail("bad vertex count line");
  }
  /* skip the n vertex-name lines (gpucpg/gpucpg.cu:475-478) */
  for (int i = 0; i < n; ++i) {
    if (getline(&line, &cap, f) < 0) break;
  }
  long long m = 0, mcap = 1024;
  int* eu = malloc(sizeof(int) * (size_t)mcap);
  int* ev = malloc(sizeof(int) * (size_t)mcap);
  float* ew = malloc(sizeof(float) * (size_t)mcap);
  int rc = 0;
  while ((len = getline(&line, &cap, f)) >= 0) {
    /* "u" -> "v"[, w];  -- fields between quotes 1-2 and 3-4 */
    char* q[4];
    char* p = line;
    int nq = 0;
    while (nq < 4 && (p = strchr(p, '"')) != NULL) { q[nq++] = p; ++p; }
    if (nq < 4) { rc = fail("malformed edge line"); break; }
    int u, v;
    if (parse_int(q[0] + 1, (size_t)(q[1] - q[0] - 1), &u) != 0 ||
        parse_int(q[2] + 1, (size_t)(q[3] - q[2] - 1), &v) != 0) {
      rc = fail("malformed vertex id"); break;
    }
    float w = 1.0f;
    if (!ignore_wgts && strchr(line, ',') != NULL) {
      /* getline(',') from after the 4th quote, then getline(';') */
      char* c = strchr(q[3] + 1, ',');
      if (!c) { rc = fail("malformed weight"); break; }
      char* s = c + 1;
      char* semi = strchr(s, ';');
      size_t wl = semi ? (size_t)(semi - s) : strlen(s);
      char buf[128];
      if (wl >= sizeof(buf)) { rc = fail("malformed weight"); break; }
      memcpy(buf, s, wl);
      buf[wl] = 0;
      char* end = NULL;
      w = strtof(buf, &end);
      if (end == buf) { rc = fail("malformed weight"); break; }
    }
    if (m == mcap) {
      mcap *= 2;
      eu = realloc(eu, sizeof(int) * (size_t)mcap);
      ev = realloc(ev, sizeof(int) * (size_t)mcap);
      ew = realloc(ew, sizeof(float) * (size_t)mcap);
    }
    eu[m] = u; ev[m] = v; ew[m] = w; ++m;
  }
  free(line);
  fclose(f);
  if (rc == 0) rc = orc_build_csr(n, m, eu, ev, ew, g);
  free(eu); free(ev); free(ew);
  return rc;
}

/* ------------------------------------------------------------------ */
/* a3: levelize, host CpGen::levelize gpucpg/gpucpg.cu:21368-21403      */
/* (same level sets as the device peel bfs_td_step_privatized,          */
/*  gpucpg/gpucpg.cu:2215-2283 driven by :17642-17681)                  */
/* ------------------------------------------------------------------ */
int orc_levelize(const orc_graph* g, int* level, int* queue, int* lvlp,
                 int* num_levels) {
  const int n = g->n;
  int* deps = malloc(sizeof(int) * (size_t)(n ? n : 1));
  if (!deps) return fail("out of memory");
  int qt = 0;
  for (int v = 0; v < n; ++v) {
    deps[v] = g->oadjp[v + 1] - g->oadjp[v];
    level[v] = -1;
    if (deps[v] == 0) { queue[qt++] = v; level[v] = 0; }
  }
  int L = 0;
  int qh = 0;
  lvlp[0] = 0;
  while (qh < qt) {
    const int lvl_end = qt;
    lvlp[L + 1] = lvl_end;
    for (; qh < lvl_end; ++qh) {
      const int v = queue[qh];
      for (int e = g->iadjp[v]; e < g->iadjp[v + 1]; ++e) {
        const int x = g->iadj[e];
        if (--deps[x] == 0) { queue[qt++] = x; level[x] = L + 1; }
      }
    }
    ++L;
  }
  *num_levels = L;
  free(deps);
  return qt;  /* number of levelized vertices */
}

/* a4: relax_bu_step gpucpg/gpucpg.cu:1167-1187, loop :17933-17961,
 * init :17035-17039 (sinks 0, others INT_MAX).  With sink_mask (n bytes,
 * an endpoint subset: an extension the reference does not have), only
 * the chosen sinks start at 0; a fanout at INT_MAX offers no candidate,
 * so a vertex reaching no chosen sink stays at INT_MAX. */
void orc_relax_masked(const orc_graph* g, const int* queue, const int* lvlp,
                      int num_levels, const unsigned char* sink_mask, int* dist) {
  for (int v = 0; v < g->n; ++v) dist[v] = ORC_INT_MAX;
  if (num_levels > 0)
    for (int i = lvlp[0]; i < lvlp[1]; ++i)
      if (!sink_mask || sink_mask[queue[i]]) dist[queue[i]] = 0;
  for (int d = 1; d < num_levels; ++d) {
    for (int i = lvlp[d]; i < lvlp[d + 1]; ++i) {
      const int v = queue[i];
      int best = ORC_INT_MAX;
      for (int e = g->oadjp[v]; e < g->oadjp[v + 1]; ++e) {
        const int w = g->oadj[e];
        if (dist[w] == ORC_INT_MAX) continue;
        const int wgt = (int)(g->owgt[e] * (float)ORC_SCALE_UP);
        const long long cand = (long long)dist[w] + wgt;
        const int c = (int)cand;
        if (c < best) best = c;
      }
      dist[v] = best;
    }
  }
}

void orc_relax(const orc_graph* g, const int* queue, const int* lvlp,
               int num_levels, int* dist) {
  orc_relax_masked(g, queue, lvlp, num_levels, NULL, dist);
}

/* a5: update_successors_bu gpucpg/gpucpg.cu:2606-2628 -- the largest
 * neighbour id that achieves dist equality, -1 if none. */
void orc_successors(const orc_graph* g, const int* dist, int* succ) {
  for (int v = 0; v < g->n; ++v) {
    int s = -1;
    for (int e = g->oadjp[v]; e < g->oadjp[v + 1]; ++e) {
      const int w = g->oadj[e];
      if (dist[w] == ORC_INT_MAX) continue;
      const int wgt = (int)(g->owgt[e] * (float)ORC_SCALE_UP);
      if ((long long)dist[v] == (long long)dist[w] + wgt && w > s) s = w;
    }
    succ[v] = s;
  }
}

/* a6: (float)dist / SCALE_UP at every use, e.g. gpucpg/gpucpg.cu:2672-2673,
 * :18508, :21275-21276 -- correctly rounded IEEE division. */
void orc_distf(int n, const int* dist, float* distf) {
  for (int v = 0; v < n; ++v) distf[v] = (float)dist[v] / (float)ORC_SCALE_UP;
}

int orc_sfxt_masked(const orc_graph* g, const unsigned char* sink_mask, int* dist,
                    int* succ, int* level, int* num_levels) {
  const int n = g->n;
  int* queue = malloc(sizeof(int) * (size_t)(n ? n : 1));
  int* lvlp = malloc(sizeof(int) * ((size_t)n + 2));
  if (!queue || !lvlp) { free(queue); free(lvlp); return fail("out of memory"); }
  int cnt = orc_levelize(g, level, queue, lvlp, num_levels);
  orc_relax_masked(g, queue, lvlp, *num_levels, sink_mask, dist);
  orc_successors(g, dist, succ);
  free(queue);
  free(lvlp);
  return cnt;
}

int orc_sfxt(const orc_graph* g, int* dist, int* succ, int* level,
             int* num_levels) {
  return orc_sfxt_masked(g, NULL, dist, succ, level, num_levels);
}

/* a17: candidate_slack gpucpg/gpucpg/tc_pfxt_candidates.cuh:651-660;
 * the same left-to-right order as compute_short_long_path_counts
 * gpucpg/gpucpg.cu:2672-2676. */
float orc_candidate_cost(float parent, int src_dist, int dst_dist, float wgt) {
  const float dn = (float)dst_dist / (float)ORC_SCALE_UP;
  const float dv = (float)src_dist / (float)ORC_SCALE_UP;
  return parent + dn + wgt - dv;
}

/* a9: compute_split_inc_amount gpucpg/gpucpg.cuh:292-299 */
float orc_split_inc_amount(float avg_deg) {
  const float mn = 0.1f, mx = 100.0f, d0 = 3.0f, k = 0.8f;
  return mn + (mx - mn) / (1 + expf(k * (avg_deg - d0)));
}

int orc_classify(float slack, float split, float final_split, int use_final_split,
                 int skip_long_paths) {
  if (slack <= split) return 0;
  if (!skip_long_paths && (!use_final_split || slack <= final_split)) return 1;
  return 2;
}

long long orc_deviations(int n, const int* adjp, const int* adj, const float* wgt,
                         const int* succ, const int* dist, int* offsets, int* dsts,
                         float* deltas) {
  long long c = 0;
  offsets[0] = 0;
  for (int u = 0; u < n; ++u) {
    for (int e = adjp[u]; e < adjp[u + 1]; ++e) {
      const int w = adj[e];
      if (w == succ[u] || dist[w] == ORC_INT_MAX) continue;
      dsts[c] = w;
      deltas[c] = (float)dist[w] / (float)ORC_SCALE_UP + wgt[e] - (float)dist[u] / (float)ORC_SCALE_UP;
      ++c;
    }
    offsets[u + 1] = (int)c;
  }
  return c;
}

/* ------------------------------------------------------------------ */
/* heaps                                                              */
/* ------------------------------------------------------------------ */
typedef struct { float cost; int to; int level; } pq_item;

typedef struct { pq_item* a; long long n, cap; } minheap;
static int mh_push(minheap* h, pq_item x) {
  if (h->n == h->cap) {
    long long nc = h->cap ? h->cap * 2 : 1024;
    pq_item* na = realloc(h->a, sizeof(pq_item) * (size_t)nc);
    if (!na) return -1;
    h->a = na; h->cap = nc;
  }
  long long i = h->n++;
  while (i > 0) {
    long long p = (i - 1) / 2;
    if (h->a[p].cost <= x.cost) break;
    h->a[i] = h->a[p];
    i = p;
  }
  h->a[i] = x;
  return 0;
}
static pq_item mh_pop(minheap* h) {
  pq_item top = h->a[0];
  pq_item x = h->a[--h->n];
  long long i = 0;
  for (;;) {
    long long l = 2 * i + 1, r = l + 1, s = i;
    float sc = x.cost;
    if (l < h->n && h->a[l].cost < sc) { s = l; sc = h->a[l].cost; }
    if (r < h->n && h->a[r].cost < sc) { s = r; }
    if (s == i) break;
    h->a[i] = h->a[s];
    i = s;
  }
  if (h->n > 0) h->a[i] = x;
  return top;
}

/* max-heap of floats holding the k smallest nonzero costs seen so far */
typedef struct { float* a; long long n, k; } kheap;
static void kh_offer(kheap* h, float c) {
  if (h->n < h->k) {
    long long i = h->n++;
    while (i > 0) {
      long long p = (i - 1) / 2;
      if (h->a[p] >= c) break;
      h->a[i] = h->a[p];
      i = p;
    }
    h->a[i] = c;
    return;
  }
  if (!(c < h->a[0])) return;
  long long i = 0;
  for (;;) {
    long long l = 2 * i + 1, r = l + 1, s = i;
    float sc = c;
    if (l < h->n && h->a[l] > sc) { s = l; sc = h->a[l]; }
    if (r < h->n && h->a[r] > sc) { s = r; }
    if (s == i) break;
    h->a[i] = h->a[s];
    i = s;
  }
  h->a[i] = c;
}

static int cmp_float(const void* x, const void* y) {
  float a = *(const float*)x, b = *(const float*)y;
  return (a > b) - (a < b);
}

/* Roots: sources (in-degree 0, ascending id) with finite dist,
 * gpucpg/gpucpg.cu:18502-18513.  With root_tie_quirk, apply the
 * initial-split truncation of gpucpg/gpucpg.cu:19134-19180: the short
 * pile keeps the first (int)(0.005f*n+1) roots (ascending id) whose cost
 * is <= split = sorted[(size_t)(0.005f*n)], roots tied at <= split beyond
 * that count are lost, and every root above split goes to the long pile. */
static long long make_roots(const orc_graph* g, const int* dist,
                            int root_tie_quirk, pq_item** out,
                            orc_stats* st) {
  const int n = g->n;
  pq_item* r = malloc(sizeof(pq_item) * (size_t)(n ? n : 1));
  long long nr = 0;
  for (int v = 0; v < n; ++v) {
    if (g->iadjp[v + 1] - g->iadjp[v] != 0) continue;
    if (dist[v] == ORC_INT_MAX) continue;
    r[nr].cost = (float)dist[v] / (float)ORC_SCALE_UP;
    r[nr].to = v;
    r[nr].level = 0;
    ++nr;
  }
  if (st) st->roots = nr;
  if (root_tie_quirk && nr > 0) {
    float* s = malloc(sizeof(float) * (size_t)nr);
    for (long long i = 0; i < nr; ++i) s[i] = r[i].cost;
    qsort(s, (size_t)nr, sizeof(float), cmp_float);
    const float pn = 0.005f * (float)nr;
    const size_t idx = (size_t)pn;
    const float split = s[idx];
    const int num_short = (int)(pn + 1);
    free(s);
    if (st) st->split_index_mismatch = ((long long)num_short != (long long)idx + 1);
    long long kept_le = 0, out_n = 0, n_gt = 0;
    for (long long i = 0; i < nr; ++i) {
      if (r[i].cost > split) { r[out_n++] = r[i]; ++n_gt; continue; }
      if (kept_le < num_short) { r[out_n++] = r[i]; ++kept_le; }
    }
    if (st) {
      st->roots_dropped = nr - out_n;
      long long long_size = nr - num_short;
      st->long_dummies = long_size > n_gt ? long_size - n_gt : 0;
    }
    nr = out_n;
  }
  *out = r;
  return nr;
}

/* ------------------------------------------------------------------ */
/* a7-a15 golden: exact best-first top-K (the SEQUENTIAL method of      */
/* gpucpg/gpucpg.cu:21230-21301 with the GPG root filter/quirk and the  */
/* GPG zero-cost erase of :21338-21358), K-capped so that K=1e6 fits.   */
/* A child whose cost exceeds the K-th smallest nonzero cost pushed so  */
/* far can never be among the first K nonzero pops, so it is not pushed. */
/* ------------------------------------------------------------------ */
long long orc_topk(const orc_graph* g, const int* dist, const int* succ,
                   long long k, int root_tie_quirk, float* out_costs,
                   orc_stats* st) {
  orc_stats local;
  if (!st) st = &local;
  memset(st, 0, sizeof(*st));
  if (k <= 0) return 0;
  const int n = g->n;
  float* distf = malloc(sizeof(float) * (size_t)(n ? n : 1));
  orc_distf(n, dist, distf);
  pq_item* roots = NULL;
  long long nr = make_roots(g, dist, root_tie_quirk, &roots, st);
  minheap pq = {0};
  kheap kh = {malloc(sizeof(float) * (size_t)k), 0, k};
  for (long long i = 0; i < nr; ++i) {
    mh_push(&pq, roots[i]);
    if (roots[i].cost != 0.0f) kh_offer(&kh, roots[i].cost);
  }
  free(roots);
  long long got = 0;
  while (got < k && pq.n > 0) {
    const pq_item top = mh_pop(&pq);
    st->pops++;
    if (top.cost == 0.0f) st->zero_cost++;
    else out_costs[got++] = top.cost;
    const float slack = top.cost;
    for (int v = top.to; v != -1; v = succ[v]) {
      st->hops++;
      const float dv = distf[v];
      const int sv = succ[v];
      for (int e = g->oadjp[v]; e < g->oadjp[v + 1]; ++e) {
        st->edges++;
        const int w = g->oadj[e];
        if (w == sv) continue;
        if (dist[w] == ORC_INT_MAX) continue;
        st->children++;
        const float c = slack + distf[w] + g->owgt[e] - dv;
        if (c < slack) st->nonmonotone++;
        if (kh.n == kh.k && c > kh.a[0]) continue;
        pq_item it = {c, w, top.level + 1};
        if (mh_push(&pq, it) != 0) { st->aborted = 1; goto done; }
        st->pushes++;
        if (c != 0.0f) kh_offer(&kh, c);
      }
    }
  }
done:
  qsort(out_costs, (size_t)got, sizeof(float), cmp_float);
  free(pq.a);
  free(kh.a);
  free(distf);
  return st->aborted ? -1 : got;
}

/* ------------------------------------------------------------------ */
/* orc_topk with vertex sequences.  Every queued path keeps its         */
/* deviation (from -> to) and the index of its popped parent, as the    */
/* reference's PfxtNode keeps {from, to, parent} (gpucpg/gpucpg.cuh:     */
/* 74-107; parent = node index at expansion, gpucpg.cu:3280-3288).  A   */
/* path's vertices are its root's suffix chain with each deviation      */
/* spliced in at its from-vertex: source, succ, ..., from, to, succ...  */
/* (the chain walk of :2654-2694 read backwards).                       */
/* ------------------------------------------------------------------ */
typedef struct { float cost; int to; int from; int level; long long parent; } pq_path;
typedef struct { pq_path* a; long long n, cap; } pathheap;
static int ph_push(pathheap* h, pq_path x) {
  if (h->n == h->cap) {
    long long nc = h->cap ? h->cap * 2 : 1024;
    pq_path* na = realloc(h->a, sizeof(pq_path) * (size_t)nc);
    if (!na) return -1;
    h->a = na; h->cap = nc;
  }
  long long i = h->n++;
  while (i > 0) {
    long long p = (i - 1) / 2;
    if (h->a[p].cost <= x.cost) break;
    h->a[i] = h->a[p];
    i = p;
  }
  h->a[i] = x;
  return 0;
}
static pq_path ph_pop(pathheap* h) {
  pq_path top = h->a[0];
  pq_path x = h->a[--h->n];
  long long i = 0;
  for (;;) {
    long long l = 2 * i + 1, r = l + 1, s = i;
    float sc = x.cost;
    if (l < h->n && h->a[l].cost < sc) { s = l; sc = h->a[l].cost; }
    if (r < h->n && h->a[r].cost < sc) { s = r; }
    if (s == i) break;
    h->a[i] = h->a[s];
    i = s;
  }
  if (h->n > 0) h->a[i] = x;
  return top;
}

typedef struct { float cost; long long pop; } cost_pop;
static int cmp_cost_pop(const void* x, const void* y) {
  const cost_pop* a = (const cost_pop*)x;
  const cost_pop* b = (const cost_pop*)y;
  if (a->cost != b->cost) return (a->cost > b->cost) - (a->cost < b->cost);
  return (a->pop > b->pop) - (a->pop < b->pop);
}

/* Writes the vertices of popped path i of pops[] into out (cap entries);
 * returns the count, or -1 when it does not fit or the chain is broken. */
static long long path_vertices(const pq_path* pops, long long i, const int* succ, int n,
                               int* devf, int* devt, int* out, long long cap) {
  long long nd = 0;
  long long r = i;
  while (pops[r].from >= 0) {
    if (nd >= n) return -1;
    devf[nd] = pops[r].from;
    devt[nd] = pops[r].to;
    ++nd;
    r = pops[r].parent;
  }
  int v = pops[r].to;
  long long len = 0, di = nd;  /* deviations in root-to-leaf order: devf[nd-1] first */
  while (v != -1) {
    if (len >= cap || len > n) return -1;
    out[len++] = v;
    if (di > 0 && devf[di - 1] == v) {
      v = devt[di - 1];
      --di;
    } else {
      v = succ[v];
    }
  }
  return di == 0 ? len : -1;
}

long long orc_topk_paths(const orc_graph* g, const int* dist, const int* succ, long long k,
                         int root_tie_quirk, float* out_costs, long long* out_vptr,
                         int* out_verts, long long vcap, orc_stats* st) {
  orc_stats local;
  if (!st) st = &local;
  memset(st, 0, sizeof(*st));
  if (out_vptr) out_vptr[0] = 0;
  if (k <= 0) return 0;
  const int n = g->n;
  float* distf = malloc(sizeof(float) * (size_t)(n ? n : 1));
  orc_distf(n, dist, distf);
  pq_item* roots = NULL;
  long long nr = make_roots(g, dist, root_tie_quirk, &roots, st);
  pathheap pq = {0};
  kheap kh = {malloc(sizeof(float) * (size_t)k), 0, k};
  for (long long i = 0; i < nr; ++i) {
    pq_path it = {roots[i].cost, roots[i].to, -1, 0, -1};
    ph_push(&pq, it);
    if (roots[i].cost != 0.0f) kh_offer(&kh, roots[i].cost);
  }
  free(roots);
  long long npop = 0, pcap = 1024, got = 0;
  pq_path* pops = malloc(sizeof(pq_path) * (size_t)pcap);
  cost_pop* rep = malloc(sizeof(cost_pop) * (size_t)k);
  while (got < k && pq.n > 0) {
    const pq_path top = ph_pop(&pq);
    if (npop == pcap) {
      pcap *= 2;
      pq_path* np = realloc(pops, sizeof(pq_path) * (size_t)pcap);
      if (!np) { st->aborted = 1; goto done; }
      pops = np;
    }
    const long long me = npop++;
    pops[me] = top;
    st->pops++;
    if (top.cost == 0.0f) st->zero_cost++;
    else { rep[got].cost = top.cost; rep[got].pop = me; ++got; }
    const float slack = top.cost;
    for (int v = top.to; v != -1; v = succ[v]) {
      st->hops++;
      const float dv = distf[v];
      const int sv = succ[v];
      for (int e = g->oadjp[v]; e < g->oadjp[v + 1]; ++e) {
        st->edges++;
        const int w = g->oadj[e];
        if (w == sv) continue;
        if (dist[w] == ORC_INT_MAX) continue;
        st->children++;
        const float c = slack + distf[w] + g->owgt[e] - dv;
        if (c < slack) st->nonmonotone++;
        if (kh.n == kh.k && c > kh.a[0]) continue;
        pq_path it = {c, w, v, top.level + 1, me};
        if (ph_push(&pq, it) != 0) { st->aborted = 1; goto done; }
        st->pushes++;
        if (c != 0.0f) kh_offer(&kh, c);
      }
    }
  }
  qsort(rep, (size_t)got, sizeof(cost_pop), cmp_cost_pop);
  {
    int* devf = malloc(sizeof(int) * (size_t)(n + 1));
    int* devt = malloc(sizeof(int) * (size_t)(n + 1));
    long long off = 0;
    for (long long i = 0; i < got; ++i) {
      out_costs[i] = rep[i].cost;
      if (!out_vptr || !out_verts) continue;
      const long long len = path_vertices(pops, rep[i].pop, succ, n, devf, devt, out_verts + off,
                                          vcap - off);
      if (len < 0) { st->aborted = 5; break; }
      off += len;
      out_vptr[i + 1] = off;
    }
    free(devf);
    free(devt);
  }
done:
  free(pq.a);
  free(kh.a);
  free(pops);
  free(rep);
  free(distf);
  return st->aborted ? -1 : got;
}

/* Faithful SEQUENTIAL, gpucpg/gpucpg.cu:21230-21301 + :21332-21358:
 * every source seeds the queue (no finite-dist filter), k pops, every
 * deviation child pushed, final sort, zero-cost nodes erased. */
long long orc_sequential(const orc_graph* g, const int* dist,
                         const int* succ, long long k, float* out_costs,
                         orc_stats* st) {
  orc_stats local;
  if (!st) st = &local;
  memset(st, 0, sizeof(*st));
  const int n = g->n;
  minheap pq = {0};
  for (int v = 0; v < n; ++v) {
    if (g->iadjp[v + 1] - g->iadjp[v] != 0) continue;
    pq_item it = {(float)dist[v] / (float)ORC_SCALE_UP, v, 0};
    mh_push(&pq, it);
    st->roots++;
  }
  long long got = 0;
  for (long long i = 0; i < k; ++i) {
    if (pq.n == 0) break;
    const pq_item top = mh_pop(&pq);
    st->pops++;
    out_costs[got++] = top.cost;
    for (int v = top.to; v != -1; v = succ[v]) {
      st->hops++;
      for (int e = g->oadjp[v]; e < g->oadjp[v + 1]; ++e) {
        st->edges++;
        const int w = g->oadj[e];
        if (w == succ[v]) continue;
        if (dist[w] == ORC_INT_MAX) continue;
        const float dn = (float)dist[w] / (float)ORC_SCALE_UP;
        const float dv = (float)dist[v] / (float)ORC_SCALE_UP;
        const float c = top.cost + dn + g->owgt[e] - dv;
        pq_item it = {c, w, top.level + 1};
        if (mh_push(&pq, it) != 0) { st->aborted = 1; free(pq.a); return -1; }
        st->pushes++;
        st->children++;
      }
    }
  }
  free(pq.a);
  qsort(out_costs, (size_t)got, sizeof(float), cmp_float);
  long long w = 0;
  for (long long i = 0; i < got; ++i) {
    if (out_costs[i] == 0.0f) { st->zero_cost++; continue; }
    out_costs[w++] = out_costs[i];
  }
  return w;
}

/* ------------------------------------------------------------------ */
/* Faithful SHORT_LONG schedule, gpucpg/gpucpg.cu:19134-21228 (GPG       */
/* branch: TC off, tile_spur off), :21323-21358 (final sort + erase).   */
/* The device kernels are replaced by sequential loops with identical   */
/* arithmetic: compute_short_long_path_counts :2630-2769,                */
/* expand_short_pile :3234-3306, _skip_long_paths :7066-7119,            */
/* _update_final_window :7121-7188.  Small inputs only.                  */
/* ------------------------------------------------------------------ */
typedef struct { orc_node* a; long long n, cap; } pile;
static int pile_reserve(pile* p, long long cap) {
  if (cap <= p->cap) return 0;
  long long nc = p->cap ? p->cap : 64;
  while (nc < cap) nc *= 2;
  orc_node* na = realloc(p->a, sizeof(orc_node) * (size_t)nc);
  if (!na) return -1;
  p->a = na; p->cap = nc;
  return 0;
}
static int pile_resize(pile* p, long long n) {
  if (pile_reserve(p, n) != 0) return -1;
  for (long long i = p->n; i < n; ++i) {
    orc_node d = {-1, -1, -1, -1, 0, 0.0f};  /* PfxtNode() defaults */
    p->a[i] = d;
  }
  p->n = n;
  return 0;
}
static int cmp_node(const void* x, const void* y) {
  float a = ((const orc_node*)x)->slack, b = ((const orc_node*)y)->slack;
  return (a > b) - (a < b);
}

typedef struct {
  const orc_graph* g;
  const int* dist;
  const int* succ;
} sl_ctx;

/* mode 0: count (<=split short, >split long)
 * mode 1: count (<=split short, (split,final] long)
 * emit 0: none; 1: expand_short_pile; 2: skip_long; 3: final window */
static void sl_walk(const sl_ctx* c, const pile* sp, long long wb, long long we,
                    float split, float final_split, int mode, int emit,
                    long long* ns, long long* nl, pile* out_s, long long* ts,
                    pile* out_l, long long* tl) {
  *ns = 0; *nl = 0;
  for (long long idx = wb; idx < we; ++idx) {
    const orc_node node = sp->a[idx];
    int v = node.to;
    while (v != -1) {
      for (int e = c->g->oadjp[v]; e < c->g->oadjp[v + 1]; ++e) {
        const int w = c->g->oadj[e];
        if (w == c->succ[v]) continue;
        if (c->dist[w] == ORC_INT_MAX) continue;
        const float dn = (float)c->dist[w] / (float)ORC_SCALE_UP;
        const float dv = (float)c->dist[v] / (float)ORC_SCALE_UP;
        const float s = node.slack + dn + c->g->owgt[e] - dv;
        orc_node nn = {node.level + 1, v, w, (int)idx, 0, s};
        if (emit == 0) {
          if (mode == 0) { if (s > split) ++*nl; else ++*ns; }
          else { if (s <= split) ++*ns; else if (s > split && s <= final_split) ++*nl; }
        } else if (emit == 1) {
          if (s <= split) out_s->a[(*ts)++] = nn; else out_l->a[(*tl)++] = nn;
        } else if (emit == 2) {
          if (s <= split) out_s->a[(*ts)++] = nn;
        } else {
          if (s <= split) out_s->a[(*ts)++] = nn;
          else if (s > split && s <= final_split) out_l->a[(*tl)++] = nn;
        }
      }
      v = c->succ[v];
    }
  }
}

long long orc_short_long(const orc_graph* g, const int* dist,
                         const int* succ, long long k, float init_split_perc,
                         int use_fixed_inc, float fixed_inc,
                         double long_pile_limit_bytes,
                         long long node_budget, float* out_costs,
                         orc_stats* st) {
  orc_stats local;
  if (!st) st = &local;
  memset(st, 0, sizeof(*st));
  sl_ctx c = {g, dist, succ};
  const int n = g->n;
  pile sp = {0}, lp = {0};
  long long result = -1;
  /* roots, :18502-18513 */
  for (int v = 0; v < n; ++v) {
    if (g->iadjp[v + 1] - g->iadjp[v] != 0) continue;
    if (dist[v] == ORC_INT_MAX) continue;
    if (pile_reserve(&sp, sp.n + 1) != 0) goto out;
    orc_node r = {0, -1, v, -1, 0, (float)dist[v] / (float)ORC_SCALE_UP};
    sp.a[sp.n++] = r;
  }
  st->roots = sp.n;
  long long short_size = sp.n;
  if (short_size == 0) { result = 0; goto out; }
  /* initial split, :19134-19180 */
  {
    orc_node* tmp = malloc(sizeof(orc_node) * (size_t)short_size);
    memcpy(tmp, sp.a, sizeof(orc_node) * (size_t)short_size);
    qsort(tmp, (size_t)short_size, sizeof(orc_node), cmp_node);
    const float pn = init_split_perc * (float)short_size;
    const size_t idx = (size_t)pn;
    float split0 = tmp[idx].slack;
    free(tmp);
    const int num_short = (int)(pn + 1);
    st->split_index_mismatch = ((long long)num_short != (long long)idx + 1);
    const long long num_long = short_size - num_short;
    if (pile_resize(&lp, num_long) != 0) goto out;
    long long wl = 0, ws = 0, n_le = 0;
    for (long long i = 0; i < short_size; ++i) {
      if (sp.a[i].slack > split0) {
        if (wl >= num_long) { st->aborted = 3; goto out; }  /* reference UB */
        lp.a[wl++] = sp.a[i];
      } else {
        sp.a[ws++] = sp.a[i];
        ++n_le;
      }
    }
    if (n_le < num_short) { st->aborted = 3; goto out; }  /* reference UB */
    st->roots_dropped = n_le - num_short;
    st->long_dummies = num_long - wl;
    sp.n = num_short;
    short_size = num_short;
    float split = split0;
    long long long_size = num_long;
    long long wstart = 0, wend = short_size;
    long long ns = 0, nl = 0;
    const float init_inc = use_fixed_inc ? fixed_inc
                                         : orc_split_inc_amount((float)g->m / (float)n);
    float inc = init_inc;
    const float limit = (float)long_pile_limit_bytes;
    long long final_window_size = 0;
    float final_split = FLT_MAX;
    long long steps = 0;
    long long guard = 0;
    for (;;) {
      const long long W = wend - wstart;
      if (W > 0) {
        sl_walk(&c, &sp, wstart, wend, split, final_split, 0, 0, &ns, &nl,
                NULL, NULL, NULL, NULL);
        if ((float)(sizeof(orc_node) * (size_t)(long_size + nl)) > limit &&
            final_window_size == 0 && short_size + ns < k) {
          const long long needed = k - (short_size + ns);
          qsort(lp.a, (size_t)lp.n, sizeof(orc_node), cmp_node);
          if (needed - 1 >= lp.n) { st->aborted = 2; goto out; }
          final_split = lp.a[needed - 1].slack;
          long long w2 = 0;
          for (long long i = 0; i < lp.n; ++i)
            if (!(lp.a[i].slack > final_split)) lp.a[w2++] = lp.a[i];
          lp.n = w2;
          final_window_size = w2;
          long_size = w2;
          st->final_window = w2;
        }
        short_size += ns;
        if (short_size + long_size + nl > node_budget) { st->aborted = 1; goto out; }
        if (pile_reserve(&sp, short_size) != 0) goto out;
        long long ts = sp.n;  /* == old short size */
        if (short_size >= k) {
          lp.n = 0;
          long_size = 0;
          sl_walk(&c, &sp, wstart, wend, split, final_split, 0, 2, &ns, &nl,
                  &sp, &ts, NULL, NULL);
        } else {
          if (final_window_size > 0) {
            sl_walk(&c, &sp, wstart, wend, split, final_split, 1, 0, &ns, &nl,
                    NULL, NULL, NULL, NULL);
          }
          long_size += nl;
          long long tl = lp.n;
          if (pile_reserve(&lp, long_size) != 0) goto out;
          sl_walk(&c, &sp, wstart, wend, split, final_split, 0,
                  final_window_size > 0 ? 3 : 1, &ns, &nl, &sp, &ts, &lp, &tl);
          lp.n = tl;
        }
        /* the walks leave ns/nl as the count pass reported */
        ns = ts - sp.n;
        sp.n = ts;
        wstart += W;
        wend = short_size;
        if (short_size >= k && long_size > 0) { lp.n = 0; long_size = 0; }
        if (wstart == wend && final_window_size > 0 && final_split > split) {
          split = final_split;
          if (pile_reserve(&sp, short_size + long_size) != 0) goto out;
          memcpy(sp.a + wstart, lp.a, sizeof(orc_node) * (size_t)lp.n);
          short_size += long_size;
          sp.n = short_size;
          wend += long_size;
        }
      } else {
        ++steps;
        if (long_size == 0 || short_size >= k) break;
        while (ns == 0) {
          if (++guard > 100000000) { st->aborted = 4; goto out; }
          if (steps == 1) {
            float mn = FLT_MAX;
            for (long long i = 0; i < lp.n; ++i) if (lp.a[i].slack < mn) mn = lp.a[i].slack;
            const float a1 = mn + inc, a2 = split + inc;
            split = a1 > a2 ? a1 : a2;
          } else {
            split += inc;
          }
          long long remaining = 0;
          for (long long i = 0; i < lp.n; ++i) if (lp.a[i].slack > split) ++remaining;
          ns = (long long)lp.n - remaining;
          nl = remaining;
        }
        short_size += ns;
        if (!use_fixed_inc) {
          if ((float)short_size > 0.5f * (float)k) inc = init_inc;
          else inc *= 1.2f;
        }
        if (pile_reserve(&sp, short_size) != 0) goto out;
        long long w2 = wend;
        for (long long i = 0; i < lp.n; ++i)
          if (lp.a[i].slack <= split) sp.a[w2++] = lp.a[i];
        sp.n = short_size;
        wend += ns;
        if (short_size >= k) {
          lp.n = 0;
          long_size = 0;
        } else {
          long long w3 = 0;
          for (long long i = 0; i < lp.n; ++i)
            if (!(lp.a[i].slack <= split)) lp.a[w3++] = lp.a[i];
          lp.n = w3;
          long_size = w3;
        }
      }
    }
    st->steps = steps;
    st->short_pile = short_size;
    /* final sort + zero erase + first k, :21323-21358, :21409-21420 */
    qsort(sp.a, (size_t)sp.n, sizeof(orc_node), cmp_node);
    long long got = 0;
    for (long long i = 0; i < sp.n && got < k; ++i) {
      if (sp.a[i].slack == 0.0f) { st->zero_cost++; continue; }
      out_costs[got++] = sp.a[i].slack;
    }
    result = got;
  }
out:
  free(sp.a);
  free(lp.a);
  return result;
}

/* ------------------------------------------------------------------ */
/* Brute force: every source->sink vertex path by DFS over the fanout   */
/* CSR.  The cost of a path applies the deviation recurrence at every   */
/* step that leaves the suffix chain (next vertex != succ), in path     */
/* order; a step onto succ keeps the cost and is taken once even with   */
/* parallel edges, matching the reference's exclusion by vertex id.     */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_graph* g;
  const int* dist;
  const int* succ;
  const float* distf;
  float* out;
  long long n, cap;
} dfs_ctx;

static int dfs(dfs_ctx* c, int u, float cost) {
  const int s = c->succ[u];
  if (s == -1) {
    if (c->n >= c->cap) return -1;
    c->out[c->n++] = cost;
    return 0;
  }
  if (dfs(c, s, cost) != 0) return -1;
  for (int e = c->g->oadjp[u]; e < c->g->oadjp[u + 1]; ++e) {
    const int w = c->g->oadj[e];
    if (w == s || c->dist[w] == ORC_INT_MAX) continue;
    const float nc = cost + c->distf[w] + c->g->owgt[e] - c->distf[u];
    if (dfs(c, w, nc) != 0) return -1;
  }
  return 0;
}

long long orc_all_paths(const orc_graph* g, const int* dist,
                        const int* succ, long long cap, float* out_costs) {
  const int n = g->n;
  float* distf = malloc(sizeof(float) * (size_t)(n ? n : 1));
  orc_distf(n, dist, distf);
  dfs_ctx c = {g, dist, succ, distf, out_costs, 0, cap};
  long long rc = 0;
  for (int v = 0; v < n && rc == 0; ++v) {
    if (g->iadjp[v + 1] - g->iadjp[v] != 0) continue;
    if (dist[v] == ORC_INT_MAX) continue;
    if (dfs(&c, v, distf[v]) != 0) rc = -1;
  }
  free(distf);
  if (rc != 0) return -1;
  qsort(out_costs, (size_t)c.n, sizeof(float), cmp_float);
  long long w = 0;
  for (long long i = 0; i < c.n; ++i)
    if (out_costs[i] != 0.0f) out_costs[w++] = out_costs[i];
  return w;
}

/* Brute force with vertex sequences: the same DFS, keeping the vertex  */
/* stack and writing each complete path's vertices (CSR in out_vptr).   */
typedef struct {
  dfs_ctx c;
  int* stack;
  long long depth;
  long long* vptr;
  int* verts;
  long long vn, vcap;
} dfsv_ctx;

static int dfsv(dfsv_ctx* x, int u, float cost) {
  x->stack[x->depth++] = u;
  const int s = x->c.succ[u];
  int rc = 0;
  if (s == -1) {
    if (x->c.n >= x->c.cap || x->vn + x->depth > x->vcap) { rc = -1; goto out; }
    x->c.out[x->c.n] = cost;
    memcpy(x->verts + x->vn, x->stack, sizeof(int) * (size_t)x->depth);
    x->vn += x->depth;
    x->vptr[++x->c.n] = x->vn;
    goto out;
  }
  if (dfsv(x, s, cost) != 0) { rc = -1; goto out; }
  for (int e = x->c.g->oadjp[u]; e < x->c.g->oadjp[u + 1]; ++e) {
    const int w = x->c.g->oadj[e];
    if (w == s || x->c.dist[w] == ORC_INT_MAX) continue;
    const float nc = cost + x->c.distf[w] + x->c.g->owgt[e] - x->c.distf[u];
    if (dfsv(x, w, nc) != 0) { rc = -1; goto out; }
  }
out:
  x->depth--;
  return rc;
}

/* Every source->sink path with its vertices, in DFS order (unsorted, zero
 * costs included); returns the number of paths or -1 when a cap is hit. */
long long orc_all_paths_vertices(const orc_graph* g, const int* dist, const int* succ,
                                 long long cap, float* out_costs, long long* out_vptr,
                                 int* out_verts, long long vcap) {
  const int n = g->n;
  float* distf = malloc(sizeof(float) * (size_t)(n ? n : 1));
  orc_distf(n, dist, distf);
  dfsv_ctx x;
  memset(&x, 0, sizeof(x));
  x.c.g = g; x.c.dist = dist; x.c.succ = succ; x.c.distf = distf;
  x.c.out = out_costs; x.c.cap = cap;
  x.stack = malloc(sizeof(int) * (size_t)(n + 1));
  x.vptr = out_vptr; x.verts = out_verts; x.vcap = vcap;
  out_vptr[0] = 0;
  long long rc = 0;
  for (int v = 0; v < n && rc == 0; ++v) {
    if (g->iadjp[v + 1] - g->iadjp[v] != 0) continue;
    if (dist[v] == ORC_INT_MAX) continue;
    if (dfsv(&x, v, distf[v]) != 0) rc = -1;
  }
  free(distf);
  free(x.stack);
  return rc != 0 ? -1 : x.c.n;
}

/* ------------------------------------------------------------------ */
/* compare_prefix, examples/tc-pfxt-inprocess-common.cuh:73-102         */
/* ------------------------------------------------------------------ */
void orc_compare_prefix(const float* baseline, long long nb, const float* tc,
                        long long nt, int k, float tol, orc_compare* r) {
  memset(r, 0, sizeof(*r));
  r->baseline_count = nb;
  r->tc_count = nt;
  long long kk = k > 0 ? k : 0;
  long long mn = nb < nt ? nb : nt;
  r->compared = kk < mn ? kk : mn;
  for (long long i = 0; i < r->compared; ++i) {
    const float diff = fabsf(baseline[i] - tc[i]);
    if (diff > r->max_diff) { r->max_diff = diff; r->max_diff_rank = (int)i + 1; }
    if (r->first_mismatch_rank == 0 && diff >= tol) r->first_mismatch_rank = (int)i + 1;
  }
  r->pass = nb >= kk && nt == kk && r->compared == kk && r->first_mismatch_rank == 0;
}

long long orc_first_bit_mismatch(const float* a, const float* b, long long n) {
  for (long long i = 0; i < n; ++i) {
    uint32_t x, y;
    memcpy(&x, a + i, 4);
    memcpy(&y, b + i, 4);
    if (x != y) return i;
  }
  return -1;
}

/* ------------------------------------------------------------------ */
/* CPU baseline: host SFXT + best-first PFXT, single thread,            */
/* std::chrono::steady_clock-equivalent (CLOCK_MONOTONIC).              */
/* ------------------------------------------------------------------ */
static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

double orc_query_ms(const orc_graph* g, long long k, float* out_costs,
                    long long* out_count, double* sfxt_ms, double* pfxt_ms) {
  const int n = g->n;
  int* dist = malloc(sizeof(int) * (size_t)(n ? n : 1));
  int* succ = malloc(sizeof(int) * (size_t)(n ? n : 1));
  int* level = malloc(sizeof(int) * (size_t)(n ? n : 1));
  int L = 0;
  const double t0 = now_ms();
  orc_sfxt(g, dist, succ, level, &L);
  const double t1 = now_ms();
  long long got = orc_topk(g, dist, succ, k, 1, out_costs, NULL);
  const double t2 = now_ms();
  free(dist); free(succ); free(level);
  if (out_count) *out_count = got;
  if (sfxt_ms) *sfxt_ms = t1 - t0;
  if (pfxt_ms) *pfxt_ms = t2 - t1;
  return t2 - t0;
}
