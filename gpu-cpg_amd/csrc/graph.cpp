// Host-side graph I/O and the seeded synthetic layered-DAG generator.
#include "graph.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <stdexcept>
#include <thread>

namespace cpg {

void build_csr(const EdgeList& el, HostGraph& g) {
  const int n = el.n;
  const int64_t m = static_cast<int64_t>(el.u.size());
  if (n < 0) throw std::runtime_error("negative vertex count");
  if (m > 0x7fffffffLL) throw std::runtime_error("edge count exceeds int32 CSR offsets");
  g = HostGraph();
  g.n = n;
  g.m = m;
  g.oadjp.assign(static_cast<size_t>(n) + 1, 0);
  g.iadjp.assign(static_cast<size_t>(n) + 1, 0);
  for (int64_t e = 0; e < m; ++e) {
    const int u = el.u[e], v = el.v[e];
    if (u < 0 || u >= n || v < 0 || v >= n)
      throw std::runtime_error("vertex id out of range in edge " + std::to_string(e));
    g.oadjp[u + 1]++;
    g.iadjp[v + 1]++;
  }
  for (int i = 0; i < n; ++i) {
    g.max_out_degree = std::max(g.max_out_degree, g.oadjp[i + 1]);
    g.oadjp[i + 1] += g.oadjp[i];
    g.iadjp[i + 1] += g.iadjp[i];
  }
  g.oadj.resize(m);
  g.owgt.resize(m);
  g.iadj.resize(m);
  g.iwgt.resize(m);
  std::vector<int> oc(g.oadjp.begin(), g.oadjp.end() - 1);
  std::vector<int> ic(g.iadjp.begin(), g.iadjp.end() - 1);
  float mw = 0.0f;
  for (int64_t e = 0; e < m; ++e) {
    const int u = el.u[e], v = el.v[e];
    const float w = el.w[e];
    const int p = oc[u]++;
    g.oadj[p] = v;
    g.owgt[p] = w;
    const int q = ic[v]++;
    g.iadj[q] = u;
    g.iwgt[q] = w;
    mw = std::max(mw, std::fabs(w));
  }
  g.max_abs_wgt = mw;
  for (int i = 0; i < n; ++i) {
    if (g.oadjp[i + 1] == g.oadjp[i]) g.sinks.push_back(i);
    if (g.iadjp[i + 1] == g.iadjp[i]) g.srcs.push_back(i);
  }
}

// std::stoi semantics on [b, e): leading blanks, optional sign, digits.
static bool parse_int(const char* b, const char* e, int* out) {
  char buf[64];
  const size_t len = static_cast<size_t>(e - b);
  if (len >= sizeof(buf)) return false;
  std::memcpy(buf, b, len);
  buf[len] = 0;
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(buf, &end, 10);
  if (end == buf || errno == ERANGE || v > 2147483647L || v < -2147483647L - 1) return false;
  *out = static_cast<int>(v);
  return true;
}

namespace {

// Parses the edge lines in [b, e) (whole lines) into u/v/w, in order.  On a
// malformed line sets err and the line's index within the range.
struct EdgeChunk {
  std::vector<int> u, v;
  std::vector<float> w;
  std::string err;
  int64_t err_line = -1;  // line index within the chunk
};

void parse_edge_lines(const char* b, const char* e, bool ignore_wgts, EdgeChunk& out) {
  const char* p = b;
  int64_t li = 0;
  for (; p < e; ++li) {
    const char* lb = p;
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(e - p)));
    const char* le = nl ? nl : e;
    p = nl ? nl + 1 : e;
    const char* q[4];
    int nq = 0;
    for (const char* c = lb; c < le && nq < 4; ++c)
      if (*c == '"') q[nq++] = c;
    int u, v;
    if (nq < 4 || !parse_int(q[0] + 1, q[1], &u) || !parse_int(q[2] + 1, q[3], &v)) {
      out.err = "malformed edge";
      out.err_line = li;
      return;
    }
    float w = 1.0f;
    if (!ignore_wgts && std::memchr(lb, ',', static_cast<size_t>(le - lb)) != nullptr) {
      const char* c = static_cast<const char*>(std::memchr(q[3] + 1, ',', static_cast<size_t>(le - q[3] - 1)));
      const char* s = c ? c + 1 : nullptr;
      const char* semi = s ? static_cast<const char*>(std::memchr(s, ';', static_cast<size_t>(le - s))) : nullptr;
      const char* we = semi ? semi : le;
      char wb[128];
      const size_t wl = s ? static_cast<size_t>(we - s) : 0;
      char* e2 = nullptr;
      if (s && wl < sizeof(wb)) {
        std::memcpy(wb, s, wl);
        wb[wl] = 0;
        w = std::strtof(wb, &e2);
      }
      if (!s || wl >= sizeof(wb) || e2 == wb) {
        out.err = "malformed weight";
        out.err_line = li;
        return;
      }
    }
    out.u.push_back(u);
    out.v.push_back(v);
    out.w.push_back(w);
  }
}

int64_t count_lines(const char* b, const char* e) {
  int64_t c = 0;
  for (const char* p = b; p < e; ++c) {
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(e - p)));
    p = nl ? nl + 1 : e;
  }
  return c;
}

}  // namespace

// The edge lines are cut into chunks at line breaks and parsed on up to 16
// host threads; the chunks are joined in file order, so the edge order (and
// hence the CSR order) is the file's.
void parse_text(const std::string& path, bool ignore_wgts, EdgeList& el, int threads) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("Unable to open file: " + path);
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<char> buf(static_cast<size_t>(sz > 0 ? sz : 0) + 1);
  const size_t got = sz > 0 ? std::fread(buf.data(), 1, static_cast<size_t>(sz), f) : 0;
  std::fclose(f);
  buf[got] = 0;
  const char* p = buf.data();
  const char* end = p + got;
  auto next_line = [&](const char*& lb, const char*& le) -> bool {
    if (p >= end) return false;
    lb = p;
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
    le = nl ? nl : end;
    p = nl ? nl + 1 : end;
    return true;
  };
  const char *lb, *le;
  el = EdgeList();
  int n = 0;
  if (!next_line(lb, le) || !parse_int(lb, le, &n) || n < 0)
    throw std::runtime_error("bad vertex count line in " + path);
  el.n = n;
  for (int i = 0; i < n; ++i)
    if (!next_line(lb, le)) break;
  const int64_t first_line = static_cast<int64_t>(n) + 2;  // 1-based line of the first edge
  const size_t bytes = static_cast<size_t>(end - p);
  int nt = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
  nt = std::max(1, std::min({nt, 16, static_cast<int>(bytes / (1u << 20)) + 1}));
  std::vector<const char*> cut(static_cast<size_t>(nt) + 1, end);
  cut[0] = p;
  for (int t = 1; t < nt; ++t) {
    const char* c = p + bytes * static_cast<size_t>(t) / static_cast<size_t>(nt);
    if (c < cut[t - 1]) c = cut[t - 1];
    const char* nl = c < end ? static_cast<const char*>(std::memchr(c, '\n', static_cast<size_t>(end - c))) : nullptr;
    cut[t] = nl ? nl + 1 : end;
  }
  std::vector<EdgeChunk> ch(static_cast<size_t>(nt));
  if (nt == 1) {
    parse_edge_lines(cut[0], cut[1], ignore_wgts, ch[0]);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
      th.emplace_back([&, t]() { parse_edge_lines(cut[t], cut[t + 1], ignore_wgts, ch[t]); });
    for (auto& x : th) x.join();
  }
  for (int t = 0; t < nt; ++t)
    if (!ch[t].err.empty()) {
      const int64_t line = first_line + count_lines(cut[0], cut[t]) + ch[t].err_line;
      throw std::runtime_error(ch[t].err + " at line " + std::to_string(line));
    }
  size_t m = 0;
  for (const auto& c : ch) m += c.u.size();
  el.u.reserve(m);
  el.v.reserve(m);
  el.w.reserve(m);
  for (auto& c : ch) {
    el.u.insert(el.u.end(), c.u.begin(), c.u.end());
    el.v.insert(el.v.end(), c.v.begin(), c.v.end());
    el.w.insert(el.w.end(), c.w.begin(), c.w.end());
  }
}

// Binary CSR (gpucpg/graph.h:24-55): N, M, offsets[N+1], adjncy[M],
// wgts[M], each a native-endian 4-byte word.
void write_bin(const HostGraph& g, const std::string& path) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("Unable to open file: " + path);
  const int32_t hdr[2] = {g.n, static_cast<int32_t>(g.m)};
  bool ok = std::fwrite(hdr, 4, 2, f) == 2;
  ok = ok && std::fwrite(g.oadjp.data(), 4, g.oadjp.size(), f) == g.oadjp.size();
  ok = ok && std::fwrite(g.oadj.data(), 4, g.oadj.size(), f) == g.oadj.size();
  ok = ok && std::fwrite(g.owgt.data(), 4, g.owgt.size(), f) == g.owgt.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error("write failed: " + path);
}

void read_bin(const std::string& path, EdgeList& el) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("Unable to open file: " + path);
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  int32_t hdr[2] = {0, 0};
  if (sz < 8 || std::fread(hdr, 4, 2, f) != 2) {
    std::fclose(f);
    throw std::runtime_error("binary CSR too short: " + path);
  }
  const int64_t n = hdr[0], m = hdr[1];
  if (n < 0 || m < 0 || 8 + 4 * (n + 1) + 8 * m != static_cast<int64_t>(sz)) {
    std::fclose(f);
    throw std::runtime_error("binary CSR size does not match its header: " + path);
  }
  std::vector<int> adjp(static_cast<size_t>(n) + 1);
  el = EdgeList();
  el.n = static_cast<int>(n);
  el.v.resize(static_cast<size_t>(m));
  el.w.resize(static_cast<size_t>(m));
  bool ok = std::fread(adjp.data(), 4, adjp.size(), f) == adjp.size();
  ok = ok && std::fread(el.v.data(), 4, el.v.size(), f) == el.v.size();
  ok = ok && std::fread(el.w.data(), 4, el.w.size(), f) == el.w.size();
  std::fclose(f);
  if (!ok) throw std::runtime_error("binary CSR read failed: " + path);
  if (adjp[0] != 0 || adjp[static_cast<size_t>(n)] != m)
    throw std::runtime_error("binary CSR offsets do not span the edges");
  el.u.resize(static_cast<size_t>(m));
  for (int64_t u = 0; u < n; ++u) {
    if (adjp[u + 1] < adjp[u]) throw std::runtime_error("binary CSR offsets decrease");
    for (int e = adjp[u]; e < adjp[u + 1]; ++e) el.u[e] = static_cast<int>(u);
  }
}

void write_text(const EdgeList& el, const std::string& path) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("Unable to open file: " + path);
  std::fprintf(f, "%d\n", el.n);
  for (int i = 0; i < el.n; ++i) std::fprintf(f, "\"%d\";\n", i);
  for (size_t e = 0; e < el.u.size(); ++e)
    std::fprintf(f, "\"%d\" -> \"%d\", %.9g;\n", el.u[e], el.v[e], static_cast<double>(el.w[e]));
  std::fclose(f);
}

// ---- seeded samplers over raw std::mt19937_64 output (no libstdc++
// distribution objects, so the stream is fixed by this file alone) ----
namespace {
struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t s) : g(s) {}
  double u01() { return (static_cast<double>(g() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  int uni(int a, int b) {  // inclusive
    const int r = a + static_cast<int>(u01() * (static_cast<double>(b) - a + 1.0));
    return r > b ? b : r;
  }
  double normal() {
    const double u1 = u01(), u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int geometric(double mean, int cap) {
    if (mean <= 1.0) return 1;
    const double p = 1.0 / mean;
    const double k = 1.0 + std::floor(std::log(u01()) / std::log1p(-p));
    return k > cap ? cap : static_cast<int>(k);
  }
};
}  // namespace

void generate(const GenParams& P, EdgeList& el) {
  if (P.n < 2 || P.levels < 2) throw std::runtime_error("generator needs n >= 2 and levels >= 2");
  Rng rng(P.seed);
  const int n = P.n, L = P.levels;
  const int S = std::max(1, static_cast<int>(std::llround(P.src_frac * n)));
  const int T = std::max(1, static_cast<int>(std::llround(P.sink_frac * n)));
  const int middle = n - S - T;
  if (middle < 0 || (L == 2 && middle != 0) || (L > 2 && middle < L - 2))
    throw std::runtime_error("generator: level sizes do not fit n");
  std::vector<int> lstart(L + 1, 0), lsize(L, 0);
  lsize[0] = S;
  lsize[L - 1] = T;
  if (L > 2) {
    const int base = middle / (L - 2), rem = middle % (L - 2);
    for (int l = 1; l < L - 1; ++l) lsize[l] = base + (l - 1 < rem ? 1 : 0);
  }
  for (int l = 0; l < L; ++l) lstart[l + 1] = lstart[l] + lsize[l];
  std::vector<int> lvl_of(n);
  for (int l = 0; l < L; ++l)
    for (int i = lstart[l]; i < lstart[l + 1]; ++i) lvl_of[i] = l;
  // level-order index -> vertex id
  std::vector<int> id(n);
  for (int i = 0; i < n; ++i) id[i] = i;
  if (P.permute_ids) {
    Rng pr(P.seed ^ 0x9E3779B97F4A7C15ULL);
    for (int i = n - 1; i > 0; --i) std::swap(id[i], id[pr.uni(0, i)]);
  }
  auto weight = [&](Rng& r) -> float {
    // the draw is made even for unit weights so that the unit-weight
    // variant has exactly the same edges as the weighted one
    float w;
    if (P.wgt_kind == 1) w = static_cast<float>(std::exp(P.mu + P.sigma * r.normal()));
    else w = static_cast<float>(P.w_lo + (P.w_hi - P.w_lo) * r.u01());
    return P.unit_wgts ? 1.0f : w;
  };
  auto span = [&](Rng& r) -> int {
    if (P.span_kind == 0) return r.uni(1, P.span_hi);
    return r.u01() < P.span_p1 ? 1 : r.uni(2, P.span_hi);
  };
  std::vector<std::vector<int>> tg(n);  // targets per level-order index
  std::vector<int> eu, ev;
  std::vector<float> ew;
  std::vector<int> indeg(n, 0);
  auto add = [&](int a, int b, float w) {
    tg[a].push_back(b);
    eu.push_back(a);
    ev.push_back(b);
    ew.push_back(w);
    indeg[b]++;
  };
  auto has = [&](int a, int b) {
    for (int x : tg[a]) if (x == b) return true;
    return false;
  };
  for (int a = 0; a < lstart[L - 1]; ++a) {
    const int l = lvl_of[a];
    const int deg = P.deg_kind == 0 ? rng.uni(P.deg_lo, P.deg_hi) : rng.geometric(P.deg_mean, P.deg_cap);
    for (int t = 0; t < deg; ++t) {
      for (int tries = 0; tries < 64; ++tries) {
        const int tl = std::min(l + span(rng), L - 1);
        const int b = lstart[tl] + rng.uni(0, lsize[tl] - 1);
        if (has(a, b)) continue;
        add(a, b, weight(rng));
        break;
      }
    }
    if (P.sink_edge) {
      for (int tries = 0; tries < 64; ++tries) {
        const int b = lstart[L - 1] + rng.uni(0, T - 1);
        if (has(a, b)) continue;
        add(a, b, weight(rng));
        break;
      }
    }
  }
  // every non-source vertex needs a fanin: an edge from the previous level
  for (int b = lstart[1]; b < n; ++b) {
    if (indeg[b] != 0) continue;
    const int l = lvl_of[b];
    const int a = lstart[l - 1] + rng.uni(0, lsize[l - 1] - 1);
    add(a, b, weight(rng));
  }
  if (P.dense_mean > 0.0) {
    Rng dr(P.dense_seed);
    const int64_t target = std::llround(P.dense_mean * n);
    const int nonsink = lstart[L - 1];
    int64_t misses = 0;
    while (static_cast<int64_t>(eu.size()) < target && misses < 1000000) {
      const int a = dr.uni(0, nonsink - 1);
      const int l = lvl_of[a];
      const int tl = std::min(l + dr.uni(1, P.dense_span_hi), L - 1);
      const int b = lstart[tl] + dr.uni(0, lsize[tl] - 1);
      const float w = weight(dr);
      if (has(a, b)) { ++misses; continue; }
      misses = 0;
      add(a, b, w);
    }
  }
  el = EdgeList();
  el.n = n;
  el.u.resize(eu.size());
  el.v.resize(ev.size());
  el.w = std::move(ew);
  for (size_t e = 0; e < eu.size(); ++e) {
    el.u[e] = id[eu[e]];
    el.v[e] = id[ev[e]];
  }
}

GenParams preset(int which, int index, int unit_wgts) {
  GenParams p;
  switch (which) {
    case 1:  // 50K vertices, 120 levels, out-degree U{1..4} + sink edge
      p.seed = 1; p.n = 50000; p.levels = 120; p.src_frac = 0.01; p.sink_frac = 0.01;
      p.deg_kind = 0; p.deg_lo = 1; p.deg_hi = 4; p.span_kind = 0; p.span_hi = 3;
      p.sink_edge = 1; p.wgt_kind = 0; p.w_lo = 0.01; p.w_hi = 1.0;
      break;
    case 2:
    case 3:  // 1.5M vertices, 300 levels, geometric out-degree mean 2.7
      p.seed = 2; p.n = 1500000; p.levels = 300; p.src_frac = 0.02; p.sink_frac = 0.02;
      p.deg_kind = 1; p.deg_mean = 2.7; p.deg_cap = 32; p.span_kind = 1; p.span_p1 = 0.9;
      p.span_hi = 8; p.sink_edge = 0; p.wgt_kind = 1; p.mu = -2.0; p.sigma = 0.7;
      if (which == 3) { p.dense_mean = 20.0; p.dense_seed = 3; p.dense_span_hi = 5; }
      break;
    case 4:
    case 5: {
      const uint64_t seed = (which == 4 ? 4000ULL : 5000ULL) + static_cast<uint64_t>(index);
      Rng r(seed);
      const double lo = which == 4 ? 5000.0 : 200000.0, hi = which == 4 ? 200000.0 : 3000000.0;
      const int n = static_cast<int>(std::llround(std::exp(std::log(lo) + r.u01() * (std::log(hi) - std::log(lo)))));
      const double dm = which == 4 ? 2.0 + 8.0 * r.u01() : 2.5 + 27.5 * r.u01();
      p.seed = seed * 7919ULL + 17ULL; p.n = n;
      p.levels = std::max(3, static_cast<int>(std::llround(std::sqrt(static_cast<double>(n)))));
      p.src_frac = 0.02; p.sink_frac = 0.02;
      p.deg_kind = 1; p.deg_mean = dm; p.deg_cap = std::max(32, static_cast<int>(4 * dm));
      p.span_kind = 1; p.span_p1 = 0.9; p.span_hi = 8; p.sink_edge = 0;
      p.wgt_kind = 1; p.mu = -2.0; p.sigma = 0.7;
      break;
    }
    default:
      throw std::runtime_error("unknown preset");
  }
  p.unit_wgts = unit_wgts ? 1 : 0;
  return p;
}

int64_t preset_k(int which, int index) {
  if (which == 1) return 10000;
  if (which == 2 || which == 3) return 1000000;
  if (which == 4) {
    Rng r(4000ULL + static_cast<uint64_t>(index));
    r.u01(); r.u01();  // n, mean degree
    return std::llround(std::exp(std::log(100.0) + r.u01() * (std::log(100000.0) - std::log(100.0))));
  }
  throw std::runtime_error("no per-graph K for this preset");
}

}  // namespace cpg
