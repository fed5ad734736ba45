"""Pins the oracle (CPU restatement) before it is trusted as the checker.

Every oracle function is checked against something it does not compute
itself: the reference's fixtures and unit-test known answers, closed forms,
and independent brute force written here in plain Python/numpy.
"""
import itertools
import os

import numpy as np
import pytest

import pyoracle as po

F32 = np.float32
INT_MAX = 2147483647


# ----------------------------------------------------------------- helpers
def scaled(w):
    """(int)(w * 10000.0f): binary32 multiply, truncation toward zero."""
    return int(np.trunc(F32(w) * F32(10000)))


def csr_from_edges(n, edges):
    """Per-vertex input order, vertex-major: the reference's CSR layout."""
    out = [[] for _ in range(n)]
    inn = [[] for _ in range(n)]
    for (u, v, w) in edges:
        out[u].append((v, F32(w)))
        inn[v].append((u, F32(w)))
    adjp = np.zeros(n + 1, np.int32)
    iadjp = np.zeros(n + 1, np.int32)
    for i in range(n):
        adjp[i + 1] = adjp[i] + len(out[i])
        iadjp[i + 1] = iadjp[i] + len(inn[i])
    adj = np.array([v for i in range(n) for v, _ in out[i]], np.int32)
    wgt = np.array([w for i in range(n) for _, w in out[i]], np.float32)
    iadj = np.array([u for i in range(n) for u, _ in inn[i]], np.int32)
    return adjp, adj, wgt, iadjp, iadj


def py_levels(n, adjp, adj):
    """Closed form of the peel: longest hop count to a sink."""
    memo = {}

    def lvl(v):
        if v in memo:
            return memo[v]
        nb = adj[adjp[v]:adjp[v + 1]]
        r = 0 if len(nb) == 0 else 1 + max(lvl(int(w)) for w in nb)
        memo[v] = r
        return r
    return np.array([lvl(v) for v in range(n)], np.int32)


def py_dist(n, adjp, adj, wgt, sink_mask=None):
    """Shortest int-scaled distance to any sink (Dijkstra on the reverse);
    with sink_mask, to any chosen sink."""
    import heapq
    radj = [[] for _ in range(n)]
    for u in range(n):
        for e in range(adjp[u], adjp[u + 1]):
            radj[int(adj[e])].append((u, scaled(wgt[e])))
    dist = [INT_MAX] * n
    pq = []
    for v in range(n):
        if adjp[v + 1] == adjp[v] and (sink_mask is None or sink_mask[v]):
            dist[v] = 0
            pq.append((0, v))
    heapq.heapify(pq)
    while pq:
        d, v = heapq.heappop(pq)
        if d != dist[v]:
            continue
        for u, w in radj[v]:
            if d + w < dist[u]:
                dist[u] = d + w
                heapq.heappush(pq, (d + w, u))
    return np.array(dist, np.int64)


def py_succ(n, adjp, adj, wgt, dist):
    s = np.full(n, -1, np.int32)
    for u in range(n):
        for e in range(adjp[u], adjp[u + 1]):
            w = int(adj[e])
            if dist[w] == INT_MAX:
                continue
            if dist[u] == dist[w] + scaled(wgt[e]) and w > s[u]:
                s[u] = w
    return s


def py_all_paths(n, adjp, adj, wgt, iadjp, dist, succ, drop_zero=True):
    """Every source->sink vertex path by DFS; cost by the deviation
    recurrence ((c + distf[w]) + w_e) - distf[u] at every step off the
    suffix chain; zero-cost paths dropped; ascending."""
    distf = (np.asarray(dist, np.int64).astype(np.int32).astype(F32) / F32(10000)).astype(F32)
    out = []

    def dfs(u, c):
        s = int(succ[u])
        if s == -1:
            out.append(c)
            return
        dfs(s, c)
        for e in range(adjp[u], adjp[u + 1]):
            w = int(adj[e])
            if w == s or dist[w] == INT_MAX:
                continue
            dfs(w, F32(F32(F32(c + distf[w]) + wgt[e]) - distf[u]))
    for v in range(n):
        if iadjp[v + 1] == iadjp[v] and dist[v] != INT_MAX:
            dfs(v, distf[v])
    arr = np.sort(np.array(out, np.float32))
    return arr[arr != 0.0] if drop_zero else arr


def random_dag(seed, n, m, wset=None, parallel=True):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    edges = []
    for _ in range(m):
        a, b = sorted(rng.choice(n, 2, replace=False))
        w = rng.choice(wset) if wset is not None else rng.uniform(0.01, 1.0)
        edges.append((int(perm[a]), int(perm[b]), float(F32(w))))
        if parallel and rng.random() < 0.05:
            edges.append((int(perm[a]), int(perm[b]), float(F32(rng.uniform(0.01, 1.0)))))
    order = rng.permutation(len(edges))
    return [edges[i] for i in order]


def og_from(n, edges):
    u = np.array([e[0] for e in edges], np.int32)
    v = np.array([e[1] for e in edges], np.int32)
    w = np.array([e[2] for e in edges], np.float32)
    return po.OracleGraph.from_edges(n, u, v, w)


def fixture(golden_dir, name):
    return os.path.join(golden_dir, name)


# ------------------------------------------------------------ a1 parsing
def test_parse_simple_fixture(golden_dir):
    g = po.OracleGraph.from_text(fixture(golden_dir, "simple.txt"))
    a = g.arrays()
    # hand-read from benchmarks/simple.txt (17 vertices, 17 edges)
    edges = [(0, 1, 1.1), (0, 15, 2.3), (1, 2, 1), (1, 13, 1), (2, 3, 1), (3, 4, 11.2), (3, 6, 1),
             (4, 5, 0.45), (6, 7, 1), (8, 9, 0.22), (9, 10, 1), (10, 11, 3.45), (11, 12, 1),
             (12, 14, 1), (13, 14, 1), (14, 15, 8.2), (15, 16, 1)]
    adjp, adj, wgt, iadjp, iadj = csr_from_edges(17, edges)
    np.testing.assert_array_equal(a["oadjp"], adjp)
    np.testing.assert_array_equal(a["oadj"], adj)
    np.testing.assert_array_equal(a["owgt"], wgt)
    np.testing.assert_array_equal(a["iadjp"], iadjp)
    np.testing.assert_array_equal(a["iadj"], iadj)


def test_parse_dac21_and_tv80(golden_dir):
    g = po.OracleGraph.from_text(fixture(golden_dir, "dac21_paper_example.txt"))
    assert (g.n, g.m) == (12, 14)
    a = g.arrays()
    assert a["owgt"].tolist() == [4.0, 1.0, 2.0, 4.0, 3.0, 3.0, 4.0, 5.0, 6.0, 2.0, 3.0, 1.0, 2.0, 1.0]
    t = po.OracleGraph.from_text(fixture(golden_dir, "tv80_unit_wgts.txt"))
    assert (t.n, t.m) == (17038, 23087)
    assert np.all(t.arrays()["owgt"] == 1.0)
    # ignore_wgts forces 1.0f (gpucpg/gpucpg.cu:495)
    s = po.OracleGraph.from_text(fixture(golden_dir, "simple.txt"), ignore_wgts=True)
    assert np.all(s.arrays()["owgt"] == 1.0)


def test_parse_inline_inprocess_fixture(tmp_path):
    # unittests/tc_pfxt_inprocess.cu:56-67
    p = tmp_path / "inline.gpath.txt"
    p.write_text("5\n" + "".join('"%d"\n' % i for i in range(5)) +
                 '"0" -> "1", 0.10;\n"0" -> "2", 0.20;\n"1" -> "3", 0.10;\n'
                 '"2" -> "3", 0.05;\n"3" -> "4", 0.10;\n')
    g = po.OracleGraph.from_text(str(p))
    a = g.arrays()
    assert a["oadjp"].tolist() == [0, 2, 3, 4, 5, 5]
    assert a["oadj"].tolist() == [1, 2, 3, 3, 4]
    assert a["owgt"].tolist() == [F32(0.1), F32(0.2), F32(0.1), F32(0.05), F32(0.1)]
    dist, succ, level, L = g.sfxt()
    assert dist.tolist() == [3000, 2000, 1500, 1000, 0]
    assert succ.tolist() == [1, 3, 3, 4, -1]
    costs, _ = g.topk(dist, succ, 2)
    d0, d2 = F32(3000) / F32(10000), F32(1500) / F32(10000)
    assert costs.tolist() == [d0, F32(F32(F32(d0 + d2) + F32(0.2)) - d0)]


def test_parse_rejects_malformed(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text('2\n"0"\n"1"\n"0" -> 1;\n')
    with pytest.raises(RuntimeError):
        po.OracleGraph.from_text(str(p))
    p.write_text('2\n"0"\n"1"\n"0" -> "5";\n')
    with pytest.raises(RuntimeError):
        po.OracleGraph.from_text(str(p))


# ------------------------------------------------------------ a3-a6 SFXT
def test_levelize_known_answer_8_vertex():
    # unittests/tc_sfxt_bvss.cu:153-172 (levels {3,3,2,1,2,1,0,0})
    edges = [(0, 2, 1), (0, 4, 1), (1, 2, 1), (2, 3, 1), (3, 6, 1), (4, 5, 1), (5, 6, 1)]
    g = og_from(8, edges)
    level, queue, lvlp, L = g.levelize()
    assert level.tolist() == [3, 3, 2, 1, 2, 1, 0, 0]
    assert L == 4
    assert lvlp.tolist() == [0, 2, 4, 6, 8]


@pytest.mark.parametrize("seed", range(8))
def test_sfxt_matches_closed_forms(seed):
    n = 40
    edges = random_dag(seed, n, 90)
    adjp, adj, wgt, iadjp, iadj = csr_from_edges(n, edges)
    g = og_from(n, edges)
    dist, succ, level, L = g.sfxt()
    np.testing.assert_array_equal(level, py_levels(n, adjp, adj))
    assert L == int(py_levels(n, adjp, adj).max()) + 1
    pd = py_dist(n, adjp, adj, wgt)
    np.testing.assert_array_equal(dist.astype(np.int64), pd)
    np.testing.assert_array_equal(succ, py_succ(n, adjp, adj, wgt, pd))
    np.testing.assert_array_equal(po.distf(dist), (dist.astype(F32) / F32(10000)).astype(F32))


def random_sink_mask(rng, n, adjp, frac):
    sinks = [v for v in range(n) if adjp[v + 1] == adjp[v]]
    pick = rng.choice(sinks, size=max(1, int(round(frac * len(sinks)))), replace=False)
    m = np.zeros(n, np.uint8)
    m[pick] = 1
    return m


@pytest.mark.parametrize("seed", range(8))
def test_masked_sfxt_matches_closed_forms(seed):
    """Endpoint subset: Dijkstra from the chosen sinks only; the successor
    rule over reachable fanouts; a mask of every sink equals no mask."""
    n = 40
    rng = np.random.default_rng(500 + seed)
    edges = random_dag(seed, n, 90)
    adjp, adj, wgt, iadjp, iadj = csr_from_edges(n, edges)
    g = og_from(n, edges)
    full = g.sfxt()
    allm = np.array([1 if adjp[v + 1] == adjp[v] else 0 for v in range(n)], np.uint8)
    for a, b in zip(full, g.sfxt(allm)):
        np.testing.assert_array_equal(a, b)
    mask = random_sink_mask(rng, n, adjp, [0.1, 0.3, 0.6][seed % 3])
    dist, succ, level, L = g.sfxt(mask)
    np.testing.assert_array_equal(level, full[2])
    pd = py_dist(n, adjp, adj, wgt, mask)
    np.testing.assert_array_equal(dist.astype(np.int64), pd)
    np.testing.assert_array_equal(succ, py_succ(n, adjp, adj, wgt, pd))
    assert (dist == INT_MAX).any()


@pytest.mark.parametrize("seed", range(6))
def test_masked_topk_paths_end_at_chosen_sinks(seed):
    """The masked top-K equals the brute force over the masked SFXT, and
    its costs are the weight sums of exactly the source -> chosen-sink
    paths (equal up to the 1e-4 truncation of the int-scaled distances
    and binary32 rounding)."""
    n = 14
    rng = np.random.default_rng(900 + seed)
    seen, edges = set(), []
    for e in random_dag(300 + seed, n, 32, parallel=False):  # no parallel edges
        if (e[0], e[1]) not in seen:
            seen.add((e[0], e[1]))
            edges.append(e)
    adjp, adj, wgt, iadjp, iadj = csr_from_edges(n, edges)
    g = og_from(n, edges)
    mask = random_sink_mask(rng, n, adjp, 0.5)
    dist, succ, _, _ = g.sfxt(mask)
    allp = py_all_paths(n, adjp, adj, wgt, iadjp, dist.astype(np.int64), succ)
    got, _ = g.topk(dist, succ, len(allp) + 3, root_tie_quirk=False)
    np.testing.assert_array_equal(got, allp)
    np.testing.assert_array_equal(g.all_paths(dist, succ), allp)
    sums = []

    def walk(u, acc, hops):
        if adjp[u + 1] == adjp[u]:
            if mask[u]:
                sums.append((acc, hops))
            return
        for e in range(adjp[u], adjp[u + 1]):
            walk(int(adj[e]), acc + float(wgt[e]), hops + 1)
    for v in range(n):
        if iadjp[v + 1] == iadjp[v]:
            walk(v, 0.0, 0)
    sums.sort()
    assert len(sums) == len(allp)
    tol = np.array([2e-4 * h + 1e-5 for _, h in sums])
    np.testing.assert_array_less(np.abs(np.array([x for x, _ in sums]) - allp), tol + 1e-4)


def test_sfxt_fixtures_closed_forms(golden_dir):
    for name in ("simple.txt", "dac21_paper_example.txt", "tv80_unit_wgts.txt"):
        g = po.OracleGraph.from_text(fixture(golden_dir, name))
        a = g.arrays()
        dist, succ, level, L = g.sfxt()
        if g.n < 1000:
            np.testing.assert_array_equal(level, py_levels(g.n, a["oadjp"], a["oadj"]))
        pd = py_dist(g.n, a["oadjp"], a["oadj"], a["owgt"])
        np.testing.assert_array_equal(dist.astype(np.int64), pd)
        np.testing.assert_array_equal(succ, py_succ(g.n, a["oadjp"], a["oadj"], a["owgt"], pd))


# ------------------------------------------------ a10/a17 known answers
def test_candidate_cost_known_answer():
    # unittests/tc_pfxt_candidates.cu:85-90
    assert po.candidate_cost(0.5, 25000, 10000, 0.25) == F32(-0.75)
    # order: ((parent + dst) + w) - src
    p, s, d, w = F32(0.1), 30001, 7, F32(1e-3)
    want = F32(F32(F32(p + F32(d) / F32(10000)) + w) - F32(s) / F32(10000))
    assert po.candidate_cost(float(p), s, d, float(w)) == want


def test_classify_known_answers():
    # unittests/tc_pfxt_candidates.cu:69-83
    assert po.classify(1.0, 1.0, 2.0, True, False) == 0
    assert po.classify(1.5, 1.0, 2.0, True, False) == 1
    assert po.classify(2.0, 1.0, 2.0, True, False) == 1
    assert po.classify(2.1, 1.0, 2.0, True, False) == 2
    assert po.classify(1.5, 1.0, 2.0, True, True) == 2


def test_deviation_csr_known_answer():
    # unittests/tc_pfxt_candidates.cu:386-408
    offs, dsts, deltas = po.deviations([0, 4, 6, 7, 7], [1, 2, 2, 3, 2, 0, 1],
                                       [0.1, 0.2, 0.25, 0.4, 0.5, 0.6, 0.7], [1, -1, 1, -1],
                                       [1000, 2000, INT_MAX, 3000])
    assert offs.tolist() == [0, 1, 2, 2, 2]
    assert dsts.tolist() == [3, 0]
    np.testing.assert_allclose(deltas, [0.6, 0.5], rtol=1e-6)


def test_split_inc_amount_closed_form():
    # gpucpg/gpucpg.cuh:292-299: 0.1 + 99.9 / (1 + exp(0.8 (d - 3)))
    assert abs(po.split_inc_amount(3.0) - 50.05) < 1e-4
    for d in (1.0, 2.7, 3.5, 20.0):
        closed = 0.1 + 99.9 / (1 + np.exp(0.8 * (d - 3)))
        assert abs(po.split_inc_amount(d) - closed) < 1e-5 * max(1.0, closed)


def test_compare_prefix_known_answers():
    # unittests/tc_pfxt_inprocess.cu:26-51
    r = po.compare_prefix([1.0, 2.0, 3.0, 4.0], [1.0, 2.000001, 3.0], 3)
    assert r["baseline_count"] == 4 and r["tc_count"] == 3 and r["compared"] == 3
    assert r["passed"] and r["max_diff_rank"] == 2 and r["first_mismatch_rank"] == 0
    r = po.compare_prefix([1.0, 2.0, 3.0], [1.0, 2.2, 3.0], 3, 1e-3)
    assert not r["passed"] and r["max_diff_rank"] == 2 and r["first_mismatch_rank"] == 2


# ------------------------------------------------------------ a7-a15 PFXT
def brute(g, n, edges, drop_zero=True):
    adjp, adj, wgt, iadjp, iadj = csr_from_edges(n, edges)
    dist, succ, level, L = g.sfxt()
    return py_all_paths(n, adjp, adj, wgt, iadjp, dist.astype(np.int64), succ, drop_zero), dist, succ


@pytest.mark.parametrize("seed", range(12))
def test_topk_equals_brute_force(seed):
    n = 14
    wset = [0.1, 0.2, 0.25, 0.5, 1.0] if seed % 2 else None
    edges = random_dag(100 + seed, n, 30, wset=wset)
    g = og_from(n, edges)
    allp, dist, succ = brute(g, n, edges)
    allz, _, _ = brute(g, n, edges, drop_zero=False)
    assert len(allp) > 0
    np.testing.assert_array_equal(g.all_paths(dist, succ), allp)
    for k in sorted({1, 3, 10, len(allp), len(allp) + 5}):
        # GPG semantics: the k smallest nonzero costs (gpucpg.cu:21338-21358)
        got, st = g.topk(dist, succ, k, root_tie_quirk=False)
        np.testing.assert_array_equal(got, allp[:k])
        # SEQUENTIAL pops k paths and erases zero-cost ones afterwards
        seq, _ = g.sequential(dist, succ, k)
        want = allz[:k]
        np.testing.assert_array_equal(seq, want[want != 0.0])


def test_topk_fixture_brute_force(golden_dir):
    for name in ("simple.txt", "dac21_paper_example.txt"):
        g = po.OracleGraph.from_text(fixture(golden_dir, name))
        a = g.arrays()
        dist, succ, level, L = g.sfxt()
        allp = py_all_paths(g.n, a["oadjp"], a["oadj"], a["owgt"], a["iadjp"],
                            dist.astype(np.int64), succ)
        got, _ = g.topk(dist, succ, 100)
        np.testing.assert_array_equal(got, allp[:100])
    # dac21: 13 paths, smallest 7 = 1+4+2 (0->4->7->10)
    g = po.OracleGraph.from_text(fixture(golden_dir, "dac21_paper_example.txt"))
    dist, succ, _, _ = g.sfxt()
    got, _ = g.topk(dist, succ, 100)
    assert len(got) == 13 and got[0] == 7.0


@pytest.mark.parametrize("seed", range(6))
def test_short_long_schedule_equals_topk(seed):
    """Faithful SHORT_LONG transliteration == exact best-first, incl. a tiny
    long-pile limit that forces the final-window prefetch branch."""
    rng = np.random.default_rng(seed)
    n = 400
    edges = random_dag(1000 + seed, n, 1600)
    g = og_from(n, edges)
    dist, succ, _, _ = g.sfxt()
    used_final_window = 0
    for k in (10, 200, 1000):
        want, st = g.topk(dist, succ, k)
        assert st["nonmonotone"] == 0
        for limit in (2e9, 24.0 * 1500, 24.0 * 1000, 24.0 * 500, 24.0 * 200):
            try:
                got, sst = g.short_long(dist, succ, k, fixed_inc=float(rng.uniform(0.05, 0.5)),
                                        long_pile_limit_bytes=limit)
            except po.ShortLongAbort as e:
                # the reference would index past its long pile here
                assert e.stats["aborted"] == 2 and limit < 2e9
                continue
            used_final_window += sst["final_window"] > 0
            np.testing.assert_array_equal(got, want)
    assert used_final_window > 0


def test_short_long_root_tie_quirk_tv80(golden_dir):
    """tv80 (unit weights) has 19 roots tied at the initial split beyond
    (int)(0.005 n + 1): the schedule loses them, and so does the quirk mode
    of the exact oracle."""
    g = po.OracleGraph.from_text(fixture(golden_dir, "tv80_unit_wgts.txt"))
    dist, succ, _, _ = g.sfxt()
    for k in (100, 1000, 3000):
        want, st = g.topk(dist, succ, k, root_tie_quirk=True)
        assert st["roots_dropped"] == 19
        got, sst = g.short_long(dist, succ, k)
        assert sst["roots_dropped"] == 19
        np.testing.assert_array_equal(got, want)
        nq, _ = g.topk(dist, succ, k, root_tie_quirk=False)
        assert len(nq) == k


def test_zero_cost_paths_are_erased():
    # 0 -> 1 with weight 0: a zero-cost root path; 2 isolated (source and sink)
    edges = [(0, 1, 0.0), (0, 3, 0.5), (3, 1, 0.5)]
    g = og_from(4, edges)
    dist, succ, _, _ = g.sfxt()
    got, st = g.topk(dist, succ, 10, root_tie_quirk=False)
    assert got.tolist() == [1.0]
    assert st["zero_cost"] == 2
    seq, _ = g.sequential(dist, succ, 10)
    assert seq.tolist() == [1.0]


# ------------------------------------------- vertex sequences (f1 row 1)
def py_edge_paths(n, adjp, adj, wgt, iadjp, dist, succ):
    """Every source->sink walk of the fanout CSR, edge by edge (parallel
    edges are distinct walks), reduced to the prefix-tree's path identity:
    a step onto succ[u] is on the suffix chain whatever edge it uses, a step
    to any other w (finite dist) is a deviation through that edge.  Returns
    {(cost bits, vertex tuple, deviation edge tuple)}: each entry one path
    of the reference's enumeration, with its recurrence cost."""
    distf = (np.asarray(dist, np.int64).astype(np.int32).astype(F32) / F32(10000)).astype(F32)
    found = set()

    def walk(u, c, verts, devs):
        if adjp[u + 1] == adjp[u]:
            if dist[u] != INT_MAX:
                found.add((int(np.float32(c).view(np.uint32)), tuple(verts), tuple(devs)))
            return
        for e in range(adjp[u], adjp[u + 1]):
            w = int(adj[e])
            if dist[w] == INT_MAX:
                continue
            if w == succ[u]:
                walk(w, c, verts + [w], devs)
            else:
                walk(w, F32(F32(F32(c + distf[w]) + wgt[e]) - distf[u]), verts + [w], devs + [e])
    for v in range(n):
        if iadjp[v + 1] == iadjp[v] and dist[v] != INT_MAX:
            walk(v, distf[v], [v], [])
    return found


def class_multiset(costs, paths):
    from collections import Counter
    return Counter((int(np.float32(c).view(np.uint32)), tuple(int(x) for x in p))
                   for c, p in zip(costs, paths))


@pytest.mark.parametrize("seed", range(12))
def test_topk_paths_equal_edge_walk_brute_force(seed):
    from collections import Counter
    n = 14
    wset = [0.1, 0.2, 0.25, 0.5, 1.0] if seed % 2 else None
    edges = random_dag(300 + seed, n, 30, wset=wset)
    g = og_from(n, edges)
    adjp, adj, wgt, iadjp, iadj = csr_from_edges(n, edges)
    dist, succ, _, _ = g.sfxt()
    brute_set = py_edge_paths(n, adjp, adj, wgt, iadjp, dist.astype(np.int64), succ)
    nonzero = [(cb, vs) for cb, vs, _ in brute_set if cb not in (0, 0x80000000)]
    want_all = Counter(nonzero)
    assert len(nonzero) > 0
    costs, paths, st = g.topk_paths(dist, succ, len(nonzero), root_tie_quirk=False)
    # every nonzero path, each with its recurrence cost and vertices
    assert class_multiset(costs, paths) == want_all
    ref_costs, _ = g.topk(dist, succ, len(nonzero), root_tie_quirk=False)
    np.testing.assert_array_equal(costs, ref_costs)
    for k in (1, 3, 7):
        if k > len(nonzero):
            continue
        ck, pk, _ = g.topk_paths(dist, succ, k, root_tie_quirk=False)
        np.testing.assert_array_equal(ck, ref_costs[:k])
        got = class_multiset(ck, pk)
        for key, cnt in got.items():
            assert want_all[key] >= cnt
        # cost classes strictly below the k-th cost are complete
        kth = int(np.float32(ck[-1]).view(np.uint32))
        below = Counter({key: c for key, c in want_all.items()
                         if np.uint32(key[0]).view(np.float32) < ck[-1]})
        assert Counter({key: c for key, c in got.items() if key[0] != kth}) == below


def test_topk_paths_fixtures_brute_force(golden_dir):
    for name in ("simple.txt", "dac21_paper_example.txt"):
        g = po.OracleGraph.from_text(fixture(golden_dir, name))
        a = g.arrays()
        dist, succ, _, _ = g.sfxt()
        brute_set = py_edge_paths(g.n, a["oadjp"], a["oadj"], a["owgt"], a["iadjp"],
                                  dist.astype(np.int64), succ)
        from collections import Counter
        want = Counter((cb, vs) for cb, vs, _ in brute_set if cb not in (0, 0x80000000))
        costs, paths, _ = g.topk_paths(dist, succ, len(want) + 3)
        assert class_multiset(costs, paths) == want
    # dac21: the cheapest path is 0 -> 4 -> 7 -> 10 (cost 1 + 4 + 2)
    g = po.OracleGraph.from_text(fixture(golden_dir, "dac21_paper_example.txt"))
    dist, succ, _, _ = g.sfxt()
    costs, paths, _ = g.topk_paths(dist, succ, 1)
    assert costs[0] == 7.0 and paths[0].tolist() == [0, 4, 7, 10]


def test_topk_paths_through_zero_cost_parent():
    # the only nonzero path 0 -> 3 -> 1 is a child of the zero-cost root 0 -> 1
    edges = [(0, 1, 0.0), (0, 3, 0.5), (3, 1, 0.5)]
    g = og_from(4, edges)
    dist, succ, _, _ = g.sfxt()
    costs, paths, st = g.topk_paths(dist, succ, 10, root_tie_quirk=False)
    assert costs.tolist() == [1.0]
    assert [p.tolist() for p in paths] == [[0, 3, 1]]
    assert st["zero_cost"] == 2


@pytest.mark.parametrize("seed", range(6))
def test_all_paths_vertices_brute_force(seed):
    """The C brute force with vertices: the same (cost, vertices) multiset
    as the edge-walk enumeration, and as many paths as the path-count DP
    P(v) = 1 at a chain end, else P(succ v) + sum of P(w) over deviations."""
    from collections import Counter
    n = 13
    edges = random_dag(500 + seed, n, 28)
    g = og_from(n, edges)
    adjp, adj, wgt, iadjp, iadj = csr_from_edges(n, edges)
    dist, succ, _, _ = g.sfxt()
    costs, paths = g.all_paths_vertices(dist, succ)
    brute_set = py_edge_paths(n, adjp, adj, wgt, iadjp, dist.astype(np.int64), succ)
    assert class_multiset(costs, paths) == Counter((cb, vs) for cb, vs, _ in brute_set)
    memo = {}

    def count(v):
        if v not in memo:
            s = int(succ[v])
            if s == -1:
                memo[v] = 1
            else:
                memo[v] = count(s) + sum(count(int(adj[e])) for e in range(adjp[v], adjp[v + 1])
                                         if int(adj[e]) != s and dist[int(adj[e])] != INT_MAX)
        return memo[v]
    total = sum(count(v) for v in range(n) if iadjp[v + 1] == iadjp[v] and dist[v] != INT_MAX)
    assert len(costs) == total
