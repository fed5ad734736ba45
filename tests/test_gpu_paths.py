"""GPU vertex reports: cpg_get_path_vertices and cpg_get_nodes.

Each reported path's vertices must be a walk of the fanout CSR from a
source to a chosen sink whose recurrence cost (root distance, then
((c + distf[w]) + w_e) - distf[u] at every step off the suffix chain) equals
the reported cost bit for bit, and within every cost class below the K-th
cost the multiset of vertex sequences must equal the oracle's (the boundary
class may hold a different choice among equal-cost paths).  Node parents
must point at earlier ranks.
"""
import os
import sys
from collections import Counter

import numpy as np
import pytest

import cpgamd
import pyoracle

pytestmark = pytest.mark.gpu
F32 = np.float32
INT_MAX = 2147483647


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def recurrence_costs(verts, adjp, adj, wgt, dist, succ, distf):
    """Every cost the vertex sequence can carry (parallel deviation edges
    give one cost each), in binary32 as the recurrence evaluates it."""
    costs = {F32(distf[verts[0]])}
    for a, b in zip(verts[:-1], verts[1:]):
        nb = adj[adjp[a]:adjp[a + 1]]
        ws = wgt[adjp[a]:adjp[a + 1]][nb == b]
        assert len(ws) > 0, f"{a}->{b} is not an edge"
        if b == succ[a]:
            continue
        assert dist[b] != INT_MAX
        costs = {F32(F32(F32(c + distf[b]) + F32(w)) - distf[a]) for c in costs for w in ws}
    return costs


def check_report(g, og, k, dist, succ, want_costs, want_paths, what):
    a = og.arrays()
    adjp, adj, wgt, iadjp = a["oadjp"], a["oadj"], a["owgt"], a["iadjp"]
    distf = pyoracle.distf(dist)
    got = g.get_slacks(k)
    assert len(got) == len(want_costs), what
    assert pyoracle.first_bit_mismatch(want_costs, got) == -1, what
    nodes = g.get_pfxt_nodes(k)
    paths = [g.path_vertices(i) for i in range(len(got))]
    for i, (c, p) in enumerate(zip(got, paths)):
        assert iadjp[p[0] + 1] == iadjp[p[0]], f"{what} path {i} does not start at a source"
        assert adjp[p[-1] + 1] == adjp[p[-1]] and dist[p[-1]] == 0, f"{what} path {i} does not end at a sink"
        assert F32(c) in recurrence_costs(p, adjp, adj, wgt, dist, succ, distf), \
            f"{what} path {i}: cost {c!r} is not its vertices' recurrence cost"
        nd = nodes[i]
        assert nd["level"] == 0 or nd["from_"] in p.tolist(), f"{what} node {i} deviation not on the path"
        assert nd["to"] in p.tolist()
        if nd["level"] == 0:
            assert nd["from_"] == -1 and p[0] == nd["to"] and nd["parent"] == -1
        else:
            pos = p.tolist().index(nd["from_"])
            assert p[pos + 1] == nd["to"], f"{what} node {i} deviation edge not taken"
        if nd["parent"] >= 0:
            assert nd["parent"] < i and nodes[nd["parent"]]["level"] == nd["level"] - 1
    kth = F32(got[-1])
    gk = Counter((int(F32(c).view(np.uint32)), tuple(p.tolist())) for c, p in zip(got, paths) if F32(c) < kth)
    wk = Counter((int(F32(c).view(np.uint32)), tuple(p.tolist())) for c, p in zip(want_costs, want_paths)
                 if F32(c) < kth)
    assert gk == wk, f"{what}: vertex sequences differ below the K-th cost"


@pytest.mark.parametrize("name", ["simple.txt", "dac21_paper_example.txt", "tv80_unit_wgts.txt"])
def test_reference_benchmarks_vertex_reports(golden_dir, name):
    path = os.path.join(golden_dir, name)
    say(f"case vertex reports {name}")
    og = pyoracle.OracleGraph.from_text(path)
    dist, succ, level, L = og.sfxt()
    g = cpgamd.CpGen()
    g.read_input(path)
    for k in (1, 10, 500):
        wc, wp, _ = og.topk_paths(dist, succ, k, max_len=L + 2)
        g.reset()
        g.report_paths(k)
        check_report(g, og, k, dist, succ, wc, wp, f"{name} K={k}")


def test_vertex_report_through_zero_cost_parent():
    """The only nonzero path 0 -> 3 -> 1 is a child of the zero-cost root
    0 -> 1, which is expanded but not reported."""
    u = np.array([0, 0, 3], np.int32)
    v = np.array([1, 3, 1], np.int32)
    w = np.array([0.0, 0.5, 0.5], np.float32)
    say("case zero-cost parent")
    g = cpgamd.CpGen()
    g.load_edges(4, u, v, w)
    g.report_paths(10, root_tie_quirk=False)
    assert g.get_slacks(10).tolist() == [1.0]
    assert g.path_vertices(0).tolist() == [0, 3, 1]
    nd = g.get_pfxt_nodes(10)[0]
    assert nd["parent"] == -1 and nd["level"] == 1


@pytest.mark.parametrize("which,k", [(1, 10000), (2, 10000)])
def test_preset_vertex_reports(which, k):
    n, u, v, w = cpgamd.gen_preset(which)
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, level, L = og.sfxt()
    say(f"case C{which} vertex reports K={k} (oracle)")
    wc, wp, st = og.topk_paths(dist, succ, k, max_len=L + 2)
    assert st["nonmonotone"] == 0
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    say(f"case C{which} vertex reports K={k} (GPU)")
    g.report_paths(k)
    check_report(g, og, k, dist, succ, wc, wp, f"C{which} K={k}")
