"""GPU parity of BVSS deviation discovery (wave64 bitmaps and int8 MFMA
tiles) through the C ABI.

Oracle: the set definition of gpucpg/tc_pfxt_bvss.cuh:253-274
(verify_adev_bvss_matches_csr): for every source u of a frontier, every
fanout u->v with v != succ[u], parallel edges once.  Known answers from
unittests/tc_pfxt_bvss.cu:9-72.  The bar is set equality of the sorted
(frontier, u, v) triples in every mode.
"""
import numpy as np
import pytest

import cpgamd

pytestmark = pytest.mark.gpu


def brute(adjp, adj, succ, frontiers):
    out = set()
    for f, fr in enumerate(frontiers):
        for u in set(int(x) for x in fr):
            for e in range(adjp[u], adjp[u + 1]):
                v = int(adj[e])
                if v != succ[u]:
                    out.add((f, u, v))
    out = sorted(out)
    return tuple(np.array([t[i] for t in out], np.int32) for i in range(3))


def unit_graph():
    """unittests/tc_pfxt_bvss.cu:9-27: the CSR, with weights that make the
    SFXT successors the test's succs {9, 8, 0, -1, 2, -1, ...}."""
    row_ptr = [0, 2, 3, 5, 5, 6, 6, 6, 6, 6, 6]
    col = [1, 9, 8, 0, 9, 2]
    eu = np.array([u for u in range(10) for _ in range(row_ptr[u], row_ptr[u + 1])], np.int32)
    ev = np.array(col, np.int32)
    ew = np.array([1.0, 0.5, 1.0, 0.1, 1.0, 1.0], np.float32)
    return eu, ev, ew


def test_reference_unit_answers():
    eu, ev, ew = unit_graph()
    print("case bvss unit graph", flush=True)
    g = cpgamd.CpGen()
    g.load_edges(10, eu, ev, ew)
    _, succ, _ = g.sfxt()
    assert succ.tolist() == [9, 8, 0, -1, 2, -1, -1, -1, -1, -1]
    g.bvss_build()
    info = g.bvss_info()
    # two (row, interval) slices with one source bit each; with 8-source
    # intervals the reference counts the same 2 slices and 2 bits
    assert info["entries"] == 2 and info["set_bits"] == 2
    assert g.bvss_row_sources(0).tolist() == []
    assert g.bvss_row_sources(1).tolist() == [0]
    assert g.bvss_row_sources(2).tolist() == []
    assert g.bvss_row_sources(9).tolist() == [2]
    for mode in (0, 1, 2):
        f, u, v = g.bvss_discover([np.array([0, 1, 2])], mode=mode)
        assert list(zip(u.tolist(), v.tolist())) == [(0, 1), (2, 9)]
        assert f.tolist() == [0, 0]


@pytest.mark.parametrize("nf,size", [(1, 50), (5, 2000), (16, 5000), (20, 300), (33, 64)])
def test_c1_frontier_batches_all_modes(nf, size):
    rng = np.random.default_rng(nf * 1000 + size)
    n, u, v, w = cpgamd.gen_preset(1)
    print(f"case bvss C1 nf={nf} size={size}", flush=True)
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    _, succ, _ = g.sfxt()
    adjp, adj, _ = g.csr()
    frontiers = [rng.integers(0, n, size=size) for _ in range(nf)]
    want = brute(adjp, adj, succ, frontiers)
    for mode in (0, 1, 2):
        got = g.bvss_discover(frontiers, mode=mode, dense_min=2)
        for a, b in zip(want, got):
            np.testing.assert_array_equal(a, b)
        if mode == 2:
            assert g.timings()["bvss_tiles"] > 0
        if mode == 1:
            assert g.timings()["bvss_tiles"] == 0


def test_dense_rows_mfma_counts_match_bitmaps():
    """Every source of a 64-source interval in every frontier: full 16x16
    tiles whose MFMA counts are checked against the decoded words (the
    kernel raises an error on any disagreement)."""
    rng = np.random.default_rng(3)
    n = 4096
    src = rng.integers(0, n // 2, size=60000)
    dst = n // 2 + rng.integers(0, n // 2, size=60000)
    w = rng.uniform(0.01, 1.0, size=60000).astype(np.float32)
    print("case bvss dense bipartite", flush=True)
    g = cpgamd.CpGen()
    g.load_edges(n, src.astype(np.int32), dst.astype(np.int32), w)
    _, succ, _ = g.sfxt()
    adjp, adj, _ = g.csr()
    frontiers = [np.arange(0, n // 2)] * 16
    want = brute(adjp, adj, succ, frontiers)
    got = g.bvss_discover(frontiers, mode=2)
    for a, b in zip(want, got):
        np.testing.assert_array_equal(a, b)
