"""GPU parity with the CSR reorder on (cpg_set_csr_reorder(1), the
reference's E_ORIENTED level-order relabelling, gpucpg/gpucpg.cu:2048-2079).

The relabelling is a layout change only: the SFXT arrays, the top-K costs
at every prefix and the vertex reports (in original ids) must equal the
oracle's exactly as with the reorder off, including endpoint subsets and
the unit-weight benchmark's ties (successors break ties on original ids)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import cpgamd
import pyoracle
from test_gpu_paths import check_report

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relabelled(n, u, v, w):
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    g.set_csr_reorder(1)
    return g


@pytest.mark.parametrize("name", ["simple.txt", "dac21_paper_example.txt", "tv80_unit_wgts.txt"])
def test_reference_benchmarks_relabelled(golden_dir, name):
    path = os.path.join(golden_dir, name)
    say(f"case reorder {name}")
    og = pyoracle.OracleGraph.from_text(path)
    dist, succ, level, L = og.sfxt()
    g = cpgamd.CpGen()
    g.read_input(path)
    g.set_csr_reorder(1)
    try:
        for k in (1, 10, 500):
            wc, wp, _ = og.topk_paths(dist, succ, k, max_len=L + 2)
            g.reset()
            g.report_paths(k)
            gd, gs, gl = g.sfxt()
            np.testing.assert_array_equal(gd, dist)
            np.testing.assert_array_equal(gs, succ)
            np.testing.assert_array_equal(gl, level)
            check_report(g, og, k, dist, succ, wc, wp, f"{name} K={k} reordered")
    finally:
        g.close()


def test_c1_relabelled_reports_and_masks():
    n, u, v, w = cpgamd.gen_preset(1)
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, level, L = og.sfxt()
    g = relabelled(n, u, v, w)
    try:
        say("case C1 reorder K=10000")
        wc, wp, _ = og.topk_paths(dist, succ, 10000, max_len=L + 2)
        g.report_paths(10000)
        assert g.reorder_ms() > 0.0
        gd, gs, gl = g.sfxt()
        np.testing.assert_array_equal(gd, dist)
        np.testing.assert_array_equal(gs, succ)
        np.testing.assert_array_equal(gl, level)
        check_report(g, og, 10000, dist, succ, wc, wp, "C1 K=10000 reordered")
        rng = np.random.default_rng(21)
        sinks = np.nonzero(np.bincount(u, minlength=n) == 0)[0]
        for frac in (0.01, 0.3):
            m = np.zeros(n, np.uint8)
            m[rng.choice(sinks, max(1, int(frac * len(sinks))), replace=False)] = 1
            say(f"case C1 reorder masked {frac}")
            d2, s2, _, _ = og.sfxt(m)
            want, _ = og.topk(d2, s2, 5000)
            g.reset()
            g.report_paths(5000, endpoint_mask=m)
            got = g.get_slacks(5000)
            assert len(got) == len(want) and pyoracle.first_bit_mismatch(want, got) == -1
        g.set_csr_reorder(0)  # back to the original layout: same answer
        g.reset()
        g.report_paths(10000)
        assert pyoracle.first_bit_mismatch(wc, g.get_slacks(10000)) == -1
    finally:
        g.close()


def test_c2_relabelled_k1m_equals_golden(golden_dir):
    """The 1.5M-vertex DAG, K = 1,000,000 on the relabelled graph: the cost
    list's SHA-256 equals the committed golden, and the prefixes equal the
    oracle's ranks."""
    with open(os.path.join(golden_dir, "synthetic_golden.json")) as f:
        gold = json.load(f)["c2"]
    n, u, v, w = cpgamd.gen_preset(2)
    g = relabelled(n, u, v, w)
    try:
        say("case C2 reorder K=1M")
        g.report_paths(gold["k"])
        say(f"C2 reorder {g.reorder_ms():.1f} ms, query {g.timings()['query_ms']:.3f} ms")
        got = g.get_slacks(gold["k"])
        assert len(got) == gold["count"]
        assert hashlib.sha256(got.astype(np.float32).tobytes()).hexdigest() == gold["costs_sha256"]
        for r, c in gold["ranks"].items():
            assert np.float32(got[int(r)]) == np.float32(c)
    finally:
        g.close()


def test_c3_unit_weight_relabelled_prefix_multisets():
    """Unit weights tie everywhere: the sorted cost multiset per prefix is
    unchanged by the relabelling."""
    n, u, v, w = cpgamd.gen_preset(3, 0, unit_wgts=True)
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, _, _ = og.sfxt()
    want, _ = og.topk(dist, succ, 100000)
    g = relabelled(n, u, v, w)
    try:
        say("case C3 unit weights reorder K=100000")
        g.report_paths(100000)
        got = g.get_slacks(100000)
        gd, gs, _ = g.sfxt()
        np.testing.assert_array_equal(gd, dist)
        np.testing.assert_array_equal(gs, succ)
        for p in (1000, 10000, 50000, 100000):
            np.testing.assert_array_equal(np.sort(got[:p]), np.sort(want[:p]))
    finally:
        g.close()
