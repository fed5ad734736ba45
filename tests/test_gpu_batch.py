"""GPU parity of batched designs (cpg_batch_*): every query of a batch
equals the oracle's top-K of its design alone (bit-exact costs; sorted
multiset for the unit-weight benchmark)."""
import os

import numpy as np
import pytest

import cpgamd
import pyoracle

pytestmark = pytest.mark.gpu


def design_from_text(path):
    g = cpgamd.CpGen()
    g.read_input(path)
    adjp, adj, w = g.csr()
    n = g.num_verts()
    u = np.repeat(np.arange(n, dtype=np.int32), np.diff(adjp))
    return n, u, adj, w


def test_batch_of_uneven_designs_equals_oracle(golden_dir):
    designs = []
    for name, k in (("simple.txt", 5), ("dac21_paper_example.txt", 100),
                    ("tv80_unit_wgts.txt", 1000)):
        n, u, v, w = design_from_text(os.path.join(golden_dir, name))
        designs.append((name, n, u, v, w, k))
    n, u, v, w = cpgamd.gen_preset(1)
    designs.append(("C1", n, u, v, w, 10000))
    for i in range(10):  # the first designs of the 256-design batch workload
        n, u, v, w = cpgamd.gen_preset(4, i)
        designs.append((f"batch{i}", n, u, v, w, int(cpgamd.preset_k(4, i))))
    print(f"case batch of {len(designs)} designs", flush=True)
    b = cpgamd.Batch()
    for d in designs:
        b.add(d[1], d[2], d[3], d[4], d[5])
    b.run()
    print(f"batch timings {b.timings()}", flush=True)
    for q, (name, n, u, v, w, k) in enumerate(designs):
        og = pyoracle.OracleGraph.from_edges(n, u, v, w)
        dist, succ, level, L = og.sfxt()
        want, _ = og.topk(dist, succ, k)
        got = b.get_slacks(q, k)
        assert len(got) == len(want), f"{name}: {len(got)} vs {len(want)}"
        if name.startswith("tv80"):
            np.testing.assert_array_equal(np.sort(got), np.sort(want))
        else:
            assert pyoracle.first_bit_mismatch(want, got) == -1, name


def test_batch_result_equals_plain_query_and_rerun():
    """A design's result is the same alone (cpg_query) and in the batch,
    and a second run of the same batch gives the same results."""
    designs = [cpgamd.gen_preset(4, i) + (int(cpgamd.preset_k(4, i)),) for i in range(10, 16)]
    print("case batch vs plain queries", flush=True)
    b = cpgamd.Batch()
    for n, u, v, w, k in designs:
        b.add(n, u, v, w, k)
    b.run()
    first = [b.get_slacks(q, d[4]) for q, d in enumerate(designs)]
    b.run()
    for q, (n, u, v, w, k) in enumerate(designs):
        g = cpgamd.CpGen()
        g.load_edges(n, u, v, w)
        g.report_paths(k)
        plain = g.get_slacks(k)
        assert pyoracle.first_bit_mismatch(plain, first[q]) == -1 and len(plain) == len(first[q])
        again = b.get_slacks(q, k)
        assert pyoracle.first_bit_mismatch(first[q], again) == -1 and len(again) == len(first[q])


def test_full_256_design_batch_equals_oracle():
    """The whole 256-design batch workload (5K-200K vertices each, K
    100-100K) in one Batch: every query equals the oracle's top-K of its
    design alone, bit for bit."""
    designs = []
    for i in range(256):
        n, u, v, w = cpgamd.gen_preset(4, i)
        designs.append((n, u, v, w, int(cpgamd.preset_k(4, i))))
    print(f"case full batch of {len(designs)} designs", flush=True)
    b = cpgamd.Batch()
    try:
        for n, u, v, w, k in designs:
            b.add(n, u, v, w, k)
        b.run()
        print(f"batch timings {b.timings()}", flush=True)
        bad = []
        for q, (n, u, v, w, k) in enumerate(designs):
            og = pyoracle.OracleGraph.from_edges(n, u, v, w)
            dist, succ, _, _ = og.sfxt()
            want, _ = og.topk(dist, succ, k)
            got = b.get_slacks(q, k)
            if len(got) != len(want) or pyoracle.first_bit_mismatch(want, got) != -1:
                bad.append((q, len(got), len(want), pyoracle.first_bit_mismatch(want, got)))
        assert not bad, bad[:8]
    finally:
        b.close()


def test_batch_autotune_keeps_results_exact():
    """cpg_batch_autotune times whole batch runs over the round shapes and
    keeps the fastest; the results under the chosen shapes still equal
    the oracle's, and out-of-range shapes are refused."""
    designs = [cpgamd.gen_preset(4, i) + (int(cpgamd.preset_k(4, i)),) for i in range(16, 40)]
    print(f"case batch autotune over {len(designs)} designs", flush=True)
    b = cpgamd.Batch()
    try:
        for n, u, v, w, k in designs:
            b.add(n, u, v, w, k)
        t = b.autotune(max_trials=8)
        print(f"tuning {t}", flush=True)
        assert 1 <= t["trials"] <= 8 and t["best_ms"] > 0.0 and t["best_ms"] <= t["start_ms"] * 1.0001
        assert t["group"] in (2, 4, 8, 16) and t["grid_mult"] in (4, 8, 16)
        for q, (n, u, v, w, k) in enumerate(designs):
            og = pyoracle.OracleGraph.from_edges(n, u, v, w)
            dist, succ, _, _ = og.sfxt()
            want, _ = og.topk(dist, succ, k)
            got = b.get_slacks(q, k)
            assert len(got) == len(want) and pyoracle.first_bit_mismatch(want, got) == -1, q
        b.set_tuning(group=2, grid_mult=16, target_div=(2, 32, 4))
        b.run()
        for q, (n, u, v, w, k) in enumerate(designs[:6]):
            og = pyoracle.OracleGraph.from_edges(n, u, v, w)
            dist, succ, _, _ = og.sfxt()
            want, _ = og.topk(dist, succ, k)
            assert pyoracle.first_bit_mismatch(want, b.get_slacks(q, k)) == -1
        with pytest.raises(RuntimeError):
            b.set_tuning(group=0)
    finally:
        b.close()
