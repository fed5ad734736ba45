"""GPU parity: the HIP engine against the oracle, through the C ABI.

Bar: SFXT dist/succ/level equal element-for-element; path costs equal
bit-for-bit at every prefix (sorted multiset for unit weights, where
equal-cost ties make path identity ambiguous).
"""
import os
import sys

import numpy as np
import pytest

import cpgamd
import pyoracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


_CACHE = {}


def preset_graph(which, unit=False):
    key = (which, unit)
    if key not in _CACHE:
        n, u, v, w = cpgamd.gen_preset(which, 0, unit)
        og = pyoracle.OracleGraph.from_edges(n, u, v, w)
        dist, succ, level, L = og.sfxt()
        _CACHE[key] = (n, u, v, w, og, dist, succ, level, L)
    return _CACHE[key]


def engine_for(n, u, v, w):
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    return g


def assert_bits_equal(want, got, what):
    assert len(got) == len(want), f"{what}: {len(got)} paths, oracle {len(want)}"
    mm = pyoracle.first_bit_mismatch(want, got)
    assert mm == -1, f"{what}: first mismatch at rank {mm}: {got[mm]!r} vs {want[mm]!r}"


@pytest.mark.parametrize("name", ["simple.txt", "dac21_paper_example.txt", "tv80_unit_wgts.txt"])
def test_reference_benchmarks_sfxt_and_topk(golden_dir, name):
    path = os.path.join(golden_dir, name)
    say(f"case benchmark {name}")
    og = pyoracle.OracleGraph.from_text(path)
    dist, succ, level, L = og.sfxt()
    g = cpgamd.CpGen()
    g.read_input(path)
    gd, gs, gl = g.sfxt()
    np.testing.assert_array_equal(gd, dist)
    np.testing.assert_array_equal(gs, succ)
    np.testing.assert_array_equal(gl, level)
    assert g.graph_diameter == L
    for k in (1, 2, 5, 10, 100, 1000, 5000):
        want, _ = og.topk(dist, succ, k)
        g.reset()
        g.report_paths(k)
        got = g.get_slacks(k)
        if name.startswith("tv80"):
            np.testing.assert_array_equal(np.sort(got), np.sort(want))
        else:
            assert_bits_equal(want, got, f"{name} K={k}")


def test_c1_sfxt_parity():
    n, u, v, w, og, dist, succ, level, L = preset_graph(1)
    say("case C1 sfxt")
    g = engine_for(n, u, v, w)
    gd, gs, gl = g.sfxt()
    np.testing.assert_array_equal(gd, dist)
    np.testing.assert_array_equal(gs, succ)
    np.testing.assert_array_equal(gl, level)
    assert g.graph_diameter == L


@pytest.mark.parametrize("k", [1, 100, 10000, 50000])
def test_c1_topk_bit_exact(k):
    n, u, v, w, og, dist, succ, level, L = preset_graph(1)
    say(f"case C1 topk K={k}")
    want, st = og.topk(dist, succ, k)
    assert st["nonmonotone"] == 0
    g = engine_for(n, u, v, w)
    g.report_paths(k)
    got = g.get_slacks(k)
    assert_bits_equal(want, got, f"C1 K={k}")


def test_c2_sfxt_parity():
    n, u, v, w, og, dist, succ, level, L = preset_graph(2)
    say("case C2 sfxt")
    g = engine_for(n, u, v, w)
    gd, gs, gl = g.sfxt()
    np.testing.assert_array_equal(gd, dist)
    np.testing.assert_array_equal(gs, succ)
    np.testing.assert_array_equal(gl, level)
    assert g.graph_diameter == L


def test_c2_topk_prefixes_bit_exact():
    """K=1M on the 1.5M-vertex DAG; every prefix 1K/10K/50K/100K/1M."""
    n, u, v, w, og, dist, succ, level, L = preset_graph(2)
    say("case C2 oracle K=1M")
    want, st = og.topk(dist, succ, 1_000_000)
    assert st["nonmonotone"] == 0
    say("case C2 GPU K=1M")
    g = engine_for(n, u, v, w)
    g.report_paths(1_000_000)
    got = g.get_slacks(1_000_000)
    say(f"C2 timings {g.timings()}")
    for kp in (1000, 10_000, 50_000, 100_000, 1_000_000):
        assert_bits_equal(want[:kp], got[:kp], f"C2 prefix {kp}")


def wide_fanout_graph():
    """Two wide vertices: one with 3000 fanouts (longer than the 2048-record
    LDS sort, so its deviation list is sorted through the second buffer) and
    one with 40 (sorted in LDS), plus tied weights for the stable order."""
    rng = np.random.default_rng(7)
    f1, f2, ns, mid0 = 3000, 40, 3, 2
    n = mid0 + f1 + ns
    u, v, w = [], [], []
    for i in range(f1):
        u.append(0); v.append(mid0 + i); w.append(rng.uniform(0.01, 1))
    for i in range(f2):
        u.append(1); v.append(mid0 + i * 7); w.append(rng.uniform(0.01, 1))
    for i in range(f1):
        for s in range(ns):
            if rng.random() < 0.7 or s == 0:
                u.append(mid0 + i); v.append(mid0 + f1 + s); w.append(rng.uniform(0.01, 1))
    w[5] = w[6] = w[7]
    return (n, np.array(u, np.int32), np.array(v, np.int32), np.array(w, np.float32))


def test_wide_fanout_deviation_sort():
    n, u, v, w = wide_fanout_graph()
    say("case wide fanout 3000/40")
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, level, L = og.sfxt()
    g = engine_for(n, u, v, w)
    gd, gs, gl = g.sfxt()
    np.testing.assert_array_equal(gd, dist)
    np.testing.assert_array_equal(gs, succ)
    np.testing.assert_array_equal(gl, level)
    for k in (1, 10, 1000, 5000):
        want, _ = og.topk(dist, succ, k)
        g.reset()
        g.report_paths(k)
        assert_bits_equal(want, g.get_slacks(k), f"wide fanout K={k}")


@pytest.mark.parametrize("blocks", [1, 16, 256])
def test_c1_level_loop_grid_does_not_change_results(blocks):
    n, u, v, w, og, dist, succ, level, L = preset_graph(1)
    say(f"case C1 level-loop grid {blocks}")
    g = engine_for(n, u, v, w)
    g.report_paths(10000, sfxt_blocks=blocks)
    assert g.timings()["sfxt_blocks"] <= blocks
    gd, gs, gl = g.sfxt()
    np.testing.assert_array_equal(gd, dist)
    np.testing.assert_array_equal(gs, succ)
    np.testing.assert_array_equal(gl, level)
    want, _ = og.topk(dist, succ, 10000)
    assert_bits_equal(want, g.get_slacks(10000), f"C1 grid {blocks}")


@pytest.mark.parametrize("k,max_nodes", [(10000, 12000), (50000, 60000)])
def test_c1_node_store_overflow_restarts_exactly(k, max_nodes):
    """A node store too small for the single-walk rounds overflows; the
    query restarts with a larger store and a count pass every round, and
    the result is still exact."""
    n, u, v, w, og, dist, succ, level, L = preset_graph(1)
    say(f"case C1 overflow K={k} max_nodes={max_nodes}")
    g = engine_for(n, u, v, w)
    g.report_paths(k, max_nodes=max_nodes)
    assert g.timings()["retries"] >= 1
    want, _ = og.topk(dist, succ, k)
    assert_bits_equal(want, g.get_slacks(k), f"C1 overflow K={k}")


@pytest.mark.parametrize("design", [2, 9])
def test_batch_designs_with_store_overflow(design):
    """Regression: designs of the 256-design batch workload whose single
    query overflows the default node store inside a group of rounds.  A
    later round of that group that finds nothing left to expand
    histogrammed the children of the round from fill_node_base (then past
    the store's capacity) and read one group of keys beyond the key buffer;
    the range is now clamped to the store."""
    n, u, v, w = cpgamd.gen_preset(4, design)
    k = int(cpgamd.preset_k(4, design))
    say(f"case batch design {design} single query K={k} (store overflow)")
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, _, _ = og.sfxt()
    want, _ = og.topk(dist, succ, k)
    for rep in range(3):
        g = engine_for(n, u, v, w)
        g.report_paths(k)
        assert g.timings()["retries"] >= 1
        assert_bits_equal(want, g.get_slacks(k), f"batch design {design} rep {rep}")


@pytest.mark.parametrize("probe", [1, -1])
@pytest.mark.parametrize("name", ["simple.txt", "dac21_paper_example.txt", "tv80_unit_wgts.txt"])
def test_reference_benchmarks_bound_probe_modes(golden_dir, name, probe):
    """The first bound from pruned count passes (probe=1) or from an
    unbounded first round (probe=-1) gives the same paths."""
    path = os.path.join(golden_dir, name)
    say(f"case benchmark {name} bound_probe={probe}")
    og = pyoracle.OracleGraph.from_text(path)
    dist, succ, level, L = og.sfxt()
    g = cpgamd.CpGen()
    g.read_input(path)
    for k in (1, 5, 1000):
        want, _ = og.topk(dist, succ, k)
        g.reset()
        g.report_paths(k, bound_probe=probe)
        got = g.get_slacks(k)
        if name.startswith("tv80"):
            np.testing.assert_array_equal(np.sort(got), np.sort(want))
        else:
            assert_bits_equal(want, got, f"{name} K={k} probe={probe}")


@pytest.mark.parametrize("probe", [1, -1])
@pytest.mark.parametrize("k", [100, 10000])
def test_c1_bound_probe_modes(k, probe):
    n, u, v, w, og, dist, succ, level, L = preset_graph(1)
    say(f"case C1 K={k} bound_probe={probe}")
    g = engine_for(n, u, v, w)
    g.report_paths(k, bound_probe=probe)
    if probe < 0:
        assert g.timings()["probes"] == 0
    want, _ = og.topk(dist, succ, k)
    assert_bits_equal(want, g.get_slacks(k), f"C1 K={k} probe={probe}")


def assert_multiset_prefixes(want, got, what, prefixes):
    assert len(got) == len(want), f"{what}: {len(got)} paths, oracle {len(want)}"
    for kp in prefixes:
        if kp > len(want):
            continue
        np.testing.assert_array_equal(np.sort(got[:kp]), np.sort(want[:kp]),
                                      err_msg=f"{what} prefix {kp}")


@pytest.mark.parametrize("k", [10000, 50000])
def test_c1_unit_weights_sorted_multiset(k):
    """Unit weights: massive equal-cost ties; judged on the sorted multiset."""
    n, u, v, w, og, dist, succ, level, L = preset_graph(1, unit=True)
    say(f"case C1 unit weights K={k}")
    g = engine_for(n, u, v, w)
    g.report_paths(k)
    want, _ = og.topk(dist, succ, k)
    assert_multiset_prefixes(want, g.get_slacks(k), f"C1 unit K={k}", (100, 1000, k))


def test_c3_densified_topk_prefixes_bit_exact():
    """The 1.5M-vertex DAG densified to mean out-degree 20 (30M edges), K=1M."""
    n, u, v, w, og, dist, succ, level, L = preset_graph(3)
    say("case C3 sfxt + oracle K=1M")
    want, st = og.topk(dist, succ, 1_000_000)
    assert st["nonmonotone"] == 0
    g = engine_for(n, u, v, w)
    gd, gs, gl = g.sfxt()
    np.testing.assert_array_equal(gd, dist)
    np.testing.assert_array_equal(gs, succ)
    np.testing.assert_array_equal(gl, level)
    say("case C3 GPU K=1M")
    g.report_paths(1_000_000)
    got = g.get_slacks(1_000_000)
    say(f"C3 timings {g.timings()}")
    for kp in (1000, 10_000, 50_000, 100_000, 1_000_000):
        assert_bits_equal(want[:kp], got[:kp], f"C3 prefix {kp}")


def test_c3_unit_weights_sorted_multiset_prefixes():
    """C3 with every delay 1.0f: equal-cost ties everywhere; sorted multiset
    per K prefix."""
    n, u, v, w, og, dist, succ, level, L = preset_graph(3, unit=True)
    say("case C3 unit oracle K=1M")
    want, _ = og.topk(dist, succ, 1_000_000)
    g = engine_for(n, u, v, w)
    say("case C3 unit GPU K=1M")
    g.report_paths(1_000_000)
    got = g.get_slacks(1_000_000)
    say(f"C3 unit timings {g.timings()}")
    assert_multiset_prefixes(want, got, "C3 unit", (1000, 10_000, 50_000, 100_000, 1_000_000))


def medium_fanout_graph():
    """3000 vertices in 20 levels, out-degree up to ~90 (distinct targets
    1-3 levels ahead), weights mixing exact ties (0.5, 1.0) and uniform
    draws: deviation lists of every length class (thread, wave, block
    sort)."""
    rng = np.random.default_rng(11)
    levels, per = 20, 150
    n = levels * per
    u, v, w = [], [], []
    for lv in range(levels - 1):
        for i in range(per):
            x = lv * per + i
            deg = int(rng.integers(1, 90))
            tl = rng.integers(lv + 1, min(levels, lv + 4), size=deg)
            tg = tl * per + rng.integers(0, per, size=deg)
            for t in np.unique(tg):
                u.append(x); v.append(int(t)); w.append(float(rng.choice([0.5, 1.0, rng.uniform(0.01, 1)])))
    return n, np.array(u, np.int32), np.array(v, np.int32), np.array(w, np.float32)


def test_medium_fanout_deviation_sorts():
    n, u, v, w = medium_fanout_graph()
    say("case medium fanout (<= 90)")
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, level, L = og.sfxt()
    g = engine_for(n, u, v, w)
    gd, gs, gl = g.sfxt()
    np.testing.assert_array_equal(gd, dist)
    np.testing.assert_array_equal(gs, succ)
    np.testing.assert_array_equal(gl, level)
    for k in (10, 1000, 20000):
        want, _ = og.topk(dist, succ, k)
        g.reset()
        g.report_paths(k)
        assert_bits_equal(want, g.get_slacks(k), f"medium fanout K={k}")


def star_graph(fan=5000, seed=5):
    """A sink with `fan` fanins, each fed by a source and by one hub:
    settling the sink releases `fan` vertices at once (more than one block's
    LDS pool holds), so the dataflow loop spills to its global queue."""
    rng = np.random.default_rng(seed)
    n = 2 + 2 * fan
    u, v = [], []
    for i in range(fan):
        u += [2 + i, 2 + fan + i, 1]
        v += [0, 2 + i, 2 + i]
    w = rng.uniform(0.01, 1.0, len(u)).astype(np.float32)
    return n, np.array(u, np.int32), np.array(v, np.int32), w


@pytest.mark.parametrize("method,blocks", [(2, 1), (2, 2), (2, 0), (3, 1), (3, 2), (3, 0), (4, 1), (5, 1), (5, 2), (5, 0)])
def test_sfxt_lds_queue_overflow(method, blocks):
    """The dataflow loop's LDS pool, the local-queue loop's LDS level
    buffer and the level loops' LDS push queues overflow into their global
    queues; results are unchanged.  The edge-parallel loop also takes the
    3000-record vertex over many wave steps."""
    n, u, v, w = star_graph()
    say(f"case star graph sfxt method {method} blocks={blocks}")
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, level, L = og.sfxt()
    g = engine_for(n, u, v, w)
    g.report_paths(1000, sfxt_blocks=blocks, sfxt_method=method)
    gd, gs, gl = g.sfxt()
    np.testing.assert_array_equal(gd, dist)
    np.testing.assert_array_equal(gs, succ)
    np.testing.assert_array_equal(gl, level)
    assert g.graph_diameter == L
    want, _ = og.topk(dist, succ, 1000)
    assert_bits_equal(want, g.get_slacks(1000), f"star blocks={blocks}")


@pytest.mark.parametrize("which", [1, 2, 3])
def test_sfxt_methods_agree(which):
    """The five SFXT loops (level loop with a global queue, dataflow loop,
    local-queue level loop, fence-free level loop, edge-parallel fence-free
    level loop) give the oracle's dist/succ/level on C1, C2 and the
    densified C3."""
    n, u, v, w, og, dist, succ, level, L = preset_graph(which)
    g = engine_for(n, u, v, w)
    for method in (1, 2, 3, 4, 5):
        say(f"case C{which} sfxt method {method}")
        g.reset()
        g.report_paths(10, sfxt_method=method, kernel_timing=True)
        gd, gs, gl = g.sfxt()
        np.testing.assert_array_equal(gd, dist)
        np.testing.assert_array_equal(gs, succ)
        np.testing.assert_array_equal(gl, level)
        assert g.graph_diameter == L
        say(f"C{which} method {method} levels_ms {g.timings()['sfxt_levels_ms']:.3f}")


def test_sfxt_dataflow_cycle_terminates():
    """A cycle never releases its vertices: the loop ends by quiescence and
    leaves them unsettled, as the level loop does."""
    u = np.array([0, 1, 2, 2, 4], np.int32)
    v = np.array([1, 2, 1, 3, 3], np.int32)
    w = np.array([1, 2, 3, 4, 5], np.float32)
    say("case cycle graph")
    res = []
    for method in (1, 2, 3, 4, 5):
        g = engine_for(5, u, v, w)
        g.report_paths(10, sfxt_method=method)
        res.append(g.sfxt() + (g.get_slacks(10),))
    for other in res[1:]:
        for a, b in zip(res[0], other):
            np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(res[0][2], [-1, -1, -1, 0, 1])


NM_GRID = [0.1, 0.2, 0.3, 0.7, 0.05, 0.15, 0.25, 0.35, 0.45, 0.6, 1.1, 1.3]


def decimal_grid_dag(seed, n=10, m=22):
    """Tiny DAG with weights on a decimal grid: int-scaled ties are common
    and binary32 rounds some deviations below their parent's cost (e.g.
    distf 0.1 + weight 0.2 < distf 0.3 in binary32)."""
    rng = np.random.default_rng(seed)
    edges = []
    for _ in range(m):
        a, b = sorted(rng.choice(n, 2, replace=False))
        edges.append((int(a), int(b), float(np.float32(rng.choice(NM_GRID)))))
    return (n, np.array([e[0] for e in edges], np.int32), np.array([e[1] for e in edges], np.int32),
            np.array([e[2] for e in edges], np.float32))


@pytest.mark.parametrize("seed", [5, 12, 20, 31, 57])
def test_nonmonotone_children_counted_and_checked(seed):
    """With K past the number of paths every path is expanded: the engine
    finds every path (the brute-force list) and counts exactly the children
    the oracle counts below their parent; monotone_check turns that into an
    error."""
    n, u, v, w = decimal_grid_dag(seed)
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, _, _ = og.sfxt()
    allp = og.all_paths(dist, succ)
    k = len(allp) + 5
    want, st = og.topk(dist, succ, k, root_tie_quirk=False)
    assert st["nonmonotone"] > 0
    np.testing.assert_array_equal(want, allp)
    say(f"case non-monotone seed {seed}: {st['nonmonotone']} children, {len(allp)} paths")
    g = engine_for(n, u, v, w)
    g.report_paths(k, root_tie_quirk=False)
    assert_bits_equal(allp, g.get_slacks(k), f"non-monotone seed {seed}")
    assert g.timings()["nonmonotone"] == st["nonmonotone"]
    with pytest.raises(cpgamd.CpgError, match="cost less than its parent"):
        g.report_paths(k, root_tie_quirk=False, monotone_check=True)


def test_presets_have_no_nonmonotone_children():
    """The counter is 0 on the K=10000 query of the 50K-vertex DAG, and the
    check passes there."""
    n, u, v, w, og, dist, succ, level, L = preset_graph(1)
    g = engine_for(n, u, v, w)
    g.report_paths(10000, monotone_check=True)
    assert g.timings()["nonmonotone"] == 0


def test_mem_budget_bytes_bounds_the_node_store():
    """mem_budget_bytes: a budget with room for the store gives the exact
    answer within it; one with no room, or one the store must outgrow after
    an overflow, fails with an error naming the budget."""
    n, u, v, w, og, dist, succ, level, L = preset_graph(1)
    g = engine_for(n, u, v, w)
    g.report_paths(10)
    t = g.timings()
    sfxt_bytes = t["device_bytes"] - 61 * t["node_slots"]  # graph + SFXT
    say(f"case C1 memory budget (graph + SFXT {sfxt_bytes} bytes, store {t['node_slots']} slots)")
    want, _ = og.topk(dist, succ, 10000)
    budget = sfxt_bytes + 61 * 400_000
    assert budget < t["device_bytes"]
    g.report_paths(10000, mem_budget_bytes=budget)
    assert_bits_equal(want, g.get_slacks(10000), "C1 K=10000 within the budget")
    assert g.timings()["device_bytes"] <= budget and g.timings()["node_slots"] <= 400_000
    with pytest.raises(cpgamd.CpgError, match="mem_budget_bytes"):
        g.report_paths(10000, mem_budget_bytes=sfxt_bytes + 61 * 100)
    # 60000 slots overflow at K=50000 (see the overflow test above); the
    # budget lets the store grow to 61000 slots only
    with pytest.raises(cpgamd.CpgError, match="mem_budget_bytes"):
        g.report_paths(50000, max_nodes=60000, mem_budget_bytes=sfxt_bytes + 61 * 61000)
    g.report_paths(50000, max_nodes=60000)
    assert g.timings()["retries"] >= 1


def test_costs_into_pinned_host_array():
    """get_slacks into a pinned host array (cpg_host_alloc) gives the same
    costs as into an ordinary one, for a prefix and for all K; an array
    shorter than K is refused."""
    n, u, v, w = cpgamd.gen_preset(1)
    g = engine_for(n, u, v, w)
    say("case costs into pinned host memory")
    g.report_paths(20000)
    want = g.get_slacks(20000)
    with cpgamd.HostArray(20000) as host:
        got = g.get_slacks(20000, out=host.array)
        np.testing.assert_array_equal(got, want)
        got = g.get_slacks(5000, out=host.array)
        np.testing.assert_array_equal(got, want[:5000])
        with pytest.raises(ValueError):
            g.get_slacks(20001, out=host.array)
