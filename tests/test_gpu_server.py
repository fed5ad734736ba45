"""GPU parity of served queries (cpg_server_*): resident graphs, a queue of
top-K queries with endpoint subsets, PFXT rounds of the in-flight queries
batched together.  Every result equals the oracle's masked top-K for its
graph, K and endpoint subset (bit-exact), whatever else is in flight."""
import sys

import numpy as np
import pytest

import cpgamd
import pyoracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sink_mask(rng, n, u, frac):
    sinks = np.nonzero(np.bincount(u, minlength=n) == 0)[0]
    m = np.zeros(n, np.uint8)
    m[rng.choice(sinks, max(1, int(frac * len(sinks))), replace=False)] = 1
    return m


def run_stream(server, max_steps=100000):
    steps = 0
    while server.step() > 0:
        steps += 1
        assert steps < max_steps
    return steps


@pytest.mark.parametrize("slots", [3, 64])
def test_served_queries_equal_oracle(slots):
    graphs = [cpgamd.gen_preset(1)] + [cpgamd.gen_preset(4, i) for i in (0, 2, 3)]
    S = cpgamd.Server(max_in_flight=slots)
    ogs = []
    for n, u, v, w in graphs:
        S.add_graph(n, u, v, w)
        ogs.append(pyoracle.OracleGraph.from_edges(n, u, v, w))
    rng = np.random.default_rng(7 + slots)
    sub = {}
    for _ in range(16):
        gi = int(rng.integers(0, len(graphs)))
        n, u, v, w = graphs[gi]
        k = int(rng.choice([50, 1000, 20000, 60000]))
        frac = float(rng.choice([0.01, 0.2, 0.7, 1.0]))
        m = None if frac == 1.0 else sink_mask(rng, n, u, frac)
        sub[S.submit(gi, k, m)] = (gi, k, m)
    say(f"case server slots={slots} queries={len(sub)}")
    run_stream(S)
    st = S.stats()
    say(f"server stats {st}")
    done = S.poll()
    assert sorted(done) == sorted(sub)
    for qid in done:
        gi, k, m = sub[qid]
        d, s, _, _ = ogs[gi].sfxt(m)
        want, _ = ogs[gi].topk(d, s, k)
        got = S.get_slacks(qid, k)
        assert len(got) == len(want), qid
        assert pyoracle.first_bit_mismatch(want, got) == -1, qid
        sub_t, adm_t, done_t = S.times(qid)
        assert sub_t <= adm_t <= done_t
        S.release(qid)
    if slots == 3:
        assert st["compactions"] >= 1


def test_served_result_equals_single_query_alone():
    """Isolation: a query's result in a mixed stream equals cpg_query's
    for the same graph, K and endpoint subset on its own handle."""
    n, u, v, w = cpgamd.gen_preset(4, 1)
    rng = np.random.default_rng(3)
    masks = [sink_mask(rng, n, u, f) for f in (0.05, 0.5)] + [None]
    S = cpgamd.Server(max_in_flight=8)
    S.add_graph(n, u, v, w)
    c1 = cpgamd.gen_preset(1)
    S.add_graph(*c1)
    ids = [S.submit(0, 30000, m) for m in masks] + [S.submit(1, 5000, None)]
    say("case server isolation")
    run_stream(S)
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    for qid, m in zip(ids, masks):
        g.reset()
        g.report_paths(30000, endpoint_mask=m)
        want = g.get_slacks(30000)
        got = S.get_slacks(qid, 30000)
        np.testing.assert_array_equal(want.view(np.uint32), got.view(np.uint32))


def test_server_rejects_bad_input():
    n, u, v, w = cpgamd.gen_preset(1)
    S = cpgamd.Server(max_in_flight=2)
    S.add_graph(n, u, v, w)
    with pytest.raises(cpgamd.CpgError):
        S.submit(5, 10)
    with pytest.raises(ValueError):
        S.submit(0, 10, np.ones(n - 1, np.uint8))
    qid = S.submit(0, 10)
    with pytest.raises(cpgamd.CpgError):
        S.get_slacks(qid, 10)  # not finished yet
    run_stream(S)
    assert len(S.get_slacks(qid, 10)) == 10


@pytest.mark.parametrize("nq", [9, 16])
def test_full_admission_group_on_one_graph(nq):
    """nq queries on one graph are admitted as one group: one level loop
    relaxes their endpoint subsets together (16 subsets, padded with empty
    ones for 9), and every result equals the oracle's."""
    n, u, v, w = cpgamd.gen_preset(4, 38)  # mean fanout 9: the wave deviation pass too
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    rng = np.random.default_rng(100 + nq)
    S = cpgamd.Server(max_in_flight=64)
    S.add_graph(n, u, v, w)
    sub = {}
    for i in range(nq):
        k = int(rng.choice([100, 3000, 30000]))
        m = None if i == 0 else sink_mask(rng, n, u, float(rng.choice([0.01, 0.3, 0.9])))
        sub[S.submit(0, k, m)] = (k, m)
    say(f"case server group of {nq} on one graph")
    run_stream(S)
    assert S.stats()["sfxt_groups"] == 1
    for qid, (k, m) in sub.items():
        d, s, _, _ = og.sfxt(m)
        want, _ = og.topk(d, s, k)
        got = S.get_slacks(qid, k)
        assert len(got) == len(want) and pyoracle.first_bit_mismatch(want, got) == -1, qid


def test_served_stream_graphs_equal_oracle():
    """Queries on the served-stream graphs themselves (200K-2.6M vertices,
    4M-70M edges, mean fanout 2.5-30): endpoint subsets of 1-100% of the
    sinks and K from 1K to 100K, all in flight together; every result equals
    the oracle's masked top-K bit for bit."""
    gids = [0, 2, 3, 4, 5, 7]
    graphs = {i: cpgamd.gen_preset(5, i) for i in gids}
    S = cpgamd.Server(max_in_flight=64)
    try:
        slot = {}
        for i in gids:
            slot[i] = S.add_graph(*graphs[i])
        rng = np.random.default_rng(555)
        plan = [(7, 1.0, 100000), (7, 0.05, 20000)]
        for _ in range(12):
            gi = int(rng.choice(gids[:-1]))
            frac = float(rng.choice([0.01, 0.1, 0.4, 0.8, 1.0]))
            k = int(rng.choice([1000, 10000, 50000, 100000]))
            plan.append((gi, frac, k))
        sub = {}
        for gi, frac, k in plan:
            n, u, v, w = graphs[gi]
            m = None if frac == 1.0 else sink_mask(rng, n, u, frac)
            sub[S.submit(slot[gi], k, m)] = (gi, k, m, frac)
        say(f"case served-stream graphs, {len(sub)} queries")
        run_stream(S)
        say(f"server stats {S.stats()}")
        bad = []
        for gi in gids:
            og = pyoracle.OracleGraph.from_edges(*graphs[gi])
            for qid, (g2, k, m, frac) in sub.items():
                if g2 != gi:
                    continue
                d, s, _, _ = og.sfxt(m)
                want, _ = og.topk(d, s, k)
                got = S.get_slacks(qid, k)
                if len(got) != len(want) or pyoracle.first_bit_mismatch(want, got) != -1:
                    bad.append((gi, k, frac, len(got), len(want)))
            del og
        assert not bad, bad
    finally:
        S.close()
