"""GPU parity of endpoint-subset queries (paths ending at chosen sinks).

The endpoint subset is an extension of this engine (the reference has no
such option): only the chosen sinks start the SFXT at distance 0, every
other vertex that reaches none of them is unreachable.  Oracle: the
masked SFXT of oracle/oracle.c (orc_sfxt_masked, pinned against an
independent Dijkstra from the chosen sinks and against the weight sums of
the source -> chosen-sink paths in tests/test_oracle.py) and its exact
top-K.  Bar: SFXT dist/succ equal element-for-element, costs bit-exact.
"""
import sys

import numpy as np
import pytest

import cpgamd
import pyoracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sink_mask(n, u, frac, seed):
    outdeg = np.bincount(u, minlength=n)
    sinks = np.nonzero(outdeg == 0)[0]
    rng = np.random.default_rng(seed)
    m = np.zeros(n, np.uint8)
    m[rng.choice(sinks, max(1, int(round(frac * len(sinks)))), replace=False)] = 1
    return m


@pytest.mark.parametrize("frac", [0.002, 0.01, 0.3, 1.0])
def test_c1_masked_query_matches_oracle(frac):
    n, u, v, w = cpgamd.gen_preset(1)
    mask = sink_mask(n, u, frac, 11)
    say(f"case endpoints C1 frac={frac} chosen={int(mask.sum())}")
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    d, s, _, _ = og.sfxt(mask)
    want, _ = og.topk(d, s, 10000)
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    g.reset()
    g.report_paths(10000, endpoint_mask=mask)
    got = g.get_slacks(10000)
    assert len(got) == len(want)
    assert pyoracle.first_bit_mismatch(want, got) == -1
    gd, gs, _ = g.sfxt()  # the SFXT of the last query: the masked one
    np.testing.assert_array_equal(gd, d)
    np.testing.assert_array_equal(gs, s)


def test_mask_changes_with_cache_enabled():
    """The cache holds the last SFXT with its endpoint subset: a query with
    the same subset reuses it, another subset recomputes it, and an
    all-sinks mask is the unmasked query."""
    n, u, v, w = cpgamd.gen_preset(1)
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    g.enable_tc_pfxt_static_cache(True)
    k = 3000
    masks = [sink_mask(n, u, 0.05, 1), sink_mask(n, u, 0.05, 1), sink_mask(n, u, 0.05, 2),
             None, sink_mask(n, u, 1.0, 3)]
    wants = []
    for m in masks:
        d, s, _, _ = og.sfxt(m)
        wants.append(og.topk(d, s, k)[0])
    for i, (m, want) in enumerate(zip(masks, wants)):
        say(f"case endpoints cache step {i}")
        g.reset()
        g.report_paths(k, endpoint_mask=m)
        got = g.get_slacks(k)
        assert len(got) == len(want) and pyoracle.first_bit_mismatch(want, got) == -1
    np.testing.assert_array_equal(wants[3], wants[4])  # all sinks == no mask
    assert g.tc_pfxt_static_cache_hits() == 2  # steps 1 and 4
    assert g.tc_pfxt_static_cache_misses() == 3


def test_c5_graph_masked_small_k():
    """A served-stream graph (the first of the eight resident DAGs) with a
    1% endpoint subset."""
    n, u, v, w = cpgamd.gen_preset(5, 0)
    mask = sink_mask(n, u, 0.01, 5)
    say(f"case endpoints C5 graph 0 n={n} m={len(u)} chosen={int(mask.sum())}")
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    d, s, _, _ = og.sfxt(mask)
    want, _ = og.topk(d, s, 20000)
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    g.reset()
    g.report_paths(20000, endpoint_mask=mask)
    got = g.get_slacks(20000)
    assert len(got) == len(want) and pyoracle.first_bit_mismatch(want, got) == -1


def test_bad_mask_length_is_rejected():
    n, u, v, w = cpgamd.gen_preset(1)
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    with pytest.raises(ValueError):
        g.report_paths(10, endpoint_mask=np.ones(n - 1, np.uint8))
