"""The engine's unchanged HIP sources on the CPU SIMT emulator (tests/simt):
blocks run as OS threads, waves as fibers, cooperative launches with every
block resident.  These runs check on the CPU what the GPU suite checks on
the device: both SFXT loops give the oracle's dist/succ/level, a query
gives the oracle's costs bit for bit, a node-store overflow restarts
exactly, and the dataflow loop's LDS pool spills to its global queue.
SIMT_GUARD=1 places every device buffer against an inaccessible page, so
an access past a buffer's end faults here instead of on a GPU."""
import os
import subprocess

import numpy as np
import pytest

import cpgamd
import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIMT = os.path.join(ROOT, "tests", "simt")
LIB = os.path.join(SIMT, "libcpgamd_simt.so")


@pytest.fixture(scope="module")
def emu():
    """The emulator build of the C ABI, loaded in place of the HIP one for
    this module's tests only."""
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-s", "-j", jobs, "-C", SIMT], check=True)
    os.environ.setdefault("SIMT_CUS", "4")
    os.environ.setdefault("SIMT_GUARD", "1")
    saved = cpgamd._lib
    cpgamd._lib = cpgamd.load_library(LIB)
    try:
        yield cpgamd
    finally:
        cpgamd._lib = saved


def bits_equal(want, got):
    return len(want) == len(got) and pyoracle.first_bit_mismatch(want, got) == -1


@pytest.fixture(scope="module")
def c1():
    n, u, v, w = cpgamd.gen_preset(1)
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, level, L = og.sfxt()
    return n, u, v, w, og, dist, succ, level, L


@pytest.mark.parametrize("method", [1, 2, 3, 4, 5])
def test_emulated_c1_query_equals_oracle(emu, c1, method):
    n, u, v, w, og, dist, succ, level, L = c1
    g = emu.CpGen()
    try:
        g.load_edges(n, u, v, w)
        g.report_paths(100, sfxt_method=method)
        gd, gs, gl = g.sfxt()
        np.testing.assert_array_equal(gd, dist)
        np.testing.assert_array_equal(gs, succ)
        np.testing.assert_array_equal(gl, level)
        assert g.graph_diameter == L
        want, _ = og.topk(dist, succ, 100)
        assert bits_equal(want, g.get_slacks(100))
    finally:
        g.close()


def test_emulated_relabelled_c1_equals_oracle(emu, c1):
    """The level-order CSR relabelling: dist/succ/level (in original ids)
    and the costs equal the oracle's, with and without an endpoint
    subset."""
    n, u, v, w, og, dist, succ, level, L = c1
    g = emu.CpGen()
    try:
        g.load_edges(n, u, v, w)
        g.set_csr_reorder(1)
        g.report_paths(100)
        gd, gs, gl = g.sfxt()
        np.testing.assert_array_equal(gd, dist)
        np.testing.assert_array_equal(gs, succ)
        np.testing.assert_array_equal(gl, level)
        want, _ = og.topk(dist, succ, 100)
        assert bits_equal(want, g.get_slacks(100))
        sinks = np.nonzero(np.bincount(u, minlength=n) == 0)[0]
        m = np.zeros(n, np.uint8)
        m[np.random.default_rng(5).choice(sinks, len(sinks) // 3, replace=False)] = 1
        d2, s2, _, _ = og.sfxt(m)
        want2, _ = og.topk(d2, s2, 100)
        g.reset()
        g.report_paths(100, endpoint_mask=m)
        assert bits_equal(want2, g.get_slacks(100))
    finally:
        g.close()


def test_emulated_batch_equals_oracle(emu):
    """A batch of six small seeded designs (5K-9K vertices, K 112 to 10217):
    the multi-query engine's rounds (the later ones replayed from a captured
    HIP graph) give every design the oracle's best-first top K bit for
    bit."""
    picks = [12, 20, 24, 33, 36, 40]
    b = emu.Batch()
    want = []
    try:
        for i in picks:
            n, u, v, w = emu.gen_preset(4, i)
            k = int(emu.preset_k(4, i))
            b.add(n, u, v, w, k)
            og = pyoracle.OracleGraph.from_edges(n, u, v, w)
            dist, succ, _, _ = og.sfxt()
            want.append((og.topk(dist, succ, k)[0], k))
        b.run()
        assert b.timings()["rounds"] > 1
        for q, (wc, k) in enumerate(want):
            assert bits_equal(wc, b.get_slacks(q, k)), f"design {picks[q]}"
    finally:
        b.close()


def test_emulated_node_store_overflow_restarts_exactly(emu, c1):
    n, u, v, w, og, dist, succ, level, L = c1
    g = emu.CpGen()
    try:
        g.load_edges(n, u, v, w)
        g.report_paths(2000, max_nodes=2500)
        assert g.timings()["retries"] >= 1
        want, _ = og.topk(dist, succ, 2000)
        assert bits_equal(want, g.get_slacks(2000))
    finally:
        g.close()


@pytest.mark.parametrize("method", [2, 3, 5])
def test_emulated_lds_queue_spill(emu, method):
    """One sink with 3000 fanins: settling it releases more vertices than a
    block's LDS queue holds (the dataflow loop's pool, the local-queue
    loop's level buffer, the edge-parallel loop's level queue), so each
    spills to its global queue; the edge-parallel loop also takes that
    sink's 3000 records over many wave steps."""
    fan = 3000
    rng = np.random.default_rng(11)
    u, v = [], []
    for i in range(fan):
        u += [2 + i, 2 + fan + i, 1]
        v += [0, 2 + i, 2 + i]
    n = 2 + 2 * fan
    u = np.array(u, np.int32)
    v = np.array(v, np.int32)
    w = rng.uniform(0.01, 1.0, len(u)).astype(np.float32)
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, level, L = og.sfxt()
    g = emu.CpGen()
    try:
        g.load_edges(n, u, v, w)
        g.report_paths(50, sfxt_blocks=2, sfxt_method=method)
        gd, gs, gl = g.sfxt()
        np.testing.assert_array_equal(gd, dist)
        np.testing.assert_array_equal(gs, succ)
        np.testing.assert_array_equal(gl, level)
        want, _ = og.topk(dist, succ, 50)
        assert bits_equal(want, g.get_slacks(50))
    finally:
        g.close()


def test_emulated_dense_fanouts_equal_oracle(emu):
    """A graph of mean fanout far above 8, so the deviation index is built
    one wave per vertex: fanouts of up to 64 records, of 65-128 (two per
    lane, ranked in registers) and of 129-200 (queued for the block sort),
    with delays from three values so that many deviation records tie.  The
    query's costs match the oracle's bit for bit."""
    rng = np.random.default_rng(23)
    width, levels = 200, 10
    n = width * levels
    u, v = [], []
    for x in range(n - width):
        lvl = x // width
        hi = min(n, (lvl + 4) * width)
        pool = np.arange((lvl + 1) * width, hi)
        band = rng.random()
        deg = int(rng.integers(1, 11) if band < 0.4 else rng.integers(65, 129) if band < 0.8
                  else rng.integers(129, 201))
        tgt = rng.choice(pool, min(deg, len(pool)), replace=False)
        u += [x] * len(tgt)
        v += tgt.tolist()
    u = np.array(u, np.int32)
    v = np.array(v, np.int32)
    assert len(u) >= 8 * n
    w = rng.choice(np.array([0.125, 0.25, 0.5], np.float32), len(u))
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, level, L = og.sfxt()
    g = emu.CpGen()
    try:
        g.load_edges(n, u, v, w)
        g.report_paths(3000)
        want, _ = og.topk(dist, succ, 3000)
        assert bits_equal(want, g.get_slacks(3000))
    finally:
        g.close()
