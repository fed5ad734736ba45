"""The C++ facade (include/gpucpg_amd.hpp) runs a query end to end: a
driver reading exactly the members the reference's in-process drivers
read (examples/tc-pfxt-inprocess-common.cuh:104-143) gets the oracle's
costs bit for bit and one step time and path count per expansion round."""
import os
import subprocess

import numpy as np
import pytest

import cpgamd
import pyoracle
from test_host_abi import build_facade_query_driver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [10, 10000])
def test_cpp_facade_query_matches_oracle(tmp_path, k):
    n, u, v, w = cpgamd.gen_preset(1)
    text = str(tmp_path / "c1.txt")
    cpgamd.write_text(n, u, v, w, text)
    exe = build_facade_query_driver(tmp_path)
    out = str(tmp_path / "costs.bin")
    print(f"case C++ facade driver C1 K={k}", flush=True)
    res = subprocess.run([str(exe), text, str(k), out], check=True, capture_output=True, text=True,
                         timeout=300)
    print(res.stdout, flush=True)
    f = res.stdout.split()
    vals = dict(zip(f[0::2], f[1::2]))
    got = np.fromfile(out, np.float32)
    og = pyoracle.OracleGraph.from_text(text)
    dist, succ, level, L = og.sfxt()
    want, _ = og.topk(dist, succ, k)
    assert len(got) == len(want) and pyoracle.first_bit_mismatch(want, got) == -1
    steps = int(vals["steps"])
    assert steps >= 1 and int(vals["step_times"]) == steps and int(vals["paths_per_step"]) == steps
    # paths stored by the rounds (K=10: the roots and the bound probe's
    # children already hold the top 10, so the one round may store none)
    assert 0 <= int(vals["gen"]) <= int(vals["total"])
    if k >= 10000:
        assert int(vals["gen"]) > 0
    assert float(vals["pfxt_ms"]) > 0.0 and float(vals["lvlize_ms"]) > 0.0
    assert float(vals["pfxt_ms"]) <= float(vals["expand_ms"]) + 1e-6
    assert int(vals["diameter"]) == L
