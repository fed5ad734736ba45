"""Host-side checks of the product library (no GPU needed).

The library loads, exports every function include/cpg.h declares, and its
host logic (text loader, CSR build, seeded generator, text writer) agrees
with the oracle's loader and the committed fixtures.  No call here reaches
the device.
"""
import hashlib
import json
import os
import re

import numpy as np
import pytest

import cpgamd
import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cpg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cpg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = cpgamd.lib()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(cpgamd.ABI), set(names) ^ set(cpgamd.ABI)


def host_csr(g):
    adjp, adj, w = g.csr()
    return adjp, adj, w


@pytest.mark.parametrize("name", ["simple.txt", "dac21_paper_example.txt", "tv80_unit_wgts.txt"])
def test_text_loader_matches_oracle(golden_dir, name):
    path = os.path.join(golden_dir, name)
    g = cpgamd.CpGen()
    g.read_input(path)
    og = pyoracle.OracleGraph.from_text(path)
    a = og.arrays()
    adjp, adj, w = host_csr(g)
    assert g.num_verts() == og.n and g.num_edges() == og.m
    np.testing.assert_array_equal(adjp, a["oadjp"])
    np.testing.assert_array_equal(adj, a["oadj"])
    np.testing.assert_array_equal(w, a["owgt"])
    g2 = cpgamd.CpGen()
    g2.read_input(path, ignore_wgts=True)
    assert np.all(host_csr(g2)[2] == 1.0)


def test_text_loader_errors(tmp_path):
    g = cpgamd.CpGen()
    with pytest.raises(cpgamd.CpgError):
        g.read_input(str(tmp_path / "missing.txt"))
    p = tmp_path / "bad.txt"
    p.write_text('2\n"0"\n"1"\n"0" -> 1;\n')
    with pytest.raises(cpgamd.CpgError):
        g.read_input(str(p))
    p.write_text('2\n"0"\n"1"\n"0" -> "7", 0.5;\n')
    with pytest.raises(cpgamd.CpgError, match="out of range"):
        g.read_input(str(p))
    p.write_text('2\n"0"\n"1"\n"0" -> "1", x;\n')
    with pytest.raises(cpgamd.CpgError):
        g.read_input(str(p))


def csr_sha(adjp, adj, w):
    h = hashlib.sha256()
    for a in (adjp, adj, w):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_generator_is_seeded_and_matches_fixture(golden_dir):
    meta = json.load(open(os.path.join(golden_dir, "synthetic_golden.json")))
    n, u, v, w = cpgamd.gen_preset(1)
    n2, u2, v2, w2 = cpgamd.gen_preset(1)
    assert n == n2 and np.array_equal(u, u2) and np.array_equal(v, v2) and np.array_equal(w, w2)
    g = cpgamd.CpGen()
    g.load_edges(n, u, v, w)
    assert csr_sha(*g.csr()) == meta["c1"]["csr_sha256"]
    assert g.num_verts() == 50000 and g.num_edges() == meta["c1"]["m"]
    # the configuration's shape: 1% sources, 1% sinks, weights in [0.01, 1]
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    a = og.arrays()
    outdeg = np.diff(a["oadjp"])
    indeg = np.diff(a["iadjp"])
    assert (outdeg == 0).sum() == 500 and (indeg == 0).sum() == 500
    assert w.min() >= 0.01 and w.max() <= 1.0


def test_c1_golden_costs(golden_dir):
    want = np.load(os.path.join(golden_dir, "c1_k10000_costs.npy"))
    n, u, v, w = cpgamd.gen_preset(1)
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    dist, succ, level, L = og.sfxt()
    got, st = og.topk(dist, succ, 10_000)
    assert pyoracle.first_bit_mismatch(want, got) == -1 and len(got) == len(want)
    assert st["nonmonotone"] == 0


def test_text_round_trip(tmp_path):
    n, u, v, w = cpgamd.gen_preset(1)
    path = str(tmp_path / "c1.txt")
    cpgamd.write_text(n, u, v, w, path)
    g1 = cpgamd.CpGen()
    g1.load_edges(n, u, v, w)
    g2 = cpgamd.CpGen()
    g2.read_input(path)
    for x, y in zip(g1.csr(), g2.csr()):
        np.testing.assert_array_equal(x, y)
    og = pyoracle.OracleGraph.from_text(path)
    np.testing.assert_array_equal(og.arrays()["owgt"], g1.csr()[2])


def test_load_csr_equals_load_edges():
    n, u, v, w = cpgamd.gen_preset(1)
    g1 = cpgamd.CpGen()
    g1.load_edges(n, u, v, w)
    adjp, adj, wt = g1.csr()
    g2 = cpgamd.CpGen()
    g2.load_csr(adjp, adj, wt)
    for x, y in zip(g1.csr(), g2.csr()):
        np.testing.assert_array_equal(x, y)
    with pytest.raises(cpgamd.CpgError):
        g2.load_csr(np.array([0, 5], np.int32), adj[:2], wt[:2])


def test_batch_and_served_presets_have_the_stated_shapes():
    for i in (0, 7, 255):
        k = cpgamd.preset_k(4, i)
        assert 100 <= k <= 100_000
    n, u, v, w = cpgamd.gen_preset(4, 3)
    assert 5000 <= n <= 200_000
    og = pyoracle.OracleGraph.from_edges(n, u, v, w)
    a = og.arrays()
    assert (np.diff(a["oadjp"]) == 0).sum() == round(0.02 * n)
    mean = len(u) / n
    assert 1.5 <= mean <= 11.0


def test_unit_weight_variant_keeps_edges():
    # the unit-weight densified variant must be the same graph as the weighted
    # one; checked on the 50K preset to keep the CPU suite fast
    n, u, v, w = cpgamd.gen_preset(1)
    n1, u1, v1, w1 = cpgamd.gen_preset(1, 0, True)
    assert n == n1 and np.array_equal(u, u1) and np.array_equal(v, v1)
    assert np.all(w1 == 1.0)


FACADE_DRIVER = r"""
#include <cstdio>
#include "gpucpg_amd.hpp"
int main(int argc, char** argv) {
  gpucpg::CpGen g;
  g.read_input(argv[1]);
  std::printf("%zu %zu\n", g.num_verts(), g.num_edges());
  try {
    g.read_input(std::string(argv[1]) + ".missing");
  } catch (const std::runtime_error& e) {
    std::printf("error %s\n", e.what());
  }
  return 0;
}
"""


def test_cpp_facade_compiles_links_and_reads_input(tmp_path, golden_dir):
    """include/gpucpg_amd.hpp: a reference-style driver builds with g++
    against libcpgamd.so and reads a benchmark file (host calls only)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "gpu-cpg_amd")
    src = tmp_path / "driver.cpp"
    src.write_text(FACADE_DRIVER)
    exe = tmp_path / "driver"
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", os.path.join(root, "include"), str(src),
                    "-L", libdir, "-lcpgamd", "-Wl,-rpath," + libdir, "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe), os.path.join(golden_dir, "simple.txt")], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    og = pyoracle.OracleGraph.from_text(os.path.join(golden_dir, "simple.txt"))
    n, m = og.n, og.m
    assert out[0] == f"{n} {m}"
    assert out[1].startswith("error ")
