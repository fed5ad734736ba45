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


def test_binary_csr_round_trip(tmp_path, golden_dir):
    """write_to_csr_bin / read_input_bin (gpucpg/graph.h:24-55 layout): the
    file is N, M, offsets, adjncy, wgts as 4-byte words, and reading it back
    gives the same CSR as the text loader."""
    import struct
    for name in ("simple.txt", "tv80_unit_wgts.txt"):
        g = cpgamd.CpGen()
        g.read_input(os.path.join(golden_dir, name))
        adjp, adj, w = host_csr(g)
        path = str(tmp_path / (name + ".bin"))
        g.write_to_csr_bin(path)
        raw = open(path, "rb").read()
        n, m = struct.unpack("<ii", raw[:8])
        assert (n, m) == (g.num_verts(), g.num_edges())
        assert len(raw) == 8 + 4 * (n + 1) + 8 * m
        np.testing.assert_array_equal(np.frombuffer(raw, np.int32, n + 1, 8), adjp)
        np.testing.assert_array_equal(np.frombuffer(raw, np.int32, m, 8 + 4 * (n + 1)), adj)
        np.testing.assert_array_equal(np.frombuffer(raw, np.float32, m, 8 + 4 * (n + 1) + 4 * m), w)
        g2 = cpgamd.CpGen()
        g2.read_input_bin(path)
        for a, b in zip(host_csr(g2), (adjp, adj, w)):
            np.testing.assert_array_equal(a, b)


def test_binary_csr_hand_written_and_rejects_bad_files(tmp_path):
    import struct
    # 3 vertices: 0 -> 1 (0.5), 0 -> 2 (0.25), 1 -> 2 (1.0)
    body = struct.pack("<ii", 3, 3) + struct.pack("<4i", 0, 2, 3, 3) + struct.pack("<3i", 1, 2, 2) \
        + struct.pack("<3f", 0.5, 0.25, 1.0)
    good = tmp_path / "g.bin"
    good.write_bytes(body)
    g = cpgamd.CpGen()
    g.read_input_bin(str(good))
    adjp, adj, w = host_csr(g)
    assert adjp.tolist() == [0, 2, 3, 3] and adj.tolist() == [1, 2, 2] and w.tolist() == [0.5, 0.25, 1.0]
    for bad in (body[:-4], body + b"\0\0\0\0",
                struct.pack("<ii", 3, 3) + struct.pack("<4i", 0, 2, 1, 3) + body[24:],
                struct.pack("<ii", 3, 3) + struct.pack("<4i", 0, 2, 3, 3) + struct.pack("<3i", 1, 2, 7)
                + body[-12:]):
        p = tmp_path / "bad.bin"
        p.write_bytes(bad)
        with pytest.raises(cpgamd.CpgError):
            g.read_input_bin(str(p))


def test_parallel_text_parser_matches_oracle(tmp_path):
    """A multi-megabyte file is parsed in chunks on several host threads;
    the CSR equals the oracle's single-threaded parse, and an error deep in
    the file reports its own line number."""
    n, u, v, w = cpgamd.gen_preset(1)
    path = str(tmp_path / "c1.txt")
    cpgamd.write_text(n, u, v, w, path)
    assert os.path.getsize(path) > 4 << 20
    g = cpgamd.CpGen()
    g.read_input(path)
    og = pyoracle.OracleGraph.from_text(path)
    a = og.arrays()
    adjp, adj, wg = host_csr(g)
    np.testing.assert_array_equal(adjp, a["oadjp"])
    np.testing.assert_array_equal(adj, a["oadj"])
    np.testing.assert_array_equal(wg, a["owgt"])
    lines = open(path).read().split("\n")
    bad_line = n + 1 + 150_000  # 0-based index of an edge line far from the start
    lines[bad_line] = '"1" -> x;'
    bad = str(tmp_path / "bad.txt")
    open(bad, "w").write("\n".join(lines))
    with pytest.raises(cpgamd.CpgError, match=f"line {bad_line + 1}$"):
        g.read_input(bad)


FACADE_QUERY_DRIVER = r"""
// The member set the reference's in-process drivers read
// (examples/tc-pfxt-inprocess-common.cuh:104-143, tc-pfxt-gate5.cu:55-60).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include "gpucpg_amd.hpp"
using namespace std::chrono_literals;
using gpucpg::CpGen;
using gpucpg::CsrReorderMethod;
using gpucpg::PfxtExpMethod;
using gpucpg::PropDistMethod;

static double short_long_pfxt_ms(const CpGen& cpgen) {
  double total_ms = 0.0;
  for (const auto& step_time : cpgen.short_long_step_times) total_ms += step_time / 1ms;
  return total_ms;
}

int main(int argc, char** argv) {
  const int k = std::atoi(argv[2]);
  CpGen cpgen;
  cpgen.read_input(argv[1]);
  cpgen.reset();
  cpgen.report_paths(k, 10, true, PropDistMethod::LEVELIZE_THEN_RELAX, PfxtExpMethod::SHORT_LONG,
                     false, 0.005f, 5.0f, 8, false, false, false, true, CsrReorderMethod::NONE,
                     false);
  const double pfxt_ms = short_long_pfxt_ms(cpgen);
  const std::vector<float> costs = cpgen.get_slacks(k);
  std::ofstream os(argv[3], std::ios::binary);
  os.write(reinterpret_cast<const char*>(costs.data()), costs.size() * sizeof(float));
  long long gen = 0;
  for (int p : cpgen.paths_gen_per_step) gen += p;
  std::printf("paths %zu steps %d step_times %zu paths_per_step %zu gen %lld total %d pfxt_ms %.6f "
              "expand_ms %.6f lvlize_ms %.6f relax_ms %.6f diameter %d\n",
              costs.size(), cpgen.short_long_expansion_steps, cpgen.short_long_step_times.size(),
              cpgen.paths_gen_per_step.size(), gen, cpgen.total_gen_paths, pfxt_ms,
              cpgen.expand_time / 1ms, cpgen.lvlize_time / 1ms, cpgen.relax_time / 1ms,
              cpgen.graph_diameter);
  return 0;
}
"""


def build_facade_query_driver(tmp_path):
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    libdir = os.path.join(ROOT, "gpu-cpg_amd")
    src = tmp_path / "query_driver.cpp"
    src.write_text(FACADE_QUERY_DRIVER)
    exe = tmp_path / "query_driver"
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", libdir, "-lcpgamd", "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    return exe


def test_cpp_facade_query_driver_compiles(tmp_path):
    """A driver reading short_long_step_times, paths_gen_per_step and the
    timing members compiles and links against the facade (it runs in the
    GPU suite)."""
    assert build_facade_query_driver(tmp_path).exists()
