"""cpg_sfxt_noalloc (the SFXT with every intermediate buffer in a static
device array) against cpg_sfxt + cpg_get_sfxt on the same graphs: two calls
in a row, the second on a smaller graph, each equal to the handle's
dist/succ/level exactly; a graph above the maximum is refused with an
error.  Device inputs and outputs are torch tensors (device memory only)."""
import ctypes

import numpy as np
import pytest

import cpgamd

pytestmark = pytest.mark.gpu


def device_graph(torch, n, u, v, w):
    """Fanout offsets, fanin offsets and records {fanin, weight bits} and
    the sink list, on the device."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    w = np.ascontiguousarray(w, np.float32)
    outdeg = np.bincount(u, minlength=n)
    oadjp = np.zeros(n + 1, np.int32)
    oadjp[1:] = np.cumsum(outdeg)
    order = np.argsort(v, kind="stable")
    indeg = np.bincount(v, minlength=n)
    iadjp = np.zeros(n + 1, np.int32)
    iadjp[1:] = np.cumsum(indeg)
    rec = np.zeros((len(u), 2), np.int32)
    rec[:, 0] = u[order]
    rec[:, 1] = w[order].view(np.int32)
    sinks = np.nonzero(outdeg == 0)[0].astype(np.int32)
    dev = "cuda:0"
    return (torch.from_numpy(oadjp).to(dev), torch.from_numpy(iadjp).to(dev),
            torch.from_numpy(rec.reshape(-1).copy()).to(dev), torch.from_numpy(sinks).to(dev),
            len(sinks))


def run_noalloc(torch, L, n, u, v, w):
    oadjp, iadjp, rec, sinks, ns = device_graph(torch, n, u, v, w)
    out = [torch.full((n,), -7, dtype=torch.int32, device="cuda:0") for _ in range(3)]
    torch.cuda.synchronize()
    rc = L.cpg_sfxt_noalloc(n, oadjp.data_ptr(), iadjp.data_ptr(), rec.data_ptr(), sinks.data_ptr(),
                            ns, out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(), None)
    if rc < 0:
        raise cpgamd.CpgError(L.cpg_last_error().decode())
    return rc, [o.cpu().numpy() for o in out]


def handle_sfxt(n, u, v, w):
    g = cpgamd.CpGen()
    try:
        g.load_edges(n, u, v, w)
        g.run_sfxt()
        d, s, lv = g.sfxt()
        return g.graph_diameter, d.copy(), s.copy(), lv.copy()
    finally:
        g.close()


def test_noalloc_sfxt_equals_handle_sfxt_twice():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = cpgamd.lib()
    graphs = [cpgamd.gen_preset(1), cpgamd.gen_preset(4, 0)]
    assert graphs[1][0] < graphs[0][0]  # the second call is on a smaller graph
    for n, u, v, w in graphs:
        print(f"case noalloc sfxt n={n} m={len(u)}", flush=True)
        levels, (dist, succ, level) = run_noalloc(torch, L, n, u, v, w)
        hl, hd, hs, hlv = handle_sfxt(n, u, v, w)
        assert levels == hl
        np.testing.assert_array_equal(dist, hd)
        np.testing.assert_array_equal(succ, hs)
        np.testing.assert_array_equal(level, hlv)


def test_noalloc_rejects_graphs_above_the_maximum():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = cpgamd.lib()
    big = L.cpg_sfxt_noalloc_max_verts() + 1
    d = torch.zeros(4, dtype=torch.int32, device="cuda:0")
    p = d.data_ptr()
    rc = L.cpg_sfxt_noalloc(big, p, p, p, p, 1, p, p, p, None)
    assert rc < 0 and "at most" in L.cpg_last_error().decode()
