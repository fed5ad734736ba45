"""cpg_sfxt_noalloc (the SFXT with every intermediate buffer in a static
device array) against cpg_sfxt + cpg_get_sfxt on the same graphs: two calls
in a row, the second on a smaller graph, each equal to the handle's
dist/succ/level exactly; a graph above the maximum is refused with an
error.

Device inputs and outputs are allocated with the HIP runtime the engine
library itself is linked against (looked up by its soname once the library
is loaded), so the process holds one HIP runtime whatever ran before."""
import ctypes

import numpy as np
import pytest

import cpgamd

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the library's own HIP runtime."""

    def __init__(self):
        cpgamd.lib()  # loads libamdhip64 as the library's dependency
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def ok(self, rc, what):
        assert rc == 0, f"{what} returned HIP error {rc}"

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.ok(self.rt.hipMalloc(ctypes.byref(p), max(int(nbytes), 4)), "hipMalloc")
        self.live.append(p.value)
        return p.value

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        if arr.nbytes:
            self.ok(self.rt.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")
        return p

    def get(self, p, n, dtype=np.int32):
        out = np.empty(n, dtype)
        if out.nbytes:
            self.ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free_all(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def device_graph(hip, n, u, v, w):
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
    return hip.put(oadjp), hip.put(iadjp), hip.put(rec.reshape(-1)), hip.put(sinks), len(sinks)


def run_noalloc(hip, L, n, u, v, w):
    oadjp, iadjp, rec, sinks, ns = device_graph(hip, n, u, v, w)
    out = [hip.put(np.full(n, -7, np.int32)) for _ in range(3)]
    rc = L.cpg_sfxt_noalloc(n, oadjp, iadjp, rec, sinks, ns, out[0], out[1], out[2], None)
    if rc < 0:
        raise cpgamd.CpgError(L.cpg_last_error().decode())
    return rc, [hip.get(o, n) for o in out]


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
    L = cpgamd.lib()
    hip = Hip()
    graphs = [cpgamd.gen_preset(1), cpgamd.gen_preset(4, 0)]
    assert graphs[1][0] < graphs[0][0]  # the second call is on a smaller graph
    try:
        for n, u, v, w in graphs:
            print(f"case noalloc sfxt n={n} m={len(u)}", flush=True)
            levels, (dist, succ, level) = run_noalloc(hip, L, n, u, v, w)
            hl, hd, hs, hlv = handle_sfxt(n, u, v, w)
            assert levels == hl
            np.testing.assert_array_equal(dist, hd)
            np.testing.assert_array_equal(succ, hs)
            np.testing.assert_array_equal(level, hlv)
    finally:
        hip.free_all()


def test_noalloc_rejects_graphs_above_the_maximum():
    L = cpgamd.lib()
    hip = Hip()
    try:
        p = hip.alloc(16)
        big = L.cpg_sfxt_noalloc_max_verts() + 1
        print(f"case noalloc refuses n={big}", flush=True)
        rc = L.cpg_sfxt_noalloc(big, p, p, p, p, 1, p, p, p, None)
        assert rc < 0 and "at most" in L.cpg_last_error().decode()
    finally:
        hip.free_all()
