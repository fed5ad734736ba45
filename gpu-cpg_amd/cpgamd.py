"""ctypes binding of libcpgamd.so (include/cpg.h) and a CpGen-shaped wrapper.

The wrapper mirrors gpucpg::CpGen (gpucpg/gpucpg.cuh:119-431): read_input,
reset, report_paths, get_slacks, get_pfxt_nodes, num_verts, num_edges,
graph_diameter and the static-cache controls.  Every call goes through the
C ABI into the HIP engine; there is no Python or CPU fallback, and loading
fails with an exception when the library has not been built.
"""
import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpgamd.so")


class CpgNode(ctypes.Structure):
    _fields_ = [("level", ctypes.c_int), ("from_", ctypes.c_int), ("to", ctypes.c_int),
                ("parent", ctypes.c_int), ("num_children", ctypes.c_int),
                ("cost", ctypes.c_float)]


class QueryOpts(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_int), ("init_split_perc", ctypes.c_float),
                ("root_tie_quirk", ctypes.c_int), ("batch", ctypes.c_int64),
                ("max_nodes", ctypes.c_int64), ("mem_budget_bytes", ctypes.c_int64),
                ("endpoint_mask", ctypes.c_void_p), ("collect_stats", ctypes.c_int),
                ("kernel_timing", ctypes.c_int), ("sfxt_blocks", ctypes.c_int),
                ("bound_probe", ctypes.c_int), ("sfxt_method", ctypes.c_int),
                ("step_times", ctypes.c_int), ("monotone_check", ctypes.c_int),
                ("pad", ctypes.c_int)]


class Timing(ctypes.Structure):
    _fields_ = [("sfxt_ms", ctypes.c_double), ("pfxt_ms", ctypes.c_double),
                ("select_ms", ctypes.c_double), ("query_ms", ctypes.c_double),
                ("rounds", ctypes.c_int64), ("expanded", ctypes.c_int64),
                ("generated", ctypes.c_int64), ("hops", ctypes.c_int64),
                ("dev_visits", ctypes.c_int64), ("roots", ctypes.c_int64),
                ("roots_dropped", ctypes.c_int64), ("retries", ctypes.c_int64),
                ("sfxt_kernel_ms", ctypes.c_double), ("walk_count_ms", ctypes.c_double),
                ("walk_fill_ms", ctypes.c_double), ("walk_launches", ctypes.c_int64),
                ("sfxt_levels_ms", ctypes.c_double), ("sfxt_blocks", ctypes.c_int64),
                ("probes", ctypes.c_int64), ("bvss_ms", ctypes.c_double),
                ("bvss_tiles", ctypes.c_int64), ("device_bytes", ctypes.c_int64),
                ("node_slots", ctypes.c_int64), ("nonmonotone", ctypes.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "submitted", "admitted", "completed", "rounds", "compactions", "restarts",
        "peak_nodes", "node_cap", "peak_index_bytes", "in_flight", "queued", "sfxt_groups",
        "nonmonotone", "root_bounds")] + [
        (n, ctypes.c_double) for n in ("sfxt_ms", "group_wait_ms", "probe_ms", "round_ms",
                                       "extract_ms", "compact_ms")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchTuning(ctypes.Structure):
    _fields_ = [("group", ctypes.c_int), ("grid_mult", ctypes.c_int),
                ("target_div", ctypes.c_int * 3), ("trials", ctypes.c_int),
                ("start_ms", ctypes.c_double), ("best_ms", ctypes.c_double)]

    def as_dict(self):
        return {"group": self.group, "grid_mult": self.grid_mult,
                "target_div": list(self.target_div), "trials": self.trials,
                "start_ms": round(self.start_ms, 3), "best_ms": round(self.best_ms, 3)}


class EdgeList(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("m", ctypes.c_int64),
                ("u", ctypes.POINTER(ctypes.c_int)), ("v", ctypes.POINTER(ctypes.c_int)),
                ("w", ctypes.POINTER(ctypes.c_float))]


# name -> (restype, argtypes); the list is the C ABI declared in include/cpg.h
_H = ctypes.c_void_p
_I64 = ctypes.c_int64
_PI = ctypes.POINTER(ctypes.c_int)
_PF = ctypes.POINTER(ctypes.c_float)
ABI = {
    "cpg_create": (_H, []),
    "cpg_destroy": (None, [_H]),
    "cpg_last_error": (ctypes.c_char_p, []),
    "cpg_load_text": (ctypes.c_int, [_H, ctypes.c_char_p, ctypes.c_int]),
    "cpg_load_edges": (ctypes.c_int, [_H, ctypes.c_int, _I64, _PI, _PI, _PF]),
    "cpg_load_bin": (ctypes.c_int, [_H, ctypes.c_char_p]),
    "cpg_write_bin": (ctypes.c_int, [_H, ctypes.c_char_p]),
    "cpg_load_csr": (ctypes.c_int, [_H, ctypes.c_int, _I64, _PI, _PI, _PF]),
    "cpg_get_csr": (ctypes.c_int, [_H, _PI, _PI, _PF]),
    "cpg_num_verts": (_I64, [_H]),
    "cpg_num_edges": (_I64, [_H]),
    "cpg_graph_diameter": (ctypes.c_int, [_H]),
    "cpg_reset": (ctypes.c_int, [_H]),
    "cpg_query_opts_default": (None, [ctypes.POINTER(QueryOpts)]),
    "cpg_query": (ctypes.c_int, [_H, _I64, ctypes.POINTER(QueryOpts)]),
    "cpg_num_paths": (_I64, [_H]),
    "cpg_get_costs": (_I64, [_H, _I64, _PF]),
    "cpg_host_alloc": (ctypes.c_void_p, [_I64]),
    "cpg_host_free": (ctypes.c_int, [ctypes.c_void_p]),
    "cpg_get_nodes": (_I64, [_H, _I64, ctypes.POINTER(CpgNode)]),
    "cpg_get_path_vertices": (_I64, [_H, _I64, _PI, _I64]),
    "cpg_timings": (ctypes.c_int, [_H, ctypes.POINTER(Timing)]),
    "cpg_step_times": (_I64, [_H, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I64), _I64]),
    "cpg_set_csr_reorder": (ctypes.c_int, [_H, ctypes.c_int]),
    "cpg_reorder_ms": (ctypes.c_double, [_H]),
    "cpg_sfxt": (ctypes.c_int, [_H]),
    "cpg_sfxt_noalloc": (ctypes.c_int, [ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int]
                         + [ctypes.c_void_p] * 4),
    "cpg_sfxt_noalloc_max_verts": (ctypes.c_int, []),
    "cpg_get_sfxt": (ctypes.c_int, [_H, _PI, _PI, _PI]),
    "cpg_batch_create": (_H, []),
    "cpg_batch_destroy": (None, [_H]),
    "cpg_batch_add_edges": (ctypes.c_int, [_H, ctypes.c_int, _I64, _PI, _PI, _PF, _I64]),
    "cpg_batch_num_queries": (ctypes.c_int, [_H]),
    "cpg_batch_run": (ctypes.c_int, [_H, ctypes.POINTER(QueryOpts)]),
    "cpg_batch_autotune": (ctypes.c_int, [_H, ctypes.POINTER(QueryOpts), ctypes.c_int]),
    "cpg_batch_get_tuning": (ctypes.c_int, [_H, ctypes.POINTER(BatchTuning)]),
    "cpg_batch_set_tuning": (ctypes.c_int, [_H, ctypes.POINTER(BatchTuning)]),
    "cpg_batch_get_costs": (_I64, [_H, ctypes.c_int, _I64, _PF]),
    "cpg_batch_timings": (ctypes.c_int, [_H, ctypes.POINTER(Timing)]),
    "cpg_server_create": (_H, [_I64, ctypes.c_int, _I64]),
    "cpg_server_destroy": (None, [_H]),
    "cpg_server_add_graph": (ctypes.c_int, [_H, ctypes.c_int, _I64, _PI, _PI, _PF]),
    "cpg_server_submit": (_I64, [_H, ctypes.c_int, _I64, ctypes.c_void_p]),
    "cpg_server_step": (_I64, [_H]),
    "cpg_server_poll": (_I64, [_H, ctypes.POINTER(_I64), _I64]),
    "cpg_server_get_costs": (_I64, [_H, _I64, _I64, _PF]),
    "cpg_server_query_times": (ctypes.c_int, [_H, _I64, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double)]),
    "cpg_server_release": (ctypes.c_int, [_H, _I64]),
    "cpg_server_get_stats": (ctypes.c_int, [_H, ctypes.POINTER(ServerStats)]),
    "cpg_now_ms": (ctypes.c_double, []),
    "cpg_bvss_build": (ctypes.c_int, [_H]),
    "cpg_bvss_info": (ctypes.c_int, [_H, ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                                     ctypes.POINTER(_I64)]),
    "cpg_bvss_row_sources": (_I64, [_H, ctypes.c_int, _PI, _I64]),
    "cpg_bvss_discover": (_I64, [_H, ctypes.c_int, _PI, _PI, ctypes.c_int, ctypes.c_int,
                                 _PI, _PI, _PI, _I64]),
    "cpg_cache_enable": (ctypes.c_int, [_H, ctypes.c_int]),
    "cpg_cache_clear": (ctypes.c_int, [_H]),
    "cpg_cache_hits": (ctypes.c_int, [_H]),
    "cpg_cache_misses": (ctypes.c_int, [_H]),
    "cpg_gen_preset": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(EdgeList)]),
    "cpg_edge_list_free": (None, [ctypes.POINTER(EdgeList)]),
    "cpg_preset_k": (_I64, [ctypes.c_int, ctypes.c_int]),
    "cpg_write_text": (ctypes.c_int, [ctypes.POINTER(EdgeList), ctypes.c_char_p]),
}

_lib = None


def load_library(path):
    """Loads a build of the C ABI from an explicit path and declares its
    signatures.  The default (lib()) is the HIP build next to this file."""
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    for name, (res, args) in ABI.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib():
    """The HIP engine library; raises if it is missing."""
    global _lib
    if _lib is None:
        _lib = load_library(LIB_PATH)
    return _lib


class CpgError(RuntimeError):
    pass


class HostArray:
    """A float32 numpy array in pinned host memory (cpg_host_alloc), for
    result copies that go straight from the device.  The memory lives until
    close(); the array must not be used after that."""

    def __init__(self, n, library=None):
        self.L = library or lib()
        n = max(int(n), 1)
        self.p = self.L.cpg_host_alloc(4 * n)
        if not self.p:
            raise CpgError(self.L.cpg_last_error().decode())
        self.array = np.ctypeslib.as_array((ctypes.c_float * n).from_address(self.p))

    def close(self):
        if self.p:
            self.array = None
            self.L.cpg_host_free(self.p)
            self.p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if sys.is_finalizing():
            return
        self.close()


def _check(rc, L=None):
    if rc is None or (isinstance(rc, int) and rc < 0):
        raise CpgError((L or lib()).cpg_last_error().decode())
    return rc


def _iptr(a):
    return a.ctypes.data_as(_PI)


def _fptr(a):
    return a.ctypes.data_as(_PF)


def gen_preset(which, index=0, unit_wgts=False):
    """Seeded synthetic workload as (n, u, v, w) numpy arrays (host only)."""
    el = EdgeList()
    _check(lib().cpg_gen_preset(which, index, 1 if unit_wgts else 0, ctypes.byref(el)))
    try:
        m = el.m
        u = np.ctypeslib.as_array(el.u, shape=(m,)).copy() if m else np.zeros(0, np.int32)
        v = np.ctypeslib.as_array(el.v, shape=(m,)).copy() if m else np.zeros(0, np.int32)
        w = np.ctypeslib.as_array(el.w, shape=(m,)).copy() if m else np.zeros(0, np.float32)
        return el.n, u, v, w
    finally:
        lib().cpg_edge_list_free(ctypes.byref(el))


def preset_k(which, index=0):
    return _check(lib().cpg_preset_k(which, index))


def write_text(n, u, v, w, path):
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    w = np.ascontiguousarray(w, np.float32)
    el = EdgeList(n, len(u), _iptr(u), _iptr(v), _fptr(w))
    _check(lib().cpg_write_text(ctypes.byref(el), path.encode()))


def _c(g, rc):
    return _check(rc, g.L)


class CpGen:
    """gpucpg::CpGen-shaped handle over the C ABI."""

    def __init__(self, library=None):
        self.L = library or lib()
        self.h = self.L.cpg_create()
        if not self.h:
            raise CpgError(self.L.cpg_last_error().decode())
        self._k = 0

    def close(self):
        """Frees the handle's device memory, stream and events now.  Objects
        still open when the interpreter shuts down are left to the process
        exit: no HIP call is made from a finalizer then (the HIP runtime or
        a profiler attached to it may already be torn down)."""
        if self.h:
            self.L.cpg_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    # -- input (gpucpg/gpucpg.cu:453-548)
    def read_input(self, filename, ignore_wgts=False):
        _c(self, self.L.cpg_load_text(self.h, filename.encode(), 1 if ignore_wgts else 0))

    def read_input_bin(self, filename):
        """Binary fanout CSR (gpucpg/graph.h:24-55)."""
        _c(self, self.L.cpg_load_bin(self.h, filename.encode()))

    def write_to_csr_bin(self, filename):
        _c(self, self.L.cpg_write_bin(self.h, filename.encode()))

    def load_edges(self, n, u, v, w):
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        w = np.ascontiguousarray(w, np.float32)
        _c(self, self.L.cpg_load_edges(self.h, int(n), len(u), _iptr(u), _iptr(v), _fptr(w)))

    def load_csr(self, adjp, adjncy, wgts):
        adjp = np.ascontiguousarray(adjp, np.int32)
        adjncy = np.ascontiguousarray(adjncy, np.int32)
        wgts = np.ascontiguousarray(wgts, np.float32)
        _c(self, self.L.cpg_load_csr(self.h, len(adjp) - 1, len(adjncy), _iptr(adjp),
                                  _iptr(adjncy), _fptr(wgts)))

    def csr(self):
        n, m = self.num_verts(), self.num_edges()
        adjp = np.zeros(n + 1, np.int32)
        adj = np.zeros(max(m, 1), np.int32)
        w = np.zeros(max(m, 1), np.float32)
        _c(self, self.L.cpg_get_csr(self.h, _iptr(adjp), _iptr(adj), _fptr(w)))
        return adjp, adj[:m], w[:m]

    def num_verts(self):
        return self.L.cpg_num_verts(self.h)

    def num_edges(self):
        return self.L.cpg_num_edges(self.h)

    @property
    def graph_diameter(self):
        return self.L.cpg_graph_diameter(self.h)

    # -- query (gpucpg/gpucpg.cuh:211-250, gpucpg/gpucpg.cu:16910-21366)
    def reset(self):
        _c(self, self.L.cpg_reset(self.h))

    def report_paths(self, k, *args, root_tie_quirk=True, batch=0, max_nodes=0,
                     collect_stats=False, init_split_perc=0.005, kernel_timing=False,
                     sfxt_blocks=0, bound_probe=0, endpoint_mask=None, sfxt_method=0,
                     step_times=False, monotone_check=False, mem_budget_bytes=0, **kwargs):
        """Runs the GPG query for the k lowest-cost paths.  Positional
        arguments after k (max_dev_lvls, enable_compress, methods, ...) are
        accepted for call compatibility with the reference; the engine always
        runs its exact GPU path.  endpoint_mask: optional array of n bytes
        (nonzero = chosen sink) restricting the paths to those ending at the
        chosen sinks."""
        o = QueryOpts()
        self.L.cpg_query_opts_default(ctypes.byref(o))
        o.root_tie_quirk = 1 if root_tie_quirk else 0
        o.batch = int(batch)
        o.max_nodes = int(max_nodes)
        o.collect_stats = 1 if collect_stats else 0
        o.init_split_perc = float(init_split_perc)
        o.kernel_timing = 1 if kernel_timing else 0
        o.sfxt_blocks = int(sfxt_blocks)
        o.bound_probe = int(bound_probe)
        o.sfxt_method = int(sfxt_method)
        o.step_times = 1 if step_times else 0
        o.monotone_check = 1 if monotone_check else 0
        o.mem_budget_bytes = int(mem_budget_bytes)
        mask = None
        if endpoint_mask is not None:
            mask = np.ascontiguousarray(endpoint_mask, dtype=np.uint8)
            if mask.shape != (self.num_verts(),):
                raise ValueError("endpoint_mask must hold one byte per vertex")
            o.endpoint_mask = mask.ctypes.data
        _c(self, self.L.cpg_query(self.h, int(k), ctypes.byref(o)))
        self._k = int(k)

    def num_paths(self):
        return self.L.cpg_num_paths(self.h)

    def get_slacks(self, k, out=None):
        """First k costs (ascending).  out: a float32 array of at least k
        elements to fill (a HostArray's, pinned, avoids staging); the
        returned array is a view of it."""
        if out is None:
            out = np.empty(max(int(k), 1), np.float32)
        elif out.dtype != np.float32 or not out.flags.c_contiguous or len(out) < max(int(k), 1):
            raise ValueError("out must be a contiguous float32 array of at least k elements")
        n = _c(self, self.L.cpg_get_costs(self.h, int(k), _fptr(out)))
        return out[:n]

    def get_pfxt_nodes(self, k):
        k = max(int(k), 0)
        buf = (CpgNode * max(k, 1))()
        n = _c(self, self.L.cpg_get_nodes(self.h, k, buf))
        arr = np.ctypeslib.as_array(buf)[:n]
        return np.array(arr, copy=True)

    def path_vertices(self, i):
        cap = 1 << 16
        while True:
            out = np.zeros(cap, np.int32)
            n = self.L.cpg_get_path_vertices(self.h, int(i), _iptr(out), cap)
            if n >= 0:
                return out[:n]
            if n <= -1000:
                cap = -(n + 1000) + 1
                continue
            raise CpgError(self.L.cpg_last_error().decode())

    def timings(self):
        t = Timing()
        _c(self, self.L.cpg_timings(self.h, ctypes.byref(t)))
        return t.as_dict()

    def step_times(self):
        """(per-round device ms, per-round stored paths) of the last query
        run with step_times=True."""
        n = _c(self, self.L.cpg_step_times(self.h, None, None, 0))
        ms = (ctypes.c_double * max(n, 1))()
        paths = (_I64 * max(n, 1))()
        n = _c(self, self.L.cpg_step_times(self.h, ms, paths, n))
        return [ms[i] for i in range(n)], [paths[i] for i in range(n)]

    def sfxt(self):
        n = self.num_verts()
        dist = np.zeros(max(n, 1), np.int32)
        succ = np.zeros(max(n, 1), np.int32)
        level = np.zeros(max(n, 1), np.int32)
        _c(self, self.L.cpg_get_sfxt(self.h, _iptr(dist), _iptr(succ), _iptr(level)))
        return dist[:n], succ[:n], level[:n]

    def run_sfxt(self):
        _c(self, self.L.cpg_sfxt(self.h))

    def set_csr_reorder(self, method):
        """CsrReorderMethod of the resident graph: 0 NONE, 1 E_ORIENTED
        (level-order relabelling on the device; results unchanged)."""
        _c(self, self.L.cpg_set_csr_reorder(self.h, int(method)))

    def reorder_ms(self):
        return float(self.L.cpg_reorder_ms(self.h))

    # -- BVSS deviation discovery (gpucpg/tc_pfxt_bvss.cuh)
    def bvss_build(self):
        _c(self, self.L.cpg_bvss_build(self.h))

    def bvss_info(self):
        a, b, c = _I64(), _I64(), _I64()
        _c(self, self.L.cpg_bvss_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"intervals": a.value, "entries": b.value, "set_bits": c.value}

    def bvss_row_sources(self, row):
        cap = 64
        while True:
            out = np.zeros(cap, np.int32)
            n = self.L.cpg_bvss_row_sources(self.h, int(row), _iptr(out), cap)
            if n >= 0:
                return out[:n]
            if n <= -1000:
                cap = -(n + 1000)
                continue
            _c(self, -1)

    def bvss_discover(self, frontiers, mode=0, dense_min=4):
        """Deviation edges leaving each frontier (a list of source-vertex
        arrays); returns sorted (frontier, u, v) arrays."""
        fptr = np.zeros(len(frontiers) + 1, np.int32)
        for i, fr in enumerate(frontiers):
            fptr[i + 1] = fptr[i] + len(fr)
        fsrc = np.concatenate([np.asarray(fr, np.int32) for fr in frontiers]) if frontiers \
            else np.zeros(0, np.int32)
        fsrc = np.ascontiguousarray(fsrc if len(fsrc) else np.zeros(1, np.int32), np.int32)
        cap = 1 << 16
        while True:
            of = np.zeros(cap, np.int32)
            ou = np.zeros(cap, np.int32)
            ov = np.zeros(cap, np.int32)
            n = self.L.cpg_bvss_discover(self.h, len(frontiers), _iptr(fptr), _iptr(fsrc), int(mode),
                                         int(dense_min), _iptr(of), _iptr(ou), _iptr(ov), cap)
            if n >= 0:
                return of[:n], ou[:n], ov[:n]
            if n <= -1000:
                cap = -(n + 1000)
                continue
            _c(self, -1)

    # -- static cache (gpucpg/gpucpg.cuh:186-189)
    def enable_tc_pfxt_static_cache(self, enabled):
        _c(self, self.L.cpg_cache_enable(self.h, 1 if enabled else 0))

    def clear_tc_pfxt_static_cache(self):
        _c(self, self.L.cpg_cache_clear(self.h))

    def tc_pfxt_static_cache_hits(self):
        return self.L.cpg_cache_hits(self.h)

    def tc_pfxt_static_cache_misses(self):
        return self.L.cpg_cache_misses(self.h)


class Batch:
    """Many designs, one K-path query each, solved together (cpg_batch_*)."""

    def __init__(self, library=None):
        self.L = library or lib()
        self.h = self.L.cpg_batch_create()
        if not self.h:
            raise RuntimeError(self.L.cpg_last_error().decode())

    def close(self):
        if self.h:
            self.L.cpg_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def add(self, n, u, v, w, k):
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        w = np.ascontiguousarray(w, np.float32)
        return _check(self.L.cpg_batch_add_edges(self.h, int(n), len(u), _iptr(u), _iptr(v),
                                                 _fptr(w), int(k)), self.L)

    def autotune(self, max_trials=16):
        """Chooses the round shapes by timing whole batch runs
        (cpg_batch_autotune); returns the chosen shapes."""
        o = QueryOpts()
        self.L.cpg_query_opts_default(ctypes.byref(o))
        _check(self.L.cpg_batch_autotune(self.h, ctypes.byref(o), int(max_trials)), self.L)
        return self.tuning()

    def tuning(self):
        t = BatchTuning()
        _check(self.L.cpg_batch_get_tuning(self.h, ctypes.byref(t)), self.L)
        return t.as_dict()

    def set_tuning(self, group=4, grid_mult=8, target_div=(8, 8, 8)):
        t = BatchTuning()
        t.group = int(group)
        t.grid_mult = int(grid_mult)
        for i in range(3):
            t.target_div[i] = int(target_div[i])
        _check(self.L.cpg_batch_set_tuning(self.h, ctypes.byref(t)), self.L)

    def run(self, root_tie_quirk=True, init_split_perc=0.005, sfxt_blocks=0, sfxt_method=0):
        o = QueryOpts()
        self.L.cpg_query_opts_default(ctypes.byref(o))
        o.root_tie_quirk = 1 if root_tie_quirk else 0
        o.init_split_perc = float(init_split_perc)
        o.sfxt_blocks = int(sfxt_blocks)
        o.sfxt_method = int(sfxt_method)
        _check(self.L.cpg_batch_run(self.h, ctypes.byref(o)), self.L)

    def get_slacks(self, q, k):
        out = np.zeros(max(int(k), 1), np.float32)
        n = _check(self.L.cpg_batch_get_costs(self.h, int(q), int(k), _fptr(out)), self.L)
        return out[:n]

    def timings(self):
        t = Timing()
        _check(self.L.cpg_batch_timings(self.h, ctypes.byref(t)), self.L)
        return t.as_dict()


class Server:
    """Resident graphs and a stream of top-K queries with endpoint subsets,
    their PFXT rounds batched continuously (cpg_server_*)."""

    def __init__(self, budget_bytes=0, max_in_flight=0, initial_nodes=0, library=None):
        self.L = library or lib()
        self.h = self.L.cpg_server_create(int(budget_bytes), int(max_in_flight), int(initial_nodes))
        if not self.h:
            raise CpgError(self.L.cpg_last_error().decode())
        self._n = []

    def close(self):
        if self.h:
            self.L.cpg_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def add_graph(self, n, u, v, w):
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        w = np.ascontiguousarray(w, np.float32)
        gid = _check(self.L.cpg_server_add_graph(self.h, int(n), len(u), _iptr(u), _iptr(v),
                                                 _fptr(w)), self.L)
        self._n.append(int(n))
        return gid

    def submit(self, graph, k, endpoint_mask=None):
        ptr = None
        if endpoint_mask is not None:
            m = np.ascontiguousarray(endpoint_mask, dtype=np.uint8)
            if m.shape != (self._n[graph],):
                raise ValueError("endpoint_mask must hold one byte per vertex")
            ptr = m.ctypes.data
        return _check(self.L.cpg_server_submit(self.h, int(graph), int(k), ptr), self.L)

    def step(self):
        return _check(self.L.cpg_server_step(self.h), self.L)

    def poll(self, cap=4096):
        ids = (ctypes.c_int64 * cap)()
        n = _check(self.L.cpg_server_poll(self.h, ids, cap), self.L)
        return [ids[i] for i in range(n)]

    def get_slacks(self, qid, k):
        out = np.zeros(max(int(k), 1), np.float32)
        n = _check(self.L.cpg_server_get_costs(self.h, int(qid), int(k), _fptr(out)), self.L)
        return out[:n]

    def times(self, qid):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(self.L.cpg_server_query_times(self.h, int(qid), ctypes.byref(a), ctypes.byref(b),
                                             ctypes.byref(c)), self.L)
        return a.value, b.value, c.value

    def release(self, qid):
        _check(self.L.cpg_server_release(self.h, int(qid)), self.L)

    def stats(self):
        st = ServerStats()
        _check(self.L.cpg_server_get_stats(self.h, ctypes.byref(st)), self.L)
        return st.as_dict()


def now_ms():
    return lib().cpg_now_ms()
