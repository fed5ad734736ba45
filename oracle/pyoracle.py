"""ctypes binding of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the checker for the GPU engine: it is imported by tests/,
by __graft_entry__.smoke() and by the CPU-baseline leg of bench.py, and
never by the product package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class Graph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("m", ctypes.c_longlong),
                ("oadjp", ctypes.POINTER(ctypes.c_int)), ("oadj", ctypes.POINTER(ctypes.c_int)),
                ("owgt", ctypes.POINTER(ctypes.c_float)),
                ("iadjp", ctypes.POINTER(ctypes.c_int)), ("iadj", ctypes.POINTER(ctypes.c_int)),
                ("iwgt", ctypes.POINTER(ctypes.c_float))]


class Stats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_longlong) for name in (
        "pops", "pushes", "hops", "edges", "children", "nonmonotone", "zero_cost", "roots",
        "roots_dropped", "long_dummies", "steps", "short_pile", "final_window")] + [
        ("split_index_mismatch", ctypes.c_int), ("aborted", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Compare(ctypes.Structure):
    _fields_ = [("baseline_count", ctypes.c_longlong), ("tc_count", ctypes.c_longlong),
                ("compared", ctypes.c_longlong), ("max_diff", ctypes.c_float),
                ("max_diff_rank", ctypes.c_int), ("first_mismatch_rank", ctypes.c_int),
                ("pass_", ctypes.c_int)]


_PI = ctypes.POINTER(ctypes.c_int)
_PF = ctypes.POINTER(ctypes.c_float)
_PG = ctypes.POINTER(Graph)
_PS = ctypes.POINTER(Stats)
_LL = ctypes.c_longlong
_SIGS = {
    "orc_parse_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _PG]),
    "orc_build_csr": (ctypes.c_int, [ctypes.c_int, _LL, _PI, _PI, _PF, _PG]),
    "orc_graph_free": (None, [_PG]),
    "orc_last_error": (ctypes.c_char_p, []),
    "orc_levelize": (ctypes.c_int, [_PG, _PI, _PI, _PI, _PI]),
    "orc_relax": (None, [_PG, _PI, _PI, ctypes.c_int, _PI]),
    "orc_successors": (None, [_PG, _PI, _PI]),
    "orc_distf": (None, [ctypes.c_int, _PI, _PF]),
    "orc_sfxt": (ctypes.c_int, [_PG, _PI, _PI, _PI, _PI]),
    "orc_sfxt_masked": (ctypes.c_int, [_PG, ctypes.c_void_p, _PI, _PI, _PI, _PI]),
    "orc_candidate_cost": (ctypes.c_float, [ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float]),
    "orc_split_inc_amount": (ctypes.c_float, [ctypes.c_float]),
    "orc_classify": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                    ctypes.c_int]),
    "orc_deviations": (_LL, [ctypes.c_int, _PI, _PI, _PF, _PI, _PI, _PI, _PI, _PF]),
    "orc_topk": (_LL, [_PG, _PI, _PI, _LL, ctypes.c_int, _PF, _PS]),
    "orc_sequential": (_LL, [_PG, _PI, _PI, _LL, _PF, _PS]),
    "orc_short_long": (_LL, [_PG, _PI, _PI, _LL, ctypes.c_float, ctypes.c_int, ctypes.c_float,
                             ctypes.c_double, _LL, _PF, _PS]),
    "orc_all_paths": (_LL, [_PG, _PI, _PI, _LL, _PF]),
    "orc_topk_paths": (_LL, [_PG, _PI, _PI, _LL, ctypes.c_int, _PF, ctypes.POINTER(_LL), _PI,
                             _LL, _PS]),
    "orc_all_paths_vertices": (_LL, [_PG, _PI, _PI, _LL, _PF, ctypes.POINTER(_LL), _PI, _LL]),
    "orc_compare_prefix": (None, [_PF, _LL, _PF, _LL, ctypes.c_int, ctypes.c_float,
                                  ctypes.POINTER(Compare)]),
    "orc_first_bit_mismatch": (_LL, [_PF, _PF, _LL]),
    "orc_query_ms": (ctypes.c_double, [_PG, _LL, _PF, ctypes.POINTER(_LL),
                                       ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]),
}
_lib = None


class ShortLongAbort(RuntimeError):
    """The schedule reached a state where the reference's own code is
    undefined (st["aborted"] says which) or the node budget ran out."""

    def __init__(self, stats):
        super().__init__("oracle short_long aborted: %r" % stats)
        self.stats = stats


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(_PI)


def _fp(a):
    return a.ctypes.data_as(_PF)


class OracleGraph:
    """Owns an orc_graph (fanout + fanin CSR in the reference's order)."""

    def __init__(self, g):
        self.g = g

    @classmethod
    def from_text(cls, path, ignore_wgts=False):
        g = Graph()
        if lib().orc_parse_text(path.encode(), 1 if ignore_wgts else 0, ctypes.byref(g)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return cls(g)

    @classmethod
    def from_edges(cls, n, u, v, w):
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        w = np.ascontiguousarray(w, np.float32)
        g = Graph()
        if lib().orc_build_csr(int(n), len(u), _ip(u), _ip(v), _fp(w), ctypes.byref(g)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return cls(g)

    def __del__(self):
        try:
            lib().orc_graph_free(ctypes.byref(self.g))
        except Exception:
            pass

    @property
    def n(self):
        return self.g.n

    @property
    def m(self):
        return self.g.m

    def arrays(self):
        n, m = self.g.n, self.g.m
        def arr(p, cnt, dt):
            return np.ctypeslib.as_array(p, shape=(cnt,)).astype(dt, copy=True) if cnt else np.zeros(0, dt)
        return dict(oadjp=arr(self.g.oadjp, n + 1, np.int32), oadj=arr(self.g.oadj, m, np.int32),
                    owgt=arr(self.g.owgt, m, np.float32), iadjp=arr(self.g.iadjp, n + 1, np.int32),
                    iadj=arr(self.g.iadj, m, np.int32), iwgt=arr(self.g.iwgt, m, np.float32))

    def levelize(self):
        n = self.g.n
        level = np.zeros(max(n, 1), np.int32)
        queue = np.zeros(max(n, 1), np.int32)
        lvlp = np.zeros(n + 2, np.int32)
        L = ctypes.c_int(0)
        cnt = lib().orc_levelize(ctypes.byref(self.g), _ip(level), _ip(queue), _ip(lvlp), ctypes.byref(L))
        return level[:n], queue[:cnt], lvlp[:L.value + 1], L.value

    def sfxt(self, sink_mask=None):
        """sink_mask: optional n-byte endpoint subset (nonzero = chosen sink)."""
        n = self.g.n
        dist = np.zeros(max(n, 1), np.int32)
        succ = np.zeros(max(n, 1), np.int32)
        level = np.zeros(max(n, 1), np.int32)
        L = ctypes.c_int(0)
        if sink_mask is None:
            lib().orc_sfxt(ctypes.byref(self.g), _ip(dist), _ip(succ), _ip(level), ctypes.byref(L))
        else:
            m = np.ascontiguousarray(sink_mask, dtype=np.uint8)
            assert m.shape == (n,)
            lib().orc_sfxt_masked(ctypes.byref(self.g), m.ctypes.data, _ip(dist), _ip(succ),
                                  _ip(level), ctypes.byref(L))
        return dist[:n], succ[:n], level[:n], L.value

    def topk(self, dist, succ, k, root_tie_quirk=True):
        out = np.zeros(max(int(k), 1), np.float32)
        st = Stats()
        got = lib().orc_topk(ctypes.byref(self.g), _ip(dist), _ip(succ), int(k),
                             1 if root_tie_quirk else 0, _fp(out), ctypes.byref(st))
        if got < 0:
            raise RuntimeError("oracle top-k aborted")
        return out[:got], st.as_dict()

    def sequential(self, dist, succ, k):
        out = np.zeros(max(int(k), 1), np.float32)
        st = Stats()
        got = lib().orc_sequential(ctypes.byref(self.g), _ip(dist), _ip(succ), int(k), _fp(out),
                                   ctypes.byref(st))
        if got < 0:
            raise RuntimeError("oracle sequential aborted")
        return out[:got], st.as_dict()

    def short_long(self, dist, succ, k, init_split_perc=0.005, fixed_inc=None,
                   long_pile_limit_bytes=2e9, node_budget=50_000_000):
        out = np.zeros(max(int(k), 1), np.float32)
        st = Stats()
        got = lib().orc_short_long(ctypes.byref(self.g), _ip(dist), _ip(succ), int(k),
                                   float(init_split_perc), 0 if fixed_inc is None else 1,
                                   0.0 if fixed_inc is None else float(fixed_inc),
                                   float(long_pile_limit_bytes), int(node_budget), _fp(out),
                                   ctypes.byref(st))
        if got < 0:
            raise ShortLongAbort(st.as_dict())
        return out[:got], st.as_dict()

    def topk_paths(self, dist, succ, k, root_tie_quirk=True, max_len=None):
        """orc_topk with each path's vertex sequence: (costs, [vertex arrays], stats)."""
        k = int(k)
        if max_len is None:
            max_len = self.g.n + 1
        vcap = max(1, k * int(max_len))
        out = np.zeros(max(k, 1), np.float32)
        vptr = np.zeros(k + 1, np.int64)
        verts = np.zeros(vcap, np.int32)
        st = Stats()
        got = lib().orc_topk_paths(ctypes.byref(self.g), _ip(dist), _ip(succ), k,
                                   1 if root_tie_quirk else 0, _fp(out),
                                   vptr.ctypes.data_as(ctypes.POINTER(_LL)), _ip(verts), vcap,
                                   ctypes.byref(st))
        if got < 0:
            raise RuntimeError("oracle top-k paths aborted: %r" % st.as_dict())
        paths = [verts[vptr[i]:vptr[i + 1]].copy() for i in range(got)]
        return out[:got], paths, st.as_dict()

    def all_paths_vertices(self, dist, succ, cap=100_000, vcap=4_000_000):
        """Brute force: every path's (cost, vertices), DFS order, zero costs kept."""
        out = np.zeros(cap, np.float32)
        vptr = np.zeros(cap + 1, np.int64)
        verts = np.zeros(vcap, np.int32)
        got = lib().orc_all_paths_vertices(ctypes.byref(self.g), _ip(dist), _ip(succ), int(cap),
                                           _fp(out), vptr.ctypes.data_as(ctypes.POINTER(_LL)),
                                           _ip(verts), int(vcap))
        if got < 0:
            raise RuntimeError("too many paths for brute force")
        return out[:got], [verts[vptr[i]:vptr[i + 1]].copy() for i in range(got)]

    def all_paths(self, dist, succ, cap=1_000_000):
        out = np.zeros(cap, np.float32)
        got = lib().orc_all_paths(ctypes.byref(self.g), _ip(dist), _ip(succ), int(cap), _fp(out))
        if got < 0:
            raise RuntimeError("too many paths for brute force")
        return out[:got]

    def query_ms(self, k):
        out = np.zeros(max(int(k), 1), np.float32)
        cnt = ctypes.c_longlong(0)
        s = ctypes.c_double(0)
        p = ctypes.c_double(0)
        total = lib().orc_query_ms(ctypes.byref(self.g), int(k), _fp(out), ctypes.byref(cnt),
                                   ctypes.byref(s), ctypes.byref(p))
        return total, s.value, p.value, out[:cnt.value]


def distf(dist):
    dist = np.ascontiguousarray(dist, np.int32)
    out = np.zeros(max(len(dist), 1), np.float32)
    lib().orc_distf(len(dist), _ip(dist), _fp(out))
    return out[:len(dist)]


def candidate_cost(parent, src_dist, dst_dist, wgt):
    return lib().orc_candidate_cost(parent, src_dist, dst_dist, wgt)


def split_inc_amount(avg_deg):
    return lib().orc_split_inc_amount(avg_deg)


def compare_prefix(baseline, tc, k, tol=1e-3):
    b = np.ascontiguousarray(baseline, np.float32)
    t = np.ascontiguousarray(tc, np.float32)
    r = Compare()
    lib().orc_compare_prefix(_fp(b), len(b), _fp(t), len(t), int(k), float(tol), ctypes.byref(r))
    return dict(baseline_count=r.baseline_count, tc_count=r.tc_count, compared=r.compared,
                max_diff=r.max_diff, max_diff_rank=r.max_diff_rank,
                first_mismatch_rank=r.first_mismatch_rank, passed=bool(r.pass_))


def first_bit_mismatch(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    n = min(len(a), len(b))
    return lib().orc_first_bit_mismatch(_fp(a), _fp(b), n)


def classify(slack, split, final_split, use_final_split, skip_long_paths):
    return lib().orc_classify(slack, split, final_split, int(use_final_split), int(skip_long_paths))


def deviations(adjp, adj, wgt, succ, dist):
    adjp = np.ascontiguousarray(adjp, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    wgt = np.ascontiguousarray(wgt, np.float32)
    succ = np.ascontiguousarray(succ, np.int32)
    dist = np.ascontiguousarray(dist, np.int32)
    n = len(adjp) - 1
    offs = np.zeros(n + 1, np.int32)
    dsts = np.zeros(max(len(adj), 1), np.int32)
    dl = np.zeros(max(len(adj), 1), np.float32)
    c = lib().orc_deviations(n, _ip(adjp), _ip(adj), _fp(wgt), _ip(succ), _ip(dist), _ip(offs),
                             _ip(dsts), _fp(dl))
    return offs, dsts[:c], dl[:c]
