"""ctypes wrapper of the CPU oracle (liboracle.so) and of the reference shim
(_ref/libcchref.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, where it is the checker or the timed CPU baseline; the product path never
uses it.  Function-by-function reference citations are in cch_oracle.c.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
INF = 2147483647
_o = None
_r = None


def _p(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def olib():
    global _o
    if _o is None:
        L = C.CDLL(str(HERE / "liboracle.so"))
        p, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        L.orc_symbolic.argtypes = [i64, p, p, p, p, p, C.POINTER(C.c_void_p)]
        L.orc_symbolic.restype = i64
        L.orc_free.argtypes = [p]
        L.orc_map_arcs.argtypes = [i64, p, p, p, p, p, p, p]
        L.orc_map_arcs.restype = C.c_int
        L.orc_triangles.argtypes = [i64, p]
        L.orc_triangles.restype = i64
        L.orc_levels.argtypes = [i64, p, p, p]
        L.orc_customize.argtypes = [i64, p, p, i64, p, p, p, p, p]
        L.orc_customize.restype = C.c_int
        L.orc_customize_par.argtypes = [i64, p, p, i64, p, p, p, p, p, C.c_int]
        L.orc_customize_par.restype = C.c_int
        L.orc_customize_stop.argtypes = [i64, p, p, i64, p, p, p, p, p]
        L.orc_customize_stop.restype = i64
        L.orc_negcycle.argtypes = [i64, p, p, p]
        L.orc_negcycle.restype = i64
        L.orc_query_many.argtypes = [i64, p, p, p, p, p, i64, p, p, p, p, C.c_int]
        L.orc_query_many.restype = C.c_int
        L.orc_upd_create.argtypes = [i64, p, p, i64, p, p]
        L.orc_upd_create.restype = C.c_void_p
        L.orc_upd_destroy.argtypes = [p]
        L.orc_update.argtypes = [p, p, p, p, p, p, i64, p, p, p, p, p]
        L.orc_update.restype = i64
        L.orc_base.argtypes = [i64, i64, p, p, p, p, p]
        L.orc_bellman_ford.argtypes = [i64, p, p, p, i32, p]
        L.orc_bellman_ford.restype = C.c_int
        L.orc_has_negative_cycle.argtypes = [i64, p, p, p]
        L.orc_has_negative_cycle.restype = C.c_int
        L.orc_query_path.argtypes = [i64, p, p, p, p, p, p, p, p, i32, i32, p, i64, p]
        L.orc_query_path.restype = i64
        L.orc_johnson.argtypes = [i64, p, p, p, p]
        L.orc_johnson.restype = C.c_int
        L.orc_dijkstra_pot.argtypes = [i64, p, p, p, p, i32, i32, p, p, p, p]
        L.orc_dijkstra_pot.restype = i64
        L.orc_restricted.argtypes = [i64, p, p, p, p, i32, i32, p]
        L.orc_restricted.restype = i64
        _o = L
    return _o


def have_ref() -> bool:
    return (HERE / "_ref" / "libcchref.so").exists()


def rlib():
    global _r
    if _r is None:
        L = C.CDLL(str(HERE / "_ref" / "libcchref.so"))
        p, i64 = C.c_void_p, C.c_int64
        L.ref_build.argtypes = [i64, p, p, p, p, p]
        L.ref_build.restype = C.c_void_p
        L.ref_destroy.argtypes = [p]
        L.ref_info.argtypes = [p, p]
        L.ref_export.argtypes = [p, p, p, p, p, p, p]
        L.ref_customize.argtypes = [p, p, p, p, C.c_int]
        L.ref_negcycle.argtypes = [p, p, p]
        L.ref_negcycle.restype = i64
        L.ref_query.argtypes = [p, p, p, i64, p, p, p]
        L.ref_query_par.argtypes = [p, p, p, i64, p, p, p, C.c_int]
        L.ref_levels.argtypes = [p, p]
        L.ref_load_dir.argtypes = [C.c_char_p, C.c_char_p]
        L.ref_load_dir.restype = C.c_void_p
        L.ref_graph_info.argtypes = [p, p]
        L.ref_graph_export.argtypes = [p, p, p, p, p, C.c_int32, p]
        L.ref_potential.argtypes = [p, p, p, p]
        _r = L
    return _r


# ------------------------------------------------------------------ structure

class Structure:
    """first/head/parent/e_arc/e_up of one graph + order, by the oracle."""

    def __init__(self, n, off, tgt, rank):
        L = olib()
        self.n = n
        self.off = np.ascontiguousarray(off, np.int64)
        self.tgt = np.ascontiguousarray(tgt, np.int32)
        self.rank = np.ascontiguousarray(rank, np.int32)
        self.first = np.empty(n + 1, np.int64)
        self.parent = np.empty(n, np.int32)
        hp = C.c_void_p()
        mp = L.orc_symbolic(n, _p(self.off), _p(self.tgt), _p(self.rank), _p(self.first),
                            _p(self.parent), C.byref(hp))
        if mp < 0:
            raise ValueError("order must be a permutation")
        self.mp = int(mp)
        self.head = np.ctypeslib.as_array(C.cast(hp, C.POINTER(C.c_int32)), shape=(max(self.mp, 1),))[:self.mp].copy()
        L.orc_free(hp)
        self.m = int(self.off[-1])
        self.e_arc = np.empty(self.m, np.int64)
        self.e_up = np.empty(self.m, np.int8)
        if L.orc_map_arcs(n, _p(self.off), _p(self.tgt), _p(self.rank), _p(self.first), _p(self.head),
                          _p(self.e_arc), _p(self.e_up)) != 0:
            raise ValueError("arc missing from the contraction")
        self.triangles = int(L.orc_triangles(n, _p(self.first)))
        self.level = np.empty(n, np.int32)
        L.orc_levels(n, _p(self.first), _p(self.head), _p(self.level))
        depth = np.zeros(n, np.int64)
        for x in range(n - 1, -1, -1):
            p = self.parent[x]
            if p >= 0:
                depth[x] = depth[p] + 1
        self.vdepth = depth
        self.depth = int(depth.max(initial=0))

    def customize(self, w):
        w = np.ascontiguousarray(w, np.int32)
        up = np.empty(self.mp, np.int32)
        dn = np.empty(self.mp, np.int32)
        ovf = olib().orc_customize(self.n, _p(self.first), _p(self.head), self.m, _p(self.e_arc),
                                   _p(self.e_up), _p(w), _p(up), _p(dn))
        return up, dn, bool(ovf)

    def customize_stop(self, w, with_state=False):
        """Early-stop witness of python/cch.py:270-278, or -1; with_state also
        returns the (up, dn) values at the stop."""
        w = np.ascontiguousarray(w, np.int32)
        up = np.empty(self.mp, np.int32)
        dn = np.empty(self.mp, np.int32)
        wit = int(olib().orc_customize_stop(self.n, _p(self.first), _p(self.head), self.m, _p(self.e_arc),
                                            _p(self.e_up), _p(w), _p(up), _p(dn)))
        return (wit, up, dn) if with_state else wit

    def customize_par(self, w, threads=8):
        w = np.ascontiguousarray(w, np.int32)
        up = np.empty(self.mp, np.int32)
        dn = np.empty(self.mp, np.int32)
        ovf = olib().orc_customize_par(self.n, _p(self.first), _p(self.head), self.m, _p(self.e_arc),
                                       _p(self.e_up), _p(w), _p(up), _p(dn), threads)
        return up, dn, bool(ovf)

    def negcycle(self, up, dn):
        fa = np.zeros(1, np.int64)
        up = np.ascontiguousarray(up, np.int32)
        dn = np.ascontiguousarray(dn, np.int32)
        assert up.shape == dn.shape
        c = olib().orc_negcycle(len(up), _p(up), _p(dn), _p(fa))
        return int(c), int(fa[0])

    def query_ranks(self, up, dn, qs, qt, threads=1):
        qs = np.ascontiguousarray(qs, np.int32)
        qt = np.ascontiguousarray(qt, np.int32)
        dist = np.empty(len(qs), np.int32)
        sc = np.empty(len(qs), np.int64)
        ovf = olib().orc_query_many(self.n, _p(self.first), _p(self.head), _p(self.parent),
                                    _p(np.ascontiguousarray(up, np.int32)),
                                    _p(np.ascontiguousarray(dn, np.int32)), len(qs), _p(qs), _p(qt),
                                    _p(dist), _p(sc), threads)
        return dist, sc, bool(ovf)

    def query(self, up, dn, s, t, threads=1):
        """Original vertex ids."""
        return self.query_ranks(up, dn, self.rank[np.asarray(s)], self.rank[np.asarray(t)], threads)

    def query_path(self, up, dn, s, t):
        """(dist, [original vertex ids]) of python/cch.py:406-507; INF, [] if unreachable;
        (dist, None) if the unpacking passed 4n + 16 vertices (a negative cycle)."""
        if not hasattr(self, "_lower"):
            tail = np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.first)).astype(np.int32)
            order = np.argsort(self.head, kind="stable")
            lnbr = np.ascontiguousarray(tail[order], np.int32)
            lfirst = np.zeros(self.n + 1, np.int64)
            np.add.at(lfirst, self.head.astype(np.int64) + 1, 1)
            self._lower = (tail, np.cumsum(lfirst), lnbr)
            self._inv = np.argsort(self.rank).astype(np.int32)
        tail, lfirst, lnbr = self._lower
        cap = 4 * self.n + 16
        path = np.empty(cap, np.int32)
        ln = np.zeros(1, np.int64)
        d = olib().orc_query_path(self.n, _p(self.first), _p(self.head), _p(self.parent), _p(tail), _p(lfirst),
                                  _p(lnbr), _p(np.ascontiguousarray(up, np.int32)),
                                  _p(np.ascontiguousarray(dn, np.int32)), int(self.rank[s]), int(self.rank[t]),
                                  _p(path), cap, _p(ln))
        L = int(ln[0])
        if L == -2:
            return int(d), None        # unpacking passed 4n + 16 vertices: a negative cycle
        return int(d), ([] if L < 0 else self._inv[path[:L]].tolist())


class Updater:
    """python/cch.py:289-402 restated; keeps w, up, dn, base_up, base_dn."""

    def __init__(self, st: Structure, w, up, dn):
        self.st = st
        self.h = olib().orc_upd_create(st.n, _p(st.first), _p(st.head), st.m, _p(st.e_arc), _p(st.e_up))
        self.w = np.array(w, np.int32)
        self.up = np.array(up, np.int32)
        self.dn = np.array(dn, np.int32)
        self.bu = np.empty(st.mp, np.int32)
        self.bd = np.empty(st.mp, np.int32)
        olib().orc_base(st.m, st.mp, _p(st.e_arc), _p(st.e_up), _p(self.w), _p(self.bu), _p(self.bd))

    def update(self, edge_idx, new_w):
        ei = np.ascontiguousarray(edge_idx, np.int64)
        nw = np.ascontiguousarray(new_w, np.int32)
        changed = np.empty(max(self.st.mp, 1), np.int32)
        nc = np.zeros(1, np.int64)
        ovf = C.c_int(0)
        ev = olib().orc_update(self.h, _p(self.w), _p(self.up), _p(self.dn), _p(self.bu), _p(self.bd),
                               len(ei), _p(ei), _p(nw), _p(changed), _p(nc), C.byref(ovf))
        return int(ev), changed[:int(nc[0])].copy(), bool(ovf.value)

    def __del__(self):
        try:
            olib().orc_upd_destroy(self.h)
        except Exception:
            pass


# ------------------------------------------------------------------ checkers

def bellman_ford(n, off, tgt, w, s):
    dist = np.empty(n, np.int64)
    neg = olib().orc_bellman_ford(n, _p(np.ascontiguousarray(off, np.int64)),
                                  _p(np.ascontiguousarray(tgt, np.int32)),
                                  _p(np.ascontiguousarray(w, np.int32)), int(s), _p(dist))
    return dist, bool(neg)


def has_negative_cycle(n, off, tgt, w):
    return bool(olib().orc_has_negative_cycle(n, _p(np.ascontiguousarray(off, np.int64)),
                                              _p(np.ascontiguousarray(tgt, np.int32)),
                                              _p(np.ascontiguousarray(w, np.int32))))


def johnson(n, off, tgt, w):
    p = np.empty(n, np.int64)
    neg = olib().orc_johnson(n, _p(np.ascontiguousarray(off, np.int64)),
                             _p(np.ascontiguousarray(tgt, np.int32)),
                             _p(np.ascontiguousarray(w, np.int32)), _p(p))
    return p, bool(neg)


class Dijkstra:
    def __init__(self, n, off, tgt, w, p):
        self.n = n
        self.off = np.ascontiguousarray(off, np.int64)
        self.tgt = np.ascontiguousarray(tgt, np.int32)
        self.w = np.ascontiguousarray(w, np.int32)
        self.p = np.ascontiguousarray(p, np.int64)
        m = len(self.tgt)
        self.dist = np.empty(n, np.int64)
        self.done = np.empty(n, np.int8)
        self.hk = np.empty(m + n + 1, np.int64)
        self.hv = np.empty(m + n + 1, np.int32)

    def __call__(self, s, t):
        return int(olib().orc_dijkstra_pot(self.n, _p(self.off), _p(self.tgt), _p(self.w), _p(self.p),
                                           int(s), int(t), _p(self.dist), _p(self.done), _p(self.hk),
                                           _p(self.hv)))


def restricted(n, off, tgt, w, rank, v, wv):
    dist = np.empty(n, np.int64)
    return int(olib().orc_restricted(n, _p(np.ascontiguousarray(off, np.int64)),
                                     _p(np.ascontiguousarray(tgt, np.int32)),
                                     _p(np.ascontiguousarray(w, np.int32)),
                                     _p(np.ascontiguousarray(rank, np.int32)), int(v), int(wv), _p(dist)))


# ------------------------------------------------------------ reference shim

class Reference:
    """The reference's own C++ (cpp/cch.cpp) on a graph + order."""

    def __init__(self, n, off, tgt, rank=None, x=None, y=None):
        R = rlib()
        self.n = n
        off = np.ascontiguousarray(off, np.int64)
        tgt = np.ascontiguousarray(tgt, np.int32)
        xs = None if x is None else np.ascontiguousarray(x, np.float64)
        ys = None if y is None else np.ascontiguousarray(y, np.float64)
        rk = None if rank is None else np.ascontiguousarray(rank, np.int32)
        self.h = R.ref_build(n, _p(off), _p(tgt), _p(xs), _p(ys), _p(rk))
        info = np.zeros(4, np.int64)
        R.ref_info(self.h, _p(info))
        self.mp, self.triangles, self.depth = int(info[0]), int(info[1]), int(info[2])
        self.m = int(off[-1])

    @classmethod
    def from_dir(cls, d, order_file):
        """The reference's own reader (Graph::load, cpp/cch.cpp:72-95) on a graph
        directory, with the order from a raw int32 file (cpp/cch.cpp:766-777)."""
        R = rlib()
        self = cls.__new__(cls)
        self.h = R.ref_load_dir(str(d).encode(), str(order_file).encode())
        info = np.zeros(4, np.int64)
        R.ref_info(self.h, _p(info))
        self.mp, self.triangles, self.depth, self.n = (int(v) for v in info)
        g = np.zeros(3, np.int64)
        R.ref_graph_info(self.h, _p(g))
        self.m, self.n_metrics = int(g[1]), int(g[2])
        return self

    def graph(self, metric_index=-1):
        """(off, tgt, x, y, w or None) as the reference read them."""
        off = np.empty(self.n + 1, np.int64)
        tgt = np.empty(self.m, np.int32)
        x = np.empty(self.n, np.float64)
        y = np.empty(self.n, np.float64)
        w = np.empty(self.m, np.float64) if metric_index >= 0 else None
        rlib().ref_graph_export(self.h, _p(off), _p(tgt), _p(x), _p(y), int(metric_index), _p(w))
        return off, tgt, x, y, w

    def export(self):
        rank = np.empty(self.n, np.int32)
        first = np.empty(self.n + 1, np.int64)
        head = np.empty(self.mp, np.int32)
        parent = np.empty(self.n, np.int32)
        e_arc = np.empty(self.m, np.int64)
        e_up = np.empty(self.m, np.int8)
        rlib().ref_export(self.h, _p(rank), _p(first), _p(head), _p(parent), _p(e_arc), _p(e_up))
        return dict(rank=rank, first=first, head=head, parent=parent, e_arc=e_arc, e_up=e_up)

    def customize(self, w, threads=0):
        up = np.empty(self.mp, np.float64)
        dn = np.empty(self.mp, np.float64)
        rlib().ref_customize(self.h, _p(np.ascontiguousarray(w, np.int32)), _p(up), _p(dn), int(threads))
        return up, dn

    def negcycle(self, up, dn):
        return int(rlib().ref_negcycle(self.h, _p(up), _p(dn)))

    def potential(self, up, dn):
        """cch_potential (cpp/cch.cpp:635-655), by original vertex id."""
        out = np.empty(self.n, np.float64)
        rlib().ref_potential(self.h, _p(np.ascontiguousarray(up, np.float64)),
                             _p(np.ascontiguousarray(dn, np.float64)), _p(out))
        return out

    def query(self, up, dn, s, t):
        s = np.ascontiguousarray(s, np.int32)
        t = np.ascontiguousarray(t, np.int32)
        d = np.empty(len(s), np.float64)
        rlib().ref_query(self.h, _p(up), _p(dn), len(s), _p(s), _p(t), _p(d))
        return d

    def query_par(self, up, dn, s, t, threads):
        s = np.ascontiguousarray(s, np.int32)
        t = np.ascontiguousarray(t, np.int32)
        d = np.empty(len(s), np.float64)
        rlib().ref_query_par(self.h, _p(up), _p(dn), len(s), _p(s), _p(t), _p(d), int(threads))
        return d

    def levels(self):
        lv = np.empty(self.n, np.int32)
        rlib().ref_levels(self.h, _p(lv))
        return lv

    def __del__(self):
        try:
            rlib().ref_destroy(self.h)
        except Exception:
            pass


def float_to_i32(a):
    """Reference float64 values (integers or inf) -> int32 with the INF sentinel."""
    a = np.asarray(a, np.float64)
    out = np.where(np.isinf(a), INF, a)
    assert np.all(np.isinf(a) | (a == np.round(a))), "non-integer reference value"
    return out.astype(np.int64).astype(np.int32)
