"""Host side of the product (no GPU): the native preprocessing equals the oracle's
independent restatement, error paths, and the C ABI exports."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import order_of, small_graphs
from nwcch import _lib, cch_nb, graphs
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    """Every function declared in include/*.h is exported by libnwcch.so."""
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(nwcch_[a-z0-9_]+)\s*\(", text))
    assert names, "no declarations found"
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.exported_symbols()) <= names


@pytest.mark.parametrize("idx", range(12))
def test_topology_equals_oracle(idx):
    """symbolic contraction, arc map, ET levels and triangle count equal the
    oracle's set-merge restatement (python/cch.py:136-204)."""
    g = small_graphs()[idx]
    rank = order_of(g)
    assert np.array_equal(np.sort(rank), np.arange(g.n))
    t = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    ex = t.export()
    st = O.Structure(g.n, g.off, g.tgt, rank)
    for k in ("first", "head", "parent", "e_arc", "e_up"):
        assert np.array_equal(ex[k], getattr(st, k)), k
    assert np.array_equal(ex["level"], st.level)
    assert np.array_equal(ex["depth"], st.vdepth)
    assert t.triangles == st.triangles == int((np.diff(st.first) * (np.diff(st.first) - 1) // 2).sum())
    assert t.depth == st.depth


def test_triangle_index_is_the_lower_triangle_set(c0):
    """The pull-form index holds exactly the lower triangles (x; y, z) of every
    arc, each once: rebuilt as (a, b, c) triples and compared with a direct
    enumeration by merging sorted rows (python/cch.py:168-186)."""
    g, rank = c0
    t = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    ex = t.export()
    first, head = ex["first"], ex["head"]
    sched, toff, tab = t.export_triangles()
    assert np.array_equal(np.sort(sched), np.arange(t.mp))
    c = np.repeat(sched, np.diff(toff))
    got = np.stack([tab[:, 0], tab[:, 1], c], 1)
    want = []
    for x in range(g.n):
        s, e = first[x], first[x + 1]
        for i in range(s, e):
            y = head[i]
            row = head[first[y]:first[y + 1]]
            for k in range(i + 1, e):
                j = first[y] + np.searchsorted(row, head[k])
                assert head[j] == head[k]
                want.append((i, k, j))
    want = np.array(want)
    key = lambda a: np.lexsort((a[:, 1], a[:, 0], a[:, 2]))
    assert np.array_equal(got[key(got)], want[key(want)])
    # schedule: arcs of one level never read an arc of the same or a higher level
    level = ex["level"]
    tail = np.repeat(np.arange(g.n), np.diff(first))
    lc = level[tail[c]]
    assert (level[tail[tab[:, 0]]] < lc).all() and (level[tail[tab[:, 1]]] < lc).all()
    # schedule order: level ascending, then triangle count descending, then arc id
    cnt = np.diff(toff)
    keys = np.stack([level[tail[sched]], -cnt, sched], 1)
    assert np.array_equal(np.lexsort(keys.T[::-1]), np.arange(t.mp))
    # inside an arc's run the lower vertex x (hence the arc a = (x, y)) ascends
    same = c[1:] == c[:-1]
    assert (tab[1:, 0][same] > tab[:-1, 0][same]).all()


def test_order_is_deterministic_and_valid():
    g = graphs.road_graph(20000, seed=9, name="r20k")
    r1 = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    r2 = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    assert np.array_equal(r1, r2)
    assert np.array_equal(np.sort(r1), np.arange(g.n))


def test_bad_order_is_rejected():
    g = graphs.ev_grid(6, 0)
    rank = np.arange(g.n, dtype=np.int32)
    rank[3] = 4
    with pytest.raises(_lib.NwcchError, match="permutation"):
        _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)


def test_device_index_flags_are_checked():
    """Unknown flags are rejected before any GPU is touched; without a GPU the
    device-built index fails loudly instead of falling back to the host."""
    g = graphs.ev_grid(6, 0)
    rank = np.arange(g.n, dtype=np.int32)
    h = ctypes.c_void_p()
    rc = _lib.lib().nwcch_topo_create_ex(g.n, _lib.ptr(g.off), _lib.ptr(g.tgt), _lib.ptr(rank), 6, 0,
                                         ctypes.byref(h))
    assert rc == -1 and b"unknown flags" in _lib.lib().nwcch_last_error()
    ex = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank).export()
    rc = _lib.lib().nwcch_topo_from_csr_ex(g.n, _lib.ptr(ex["first"]), _lib.ptr(ex["head"]), 0, None, None,
                                           _lib.NWCCH_TOPO_DEVICE_SYMBOLIC, 0, ctypes.byref(h))
    assert rc == -1 and b"unknown flags" in _lib.lib().nwcch_last_error()   # a CSR needs no contraction
    import torch
    if not torch.cuda.is_available():
        with pytest.raises(_lib.NwcchError):
            _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0)
        with pytest.raises(_lib.NwcchError):
            _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_symbolic=0)


def test_non_chordal_csr_is_rejected():
    # 0 -> {1, 2} requires arc (1, 2); it is missing
    first = np.array([0, 2, 2, 2], np.int64)
    head = np.array([1, 2], np.int32)
    with pytest.raises(_lib.NwcchError, match="chordal"):
        _lib.Topology.from_csr(3, first, head)


def test_cch_nb_host_prerequisites(c0):
    """The drop-in host functions of python/cch_nb.py return the same arrays as
    the oracle for the same order."""
    g, _ = c0
    noff, nb, rev = cch_nb.undirected_csr(g.n, g.off, g.tgt)
    assert np.array_equal(nb[rev], np.repeat(np.arange(g.n), np.diff(noff)))
    rank = cch_nb.nested_dissection(g.n, noff, nb, rev, g.x, g.y, 64)
    first, head, parent = cch_nb.symbolic(g.n, noff, nb, rank)
    e_arc, e_up = cch_nb.edge_arcs(g.n, g.off, g.tgt, rank, first, head)
    st = O.Structure(g.n, g.off, g.tgt, rank.astype(np.int32))
    assert np.array_equal(first, st.first) and np.array_equal(head, st.head)
    assert np.array_equal(parent, st.parent) and np.array_equal(e_arc, st.e_arc)
    assert np.array_equal(e_up, st.e_up)
    assert cch_nb.count_triangles(g.n, first, head) == st.triangles


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; asked for more it exits with a message and no
    JSON line."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr and r.stdout.strip() == ""


def test_host_threads_do_not_change_the_result():
    """The host build's passes on threads (nested dissection, undirected lists,
    arc map, triangle counts, schedule sorts, triangle list) give what one
    thread gives, on the 264,000-point road graph."""
    import os
    g = graphs.make_config(1)
    old = os.environ.get("NWCCH_THREADS")
    res = []
    try:
        for th in ("1", "8"):
            os.environ["NWCCH_THREADS"] = th
            rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
            t = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
            res.append((rank, t.export(), t.export_triangles()))
    finally:
        if old is None:
            os.environ.pop("NWCCH_THREADS", None)
        else:
            os.environ["NWCCH_THREADS"] = old
    (r1, e1, x1), (r8, e8, x8) = res
    assert np.array_equal(r1, r8)
    for k in e1:
        assert np.array_equal(e1[k], e8[k]), k
    for a, b in zip(x1, x8):
        assert np.array_equal(a, b)
