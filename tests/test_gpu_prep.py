"""The metric-independent index built on the GPU (NWCCH_TOPO_DEVICE_INDEX):
triangle counts, the pull-form triangle lists (python/cch.py:168-186, by_c
grouping cch.py:308) and the update side lists (python/cch.py:309-324) equal the
host-built arrays entry for entry, so customization, updates and path
unpacking give identical results either way.  The same holds for the chordal
completion and the arc map built on the GPU (NWCCH_TOPO_DEVICE_SYMBOLIC)."""
import numpy as np
import pytest
import torch

from conftest import order_of, small_graphs
from nwcch import _lib, graphs
from nwcch.device import DeviceTopology

pytestmark = pytest.mark.gpu


def _pair(g, rank):
    th = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    td = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0)
    return th, td


def _check_same(g, rank, what, updates=True):
    print("case: device index", what, g.n, flush=True)
    th, td = _pair(g, rank)
    assert (th.n, th.m, th.mp, th.triangles, th.depth, th.levels) == \
           (td.n, td.m, td.mp, td.triangles, td.depth, td.levels)
    eh, ed = th.export(), td.export()
    for k in eh:
        assert np.array_equal(eh[k], ed[k]), k
    sh, oh, tab = th.export_triangles()
    sd, od, none = td.export_triangles(with_list=False)
    assert none is None
    assert np.array_equal(sh, sd) and np.array_equal(oh, od)
    with pytest.raises(_lib.NwcchError, match="device"):
        td.export_triangles()
    dh, dd = DeviceTopology(th), DeviceTopology(td)
    assert np.array_equal(dd.export_triangles(), tab)
    assert np.array_equal(dh.export_triangles(), tab)
    # both level kernels read the device-built list
    ud_h = dh.customize(g.w)
    assert torch.equal(dd.customize(g.w), ud_h)
    dd.set_tuning("flat_min_tri", 1 << 30)
    assert torch.equal(dd.customize(g.w), ud_h)
    dd.set_tuning("flat_min_tri", 0)
    assert dd.status() == 0 and dh.status() == 0
    if not updates or th.mp == 0:
        return
    so_h, st_h = dh.export_sides()
    so_d, st_d = dd.export_sides()
    assert np.array_equal(so_h, so_d) and np.array_equal(st_h, st_d)
    # one partial update on each: identical weights and results
    rng = np.random.default_rng(5)
    k = min(10, g.m)
    idx = rng.choice(g.m, size=k, replace=False).astype(np.int64)
    neww = rng.integers(1, 100, size=k).astype(np.int32)
    wh, wd = dh.weights(g.w), dd.weights(g.w)
    ud_d = dd.customize(wd)
    rh = dh.update(wh, ud_h, idx, neww).cpu()
    rd = dd.update(wd, ud_d, idx, neww).cpu()
    assert torch.equal(rh, rd) and torch.equal(ud_h, ud_d)
    # shortest paths with vertices: ties resolve on the list order
    s, t = graphs.random_pairs(g.n, 32, seed=3)
    dist_h, len_h, path_h = (x.cpu().numpy() for x in dh.query_path(ud_h, rank[s], rank[t], max_len=4096))
    dist_d, len_d, path_d = (x.cpu().numpy() for x in dd.query_path(ud_d, rank[s], rank[t], max_len=4096))
    assert np.array_equal(dist_h, dist_d) and np.array_equal(len_h, len_d)
    for i in range(len(s)):   # entries past a path's length are not written
        assert np.array_equal(path_h[i, :max(len_h[i], 0)], path_d[i, :max(len_d[i], 0)]), i


def test_device_index_small_graphs():
    for i, g in enumerate(small_graphs()):
        _check_same(g, order_of(g), f"small {i}")


def test_device_index_config0(c0):
    g, rank = c0
    _check_same(g, rank, "config 0")


def test_device_index_road_configs():
    for idx in (1, 3):
        g = graphs.make_config(idx)
        rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
        _check_same(g, rank, f"config {idx}", updates=idx == 1)


def test_device_index_without_triangles():
    """A star has no lower triangles (T = 0); vertex 4 has no arcs."""
    g = graphs.from_arcs(5, [(0, 1, 3), (1, 0, 3), (1, 2, -1), (3, 1, 4)])
    _check_same(g, np.array([0, 4, 1, 2, 3], np.int32), "star")


def test_device_index_from_csr(c0):
    """The CSR form (the arrays cch_nb.customize takes): the device-built index
    equals the host's, and a CSR that is not a chordal completion is rejected
    by the device count check as by the host build."""
    g, rank = c0
    print("case: device index from CSR", flush=True)
    th = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    ex = th.export()
    td = _lib.Topology.from_csr(g.n, ex["first"], ex["head"], ex["e_arc"], ex["e_up"], device_index=0)
    sh, oh, tab = th.export_triangles()
    sd, od, _ = td.export_triangles(with_list=False)
    assert np.array_equal(sh, sd) and np.array_equal(oh, od)
    assert np.array_equal(DeviceTopology(td).export_triangles(), tab)
    first = np.array([0, 2, 2, 2], np.int64)   # 0 -> {1, 2} needs the arc (1, 2)
    head = np.array([1, 2], np.int32)
    with pytest.raises(_lib.NwcchError, match="chordal"):
        _lib.Topology.from_csr(3, first, head, device_index=0)


def _check_symbolic(g, rank, what, index=False):
    """NWCCH_TOPO_DEVICE_SYMBOLIC: the chordal completion and the arc map built
    on the GPU equal the host's (python/cch.py:136-166, 188-204), array for
    array, and so does everything derived from them."""
    print("case: device symbolic", what, g.n, flush=True)
    th = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    td = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_symbolic=0,
                                  device_index=0 if index else None)
    assert (th.n, th.m, th.mp, th.triangles, th.depth, th.levels, th.n_loops, th.n_mapped) == \
           (td.n, td.m, td.mp, td.triangles, td.depth, td.levels, td.n_loops, td.n_mapped)
    eh, ed = th.export(), td.export()
    for k in eh:
        assert np.array_equal(eh[k], ed[k]), k
    sh, oh, tab = th.export_triangles()
    sd, od, tabd = td.export_triangles(with_list=not index)
    assert np.array_equal(sh, sd) and np.array_equal(oh, od)
    if not index:
        assert np.array_equal(tab, tabd)
    dh, dd = DeviceTopology(th), DeviceTopology(td)
    assert torch.equal(dd.customize(g.w), dh.customize(g.w))
    assert dd.status() == 0 and dh.status() == 0


def test_device_symbolic_small_graphs():
    for i, g in enumerate(small_graphs()):
        _check_symbolic(g, order_of(g), f"small {i}")


def test_device_symbolic_configs(c0):
    g, rank = c0
    _check_symbolic(g, rank, "config 0")
    _check_symbolic(g, rank, "config 0 with the device index", index=True)
    g = graphs.make_config(1)
    _check_symbolic(g, _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64), "config 1", index=True)


def test_device_symbolic_loops_and_isolated_vertices():
    """A self-loop maps to no shortcut (e_arc = -1); vertex 4 has no arcs."""
    g = graphs.from_arcs(5, [(0, 1, 3), (1, 0, 3), (1, 2, -1), (3, 1, 4), (2, 2, 5)])
    _check_symbolic(g, np.array([0, 4, 1, 2, 3], np.int32), "star with a loop")


def test_device_symbolic_deep_trees():
    """A path under the identity order is one chain: depth n - 1.  Depth
    14,999 uses the largest LDS bitmap the kernel takes (15,000 levels of one
    vertex); depth 15,999 is past it and is contracted on the host."""
    for n in (15000, 16000):
        u = np.arange(n - 1)
        g = graphs.from_arcs(n, [(int(a), int(a) + 1, 1) for a in u] + [(int(a) + 1, int(a), 2) for a in u])
        _check_symbolic(g, np.arange(n, dtype=np.int32), f"path of {n}")


@pytest.mark.parametrize("cap", [1, 64, 1000])
def test_device_symbolic_staging_rerun(cap, c0, monkeypatch):
    """A staging buffer too small for the rows (NWCCH_SYM_STAGE_CAP) makes the
    device completion flag the overflow, skip the rows that did not fit, and
    rerun every height with a larger buffer: the result still equals the host
    build array for array."""
    g, rank = c0
    graphs_ = [(g, rank), (small_graphs()[5], order_of(small_graphs()[5]))]
    monkeypatch.setenv("NWCCH_SYM_STAGE_CAP", str(cap))
    for gg, rr in graphs_:
        print(f"case: device completion with a staging buffer of {cap} entries, {gg.name}", flush=True)
        th = _lib.Topology.from_graph(gg.n, gg.off, gg.tgt, rr)
        td = _lib.Topology.from_graph(gg.n, gg.off, gg.tgt, rr, device_symbolic=0)
        eh, ed = th.export(), td.export()
        for k in eh:
            assert np.array_equal(eh[k], ed[k]), (cap, gg.name, k)
