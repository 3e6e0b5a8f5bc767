"""The reference's on-disk graph format (python/export_graph.py:22-52,
cpp/cch.cpp:72-95, 766-777): round trips, error handling, and the reference's
own C++ reader on directories this library writes."""
import numpy as np
import pytest

from nwcch import _lib, graph_io, graphs
from oracle import oracle as O


def test_round_trip(tmp_path, c0):
    g, rank = c0
    W, _ = graphs.metric_family(g, 2, seed=5, hmax=60)
    w1 = W[1].copy()
    w1[:3] = O.INF
    d = graph_io.write_graph_dir(tmp_path / "g", g, {"ev": W[0], "alt": w1})
    graph_io.write_order(tmp_path / "order.i32", rank)
    assert graph_io.read_meta(d) == (g.n, g.m, ["ev", "alt"])
    h = graph_io.read_graph_dir(d)
    assert h.n == g.n and np.array_equal(h.off, g.off) and np.array_equal(h.tgt, g.tgt)
    assert np.array_equal(h.x, g.x) and np.array_equal(h.y, g.y) and np.array_equal(h.w, W[0])
    h2 = graph_io.read_graph_dir(d, "alt")
    assert np.array_equal(h2.w, w1)
    assert np.isinf(graph_io.read_weights(d, "alt")[:3]).all()
    assert np.array_equal(graph_io.read_order(tmp_path / "order.i32", g.n), rank)


def test_int_weights_rules():
    assert graph_io.int_weights([1.0, -7.0, np.inf]).tolist() == [1, -7, O.INF]
    with pytest.raises(ValueError):
        graph_io.int_weights([1.5])
    assert graph_io.int_weights([1.5, 2.25], scale=4).tolist() == [6, 9]
    with pytest.raises(ValueError):
        graph_io.int_weights([np.nan])
    with pytest.raises(ValueError):
        graph_io.int_weights([3e9])


def test_bad_files_are_refused(tmp_path, c0):
    g, rank = c0
    d = graph_io.write_graph_dir(tmp_path / "g", g)
    graph_io.write_order(tmp_path / "bad.i32", np.zeros(g.n, np.int32))
    with pytest.raises(ValueError):
        graph_io.read_order(tmp_path / "bad.i32")
    np.zeros(3, np.int32).tofile(d / "tgt.i32")
    with pytest.raises(ValueError):
        graph_io.read_graph_dir(d)


@pytest.mark.skipif(not O.have_ref(), reason="the reference's C++ was not built")
def test_reference_reader_reads_what_we_write(tmp_path, c0):
    """cpp/cch.cpp's Graph::load and order loading on a directory written here:
    the same arrays, the same weights, and the same hierarchy as the library's
    topology built from memory."""
    g, rank = c0
    d = graph_io.write_graph_dir(tmp_path / "g", g, {"ev": g.w})
    graph_io.write_order(tmp_path / "order.i32", rank)
    ref = O.Reference.from_dir(d, tmp_path / "order.i32")
    off, tgt, x, y, w = ref.graph(0)
    assert ref.n == g.n and ref.m == g.m and ref.n_metrics == 1
    assert np.array_equal(off, g.off) and np.array_equal(tgt, g.tgt)
    assert np.array_equal(x, g.x) and np.array_equal(y, g.y)
    assert np.array_equal(w, g.w.astype(np.float64))
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    ex = topo.export()
    rx = ref.export()
    assert ref.mp == topo.mp and ref.triangles == topo.triangles
    assert np.array_equal(rx["first"], ex["first"]) and np.array_equal(rx["head"], ex["head"])
    up, dn = ref.customize(g.w)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    u0, d0, _ = st.customize(graph_io.read_graph_dir(d).w)
    assert np.array_equal(O.float_to_i32(up), u0) and np.array_equal(O.float_to_i32(dn), d0)
