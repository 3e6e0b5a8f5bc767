"""The measured configurations as parity tests on the GPU, against the oracle
and the reference's own C++ (oracle/_ref, cpp/cch.cpp compiled from its
sources):

* the structure built on the GPU (chordal completion, arc map, triangle lists)
  against the oracle's and the reference's (configs 1 and 3);
* a graph directory in the reference's format, read by the reference's own
  loader, customized and queried on the GPU (config 1);
* config 2: all 64 batched metrics and the flagged metric;
* config 3: serving 16 metrics with 1/10/1000-arc updates;
* config 4, the full-scale graph: every arc of a customization, the scan and a
  query sample against the reference's C++.

The long host steps run with pytest's output capture off, so their own
progress lines (one per generator stage and per large nested-dissection
cell) appear as they happen."""
import os

import numpy as np
import pytest
import torch

from nwcch import _lib, graph_io, graphs, serve
from nwcch.device import DeviceTopology
from oracle import oracle as O

pytestmark = pytest.mark.gpu
INF = O.INF


def _make(cfg, capfd):
    with capfd.disabled():
        print(f"\ncase: config {cfg}: generating and ordering", flush=True)
        g = graphs.make_config(cfg, progress=lambda m: print(f"config {cfg}: {m}", flush=True))
        rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
        print(f"config {cfg}: n={g.n} m={g.m} ordered", flush=True)
    return g, rank


def lower_triangles(first, head):
    """(c, a, b) of every lower triangle from an upward CSR: for each vertex x
    and arcs i < k of row x, a = i, b = k and c = the arc (head i, head k),
    found in row head i (python/cch_nb.py:518-556).  Sorted by (c, a)."""
    first = np.asarray(first, np.int64)
    head = np.asarray(head, np.int64)
    n = len(first) - 1
    mp = len(head)
    tail = np.repeat(np.arange(n, dtype=np.int64), np.diff(first))
    row_end = first[tail + 1]
    cnt = row_end - 1 - np.arange(mp)               # arcs after i in its row
    a = np.repeat(np.arange(mp, dtype=np.int64), cnt)
    start = np.cumsum(cnt) - cnt
    b = a + 1 + (np.arange(len(a), dtype=np.int64) - np.repeat(start, cnt))
    key = tail * n + head                            # ascending: rows in order, heads ascending
    c = np.searchsorted(key, head[a] * n + head[b])
    assert np.array_equal(key[c], head[a] * n + head[b]), "not chordal"
    o = np.lexsort((a, c))
    return c[o], a[o], b[o]


def _device_triangles(topo, dt):
    sched, toff, _ = topo.export_triangles(with_list=False)
    tab = dt.export_triangles().reshape(-1, 2).astype(np.int64)
    c = np.repeat(sched.astype(np.int64), np.diff(toff))
    o = np.lexsort((tab[:, 0], c))
    return c[o], tab[o, 0], tab[o, 1]


@pytest.mark.parametrize("cfg", [1, 3])
def test_device_structure_equals_oracle_and_reference(cfg, capfd):
    """Chordal completion, arc map, ET and the triangle lists built on the GPU
    (NWCCH_TOPO_DEVICE_SYMBOLIC + NWCCH_TOPO_DEVICE_INDEX) equal the oracle's
    set-merge restatement and the reference's CCH::symbolic / map_arcs."""
    g, rank = _make(cfg, capfd)
    td = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0, device_symbolic=0)
    ex = td.export()
    st = O.Structure(g.n, g.off, g.tgt, rank)
    assert np.array_equal(ex["first"], st.first) and np.array_equal(ex["head"], st.head)
    assert np.array_equal(ex["parent"], st.parent)
    assert np.array_equal(ex["e_arc"], st.e_arc) and np.array_equal(ex["e_up"], st.e_up)
    assert np.array_equal(ex["level"], st.level) and td.triangles == st.triangles
    if O.have_ref():
        rx = O.Reference(g.n, g.off, g.tgt, rank).export()
        for k in ("first", "head", "parent", "e_arc", "e_up"):
            assert np.array_equal(rx[k], ex[k]), k
    dt = DeviceTopology(td, 0)
    c1, a1, b1 = _device_triangles(td, dt)
    c0, a0, b0 = lower_triangles(st.first, st.head)
    assert len(c1) == st.triangles
    assert np.array_equal(c1, c0) and np.array_equal(a1, a0) and np.array_equal(b1, b0)
    # and the customization on the device-built structure equals the oracle's
    ud = dt.customize(g.w).cpu().numpy()
    up, dn, _ = st.customize(g.w)
    assert np.array_equal(ud[:, 0], up) and np.array_equal(ud[:, 1], dn)
    assert dt.status() == 0


@pytest.mark.skipif(not O.have_ref(), reason="the reference's C++ was not built")
def test_reference_format_directory_on_gpu(tmp_path, capfd):
    """A graph directory in the reference's format (meta.txt, off.i64, tgt.i32,
    x.f64, y.f64, w_<metric>.f64; python/export_graph.py:22-52) and an
    order.i32: the reference's own loader (cpp/cch.cpp:72-95, 766-777) reads
    it, this library reads it, and the GPU customization and queries equal
    the reference's CCH::customize and Query::run."""
    g, rank = _make(1, capfd)
    W, _ = graphs.metric_family(g, 2, seed=11, hmax=800)
    d = graph_io.write_graph_dir(tmp_path / "g", g, {"ev": W[0], "alt": W[1]})
    graph_io.write_order(tmp_path / "order.i32", rank)
    ref = O.Reference.from_dir(d, tmp_path / "order.i32")
    h = graph_io.read_graph_dir(d, "alt")
    r2 = graph_io.read_order(tmp_path / "order.i32", h.n)
    topo = _lib.Topology.from_graph(h.n, h.off, h.tgt, r2, device_index=0, device_symbolic=0)
    dt = DeviceTopology(topo, 0)
    ud = dt.customize(h.w)
    _, _, _, _, wref = ref.graph(1)
    up, dn = ref.customize(graph_io.int_weights(wref))
    got = ud.cpu().numpy()
    assert np.array_equal(got[:, 0], O.float_to_i32(up)) and np.array_equal(got[:, 1], O.float_to_i32(dn))
    s, t = graphs.random_pairs(h.n, 2000, seed=12)
    dist, _ = dt.query(ud, r2[s], r2[t], want_scanned=False)
    assert np.array_equal(dist.cpu().numpy(), O.float_to_i32(ref.query(up, dn, s, t)))
    assert dt.status() == 0


def test_config2_all_64_metrics(capfd):
    """Config 2: 64 metrics (slopes 5-40, own height fields) in one sweep; every
    metric equals its oracle customization, the scan counts and witnesses
    equal the oracle's, and only the metric with the injected 3-arc cycle of
    total -1 is flagged."""
    g, rank = _make(2, capfd)
    K = 64
    W, _ = graphs.metric_family(g, K, seed=2, hmax=1200)
    tri = graphs.find_directed_triangle(g, np.random.default_rng(2))
    assert tri is not None
    bad = K // 3
    W[bad] = graphs.inject_negative_cycle(W[bad], tri)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0, device_symbolic=0)
    dt = DeviceTopology(topo, 0)
    UD = dt.customize_batch(W)
    sc = dt.negcycle_scan_batch(UD, K).cpu().numpy()
    st = O.Structure(g.n, g.off, g.tgt, rank)
    for k in range(K):
        up, dn, _ = st.customize(W[k])
        got = dt.metric_view(UD, k).cpu().numpy()
        assert np.array_equal(got[:, 0], up) and np.array_equal(got[:, 1], dn), k
        assert (int(sc[0, k]), int(sc[1, k])) == st.negcycle(up, dn), k
    assert np.nonzero(sc[0] > 0)[0].tolist() == [bad]
    dt.status(clear=True)


def test_config3_serving_with_updates(capfd):
    """Config 3: 16 resident metrics served for a short Poisson window (1 M
    queries/s offered, 20 % local pairs within 5 km) with 1-, 10- and 1000-arc
    updates; every answer equals its alone recompute at its metric version,
    and a sample equals the oracle's."""
    g, rank = _make(3, capfd)
    K = 16
    W, _ = graphs.metric_family(g, K, seed=3, hmax=800)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0, device_symbolic=0)
    dt = DeviceTopology(topo, 0)
    srv = serve.Server(dt, rank, metric_budget=K * 2 * (4 * topo.m + 8 * topo.mp), slot_budget=64 << 20,
                       max_batch=65536, n_streams=2, snapshot=True)
    w0 = {}
    for k in range(K):
        assert srv.add_metric(k, W[k])
        w0[k] = W[k].copy()
    reqs = serve.poisson_workload(g, K, rate=1_000_000, duration=0.1, seed=31, update_rate=100)
    reqs = serve.absolute_updates(reqs, w0)
    stats = srv.run(reqs, max_seconds=120)
    assert stats["queries"] > 50_000 and stats["updates"] >= 3
    with capfd.disabled():
        print(f"config 3 serving: {stats['queries']} queries, {stats['updates']} updates "
              f"(rejected {stats['rejected_updates']}), p99 {stats['query_p99_ms']:.2f} ms", flush=True)
    v = srv.verify_alone(reqs, w0)
    assert v["mismatches"] == 0, v
    st = O.Structure(g.n, g.off, g.tgt, rank)
    qs = [r for r in reqs if r.kind == "query"][::4001][:25]
    for r in qs:
        up, dn, _ = st.customize(srv.weights_at(r.metric, r.version, w0[r.metric]))
        d, _, _ = st.query(up, dn, [r.s], [r.t])
        assert int(d[0]) == r.dist


@pytest.mark.skipif(not O.have_ref(), reason="the reference's C++ was not built")
def test_config4_full_scale_against_reference(capfd):
    """Config 4 (23.9 M points): the structure and triangle index built on the
    GPU, one metric customized, its scan, and 4,000 random queries, each
    against the reference's own C++ on the same graph and order (every one of
    the ~88 M arcs, both directions); one batched metric of K = 8 against its
    single-metric run."""
    os.environ["NWCCH_PROGRESS"] = "1"
    try:
        g, rank = _make(4, capfd)
        with capfd.disabled():
            topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0, device_symbolic=0)
            dt = DeviceTopology(topo, 0)
            ud = dt.customize(g.w)
            scan = dt.negcycle_scan(ud).cpu().numpy()
            s, t = graphs.random_pairs(g.n, 4000, seed=44)
            dist, _ = dt.query(ud, rank[s], rank[t], want_scanned=False)
            got = ud.cpu().numpy()
            print(f"config 4: m+={topo.mp} T={topo.triangles} customized on the GPU; reference C++ next",
                  flush=True)
            ref = O.Reference(g.n, g.off, g.tgt, rank)
            assert ref.mp == topo.mp and ref.triangles == topo.triangles
            up, dn = ref.customize(g.w)
            print("config 4: reference customization done", flush=True)
    finally:
        os.environ.pop("NWCCH_PROGRESS", None)
    assert np.array_equal(got[:, 0], O.float_to_i32(up)) and np.array_equal(got[:, 1], O.float_to_i32(dn))
    assert int(scan[0]) == ref.negcycle(up, dn)
    assert np.array_equal(dist.cpu().numpy(), O.float_to_i32(ref.query(up, dn, s, t)))
    del up, dn, ref
    W, _ = graphs.metric_family(g, 8, seed=4, hmax=3000)
    UD = dt.customize_batch(W)
    k = 5
    assert torch.equal(dt.metric_view(UD, k), dt.customize(W[k]))
    assert dt.status() == 0
