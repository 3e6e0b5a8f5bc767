"""Batched multi-metric customization: every metric of one launch sweep equals
its own single-metric oracle customization, bit for bit; the per-metric scan
flags exactly the metric with an injected negative cycle; queries on any metric
of the batch equal the oracle's."""
import numpy as np
import pytest

from nwcch import _lib, graphs
from nwcch.device import DeviceTopology
from oracle import oracle as O

pytestmark = pytest.mark.gpu
INF = O.INF


@pytest.mark.parametrize("slab", [0, 8])
@pytest.mark.parametrize("K", [1, 2, 4, 5, 8, 16, 64, 70])
def test_batch_equals_single(c0, K, slab):
    g, rank = c0
    print(f"case: batch K={K} slab={slab} on config 0", flush=True)
    W, _ = graphs.metric_family(g, K, seed=K, hmax=400)
    bad = K // 2
    tri = graphs.find_directed_triangle(g)
    if tri is None:
        k = int(g.off[50])
        v = int(g.tgt[k])
        back = [j for j in range(g.off[v], g.off[v + 1]) if g.tgt[j] == 50][0]
        W[bad, k], W[bad, back] = 0, -1
    else:
        W[bad] = graphs.inject_negative_cycle(W[bad], tri)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    UD = dt.customize_batch(W, slab=slab)
    sc = dt.negcycle_scan_batch(UD, K).cpu().numpy()
    st = O.Structure(g.n, g.off, g.tgt, rank)
    assert tuple(UD.shape) == dt.batch_shape(K, slab)
    s, t = graphs.random_pairs(g.n, 300, seed=K)
    for k in range(K):
        up, dn, _ = st.customize(W[k])
        got = dt.metric_view(UD, k).cpu().numpy()
        assert np.array_equal(got[:, 0], up), k
        assert np.array_equal(got[:, 1], dn), k
        cnt, first = st.negcycle(up, dn)
        assert (sc[0, k], sc[1, k]) == (cnt, first), k
        assert (cnt > 0) == (k == bad)
        if k in (0, K - 1) and k != bad:
            d, _ = dt.query(UD, rank[s], rank[t], metric=k, want_scanned=False)
            d0, _, _ = st.query(up, dn, s, t)
            assert np.array_equal(d.cpu().numpy(), d0)
    assert dt.status() == 0


@pytest.mark.parametrize("order,xcd", [(1, 0), (0, 1), (1, 1)])
def test_batch_layouts_equal_default(order, xcd):
    """Arc-id order inside wave classes and XCD slices of the workgroups change
    which workgroup takes which arc, never a value: config 1 with 24 metrics
    (slab widths 8 and 64) equals the default layout bit for bit."""
    import torch
    g = graphs.make_config(1)
    rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    print(f"case: batch layout order={order} xcd={xcd} on config 1", flush=True)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    W, _ = graphs.metric_family(g, 24, seed=3, hmax=800)
    for slab in (8, 64):
        ref = dt.customize_batch(W, slab=slab)
        dt.set_tuning("batch_order", order)
        dt.set_tuning("batch_xcd", xcd)
        for tpl in (4, 16):   # two wave-class shapes: the id order is rebuilt for each
            dt.set_tuning("batch_tri_per_lane", tpl)
            assert torch.equal(dt.customize_batch(W, slab=slab), ref), (slab, tpl)
        dt.set_tuning("batch_tri_per_lane", 16)
        dt.set_tuning("batch_order", 0)
        dt.set_tuning("batch_xcd", 0)
    assert dt.status() == 0



@pytest.mark.parametrize("K,slab", [(8, 0), (16, 8), (5, 8)])
def test_batch_row_merge_levels(c0, K, slab):
    """Batched levels in row-merge form (forced on every level; with 8 metrics
    per slab and K a multiple of 8, else the triangle-list kernel keeps them)
    equal the triangle-list kernel bit for bit and every metric equals its
    oracle customization, on config 0 with a metric holding a negative cycle
    (saturated values and the overflow status) and on config 1."""
    import torch
    g0, rank0 = c0
    g1 = graphs.make_config(1)
    rank1 = _lib.order(g1.n, g1.off, g1.tgt, g1.x, g1.y, 64)
    for g, rank in ((g0, rank0), (g1, rank1)):
        print(f"case: batched row-merge levels K={K} slab={slab} on {g.name}", flush=True)
        W, _ = graphs.metric_family(g, K, seed=K + 1, hmax=600)
        if g is g0:
            tri = graphs.find_directed_triangle(g)
            if tri is not None:
                W[K - 1] = graphs.inject_negative_cycle(W[K - 1], tri)
        topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
        dt = DeviceTopology(topo)
        dt.set_tuning("rows_batch", 0)
        ref = dt.customize_batch(W, slab=slab)
        st_ref = dt.status(clear=True)
        dt.set_tuning("rows_batch", 1)
        dt.set_tuning("rows_batch_min_rows", 0)
        dt.set_tuning("rows_batch_max_rows", 1 << 30)
        dt.set_tuning("rows_batch_min_level_tri", 0)
        dt.set_tuning("rows_batch_min_len", 0)
        for _ in range(2):
            got = dt.customize_batch(W, slab=slab)
            assert torch.equal(got, ref)
            assert dt.status(clear=True) == st_ref
        st = O.Structure(g.n, g.off, g.tgt, rank)
        for k in (0, K // 2):
            up, dn, _ = st.customize(W[k])
            v = dt.metric_view(got, k).cpu().numpy()
            assert np.array_equal(v[:, 0], up) and np.array_equal(v[:, 1], dn), k
