"""A customization + scan captured in a HIP graph replays to the eager result,
replay after replay, with eager work and synchronisations in between (the
graph holds kernels only: no memset nodes)."""
import numpy as np
import pytest
import torch

from nwcch import _lib
from nwcch.device import DeviceTopology, StepGraph
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(c0):
    g, rank = c0
    print("case: graph replay on config 0", flush=True)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    rng = np.random.default_rng(0)
    w = dt.weights(g.w)
    ud = torch.empty((topo.mp, 2), dtype=torch.int32, device=dt.dev)
    scan = torch.empty(2, dtype=torch.int64, device=dt.dev)

    def step():
        dt.customize(w, out=ud)
        dt.negcycle_scan(ud, out=scan)

    graph = StepGraph(step)
    for rep in range(4):
        # new weights in the captured input buffer, poison the output buffer
        wh = (g.w + rng.integers(0, 30, g.m)).astype(np.int32)
        w.copy_(torch.as_tensor(wh))
        ud.fill_(-7)
        scan.fill_(-7)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        up, dn, _ = st.customize(wh)
        got = ud.cpu().numpy()
        assert np.array_equal(got[:, 0], up) and np.array_equal(got[:, 1], dn), rep
        assert tuple(scan.cpu().tolist()) == st.negcycle(up, dn)
        eager = dt.customize(w)
        assert torch.equal(eager, ud)
    assert dt.status() == 0


def test_batched_graph_replays_match_eager(c0):
    """The batched sweep (K = 12 metrics, slab 8: two slabs) and its per-metric
    scan captured in a HIP graph replay to the eager result, replay after
    replay, with new weights written between replays."""
    from nwcch import graphs
    g, rank = c0
    print("case: batched graph replay on config 0", flush=True)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    K = 12
    W0, _ = graphs.metric_family(g, K, seed=5, hmax=400)
    Wd = dt.weights_batch(W0)
    UD = torch.empty(dt.batch_shape(K, 8), dtype=torch.int32, device=dt.dev)
    sc = torch.empty((2, K), dtype=torch.int64, device=dt.dev)

    def step():
        dt.customize_batch(Wd, out=UD, slab=8)
        dt.negcycle_scan_batch(UD, K, out=sc)

    graph = StepGraph(step)
    rng = np.random.default_rng(1)
    for rep in range(3):
        W = (W0 + rng.integers(0, 20, W0.shape)).astype(np.int32)
        Wd.copy_(dt.weights_batch(W))
        UD.fill_(-7)
        sc.fill_(-7)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got, got_sc = UD.clone(), sc.clone()
        eager = dt.customize_batch(Wd, slab=8)
        for k in range(K):   # the padding metrics of the last slab are never written
            assert torch.equal(dt.metric_view(eager, k), dt.metric_view(got, k)), (rep, k)
        assert torch.equal(dt.negcycle_scan_batch(eager, K), got_sc), rep
    assert dt.status() == 0
