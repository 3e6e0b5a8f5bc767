"""The autotuner picks launch shapes without changing any result."""
import numpy as np
import pytest
import torch

from nwcch import _lib, graphs, tune
from nwcch.device import DeviceTopology
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_tuned_shapes_keep_results(c0):
    g, rank = c0
    print("case: autotuner on config 0", flush=True)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    res = tune.tune_customize(dt, g.w, candidates=(1, 4, 16), reps=2)
    assert res["tri_per_lane"] in (1, 4, 16) and dt.get_tuning("tri_per_lane") == res["tri_per_lane"]
    ud = dt.customize(g.w)
    up, dn, _ = st.customize(g.w)
    assert np.array_equal(ud[:, 0].cpu().numpy(), up) and np.array_equal(ud[:, 1].cpu().numpy(), dn)
    W, _ = graphs.metric_family(g, 16, seed=3, hmax=400)
    rb = tune.tune_batch(dt, W, slabs=(64, 8), tris=(4, 16), reps=2)
    UD = dt.customize_batch(W, slab=rb["slab"])
    for k in (0, 7, 15):
        u, d, _ = st.customize(W[k])
        got = DeviceTopology.metric_view(UD, k).cpu().numpy()
        assert np.array_equal(got[:, 0], u) and np.array_equal(got[:, 1], d)
    s, t = graphs.random_pairs(g.n, 2000, seed=5)
    rq = tune.tune_queries(dt, ud, rank[s], rank[t], reps=2)
    assert rq["packing"] in ("none", "longest_first")
    order = dt.chain_order(rank[s], rank[t])
    assert torch.equal(torch.sort(order.long()).values, torch.arange(2000, device=order.device))
    dist, _ = dt.query(ud, rank[s], rank[t], order=order)
    d0, _, _ = st.query(up, dn, s, t)
    assert np.array_equal(dist.cpu().numpy(), d0)
    assert dt.status() == 0
