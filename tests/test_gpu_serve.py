"""Serving: continuous batching of queries with interleaved updates.  Every
served answer equals the answer computed alone on a from-scratch customization
of the metric version it was served on, and matches the CPU oracle; a
cycle-creating update is rejected (dropped in snapshot mode, undone in strict
mode); admission respects the budgets."""
import numpy as np
import pytest
import torch

from nwcch import _lib, graphs, serve
from nwcch.device import DeviceTopology
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(g, rank):
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    return DeviceTopology(topo)


@pytest.mark.parametrize("snapshot", [True, False])
def test_served_answers_equal_alone(c0, snapshot):
    g, rank = c0
    print(f"case: serving on config 0, 4 metrics, snapshot={snapshot}", flush=True)
    dt = _setup(g, rank)
    K = 4
    W, _ = graphs.metric_family(g, K, seed=9, hmax=60)
    # +-500 changes on 1 m grid arcs close negative cycles now and then: those
    # updates are rejected and undone, the rest become new metric versions
    srv = serve.Server(dt, rank, metric_budget=1 << 30, slot_budget=1 << 24, max_batch=512, n_streams=2,
                       snapshot=snapshot)
    w0 = {}
    for k in range(K):
        assert srv.add_metric(k, W[k])
        w0[k] = W[k].copy()
    reqs = serve.poisson_workload(g, K, rate=200_000, duration=0.05, seed=1, local_radius=5.0,
                                  update_rate=400, update_sizes=(1, 10, 100))
    reqs = serve.absolute_updates(reqs, w0)
    stats = srv.run(reqs, max_seconds=60)
    assert stats["queries"] > 5000 and stats["updates"] > 5
    assert all(r.done is not None for r in reqs)
    v = srv.verify_alone(reqs, w0)
    assert v["mismatches"] == 0, v
    # a sample against the CPU oracle at the served version
    st = O.Structure(g.n, g.off, g.tgt, rank)
    qs = [r for r in reqs if r.kind == "query"][::97][:60]
    for r in qs:
        w = srv.weights_at(r.metric, r.version, w0[r.metric])
        up, dn, _ = st.customize(w)
        d, _, _ = st.query(up, dn, [r.s], [r.t])
        assert int(d[0]) == r.dist


@pytest.mark.parametrize("snapshot", [True, False])
def test_cycle_creating_update_is_rejected(c0, snapshot):
    g, rank = c0
    print(f"case: serving reject-and-restore, snapshot={snapshot}", flush=True)
    dt = _setup(g, rank)
    srv = serve.Server(dt, rank, metric_budget=1 << 30, slot_budget=1 << 22, max_batch=256, snapshot=snapshot)
    assert srv.add_metric(0, g.w)
    before = srv.metrics[0].ud.clone()
    k = int(g.off[100])
    v = int(g.tgt[k])
    back = [j for j in range(g.off[v], g.off[v + 1]) if g.tgt[j] == 100][0]
    s, t = graphs.random_pairs(g.n, 200, seed=3)
    reqs = [serve.Request("query", 0, 0.0, s=int(a), t=int(b)) for a, b in zip(s[:100], t[:100])]
    reqs.append(serve.Request("update", 0, 0.001, edges=np.array([k, back], np.int64),
                              weights=np.array([0, -10 ** 6], np.int32)))
    reqs += [serve.Request("query", 0, 0.002, s=int(a), t=int(b)) for a, b in zip(s[100:], t[100:])]
    srv.run(reqs)
    upd = [r for r in reqs if r.kind == "update"][0]
    assert upd.rejected and srv.metrics[0].version == 0
    assert torch.equal(srv.metrics[0].ud, before)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    up, dn, _ = st.customize(g.w)
    d, _, _ = st.query(up, dn, s, t)
    got = [r.dist for r in reqs if r.kind == "query"]
    assert got == d.tolist()


def test_admission_by_memory(c0):
    g, rank = c0
    dt = _setup(g, rank)
    one = 2 * (4 * dt.m + 8 * dt.mp)          # snapshot mode keeps two copies
    srv = serve.Server(dt, rank, metric_budget=2 * one + 10, slot_budget=1, max_batch=64)
    assert srv.metric_bytes() == one
    assert srv.add_metric(0, g.w) and srv.add_metric(1, g.w)
    assert not srv.add_metric(2, g.w)        # over the metric budget
    srv.drop_metric(1)
    assert srv.add_metric(2, g.w)
    assert len(srv.launch_bufs) == 1         # the slot budget still admits one launch
