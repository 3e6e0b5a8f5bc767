"""GPU partial re-customization against the oracle's restatement of
python/cch.py:349-402 (bitwise values, evaluated count, changed set) and the
python/test_cch.py update scenarios through the CCH drop-in."""
import numpy as np
import pytest
import torch

from conftest import order_of, small_graphs
from nwcch import _lib, graphs
from nwcch.cch import CCH
from nwcch.device import DeviceTopology
from oracle import oracle as O

pytestmark = pytest.mark.gpu
INF = O.INF


def _setup(g, rank):
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    return dt, st


@pytest.mark.parametrize("mode,graph", [(1, 1), (2, 1), (2, 0), (3, 1), (0, 1)])
@pytest.mark.parametrize("k", [1, 10, 1000])
def test_update_matches_oracle(c0, k, mode, graph):
    """Random changes of +-U(0,500) on k arcs, four rounds in a row; mode 2
    with the per-height launches replayed from captured graphs and one by
    one, mode 3 (full recustomization diffed against the old values: the same
    evaluated count, changed set and overflow status) and auto."""
    g, rank = c0
    print(f"case: config 0 update k={k} mode={mode} update_graph={graph}", flush=True)
    dt, st = _setup(g, rank)
    dt.set_tuning("update_graph", graph)
    w = dt.weights(g.w)
    ud = dt.customize(w)
    up, dn, _ = st.customize(g.w)
    upd = O.Updater(st, g.w, up, dn)
    rng = np.random.default_rng(k * 10 + mode)
    for rnd in range(4):
        idx = rng.choice(g.m, k, replace=False)
        new = (upd.w[idx] + rng.integers(-500, 501, k)).astype(np.int32)
        res = dt.update(w, ud, idx, new, mode=mode)
        ev, nch, nneg, fneg = (int(v) for v in res.cpu())
        ev0, ch0, ovf = upd.update(idx, new)
        got = ud.cpu().numpy()
        assert np.array_equal(got[:, 0], upd.up) and np.array_equal(got[:, 1], upd.dn), rnd
        assert np.array_equal(w.cpu().numpy(), upd.w)
        assert ev == ev0, (ev, ev0)
        assert dt.last_changed(nch).cpu().numpy().tolist() == ch0.tolist()
        neg0, first0 = st.negcycle(upd.up[ch0], upd.dn[ch0])
        assert nneg == neg0
        assert fneg == (int(ch0[first0]) if neg0 else -1) or nneg == 0
        # a decrease by up to 500 on 1 m grid arcs can close a negative cycle;
        # values then run away and both sides must report the int32 overflow
        assert bool(dt.status() & 1) == ovf


def test_update_equals_full_customization_road():
    g = graphs.make_config(1)
    print("case: config 1 update k=1000", flush=True)
    rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    dt, st = _setup(g, rank)
    w = dt.weights(g.w)
    ud = dt.customize(w)
    rng = np.random.default_rng(5)
    wh = g.w.copy()
    for k in (1, 10, 1000):
        idx = rng.choice(g.m, k, replace=False)
        new = (wh[idx] + rng.integers(-500, 501, k)).astype(np.int32)
        wh[idx] = new
        res = dt.update(w, ud, idx, new)
        up, dn, _ = st.customize(wh)
        got = ud.cpu().numpy()
        assert np.array_equal(got[:, 0], up) and np.array_equal(got[:, 1], dn), k
        assert int(res[1].item()) <= int(res[0].item())


def test_cch_dropin_scenario():
    """python/test_cch.py:28-106 through the CCH drop-in: queries vs Bellman-Ford,
    raise/restore rounds equal full customization, a -1e6 update creates a
    negative cycle that the changed-arc check sees, restoring removes it."""
    rng = np.random.default_rng(3)
    for g in small_graphs()[::2]:
        print("case: drop-in scenario", g.name, flush=True)
        c = CCH(g, xy=list(zip(g.x, g.y)), leaf=8, log=lambda *_: None)
        assert not c.customize(g), "false negative cycle"
        for _ in range(10):
            s, t = (int(v) for v in rng.integers(0, g.n, 2))
            truth, _ = O.bellman_ford(g.n, g.off, g.tgt, g.w, s)
            d, scanned = c.query(s, t)
            assert d == (float("inf") if truth[t] == INF else truth[t])
            assert scanned <= c.depth * (c.depth + 1)
        w = g.w.copy()
        for _ in range(3):
            picked = rng.choice(g.m, 10, replace=False)
            old = w[picked].copy()
            for new in (old + rng.integers(1, 50, 10), old):
                w[picked] = new
                c.update_edges(list(zip(picked.tolist(), new.tolist())))
                assert not c.update_created_negative_cycle()
                full = CCH(g.with_weights(w), order=c.rank, log=lambda *_: None)
                full.customize(g.with_weights(w))
                assert torch.equal(c.ud, full.ud)
        k = int(g.off[int(np.nonzero(np.diff(g.off) > 0)[0][0])])
        u = int(np.searchsorted(g.off, k, side="right") - 1)
        v = int(g.tgt[k])
        if u not in g.tgt[g.off[v]:g.off[v + 1]]:
            continue
        old = int(w[k])
        before = c.ud.clone()
        c.update_edges([(k, -10 ** 6)])
        assert c.update_created_negative_cycle() and c.has_negative_cycle()
        assert c.last_changed == sorted(c.last_changed)
        c.update_edges([(k, old)])
        assert not c.has_negative_cycle()
        assert torch.equal(c.ud, before), "reject-and-restore must restore bitwise"


def test_stop_on_negative_cycle_witness(c0):
    """customize(g, stop_on_negative_cycle=True) returns the reference's witness:
    the first negative arc of a vertex checked during elimination
    (python/cch.py:270-278), else the full list."""
    g, rank = c0
    k = int(g.off[100])
    v = int(g.tgt[k])
    back = [j for j in range(g.off[v], g.off[v + 1]) if g.tgt[j] == 100][0]
    w = g.w.copy()
    w[k], w[back] = 0, -1
    c = CCH(g.with_weights(w), order=rank, log=lambda *_: None)
    full = c.customize(g.with_weights(w))
    st = O.Structure(g.n, g.off, g.tgt, rank)
    up, dn, _ = st.customize(w)
    neg = np.nonzero((up != INF) & (dn != INF) & (up.astype(np.int64) + dn < 0))[0]
    assert full == neg.tolist() and len(full) > 0
    early = c.customize(g.with_weights(w), stop_on_negative_cycle=True)
    wit = st.customize_stop(w)   # the reference's elimination loop with early checks
    assert early == ([wit] if wit >= 0 else neg.tolist())


def _golden():
    from pathlib import Path
    p = Path(__file__).resolve().parent / "golden" / "cch_py_vectors.npz"
    with np.load(p) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("gpu_index", [False, True])
@pytest.mark.parametrize("name", ["grid", "small0", "small1", "small2"])
def test_dropin_matches_reference_python_vectors(name, gpu_index):
    """The CCH drop-in on the GPU against vectors the reference's python/cch.py
    wrote: customized values, every update's evaluated count and changed list,
    the final values, and both witness forms of a negative cycle -- with the
    triangle index and side lists built on the host or on the GPU."""
    G = _golden()
    print(f"case: golden vectors {name} gpu_index={gpu_index}", flush=True)
    off, tgt, w, rank = (G[f"{name}/{k}"] for k in ("off", "tgt", "w", "rank"))
    n = len(rank)
    g = graphs.RoadGraph(n, off, tgt, w.astype(np.int32), np.zeros(n), np.zeros(n))
    c = CCH(g, order=rank, log=lambda *_: None, gpu_index=gpu_index)
    assert c.customize(g) == []
    ud = c.ud.cpu().numpy()
    assert np.array_equal(ud[:, 0], G[f"{name}/up0"]) and np.array_equal(ud[:, 1], G[f"{name}/dn0"])
    for u in range(int(G[f"{name}/n_updates"])):
        edges, new = G[f"{name}/u{u}/edges"], G[f"{name}/u{u}/new_w"]
        ev = c.update_edges(list(zip(edges.tolist(), new.tolist())))
        assert ev == int(G[f"{name}/u{u}/evaluated"]), (u, ev)
        assert c.last_changed == G[f"{name}/u{u}/last_changed"].tolist(), u
        assert c.update_created_negative_cycle() == bool(G[f"{name}/u{u}/negative"])
    ud = c.ud.cpu().numpy()
    assert np.array_equal(ud[:, 0], G[f"{name}/up_end"]) and np.array_equal(ud[:, 1], G[f"{name}/dn_end"])
    gn = g.with_weights(G[f"{name}/neg_w"])
    c2 = CCH(gn, order=rank, log=lambda *_: None, gpu_index=gpu_index)
    assert c2.customize(gn) == G[f"{name}/neg_witnesses"].tolist()
    assert c2.customize(gn, stop_on_negative_cycle=True) == G[f"{name}/neg_stop_witness"].tolist()
    # and the values the reference leaves when it stops
    ud = c2.ud.cpu().numpy()
    assert np.array_equal(ud[:, 0], G[f"{name}/neg_stop_up"]) and np.array_equal(ud[:, 1], G[f"{name}/neg_stop_dn"])


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_updates_queued_back_to_back(c0, mode):
    """Twelve updates issued without any host read in between (more than the
    staging slots) end in the oracle's state after the same twelve updates."""
    g, rank = c0
    print(f"case: config 0, 12 queued updates, mode={mode}", flush=True)
    dt, st = _setup(g, rank)
    w = dt.weights(g.w)
    ud = dt.customize(w)
    up, dn, _ = st.customize(g.w)
    upd = O.Updater(st, g.w, up, dn)
    rng = np.random.default_rng(40 + mode)
    results = []
    for rnd in range(12):
        k = [1, 3, 40][rnd % 3]
        idx = rng.choice(g.m, k, replace=False)
        new = (upd.w[idx] + rng.integers(1, 300, k)).astype(np.int32)   # raises: no cycles
        results.append(dt.update(w, ud, idx, new, mode=mode).clone())
        ev0, _, _ = upd.update(idx, new)
        results[-1] = (results[-1], ev0)
    torch.cuda.synchronize()
    for res, ev0 in results:
        assert int(res[0].item()) == ev0
    got = ud.cpu().numpy()
    assert np.array_equal(got[:, 0], upd.up) and np.array_equal(got[:, 1], upd.dn)
    assert dt.status() == 0


def test_update_graphs_two_streams_two_metrics(c0):
    """Per-height launches replayed from captured graphs serve any buffers:
    two metrics updated alternately, each on its own stream (so each with its
    own argument record and graphs), the second stream's updates starting at
    other heights; both end in the oracle's state, and every result equals the
    oracle's evaluated count."""
    g, rank = c0
    print("case: config 0, graph-replayed updates of two metrics on two streams", flush=True)
    dt, st = _setup(g, rank)
    assert dt.get_tuning("update_graph") == 1
    W = [g.w, (g.w + np.random.default_rng(5).integers(0, 200, g.m)).astype(np.int32)]
    ws = [dt.weights(x) for x in W]
    uds = [dt.customize(x) for x in ws]
    upds = []
    for x in W:
        up, dn, _ = st.customize(x)
        upds.append(O.Updater(st, x, up, dn))
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    rng = np.random.default_rng(77)
    results = []
    for rnd in range(10):
        i = rnd % 2
        k = [20, 300, 5][rnd % 3]
        idx = rng.choice(g.m, k, replace=False)
        new = (upds[i].w[idx] + rng.integers(1, 400, k)).astype(np.int32)
        # the update workspace is shared: one update at a time per handle
        streams[i].wait_stream(streams[1 - i])
        res = dt.update(ws[i], uds[i], idx, new, mode=2, stream=streams[i])
        ev0, _, _ = upds[i].update(idx, new)
        with torch.cuda.stream(streams[i]):
            results.append((res.clone(), ev0))
    torch.cuda.synchronize()
    for res, ev0 in results:
        assert int(res[0].item()) == ev0
    for i in range(2):
        got = uds[i].cpu().numpy()
        assert np.array_equal(got[:, 0], upds[i].up) and np.array_equal(got[:, 1], upds[i].dn), i
        assert np.array_equal(ws[i].cpu().numpy(), upds[i].w), i
    assert dt.status() == 0


@pytest.mark.parametrize("idx", [0, 4, 8, 12])
def test_stop_state_equals_oracle(idx):
    """customize(g, stop_on_negative_cycle=True) through the drop-in leaves the
    oracle's restatement of the reference's stopped values (python/cch.py:
    270-278) on graphs with an injected negative cycle; a negative self-loop
    leaves the base values (python/cch.py:259-263)."""
    g = small_graphs()[idx] if idx < 12 else graphs.make_config(0)
    rank = order_of(g) if idx < 12 else _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    print(f"case: stop state, graph {idx}", flush=True)
    tri = graphs.find_directed_triangle(g, np.random.default_rng(idx))
    if tri is not None:
        wn = graphs.inject_negative_cycle(g.w, tri, total=-3)
    else:  # a grid has no triangle: a two-way street of total -3 instead
        k = int(g.off[50])
        v = int(g.tgt[k])
        back = [j for j in range(g.off[v], g.off[v + 1]) if g.tgt[j] == 50][0]
        wn = g.w.astype(np.int32).copy()
        wn[k], wn[back] = 0, -3
    gn = g.with_weights(wn)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    wit, up, dn = st.customize_stop(wn, with_state=True)
    c = CCH(gn, order=rank, log=lambda *_: None)
    got = c.customize(gn, stop_on_negative_cycle=True)
    fu, fd, _ = st.customize(wn)
    full = np.nonzero((fu != INF) & (fd != INF) & (fu.astype(np.int64) + fd < 0))[0].tolist()
    if wit >= 0:
        assert got == [wit]
        ud = c.ud.cpu().numpy()
        assert np.array_equal(ud[:, 0], up) and np.array_equal(ud[:, 1], dn)
    else:
        assert got == full


def test_negative_self_loop_leaves_base_values():
    g0 = graphs.shifted_random(80, seed=5)
    arcs = [(u, int(v), int(w)) for u in range(g0.n) for v, w in zip(g0.tgt[g0.off[u]:g0.off[u + 1]],
                                                                     g0.w[g0.off[u]:g0.off[u + 1]])]
    arcs.append((7, 7, -1))
    g = graphs.from_arcs(g0.n, arcs, xy=list(zip(g0.x, g0.y)))
    rank = order_of(g)
    c = CCH(g, order=rank, log=lambda *_: None)
    assert c.customize(g, stop_on_negative_cycle=True) == [-1]
    st = O.Structure(g.n, g.off, g.tgt, rank)
    base_up = np.full(st.mp, INF, np.int64)
    base_dn = np.full(st.mp, INF, np.int64)
    m = st.e_arc >= 0
    np.minimum.at(base_up, st.e_arc[m & (st.e_up == 1)], g.w[m & (st.e_up == 1)])
    np.minimum.at(base_dn, st.e_arc[m & (st.e_up == 0)], g.w[m & (st.e_up == 0)])
    ud = c.ud.cpu().numpy()
    assert np.array_equal(ud[:, 0], base_up) and np.array_equal(ud[:, 1], base_dn)
    assert c.customize(g) == [-1]
