"""GPU shortest paths with their vertices (python/cch.py:406-507): the reference
Python's own paths (golden vectors), the oracle's restatement on larger graphs,
and path validity (consecutive vertices joined by arcs whose weights sum to the
distance)."""
from pathlib import Path

import numpy as np
import pytest

from conftest import order_of, small_graphs
from nwcch import _lib, graphs
from nwcch.cch import CCH
from nwcch.device import DeviceTopology
from oracle import oracle as O

pytestmark = pytest.mark.gpu
INF = O.INF


def _path_weight(g, path):
    """Sum over consecutive vertices of the lightest arc between them, or None."""
    tot = 0
    for u, v in zip(path[:-1], path[1:]):
        s, e = int(g.off[u]), int(g.off[u + 1])
        ws = [int(g.w[k]) for k in range(s, e) if int(g.tgt[k]) == v]
        if not ws:
            return None
        tot += min(ws)
    return tot


@pytest.mark.parametrize("name", ["grid", "small0", "small1", "small2"])
def test_paths_equal_reference_python(name):
    with np.load(Path(__file__).resolve().parent / "golden" / "cch_py_vectors.npz") as d:
        G = {k: d[k] for k in d.files if k.startswith(name + "/")}
    print(f"case: golden paths {name}", flush=True)
    off, tgt, w, rank = (G[f"{name}/{k}"] for k in ("off", "tgt", "w", "rank"))
    n = len(rank)
    g = graphs.RoadGraph(n, off, tgt, w.astype(np.int32), np.zeros(n), np.zeros(n))
    c = CCH(g, order=rank, log=lambda *_: None)
    assert c.customize(g) == []
    po, pv = G[f"{name}/path_off"], G[f"{name}/path_v"]
    for i, (s, t) in enumerate(zip(G[f"{name}/path_s"], G[f"{name}/path_t"])):
        d, path = c.query_path(int(s), int(t))
        want = int(G[f"{name}/path_dist"][i])
        assert (d == float("inf") and want == INF) or d == want, (i, d, want)
        assert path == pv[po[i]:po[i + 1]].tolist(), i


@pytest.mark.parametrize("which", ["c0", "c1", "small"])
def test_paths_equal_oracle_and_are_valid(c0, which):
    if which == "c0":
        g, rank = c0
    elif which == "c1":
        g = graphs.make_config(1)
        rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    else:
        g = small_graphs()[9]
        rank = order_of(g)
    print("case: paths", which, g.n, flush=True)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    ud = dt.customize(g.w)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    up, dn, _ = st.customize(g.w)
    s, t = graphs.random_pairs(g.n, 64, seed=13)
    dist, plen, path = dt.query_path(ud, rank[s], rank[t], max_len=8192)
    assert dt.status() == 0
    dist, plen, path = dist.cpu().numpy(), plen.cpu().numpy(), path.cpu().numpy()
    inv = np.argsort(rank)
    d0, _, _ = st.query(up, dn, s, t)
    assert np.array_equal(dist, d0)
    for i in range(len(s)):
        dw, pw = st.query_path(up, dn, int(s[i]), int(t[i]))
        got = inv[path[i, :plen[i]]].tolist() if plen[i] >= 0 else []
        assert int(dist[i]) == dw and got == pw, i
        if got:
            assert got[0] == s[i] and got[-1] == t[i]
            assert _path_weight(g, got) == int(dist[i])


def test_path_edge_cases():
    """s == t, unreachable pairs and a path longer than max_len."""
    g = graphs.from_arcs(6, [(0, 1, 3), (1, 2, -2), (2, 3, 4), (3, 4, 1), (4, 0, 9), (5, 4, 2)])
    rank = np.array([5, 0, 3, 1, 4, 2], np.int32)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    ud = dt.customize(g.w)
    s = np.array([2, 0, 4, 0], np.int32)
    t = np.array([2, 4, 5, 4], np.int32)
    dist, plen, path = dt.query_path(ud, rank[s], rank[t], max_len=3)
    inv = np.argsort(rank)
    dist, plen, path = dist.cpu().numpy(), plen.cpu().numpy(), path.cpu().numpy()
    assert dist[0] == 0 and plen[0] == 1 and inv[path[0, 0]] == 2
    assert dist[1] == 6 and plen[1] == 5                       # 0 1 2 3 4, truncated to 3
    assert inv[path[1, :3]].tolist() == [0, 1, 2]
    assert dist[2] == INF and plen[2] == -1                    # nothing reaches 5
    assert dt.status() & 8                                      # the truncated path
    dist, plen, path = dt.query_path(ud, rank[s], rank[t], max_len=8)
    assert dt.status() == 0 and inv[path[1, :5].cpu().numpy()].tolist() == [0, 1, 2, 3, 4]


def test_path_limit_on_negative_cycle():
    """A metric with negative cycles: every unpacking ends, the ones the oracle
    stops at 4n + 16 vertices report length -2 and NWCCH_STATUS_PATH_LIMIT, the
    others equal the oracle's paths."""
    from test_oracle import negative_cycle_metric
    g, rank, w = negative_cycle_metric()
    print("case: path limit, negative-cycle metric", g.n, flush=True)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    ud = dt.customize(w)
    dt.status(clear=True)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    up, dn, _ = st.customize(w)
    s, t = graphs.random_pairs(g.n, 32, seed=3)
    dist, plen, path = (x.cpu().numpy() for x in dt.query_path(ud, rank[s], rank[t], max_len=4 * g.n + 16))
    inv = np.argsort(rank)
    stopped = 0
    for i in range(len(s)):
        dw, pw = st.query_path(up, dn, int(s[i]), int(t[i]))
        assert int(dist[i]) == dw, i
        if pw is None:
            assert plen[i] == -2, i
            stopped += 1
        else:
            assert inv[path[i, :plen[i]]].tolist() == pw if plen[i] >= 0 else pw == [], i
    assert stopped > 0 and dt.status() & 16


@pytest.mark.parametrize("form", ["chain_walk_throughput", "chain_walk_latency", "chunked_latency",
                                  "blocked_1", "blocked_2", "blocked_4", "blocked_8", "split_2", "split_8"])
@pytest.mark.parametrize("name", ["grid", "small0", "small1", "small2"])
def test_query_scanned_equals_reference_python(name, form):
    """Every query kernel form against the distances and scanned-arc counts the
    reference's python/cch.py CCH.query returned (python/cch.py:675-732)."""
    from pathlib import Path
    G = np.load(Path(__file__).resolve().parent / "golden" / "cch_py_vectors.npz")
    print(f"case: golden queries {name} {form}", flush=True)
    off, tgt, w, rank = (G[f"{name}/{k}"] for k in ("off", "tgt", "w", "rank"))
    n = len(rank)
    topo = _lib.Topology.from_graph(n, off, tgt, rank.astype(np.int32))
    dt = DeviceTopology(topo)
    ud = dt.customize(w.astype(np.int32))
    if form.startswith("blocked") or form.startswith("split"):
        dt.set_tuning("query_blk", 1)
        dt.set_tuning("query_blk_waves", int(form.split("_")[1]))
    else:
        chunk, deep = {"chain_walk_throughput": (0, 0), "chain_walk_latency": (0, 1 << 30),
                       "chunked_latency": (1, 1 << 30)}[form]
        dt.set_tuning("query_blk", 0)
        dt.set_tuning("query_chunk", chunk)
        dt.set_tuning("query_deep_max", deep)
    qs, qt = rank[G[f"{name}/q_s"]], rank[G[f"{name}/q_t"]]
    if form.startswith("split"):
        dist, sc = dt.query_split(*dt.split(ud), qs, qt)
    else:
        dist, sc = dt.query(ud, qs, qt)
    assert np.array_equal(dist.cpu().numpy().astype(np.int64), G[f"{name}/q_dist"])
    assert np.array_equal(sc.cpu().numpy().astype(np.int64), G[f"{name}/q_scanned"])
    dt.check_status()
