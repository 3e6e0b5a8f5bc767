"""GPU parity of the HIP path against the CPU oracle: customization, the
negative-cycle scan and ET queries must be bit-identical (integer work)."""
import numpy as np
import pytest
import torch

from conftest import order_of, small_graphs
from nwcch import _lib, cch_nb, graphs
from nwcch.device import DeviceTopology
from oracle import oracle as O

pytestmark = pytest.mark.gpu
INF = O.INF


FLAT_OFF = 1 << 30
# level sweep forms: team kernel everywhere, triangle-parallel everywhere, mixed
FLAT = [FLAT_OFF, 0, 64]


def _setup(g, rank, flat=FLAT_OFF):
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    dt.set_tuning("flat_min_tri", flat)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    return dt, st


def _check_customize(dt, st, w):
    ud = dt.customize(w)
    up, dn, ovf = st.customize(w)
    got = ud.cpu().numpy()
    assert np.array_equal(got[:, 0], up), "up differs"
    assert np.array_equal(got[:, 1], dn), "down differs"
    assert dt.status() == 0 and not ovf
    cnt, first = (int(v) for v in dt.negcycle_scan(ud).cpu())
    assert (cnt, first) == st.negcycle(up, dn)
    return ud, up, dn


@pytest.mark.parametrize("flat", FLAT)
def test_customize_scan_query_c0(c0, flat):
    g, rank = c0
    print("case: config 0 grid", g.n, g.m, "flat_min_tri", flat, flush=True)
    dt, st = _setup(g, rank, flat)
    ud, up, dn = _check_customize(dt, st, g.w)
    s, t = graphs.random_pairs(g.n, 1000, seed=2)
    dist, sc = dt.query(ud, rank[s], rank[t])
    d0, s0, _ = st.query(up, dn, s, t)
    assert np.array_equal(dist.cpu().numpy(), d0)
    assert np.array_equal(sc.cpu().numpy().astype(np.int64), s0)
    # against Bellman-Ford on a sample (the config-0 acceptance check)
    for i in range(0, 1000, 50):
        bf, _ = O.bellman_ford(g.n, g.off, g.tgt, g.w, int(s[i]))
        assert dist[i].item() == bf[t[i]]
    # s == t and the scan bound
    dist2, _ = dt.query(ud, rank[:64], rank[:64])
    assert (dist2 == 0).all()
    assert int(sc.max()) <= st.depth * (st.depth + 1)


@pytest.mark.parametrize("flat", [FLAT_OFF, 0])
@pytest.mark.parametrize("idx", range(12))
def test_small_family(idx, flat):
    g = small_graphs()[idx]
    print("case: small family", idx, g.name, "flat_min_tri", flat, flush=True)
    rank = order_of(g)
    dt, st = _setup(g, rank, flat)
    ud, up, dn = _check_customize(dt, st, g.w)
    qs = np.repeat(np.arange(g.n, dtype=np.int32), 3)
    qt = np.random.default_rng(idx).integers(0, g.n, len(qs)).astype(np.int32)
    dist, sc = dt.query(ud, rank[qs], rank[qt])
    d0, s0, _ = st.query(up, dn, qs, qt)
    assert np.array_equal(dist.cpu().numpy(), d0)
    assert np.array_equal(sc.cpu().numpy().astype(np.int64), s0)


def test_road_config1_bit_exact():
    g = graphs.make_config(1)
    print("case: config 1 road graph", g.n, g.m, flush=True)
    rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    dt, st = _setup(g, rank)
    ud, up, dn = _check_customize(dt, st, g.w)
    for flat in (0, 128):
        print("case: config 1 road graph, flat_min_tri", flat, flush=True)
        dt.set_tuning("flat_min_tri", flat)
        _check_customize(dt, st, g.w)
    # store mode x XCD slicing of the triangle-parallel levels
    dt.set_tuning("flat_min_tri", 0)
    for store in (0, 1, 2):
        for xcd in (0, 1):
            print(f"case: config 1, store {store} xcd {xcd}", flush=True)
            dt.set_tuning("flat_store", store)
            dt.set_tuning("flat_xcd", xcd)
            _check_customize(dt, st, g.w)
    dt.set_tuning("flat_store", 0)
    dt.set_tuning("flat_xcd", 0)
    dt.set_tuning("flat_min_tri", FLAT_OFF)
    s, t = graphs.random_pairs(g.n, 2000, seed=11)
    dist, sc = dt.query(ud, rank[s], rank[t])
    d0, s0, _ = st.query(up, dn, s, t, threads=8)
    assert np.array_equal(dist.cpu().numpy(), d0)
    assert np.array_equal(sc.cpu().numpy().astype(np.int64), s0)
    # every query kernel gives the same answers: the blocked form with 1, 2, 4
    # or 8 wavefronts per side, in small and large launches, and the chain walk
    # (deep or shallow prefetch, chunked ancestors) with 1, 2 or 4 queries per
    # workgroup
    for waves in (1, 2, 4, 8):
        dt.set_tuning("query_blk_waves", waves)
        dist1, sc1 = dt.query(ud, rank[s], rank[t])
        assert torch.equal(dist1, dist) and torch.equal(sc1, sc), waves
        d1, s1 = dt.query(ud, rank[s[:7]], rank[t[:7]])
        assert torch.equal(d1, dist[:7]) and torch.equal(s1, sc[:7]), waves
    for unroll in (4, 16):
        dt.set_tuning("query_blk_unroll", unroll)
        dist1, sc1 = dt.query(ud, rank[s], rank[t])
        assert torch.equal(dist1, dist) and torch.equal(sc1, sc), unroll
    dt.set_tuning("query_blk_unroll", 8)
    # the metric as separate up / down arrays (the reference's layout)
    up_d, dn_d = dt.split(ud)
    assert np.array_equal(up_d.cpu().numpy(), up) and np.array_equal(dn_d.cpu().numpy(), dn)
    for waves in (0, 1, 2, 4, 8):
        for unroll in (8, 16):
            dt.set_tuning("query_blk_waves", waves)
            dt.set_tuning("query_blk_unroll", unroll)
            dist1, sc1 = dt.query_split(up_d, dn_d, rank[s], rank[t])
            assert torch.equal(dist1, dist) and torch.equal(sc1, sc), ("split", waves, unroll)
    dt.set_tuning("query_blk_unroll", 8)
    dt.set_tuning("query_blk_waves", 0)
    # a launch above query_blk_small on this shallow hierarchy takes the chain
    # walk; forced onto the blocked kernel it gives the same answers
    s5, t5 = graphs.random_pairs(g.n, 5000, seed=12)
    d5, c5 = dt.query(ud, rank[s5], rank[t5])
    dt.set_tuning("query_blk_small", 1 << 30)
    d6, c6 = dt.query(ud, rank[s5], rank[t5])
    dt.set_tuning("query_blk_small", 4096)
    assert torch.equal(d5, d6) and torch.equal(c5, c6)
    dt.set_tuning("query_blk", 0)
    chunk0 = dt.get_tuning("query_chunk")
    for chunk in (0, 1):
        for deep in (32768, 0):
            for wpb in (1, 2, 4):
                dt.set_tuning("query_chunk", chunk)
                dt.set_tuning("query_deep_max", deep)
                dt.set_tuning("query_wpb", wpb)
                dist1, sc1 = dt.query(ud, rank[s], rank[t])
                assert torch.equal(dist1, dist) and torch.equal(sc1, sc), (chunk, deep, wpb)
    dt.set_tuning("query_chunk", chunk0)
    dt.set_tuning("query_deep_max", 32768)
    dt.set_tuning("query_wpb", 1)
    dt.set_tuning("query_blk", 1)
    p, neg = O.johnson(g.n, g.off, g.tgt, g.w)
    assert not neg
    dj = O.Dijkstra(g.n, g.off, g.tgt, g.w, p)
    for i in range(20):
        assert dist[i].item() == dj(s[i], t[i])


@pytest.mark.parametrize("flat", [FLAT_OFF, 0])
def test_injected_negative_cycle_is_flagged(c0, flat):
    """A 3-arc loop of total -1 (the batched config's injected cycle) is found by
    the scan, with the oracle's count and first witness."""
    g, rank = c0
    tri = graphs.find_directed_triangle(g)
    if tri is None:  # the grid has no directed triangles: use a 2-cycle
        k = int(g.off[10])
        v = int(g.tgt[k])
        back = [j for j in range(g.off[v], g.off[v + 1]) if g.tgt[j] == 10][0]
        w = g.w.copy()
        w[k], w[back] = 0, -1
    else:
        w = graphs.inject_negative_cycle(g.w, tri)
    dt, st = _setup(g, rank, flat)
    ud, up, dn = _check_customize(dt, st, w)
    cnt = int(dt.negcycle_scan(ud)[0])
    assert cnt > 0 and O.has_negative_cycle(g.n, g.off, g.tgt, w)
    lst = dt.negcycle_list(ud)
    want = np.nonzero((up != INF) & (dn != INF) & (up.astype(np.int64) + dn < 0))[0]
    assert lst == want.tolist()


def test_unreachable_pairs_and_inf():
    """One-way arcs leave pairs unreachable: INF must absorb negative arcs."""
    g = graphs.from_arcs(5, [(0, 1, 3), (1, 2, -2), (3, 4, -7), (4, 3, 9), (2, 0, 4)])
    rank = np.array([4, 0, 2, 1, 3], np.int32)
    dt, st = _setup(g, rank)
    ud, up, dn = _check_customize(dt, st, g.w)
    s = np.array([0, 0, 3, 1, 4], np.int32)
    t = np.array([2, 3, 4, 0, 0], np.int32)
    dist, _ = dt.query(ud, rank[s], rank[t])
    assert dist.cpu().tolist() == [1, INF, -7, 2, INF]


@pytest.mark.parametrize("flat", [FLAT_OFF, 0])
def test_overflow_is_reported_not_saturated(flat):
    g = graphs.from_arcs(3, [(0, 1, 2 ** 30 + 5), (1, 2, 2 ** 30 + 5)])
    rank = np.array([1, 0, 2], np.int32)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    dt.set_tuning("flat_min_tri", flat)
    dt.customize(g.w)
    assert dt.status() & 1
    st = O.Structure(g.n, g.off, g.tgt, rank)
    assert st.customize(g.w)[2]


def test_bad_query_vertex_sets_status(c0):
    g, rank = c0
    dt, st = _setup(g, rank)
    ud = dt.customize(g.w)
    dist, _ = dt.query(ud, [0, g.n + 5], [1, 2])
    assert dist[1].item() == INF
    with pytest.raises(_lib.NwcchError, match="out of range"):
        dt.check_status()


def test_cch_nb_dropin(c0):
    """The python/cch_nb.py call sequence of run_nb.py:79-141, on the GPU."""
    g, _ = c0
    noff, nbr, rev = cch_nb.undirected_csr(g.n, g.off, g.tgt)
    rank = cch_nb.nested_dissection(g.n, noff, nbr, rev, g.x, g.y, 64)
    first, head, parent = cch_nb.symbolic(g.n, noff, nbr, rank)
    e_arc, e_up = cch_nb.edge_arcs(g.n, g.off, g.tgt, rank, first, head)
    upw, dnw = cch_nb.customize(g.n, first, head, e_arc, e_up, g.w.astype(np.float64))
    assert cch_nb.negative_cycle_arcs(upw, dnw) == 0
    st = O.Structure(g.n, g.off, g.tgt, rank.astype(np.int32))
    up, dn, _ = st.customize(g.w)
    assert np.array_equal(upw.cpu().numpy(), up) and np.array_equal(dnw.cpu().numpy(), dn)
    df = np.full(g.n, np.inf)
    s, t = graphs.random_pairs(g.n, 20, seed=7)
    d0, _, _ = st.query(up, dn, s, t)
    for i in range(20):
        got = cch_nb.query(rank[s[i]], rank[t[i]], first, head, parent, upw, dnw, df, df, None, i + 1)
        assert got == (np.inf if d0[i] == INF else d0[i])
    # separate (non-view) arrays go through an interleaving copy
    assert cch_nb.negative_cycle_arcs(upw.clone(), dnw.clone()) == 0
    dist, _ = cch_nb.query_many(rank[s], rank[t], first, head, upw, dnw)
    assert np.array_equal(dist.cpu().numpy(), d0)


@pytest.mark.parametrize("which", ["c0", "c1", "small"])
def test_potential_equals_johnson(c0, which):
    """GPU CCH potential (two level sweeps) equals the oracle's Johnson
    virtual-source distances, bit for bit, and is feasible."""
    if which == "c0":
        g, rank = c0
    elif which == "c1":
        g = graphs.make_config(1)
        rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    else:
        g = small_graphs()[5]
        rank = order_of(g)
    print("case: potential", which, g.n, flush=True)
    dt, st = _setup(g, rank)
    ud = dt.customize(g.w)
    pr = dt.potential(ud).cpu().numpy()
    assert dt.status() == 0
    p = pr[rank].astype(np.int64)          # by original vertex id
    want, neg = O.johnson(g.n, g.off, g.tgt, g.w)
    assert not neg
    assert np.array_equal(p, want)
    tails = np.repeat(np.arange(g.n), np.diff(g.off))
    assert (g.w.astype(np.int64) + p[tails] - p[g.tgt] >= 0).all()
    # the replayed launch sequence with other buffers, and launch by launch
    ud2 = dt.customize(g.w)
    assert torch.equal(dt.potential(ud2), dt.potential(ud))
    dt.set_tuning("customize_graph", 0)
    assert np.array_equal(dt.potential(ud).cpu().numpy(), pr)
    dt.set_tuning("customize_graph", 1)


def test_customize_graph_paths(c0):
    """nwcch_customize replays a captured launch sequence per stream, with the
    buffers of each call passed through an argument slot: calls with changing
    weights and outputs, on two streams at once, equal the launch-by-launch
    path and the oracle."""
    g, rank = c0
    print("case: customize graph paths on config 0", flush=True)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    rng = np.random.default_rng(4)
    ws = [(g.w + rng.integers(0, 40, g.m)).astype(np.int32) for _ in range(4)]
    want = [st.customize(w)[:2] for w in ws]
    wd = [dt.weights(w) for w in ws]
    outs = [torch.full((topo.mp, 2), -5, dtype=torch.int32, device=dt.dev) for _ in range(4)]
    for rep in range(2):
        for k in range(4):                       # one stream, new buffers every call
            dt.customize(wd[k], out=outs[k])
        torch.cuda.synchronize()
        for k in range(4):
            got = outs[k].cpu().numpy()
            assert np.array_equal(got[:, 0], want[k][0]) and np.array_equal(got[:, 1], want[k][1]), (rep, k)
            outs[k].fill_(-5)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for k in range(4):                           # two streams interleaved
        dt.customize(wd[k], out=outs[k], stream=s1 if k % 2 == 0 else s2)
    torch.cuda.synchronize()
    for k in range(4):
        got = outs[k].cpu().numpy()
        assert np.array_equal(got[:, 0], want[k][0]) and np.array_equal(got[:, 1], want[k][1]), k
    dt.set_tuning("customize_graph", 0)
    assert torch.equal(dt.customize(wd[1]), outs[1])
    dt.set_tuning("customize_graph", 1)
    assert dt.status() == 0


@pytest.mark.parametrize("min_len", [0, 16, 40])
def test_row_merge_levels(min_len):
    """Levels in row-merge form (every level, or those whose suffix pieces
    average at least 16 / 40 arcs) give the oracle's customization on config 1
    and on the small family; so do plain calls replayed from the library's
    captured launch sequence."""
    graphs_ = [(graphs.make_config(1), 64)] + [(g, 8) for g in small_graphs()[::3]]
    for g, leaf in graphs_:
        print(f"case: row-merge levels, min_len {min_len}, {g.name}", flush=True)
        rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, leaf)
        dt, st = _setup(g, rank)
        dt.set_tuning("rows_min_len", min_len)
        dt.set_tuning("rows_min_rows", 0)
        dt.set_tuning("rows_max_rows", 1 << 30)
        dt.set_tuning("rows_min_level_tri", 0)
        for _ in range(2):
            _check_customize(dt, st, g.w)
        W, _ = graphs.metric_family(g, 2, seed=5, hmax=800) if g.name.startswith("c") else (None, None)
        if W is not None:
            _check_customize(dt, st, W[1])
