"""Pins of the CPU oracle (oracle/cch_oracle.c) -- no GPU needed.

Each oracle function is checked against something other than itself: the
reference's own C++ (oracle/_ref), known-answer graphs from the reference's
tests, the exhaustive 3-vertex family, Bellman-Ford / Johnson+Dijkstra, the
closed form of a customized value, and algebraic invariants.
"""
import itertools

import numpy as np
import pytest

from conftest import order_of, small_graphs
from nwcch import graphs
from oracle import oracle as O

INF = O.INF


def _graph(n, arcs):
    return graphs.from_arcs(n, arcs)


def _truth(g, s):
    dist, neg = O.bellman_ford(g.n, g.off, g.tgt, g.w, s)
    assert not neg
    return dist


# ------------------------------------------------ known answers (test_techniques)

KATS = [
    # python/test_techniques.py:171-178 (order = vertex ids)
    ("Ex.1", 3, [(0, 1, 0), (0, 2, -1), (1, 2, -2)], 0, 2, -2),
    ("Ex.2", 4, [(0, 1, 0), (0, 2, 1), (2, 3, -1), (3, 1, -2)], 0, 1, -2),
    ("Ex.3", 3, [(0, 1, -3), (0, 2, -2), (2, 1, -2)], 0, 1, -4),
    # zero-length cycles (python/test_techniques.py:193-202)
    ("zero-cycle", 3, [(0, 1, -2), (1, 2, 0), (2, 0, 2)], 0, 1, -2),
    ("all-zero", 3, [(0, 1, 0), (1, 2, 0), (2, 0, 0)], 0, 1, 0),
    # up-down path through a lower triangle (python/test_techniques.py:204-214)
    ("lower-triangle", 3, [(1, 0, 1), (0, 2, 1)], 1, 2, 2),
]


@pytest.mark.parametrize("name,n,arcs,s,t,truth", KATS, ids=[k[0] for k in KATS])
def test_known_answers(name, n, arcs, s, t, truth):
    g = _graph(n, arcs)
    st = O.Structure(g.n, g.off, g.tgt, np.arange(n, dtype=np.int32))
    up, dn, ovf = st.customize(g.w)
    assert not ovf
    assert st.negcycle(up, dn)[0] == 0
    d, _, _ = st.query(up, dn, [s], [t])
    assert d[0] == truth
    assert _truth(g, s)[t] == truth


def test_inf_is_absorbing():
    """The saturated-int32 pitfall (python/test_techniques.py:216-220): INF + (-5)
    must stay INF.  Vertex 2 is unreachable from 0 but has a negative arc out of
    it, and 1 -> 2 does not exist: a saturating add would report 0 -> 3 finite."""
    g = _graph(4, [(0, 1, 5), (2, 3, -5), (3, 2, 7)])
    st = O.Structure(g.n, g.off, g.tgt, np.array([2, 0, 1, 3], np.int32))
    up, dn, _ = st.customize(g.w)
    d, _, _ = st.query(up, dn, [0, 0, 2], [3, 2, 3])
    assert d.tolist() == [INF, INF, -5]


def test_exhaustive_three_vertices():
    """All digraphs on 3 vertices with arc weights in {none, -2..2}, order = ids
    (python/test_techniques.py:363-385): the conservative ones (14,440,
    results/results_pruning.md:10) answer all 9 queries like Bellman-Ford, and
    the scan flags exactly the others (Theorem 3, docs/technical_note.md:129)."""
    pairs = [(u, v) for u in range(3) for v in range(3) if u != v]
    conservative = 0
    for ws in itertools.product([None, -2, -1, 0, 1, 2], repeat=6):
        arcs = [(u, v, w) for (u, v), w in zip(pairs, ws) if w is not None]
        g = _graph(3, arcs)
        st = O.Structure(3, g.off, g.tgt, np.arange(3, dtype=np.int32))
        up, dn, _ = st.customize(g.w)
        neg = O.has_negative_cycle(3, g.off, g.tgt, g.w)
        assert (st.negcycle(up, dn)[0] > 0) == neg, arcs
        if neg:
            continue
        conservative += 1
        qs = np.repeat(np.arange(3), 3)
        qt = np.tile(np.arange(3), 3)
        d, _, _ = st.query(up, dn, qs, qt)
        truth = np.concatenate([_truth(g, s) for s in range(3)])
        assert np.array_equal(d.astype(np.int64), truth), arcs
    assert conservative == 14440


# --------------------------------------------------- closed form and invariants

def test_customized_value_closed_form():
    """up[(v,w)] = shortest v->w walk with inner vertices ranked below both ends
    (python/cch.py:9-12), by brute force for every shortcut of small graphs."""
    for g in [graphs.ev_grid(6, seed=1), graphs.ev_geometric(40, seed=2), graphs.shifted_random(30, seed=3)]:
        rank = order_of(g, leaf=4)
        st = O.Structure(g.n, g.off, g.tgt, rank)
        up, dn, _ = st.customize(g.w)
        inv = np.argsort(rank)
        for x in range(g.n):
            for a in range(st.first[x], st.first[x + 1]):
                v, w = int(inv[x]), int(inv[st.head[a]])
                assert up[a] == O.restricted(g.n, g.off, g.tgt, g.w, rank, v, w)
                assert dn[a] == O.restricted(g.n, g.off, g.tgt, g.w, rank, w, v)


def test_potential_shift_metamorphic():
    """Customizing w + p(u) - p(v) yields up + p(tail) - p(head) exactly
    (docs/technical_note.md:80)."""
    g = graphs.ev_geometric(300, seed=5)
    rank = order_of(g)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    up, dn, _ = st.customize(g.w)
    p = np.random.default_rng(0).integers(-1000, 1000, g.n)
    tails = g.tails()
    w2 = (g.w + p[tails] - p[g.tgt]).astype(np.int32)
    up2, dn2, _ = st.customize(w2)
    inv = np.argsort(rank)
    tail_r = np.repeat(np.arange(g.n), np.diff(st.first))
    pv, pw = p[inv[tail_r]], p[inv[st.head]]
    fin = up != INF
    assert np.array_equal(up2[fin], up[fin] + pv[fin] - pw[fin])
    assert np.array_equal(up2 == INF, up == INF)
    fin = dn != INF
    assert np.array_equal(dn2[fin], dn[fin] + pw[fin] - pv[fin])


def test_level_parallel_equals_serial():
    """cpp/cch.cpp:844-850: the level-parallel customization is bitwise equal."""
    g = graphs.make_config(0)
    rank = order_of(g, 64)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    a = st.customize(g.w)
    b = st.customize_par(g.w, threads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_early_stop_witness_properties():
    """orc_customize_stop (python/cch.py:270-278) returns the smallest arc with a
    negative 2-cycle when its tail row ends before the last triangle's first arc
    (that vertex is checked during elimination), and -1 otherwise."""
    seen = set()
    for g in small_graphs():
        rank = order_of(g)
        st = O.Structure(g.n, g.off, g.tgt, rank)
        for k in range(0, g.m, max(1, g.m // 15)):
            u = int(np.searchsorted(g.off, k, side="right") - 1)
            v = int(g.tgt[k])
            back = [j for j in range(g.off[v], g.off[v + 1]) if g.tgt[j] == u]
            if not back:
                continue
            w = g.w.copy()
            w[k], w[back[0]] = 0, -1
            up, dn, _ = st.customize(w)
            neg = np.nonzero((up != INF) & (dn != INF) & (up.astype(np.int64) + dn < 0))[0]
            assert len(neg) > 0
            wit = st.customize_stop(w)
            d = np.diff(st.first)
            a_max = int((st.first[np.nonzero(d >= 2)[0] + 1] - 2).max())
            x = int(np.searchsorted(st.first, neg[0], side="right") - 1)
            assert wit == (int(neg[0]) if st.first[x + 1] <= a_max else -1)
            seen.add(wit >= 0)
    assert seen == {True, False} or seen == {True}


def test_queries_match_bellman_ford_and_scan_bound():
    """python/test_cch.py:38-45: ET queries equal Bellman-Ford, and a query scans
    at most depth*(depth+1) arcs."""
    rng = np.random.default_rng(3)
    for g in small_graphs():
        rank = order_of(g)
        st = O.Structure(g.n, g.off, g.tgt, rank)
        up, dn, _ = st.customize(g.w)
        assert st.negcycle(up, dn)[0] == 0, "false negative cycle"
        s = rng.integers(0, g.n, 25)
        t = rng.integers(0, g.n, 25)
        d, sc, _ = st.query(up, dn, s, t)
        for i in range(25):
            assert d[i] == _truth(g, s[i])[t[i]]
        assert (sc <= st.depth * (st.depth + 1)).all()


def test_johnson_dijkstra_agrees_with_bellman_ford():
    g = graphs.ev_geometric(300, seed=1)
    p, neg = O.johnson(g.n, g.off, g.tgt, g.w)
    assert not neg
    dj = O.Dijkstra(g.n, g.off, g.tgt, g.w, p)
    for s in range(0, g.n, 37):
        truth = _truth(g, s)
        for t in range(0, g.n, 29):
            assert dj(s, t) == truth[t]


def test_negative_cycle_verdict_matches_bellman_ford():
    """python/test_cch.py:85-104: a very negative 2-cycle is flagged by the scan
    exactly when Bellman-Ford finds a negative cycle."""
    for g in small_graphs()[::3]:
        off, tgt, w = g.off, g.tgt, g.w.copy()
        u = 0
        while off[u + 1] == off[u]:
            u += 1
        k = off[u]
        v = tgt[k]
        back = [j for j in range(off[v], off[v + 1]) if tgt[j] == u]
        w[k] = -10 ** 6
        if not back:
            continue
        rank = order_of(g)
        st = O.Structure(g.n, off, tgt, rank)
        up, dn, _ = st.customize(w)
        assert st.negcycle(up, dn)[0] > 0
        assert O.has_negative_cycle(g.n, off, tgt, w)


# -------------------------------------------------------------- partial update

def _closed_form_update(st, w_old, w_new, changes):
    """evaluated and last_changed without the queue: last_changed = arcs whose
    full customization differs; evaluated = touched arcs plus the top arcs of all
    lower triangles having a changed arc as a side (Theorem 4)."""
    u0, d0, _ = st.customize(w_old)
    u1, d1, _ = st.customize(w_new)
    changed = np.nonzero((u0 != u1) | (d0 != d1))[0]
    touched = {int(st.e_arc[k]) for k in changes if st.e_arc[k] >= 0}
    ch = set(changed.tolist())
    tops = set()
    for x in range(st.n):
        s, e = st.first[x], st.first[x + 1]
        for i in range(s, e):
            y = st.head[i]
            j = st.first[y]
            for k in range(i + 1, e):
                while st.head[j] < st.head[k]:
                    j += 1
                if i in ch or k in ch:
                    tops.add(int(j))
    return len(touched | tops), changed, (u1, d1)


def test_update_matches_full_and_closed_form():
    """python/test_cch.py:47-68 and cch.py:349-402: raise 10 random arcs, then
    restore them; the partial update equals full customization bitwise, and its
    evaluated count / last_changed list equal their closed forms."""
    rng = np.random.default_rng(3)
    for g in small_graphs()[::2]:
        rank = order_of(g)
        st = O.Structure(g.n, g.off, g.tgt, rank)
        up, dn, _ = st.customize(g.w)
        upd = O.Updater(st, g.w, up, dn)
        w = g.w.copy()
        for _ in range(3):
            picked = rng.choice(g.m, 10, replace=False)
            old = w[picked].copy()
            for new in (old + rng.integers(1, 50, 10), old):
                w_before = w.copy()
                w[picked] = new
                ev, changed, ovf = upd.update(picked, new)
                assert not ovf
                ev2, changed2, (u1, d1) = _closed_form_update(st, w_before, w, picked)
                assert np.array_equal(upd.up, u1) and np.array_equal(upd.dn, d1)
                assert ev == ev2
                assert np.array_equal(changed, changed2)
                assert st.negcycle(upd.up[changed], upd.dn[changed])[0] == 0


def test_update_creates_and_clears_negative_cycle():
    """python/test_cch.py:70-83: a partial update that creates a negative cycle is
    caught from the changed arcs alone; restoring the weight clears it and
    restores the customization bitwise (reject-and-restore)."""
    g = graphs.ev_geometric(300, seed=0)
    rank = order_of(g)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    up, dn, _ = st.customize(g.w)
    upd = O.Updater(st, g.w, up, dn)
    k = int(np.nonzero(np.diff(g.off) > 0)[0][0])
    k = int(g.off[k])
    old = int(g.w[k])
    _, changed, _ = upd.update([k], [-10 ** 6])
    assert st.negcycle(upd.up[changed], upd.dn[changed])[0] > 0
    assert st.negcycle(upd.up, upd.dn)[0] > 0
    upd.update([k], [old])
    assert st.negcycle(upd.up, upd.dn)[0] == 0
    assert np.array_equal(upd.up, up) and np.array_equal(upd.dn, dn)


# ---------------------------------------------------- the reference's own code

need_ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (reference sources absent)")


@need_ref
@pytest.mark.parametrize("which", ["c0", "grid", "geo", "shift"])
def test_oracle_equals_reference_cpp(which):
    """Structure, customization, scan and queries of the oracle are bitwise equal
    to the reference's cpp/cch.cpp (CCH::symbolic/map_arcs/customize/
    negative_cycle_arcs, Query::run) on the same graph and order."""
    g = {"c0": lambda: graphs.make_config(0), "grid": lambda: graphs.ev_grid(15, 1),
         "geo": lambda: graphs.ev_geometric(300, 2), "shift": lambda: graphs.shifted_random(200, 3)}[which]()
    rank = order_of(g, 64 if which == "c0" else 8)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    ref = O.Reference(g.n, g.off, g.tgt, rank)
    ex = ref.export()
    for k in ("first", "head", "parent", "e_arc", "e_up"):
        assert np.array_equal(getattr(st, k), ex[k]), k
    assert (st.mp, st.triangles, st.depth) == (ref.mp, ref.triangles, ref.depth)
    assert np.array_equal(st.level, ref.levels())
    up, dn, _ = st.customize(g.w)
    ru, rd = ref.customize(g.w)
    assert np.array_equal(up, O.float_to_i32(ru)) and np.array_equal(dn, O.float_to_i32(rd))
    ru2, rd2 = ref.customize(g.w, threads=4)
    assert np.array_equal(ru, ru2) and np.array_equal(rd, rd2)
    assert st.negcycle(up, dn)[0] == ref.negcycle(ru, rd)
    s, t = graphs.random_pairs(g.n, 300, seed=4)
    d, _, _ = st.query(up, dn, s, t)
    assert np.array_equal(d, O.float_to_i32(ref.query(ru, rd, s, t)))


@need_ref
def test_negative_cycle_count_equals_reference_cpp():
    g = graphs.ev_geometric(300, seed=3)
    w = g.w.copy()
    k = int(g.off[5])
    v = int(g.tgt[k])
    back = [j for j in range(g.off[v], g.off[v + 1]) if g.tgt[j] == 5]
    assert back
    w[k] = -5000
    rank = order_of(g)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    ref = O.Reference(g.n, g.off, g.tgt, rank)
    up, dn, _ = st.customize(w)
    ru, rd = ref.customize(w)
    assert st.negcycle(up, dn)[0] == ref.negcycle(ru, rd) > 0


@pytest.mark.skipif(not O.have_ref(), reason="the reference's C++ was not built")
def test_johnson_potential_equals_reference_cch_potential():
    """The virtual-source distances the oracle's Johnson computes equal the
    reference's CCH potential read off its own customization
    (cpp/cch.cpp:635-655), on the small family and the 64x64 grid."""
    from conftest import order_of, small_graphs
    from nwcch import graphs
    cases = small_graphs()[::3] + [graphs.make_config(0)]
    for g in cases:
        rank = order_of(g)
        ref = O.Reference(g.n, g.off, g.tgt, rank)
        up, dn = ref.customize(g.w)
        p_ref = ref.potential(up, dn)
        p, neg = O.johnson(g.n, g.off, g.tgt, g.w)
        assert not neg
        assert np.array_equal(p_ref.astype(np.int64), p), g.name
        # feasibility: reduced weights are non-negative
        tails = np.repeat(np.arange(g.n), np.diff(g.off))
        assert (g.w.astype(np.int64) + p[tails] - p[g.tgt] >= 0).all()


def negative_cycle_metric():
    """A shifted random graph of 200 vertices whose 10 re-weighted arcs close
    negative cycles (305 arcs with up + down < 0), with its order."""
    g = small_graphs()[11]
    rng = np.random.default_rng(5)
    idx = rng.choice(g.m, size=10, replace=False)
    w = g.w.copy()
    w[idx] = rng.integers(1, 100, size=10).astype(np.int32)
    return g, order_of(g), w


def test_path_unpacking_stops_on_negative_cycle():
    """Through a negative cycle the shortcut expansion of python/cch.py:406-507
    grows exponentially with the depth (unbounded, this metric's first query
    ran past 30 s); the unpacking stops after 4n + 16 vertices and says so."""
    g, rank, w = negative_cycle_metric()
    st = O.Structure(g.n, g.off, g.tgt, rank)
    up, dn, _ = st.customize(w)
    assert (((up != O.INF) & (dn != O.INF) & (up.astype(np.int64) + dn < 0)).sum()) > 0
    s, t = graphs.random_pairs(g.n, 32, seed=3)
    stopped = 0
    for a, b in zip(s, t):
        d, path = st.query_path(up, dn, int(a), int(b))
        if path is None:
            stopped += 1
        else:
            assert len(path) <= 4 * g.n + 16 and (not path or (path[0] == a and path[-1] == b))
    assert stopped > 0
