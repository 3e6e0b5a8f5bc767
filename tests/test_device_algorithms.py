"""CPU restatements of device algorithms, checked against brute force, the
host build or the oracle: the segmented prefix minimum of the level kernel
(DPP row shifts and row broadcasts over a 64-lane wavefront), the pull-form
triangle enumeration of the GPU-built index (down lists, binary search in the
lower row, x ascending), the chordal completion by ancestor bitmaps, the
blocked ET query, the row-merge enumeration, and the outputs the
full-recustomization update driver derives.  The GPU tests check the kernels
themselves bit for bit; these pin the algorithms they implement."""
import numpy as np

from conftest import order_of, small_graphs
from nwcch import _lib

INF = 2**31 - 1


def _dpp(vals, op, row_mask):
    """One DPP move over a wave64: row_shr:n inside 16-lane rows, or
    row_bcast:15 / row_bcast:31; lanes without a source or outside row_mask
    receive INF (the identity of min)."""
    out = np.full(64, INF, np.int64)
    for lane in range(64):
        row = lane // 16
        if not (row_mask >> row) & 1:
            continue
        if op[0] == "shr":
            if lane % 16 >= op[1]:
                out[lane] = vals[lane - op[1]]
        elif op[0] == "bcast15":
            if row >= 1:
                out[lane] = vals[16 * row - 1]
        elif row >= 2:                                   # bcast31
            out[lane] = vals[31]
    return out


def _source(lane, op):
    if op[0] == "shr":
        return lane - op[1]
    return 16 * (lane // 16) - 1 if op[0] == "bcast15" else 31


STEPS = ((("shr", 1), 0xF), (("shr", 2), 0xF), (("shr", 4), 0xF), (("shr", 8), 0xF),
         (("bcast15",), 0xA), (("bcast31",), 0xC))


def test_segmented_prefix_min_network():
    """After the six steps, merging a source lane only when it lies in the same
    run, lane l holds the minimum over [start of its run, l]."""
    rng = np.random.default_rng(0)
    for trial in range(3000):
        cuts = np.sort(rng.choice(np.arange(1, 64), size=rng.integers(0, 40), replace=False))
        arc = np.zeros(64, np.int64)
        for k in cuts:
            arc[k:] += 1
        valid = rng.integers(1, 65)
        arc[valid:] = -1                                  # lanes past the level's end
        v = rng.integers(-100, 100, 64)
        v[rng.random(64) < 0.1] = INF
        v[valid:] = INF
        starts = [lane for lane in range(64) if lane == 0 or arc[lane - 1] != arc[lane]]
        s0 = np.array([max(h for h in starts if h <= lane) for lane in range(64)])
        x = v.copy()
        for op, mask in STEPS:
            moved = _dpp(x, op, mask)
            for lane in range(64):
                if _source(lane, op) >= s0[lane]:
                    x[lane] = min(x[lane], moved[lane])
        for lane in range(64):
            assert x[lane] == v[s0[lane]:lane + 1].min(), (trial, lane)


def test_pull_enumeration_equals_host_index():
    """For every arc (y, z): the entries a = (x, y) of y's down list, x
    ascending, whose row x holds z after a -- exactly the host's run of that
    arc, in the same order."""
    g = small_graphs()[0]
    t = _lib.Topology.from_graph(g.n, g.off, g.tgt, order_of(g))
    ex = t.export()
    first, head = ex["first"], ex["head"]
    tail = np.repeat(np.arange(t.n), np.diff(first))
    down = [[] for _ in range(t.n)]
    for a in range(t.mp):                                 # ascending arc id = ascending x
        down[head[a]].append(a)
    sched, toff, tab = t.export_triangles()
    for p in range(t.mp):
        j = sched[p]
        y, z = tail[j], head[j]
        run = []
        for a in down[y]:
            x = tail[a]
            row = head[a + 1:first[x + 1]]
            k = np.searchsorted(row, z)
            if k < len(row) and row[k] == z:
                run.append((a, a + 1 + k))
        assert [tuple(r) for r in tab[toff[p]:toff[p + 1]]] == run, p


def _undirected(g):
    nb = [set() for _ in range(g.n)]
    for u in range(g.n):
        for v in g.tgt[g.off[u]:g.off[u + 1]]:
            if v != u:
                nb[u].add(int(v))
                nb[int(v)].add(u)
    return nb


def test_symbolic_by_ancestor_bitmap_equals_host():
    """The device symbolic contraction's construction, restated: the elimination
    tree by Liu's union-find walk over the original graph, then by height,
    leaves first, every row as a bitmap over ancestor distances (bit i = the
    ancestor at distance i + 1) filled from the original upper neighbours and
    the children's rows minus their first entry, read out in bit order.  It
    equals the host contraction (first, head) on the reference's test family
    and on an identity order (a deep tree)."""
    cases = [(g, order_of(g)) for g in small_graphs()[::3]]
    g = small_graphs()[1]
    cases.append((g, np.arange(g.n, dtype=np.int32)))
    for g, rank in cases:
        n = g.n
        inv = np.empty(n, np.int64)
        inv[rank] = np.arange(n)
        nb = _undirected(g)
        parent = np.full(n, -1)
        anc = np.full(n, -1)
        for x in range(n):
            for u in nb[inv[x]]:
                r = rank[u]
                if r >= x:
                    continue
                while anc[r] != -1 and anc[r] != x:
                    anc[r], r = x, anc[r]
                if anc[r] == -1:
                    anc[r] = parent[r] = x
        depth = np.zeros(n, np.int64)
        height = np.zeros(n, np.int64)
        for x in range(n - 1, -1, -1):
            if parent[x] >= 0:
                depth[x] = depth[parent[x]] + 1
        for x in range(n):
            if parent[x] >= 0:
                height[parent[x]] = max(height[parent[x]], height[x] + 1)
        kids = [[] for _ in range(n)]
        for x in range(n):
            if parent[x] >= 0:
                kids[parent[x]].append(x)
        rows = [None] * n
        for x in sorted(range(n), key=lambda v: height[v]):
            D = depth[x]
            bits = np.zeros(D, bool)
            ids = np.full(D, -1)
            cand = [rank[u] for u in nb[inv[x]] if rank[u] > x]
            for c in kids[x]:
                assert rows[c][0] == x
                cand += rows[c][1:]
            for w in cand:
                b = D - 1 - depth[w]
                assert 0 <= b < D
                assert ids[b] in (-1, w)
                bits[b], ids[b] = True, w
            rows[x] = [int(w) for w in ids[bits]]
        t = _lib.Topology.from_graph(n, g.off, g.tgt, rank)
        ex = t.export()
        assert np.array_equal(ex["first"], np.concatenate([[0], np.cumsum([len(r) for r in rows])]))
        assert np.array_equal(ex["head"], np.concatenate([np.asarray(r, np.int64) for r in rows] + [[]]))
        assert np.array_equal(ex["parent"], parent)


def _iadd(a, b):
    """INF-absorbing int32 addition, saturating like the kernels (the
    overflow flag is not modelled)."""
    if a == INF or b == INF:
        return INF
    return max(min(a + b, INF - 1), -INF + 1)


def _blocked_query(first, head, parent, vdepth, up, dn, rs, rt, B=32):
    """The blocked ET query of the device (query_blk.hip), restated: each side
    walks its ancestor chain in blocks of B; arcs inside a block go to a B x B
    (min,+) matrix and are settled in chain order, arcs leaving it upward are
    relaxed from the settled labels; labels are indexed by ET depth.  Returns
    (dist, scanned)."""
    labs, chains, scanned = [], [], 0
    for x0, val in ((rs, up), (rt, dn)):
        chain = []
        x = x0
        while x >= 0:
            chain.append(x)
            x = parent[x]
        D = int(vdepth[x0])
        lab = [INF] * (D + 1)
        lab[D] = 0
        for b in range((D + B) // B):
            dlo = D - B * b
            dtop = max(dlo - (B - 1), 0)
            bsz = dlo - dtop + 1
            M = [[INF] * B for _ in range(B)]
            rows = chain[B * b:B * b + bsz]
            for j, v in enumerate(rows):
                f, deg = int(first[v]), int(first[v + 1] - first[v])
                for a in range(min(deg, bsz - 1 - j)):
                    ad = int(vdepth[head[f + a]])
                    if ad >= dtop:
                        M[dlo - ad][j] = int(val[f + a])
            L = [lab[dlo - t] for t in range(bsz)]
            for j in range(bsz):
                if L[j] != INF:
                    for t in range(bsz):
                        L[t] = min(L[t], _iadd(L[j], M[t][j]))
            for t in range(bsz):
                lab[dlo - t] = L[t]
            for j, v in enumerate(rows):
                if L[j] == INF:
                    continue
                f, deg = int(first[v]), int(first[v + 1] - first[v])
                scanned += deg
                for a in range(deg):
                    ad = int(vdepth[head[f + a]])
                    if ad < dtop:
                        lab[ad] = min(lab[ad], _iadd(L[j], int(val[f + a])))
        labs.append(lab)
        chains.append(chain)
    on_s = set(chains[0])
    lca = next((int(vdepth[x]) for x in chains[1] if x in on_s), -1)
    best = INF
    for d in range(lca + 1):
        best = min(best, _iadd(labs[0][d], labs[1][d]))
    return best, scanned


def test_blocked_query_equals_chain_sweep():
    """The blocked query restated above equals the oracle's vertex-by-vertex
    sweep (python/cch.py:675-732) in distance and `scanned`, with blocks of
    32 and, to cut many chains across several blocks, of 4; on the reference's
    test family and on a metric with a negative cycle."""
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    for gi, g in enumerate(small_graphs()[::3]):
        rank = order_of(g)
        st = O.Structure(g.n, g.off, g.tgt, rank)
        w = g.w.copy()
        if gi == 3:                       # a negative cycle: labels run far negative
            w[rng.integers(0, g.m, 12)] = -5000
        up, dn, _ = st.customize(w)
        qs = rng.integers(0, g.n, 60)
        qt = rng.integers(0, g.n, 60)
        d0, s0, _ = st.query_ranks(up, dn, rank[qs], rank[qt])
        for B in (64, 32, 4):
            for i in range(len(qs)):
                d, s = _blocked_query(st.first, st.head, st.parent, st.vdepth, up, dn,
                                      int(rank[qs[i]]), int(rank[qt[i]]), B)
                assert (d, s) == (int(d0[i]), int(s0[i])), (gi, B, i)


def _row_merge_triangles(first, head):
    """The triangles the row-merge sweep (rows.hip) forms: for every arc
    a = (x, y), the arcs b after a in row x, in pieces of at most 64, each
    improving the arc c = (y, head b) of row y.  Returns (c, a, b) sorted."""
    first = np.asarray(first, np.int64)
    head = np.asarray(head, np.int64)
    n = len(first) - 1
    tail = np.repeat(np.arange(n), np.diff(first))
    key = {(int(tail[c]), int(head[c])): c for c in range(len(head))}
    out = []
    for a in range(len(head)):
        x, y = int(tail[a]), int(head[a])
        length = int(first[x + 1]) - a - 1
        for k in range(0, length, 64):
            for lane in range(min(64, length - k)):
                b = a + 1 + k + lane
                out.append((key[(y, int(head[b]))], a, b))
    return sorted(out)


def test_row_merge_enumeration_is_the_lower_triangle_set():
    """Every lower triangle (x; y, z) is formed exactly once by the suffix
    pieces of the row-merge sweep, with the same operand arcs as the stored
    triangle list (python/cch_nb.py:543-562 enumerates x's pairs the same way)."""
    for g in small_graphs()[::4]:
        rank = order_of(g)
        topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
        ex = topo.export()
        sched, toff, tab = topo.export_triangles()
        c = np.repeat(sched.astype(np.int64), np.diff(toff))
        want = sorted(zip(c.tolist(), tab[:, 0].tolist(), tab[:, 1].tolist()))
        assert _row_merge_triangles(ex["first"], ex["head"]) == want


def _full_update_outputs(first, head, e_arc, w_new_edges, up0, dn0, up1, dn1):
    """The outputs the full-recustomization update driver (update.hip) derives
    from the values before (up0, dn0) and after (up1, dn1) a full
    customization of the new weights: the changed arcs, and the number of
    arcs the heap-ordered update would evaluate -- the touched arcs plus the
    top arc of every lower triangle one of whose sides changed (the side
    lists of python/cch.py:309-324)."""
    first = np.asarray(first, np.int64)
    head = np.asarray(head, np.int64)
    n = len(first) - 1
    tail = np.repeat(np.arange(n), np.diff(first))
    changed = np.nonzero((up0 != up1) | (dn0 != dn1))[0]
    queued = {int(e_arc[k]) for k in w_new_edges if e_arc[k] >= 0}
    for c in changed:
        x, y = int(tail[c]), int(head[c])
        for k in range(first[x], first[x + 1]):
            if k == c:
                continue
            lo, hi = min(y, int(head[k])), max(y, int(head[k]))
            row = head[first[lo]:first[lo + 1]]
            j = int(np.searchsorted(row, hi))
            assert j < len(row) and row[j] == hi, "chordal completion lacks a top arc"
            queued.add(int(first[lo]) + j)
    return changed, len(queued)


def test_full_update_derivation_equals_heap_update():
    """The full-recustomization update driver: a full customization of the
    new weights, diffed against the old values, gives the heap-ordered
    partial update's values (Theorem 4 of docs/technical_note.md), its
    changed set and its evaluated count -- over rounds of 1 to 40 changes,
    decreases included, on the reference's test family, against the oracle's
    restatement of python/cch.py:349-402."""
    from oracle import oracle as O
    rng = np.random.default_rng(9)
    for g in small_graphs()[::2]:
        rank = order_of(g)
        st = O.Structure(g.n, g.off, g.tgt, rank)
        w = g.w.copy()
        up, dn, _ = st.customize(w)
        upd = O.Updater(st, w, up, dn)
        for k in (1, 3, 10, 40):
            idx = rng.choice(g.m, size=min(k, g.m), replace=False)
            new = (upd.w[idx] + rng.integers(-300, 301, len(idx))).astype(np.int32)
            up0, dn0 = upd.up.copy(), upd.dn.copy()
            ev, ch, _ = upd.update(idx, new)
            up1, dn1, _ = st.customize(upd.w)
            assert np.array_equal(up1, upd.up) and np.array_equal(dn1, upd.dn), (g.name, k)
            changed, evaluated = _full_update_outputs(st.first, st.head, st.e_arc, idx, up0, dn0, up1, dn1)
            assert sorted(changed.tolist()) == sorted(ch.tolist()), (g.name, k)
            assert evaluated == ev, (g.name, k, evaluated, ev)
