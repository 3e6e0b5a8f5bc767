"""Seeded synthetic road graphs with int32 energy weights.

Every generator returns a :class:`RoadGraph`: a directed CSR (``off`` int64[n+1],
``tgt`` int32[m], arcs of one tail sorted by head), int32 arc weights ``w[m]`` and
vertex coordinates in metres.  Weights follow one energy formula::

    w(u, v) = round(alpha * euclid_len_m(u, v)) + slope * (h(v) - h(u))

with integer heights ``h`` in metres.  The height term is a potential difference,
so every directed cycle has length ``sum(round(alpha * len)) > 0``: the metric is
conservative while a large share of the arcs is negative.  This is the same role
as the EV-energy generators of the reference (python/graphs.py:63-132), with
integer arithmetic instead of float64.

The arc order is the reference's ``g.adj`` order restated for a CSR: arc ``k`` of
``off[u]..off[u+1]`` is the k-th out-arc of ``u`` (python/cch.py:257 enumerates
edge weights in exactly this order), so an "edge index" of
``CCH.update_edges`` (python/cch.py:349) is an index into ``tgt``/``w``.
"""
from __future__ import annotations

import dataclasses

import numpy as np

ALPHA = 40      # energy per metre of road
SLOPE = 25      # energy per metre of climb (regenerated on descent)


@dataclasses.dataclass
class RoadGraph:
    n: int
    off: np.ndarray          # int64[n+1]
    tgt: np.ndarray          # int32[m]
    w: np.ndarray            # int32[m]
    x: np.ndarray            # float64[n], metres
    y: np.ndarray            # float64[n], metres
    h: np.ndarray | None = None   # int32[n] heights, metres
    name: str = ""
    extra: dict = dataclasses.field(default_factory=dict)

    @property
    def m(self) -> int:
        return int(self.tgt.shape[0])

    def tails(self) -> np.ndarray:
        return np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.off))

    def negative_fraction(self) -> float:
        return float((self.w < 0).mean()) if self.m else 0.0

    def adj(self):
        """Adjacency lists ``[[(v, w), ...], ...]`` (the reference's ``Graph.adj``)."""
        out = []
        for u in range(self.n):
            s, e = int(self.off[u]), int(self.off[u + 1])
            out.append(list(zip(self.tgt[s:e].tolist(), self.w[s:e].tolist())))
        return out

    def with_weights(self, w: np.ndarray) -> "RoadGraph":
        w = np.asarray(w, dtype=np.int32)
        assert w.shape == self.w.shape
        return dataclasses.replace(self, w=w)


def from_arcs(n: int, arcs, xy=None, name="") -> RoadGraph:
    """Build a RoadGraph from ``[(u, v, w), ...]`` keeping the given per-tail order
    (the reference's ``make_graph`` appends arcs in list order, test_techniques.py:37-41)."""
    arcs = list(arcs)
    tails = np.array([a[0] for a in arcs], dtype=np.int64)
    order = np.argsort(tails, kind="stable")
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, tails + 1, 1)
    off = np.cumsum(off)
    tgt = np.array([arcs[i][1] for i in order], dtype=np.int32).reshape(-1)
    w = np.array([arcs[i][2] for i in order], dtype=np.int32).reshape(-1)
    if xy is None:
        x = np.arange(n, dtype=np.float64)
        y = np.zeros(n, np.float64)
    else:
        x = np.array([p[0] for p in xy], np.float64)
        y = np.array([p[1] for p in xy], np.float64)
    return RoadGraph(n, off, tgt, w, x, y, None, name)


_GEN = None


def _gen():
    """The generator's native helpers (libnwgen.so, built with the package):
    threaded sorts, Kruskal's spanning forest, Tarjan's SCC and the CSR build."""
    global _GEN
    if _GEN is None:
        import ctypes
        from pathlib import Path
        path = Path(__file__).resolve().parent / "libnwgen.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build the package first (__graft_entry__.build())")
        lib = ctypes.CDLL(str(path))
        P = ctypes.c_void_p
        I = ctypes.c_int64
        for name, args, res in (("nwgen_unique_first", [P, I, P], I), ("nwgen_unique_sorted", [P, I], I),
                                ("nwgen_mst", [I, P, P, I, P, P, P], ctypes.c_int),
                                ("nwgen_csr", [I, P, P, I, P, P], ctypes.c_int),
                                ("nwgen_largest_scc", [I, P, P, I, P], I)):
            f = getattr(lib, name)
            f.argtypes, f.restype = args, res
        _GEN = lib
    return _GEN


def _ptr(a):
    return a.ctypes.data


def _csr_from_pairs(n, u, v):
    """Sort distinct directed pairs by (u, v): off int64[n+1], tgt int32[m]."""
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    off = np.empty(n + 1, np.int64)
    tgt = np.empty(len(u), np.int32)
    _gen().nwgen_csr(n, _ptr(u), _ptr(v), len(u), _ptr(off), _ptr(tgt))
    return off, tgt


def _largest_scc(n, u, v):
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    keep8 = np.empty(n, np.uint8)
    _gen().nwgen_largest_scc(n, _ptr(u), _ptr(v), len(u), _ptr(keep8))
    keep = keep8.view(bool)
    newid = -np.ones(n, np.int64)
    newid[keep] = np.arange(int(keep.sum()))
    sel = keep[u] & keep[v]
    return keep, newid, sel


def _energy(x, y, h, u, v, alpha=ALPHA, slope=SLOPE, ln=None):
    if ln is None:
        ln = np.hypot(x[v] - x[u], y[v] - y[u])
    w = np.rint(alpha * ln).astype(np.int64) + slope * (h[v].astype(np.int64) - h[u].astype(np.int64))
    assert np.abs(w).max(initial=0) < 2 ** 30, "weights must stay well inside int32"
    return w.astype(np.int32)


# ------------------------------------------------------------------ config 0

def grid_graph(side=64, seed=0, drop=0.10, bumps=8, bump_sigma=5.0, hmax=400,
               spacing=1.0, jitter=0.3, alpha=ALPHA, slope=SLOPE) -> RoadGraph:
    """side x side 4-neighbour grid, unit spacing in metres, coordinates jittered by
    U(-jitter, jitter), 10% of the undirected edges dropped, largest SCC kept.
    Heights: sum of ``bumps`` seeded Gaussian bumps (sigma ``bump_sigma`` grid units)
    rounded to integer metres and clipped to [0, hmax]."""
    rng = np.random.default_rng(seed)
    n0 = side * side
    ii, jj = np.meshgrid(np.arange(side), np.arange(side), indexing="xy")
    x = (ii.reshape(-1) + rng.uniform(-jitter, jitter, n0)) * spacing
    y = (jj.reshape(-1) + rng.uniform(-jitter, jitter, n0)) * spacing
    ids = np.arange(n0).reshape(side, side)
    eu = np.concatenate([ids[:, :-1].reshape(-1), ids[:-1, :].reshape(-1)])
    ev = np.concatenate([ids[:, 1:].reshape(-1), ids[1:, :].reshape(-1)])
    keep = rng.random(eu.shape[0]) >= drop
    eu, ev = eu[keep], ev[keep]
    cx = rng.uniform(0, side, bumps) * spacing
    cy = rng.uniform(0, side, bumps) * spacing
    amp = rng.uniform(0.4, 1.0, bumps) * hmax
    hf = np.zeros(n0)
    for k in range(bumps):
        hf += amp[k] * np.exp(-((x - cx[k]) ** 2 + (y - cy[k]) ** 2) / (2 * (bump_sigma * spacing) ** 2))
    h = np.clip(np.rint(hf), 0, hmax).astype(np.int32)
    u = np.concatenate([eu, ev])
    v = np.concatenate([ev, eu])
    keepv, newid, sel = _largest_scc(n0, u, v)
    u, v = newid[u[sel]], newid[v[sel]]
    x, y, h = x[keepv], y[keepv], h[keepv]
    n = int(keepv.sum())
    off, tgt = _csr_from_pairs(n, u, v)
    tails = np.repeat(np.arange(n), np.diff(off))
    w = _energy(x, y, h, tails, tgt.astype(np.int64), alpha, slope)
    g = RoadGraph(n, off, tgt, w, x, y, h, f"grid{side}_s{seed}")
    g.extra.update(dict(bump_sigma=bump_sigma, spacing=spacing, drop=drop))
    return g


# ------------------------------------------------------------- configs 1-4

def fractal_height(rng, box, hmax, res=512, beta=3.6):
    """Seeded smooth fractal field on a res x res lattice over [0, box]^2 (spectral
    synthesis with a k^-beta power spectrum), scaled to [0, hmax]."""
    kx = np.fft.fftfreq(res)[:, None]
    ky = np.fft.rfftfreq(res)[None, :]
    k = np.hypot(kx, ky)
    k[0, 0] = 1.0
    amp = k ** (-beta / 2.0)
    amp[0, 0] = 0.0
    spec = amp * (rng.standard_normal(amp.shape) + 1j * rng.standard_normal(amp.shape))
    f = np.fft.irfft2(spec, s=(res, res))
    f = (f - f.min()) / (f.max() - f.min()) * hmax
    return f


def sample_field(f, box, x, y):
    res = f.shape[0]
    gx = np.clip(x / box * (res - 1), 0, res - 1 - 1e-9)
    gy = np.clip(y / box * (res - 1), 0, res - 1 - 1e-9)
    i0 = gx.astype(np.int64)
    j0 = gy.astype(np.int64)
    fx, fy = gx - i0, gy - j0
    v = (f[i0, j0] * (1 - fx) * (1 - fy) + f[i0 + 1, j0] * fx * (1 - fy)
         + f[i0, j0 + 1] * (1 - fx) * fy + f[i0 + 1, j0 + 1] * fx * fy)
    return v


_TILES = {}   # the point set and its tile index, shared with forked workers


def _tile_edges(job):
    """Delaunay edges owned by one tile, triangulated with `margin` metres of
    neighbouring points.  An owned edge (its midpoint in the tile, half-open,
    so every edge has one owner) is kept only if one of its triangles has its
    circumdisk inside the region whose points were all included (every point
    lies in the box, so only the part of the disk inside the box counts): that
    disk is then empty of every point, so the edge is an edge of the global
    Delaunay triangulation.  Every global edge with its midpoint in the tile is present
    in the local triangulation (an empty circle stays empty when points are
    removed).  Returns the sorted distinct global (i, j) pairs, i < j, or None
    when some owned edge is not certified (the caller retries with a wider
    margin).  With certify False the owned edges are returned unchecked."""
    from scipy.spatial import Delaunay
    i, j, margin, certify = job
    T = _TILES
    pts, order, bounds, tiles, w, box = T["pts"], T["order"], T["bounds"], T["tiles"], T["w"], T["box"]
    x0 = i * w if i else -np.inf
    x1 = (i + 1) * w if i < tiles - 1 else np.inf
    y0 = j * w if j else -np.inf
    y1 = (j + 1) * w if j < tiles - 1 else np.inf
    k = int(np.ceil(margin / w))
    parts = [order[bounds[a * tiles + b]:bounds[a * tiles + b + 1]]
             for a in range(max(0, i - k), min(tiles, i + k + 1)) for b in range(max(0, j - k), min(tiles, j + k + 1))]
    cand = np.sort(np.concatenate(parts))
    P = pts[cand]
    X0, X1, Y0, Y1 = x0 - margin, x1 + margin, y0 - margin, y1 + margin
    keep = (P[:, 0] >= X0) & (P[:, 0] < X1) & (P[:, 1] >= Y0) & (P[:, 1] < Y1)
    idx, P = cand[keep], P[keep]
    if len(idx) < 3:
        return np.empty((0, 2), np.int64)
    tri = Delaunay(P).simplices
    # circumcircles, in coordinates relative to each triangle's first corner
    A = P[tri[:, 0]]
    bx, by = (P[tri[:, 1]] - A).T
    cx, cy = (P[tri[:, 2]] - A).T
    d = 2.0 * (bx * cy - by * cx)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ux = (cy * (bx * bx + by * by) - by * (cx * cx + cy * cy)) / d
        uy = (bx * (cx * cx + cy * cy) - cx * (bx * bx + by * by)) / d
        r = np.hypot(ux, uy) * (1 + 1e-9) + 1e-6
        ux += A[:, 0]
        uy += A[:, 1]
        # every point lies in [0, box]^2: the disk must miss the parts of the
        # box outside the included region (up to four strips), each tested
        # exactly as disk against rectangle
        inside = np.isfinite(ux) & np.isfinite(uy) & np.isfinite(r)
        r2 = r * r
        strips = [(rx0, rx1, ry0, ry1) for rx0, rx1, ry0, ry1 in
                  ((0.0, X0, 0.0, box), (X1, box, 0.0, box), (0.0, box, 0.0, Y0), (0.0, box, Y1, box))
                  if rx1 >= rx0 and ry1 >= ry0]   # a strip is empty where the region reaches past the box
        for rx0, rx1, ry0, ry1 in strips:
            dx = np.maximum(np.maximum(rx0 - ux, ux - rx1), 0.0)
            dy = np.maximum(np.maximum(ry0 - uy, uy - ry1), 0.0)
            inside &= dx * dx + dy * dy > r2
    nt = len(tri)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    et = np.concatenate([np.arange(nt)] * 3)
    mx = 0.5 * (P[e[:, 0], 0] + P[e[:, 1], 0])
    my = 0.5 * (P[e[:, 0], 1] + P[e[:, 1], 1])
    own = (mx >= x0) & (mx < x1) & (my >= y0) & (my < y1)
    e, et = e[own], et[own]
    g = idx[e].astype(np.int64)
    g.sort(axis=1)
    key = g[:, 0] * len(pts) + g[:, 1]
    o = np.argsort(key, kind="stable")
    key = key[o]
    start = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    if not len(key):
        return np.empty((0, 2), np.int64)
    if not certify:
        return g[o[start]]
    ok = np.maximum.reduceat(inside[et[o]], start)
    # an edge of one local triangle (on the local hull): every circle through
    # its ends whose centre lies beyond the triangle's circumcentre, away from
    # the triangle, is empty of included points too; try such circles with
    # centres ever farther out (their part inside the box shrinks toward the
    # edge before it stretches along the box's border)
    one = np.diff(np.r_[start, len(key)]) == 1
    todo = np.flatnonzero(one & ~ok)
    if len(todo):
        k = o[start[todo]]
        pa, pb = P[e[k, 0]], P[e[k, 1]]
        pr = P[tri[et[k]].sum(1) - e[k, 0] - e[k, 1]]
        mid = 0.5 * (pa + pb)
        hl = 0.5 * np.hypot(*(pb - pa).T)
        nrm = np.stack([-(pb - pa)[:, 1], (pb - pa)[:, 0]], 1) / (2.0 * hl[:, None])
        nrm *= -np.sign(((pr - mid) * nrm).sum(1))[:, None]      # pointing away from the triangle
        tt = et[k]
        s1 = np.nan_to_num((np.stack([ux[tt], uy[tt]], 1) - mid) * nrm, nan=np.inf).sum(1)
        base = np.maximum(s1, 0.0)
        good = np.zeros(len(todo), bool)
        for step in range(41):
            sv = base + 2.0 * hl * (2.0 ** step - 1.0)
            live = ~good & np.isfinite(sv) & (sv < 1e12)
            if not live.any():
                break
            cx_ = mid[:, 0] + sv * nrm[:, 0]
            cy_ = mid[:, 1] + sv * nrm[:, 1]
            rr = np.hypot(hl, sv) * (1 + 1e-9) + 1e-6
            fits = live.copy()
            for rx0, rx1, ry0, ry1 in strips:
                dx = np.maximum(np.maximum(rx0 - cx_, cx_ - rx1), 0.0)
                dy = np.maximum(np.maximum(ry0 - cy_, cy_ - ry1), 0.0)
                fits &= dx * dx + dy * dy > rr * rr
            good |= fits
        ok[todo] = good
    if not ok.all():
        return None
    return g[o[start]]


def tiled_delaunay_edges(pts, box, tiles, margin=500.0, wide_margin=4000.0, progress=None, workers=16):
    """Delaunay edges of a large point set, triangulated tile by tile (tiles x
    tiles squares of the box) in a process pool.  An inner tile takes `margin`
    metres of neighbouring points and keeps its edges if all are certified
    (see _tile_edges).  Tiles on the box's border (whose hull edges along a
    clipped, nearly straight border cannot be certified from a neighbourhood)
    and inner tiles that fail take `wide_margin` metres, and their owned edges
    are kept as they come."""
    import concurrent.futures as cf
    import multiprocessing as mp
    import os
    w = box / tiles
    tx = np.minimum((pts[:, 0] / w).astype(np.int64), tiles - 1)
    ty = np.minimum((pts[:, 1] / w).astype(np.int64), tiles - 1)
    tid = tx * tiles + ty
    order = np.argsort(tid, kind="stable")
    bounds = np.searchsorted(tid[order], np.arange(tiles * tiles + 1))
    _TILES.update(pts=pts, order=order, bounds=bounds, tiles=tiles, w=w, box=box)
    workers = max(1, min(workers, os.cpu_count() or 1))
    out = {}
    inner = [(i, j) for i in range(1, tiles - 1) for j in range(1, tiles - 1)]
    wide = [(i, j) for i in range(tiles) for j in range(tiles) if (i, j) not in set(inner)]
    try:
        with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork")) as ex:
            jobs = [(i, j, margin, True) for i, j in inner] + [(i, j, wide_margin, False) for i, j in wide]
            # the wide border tiles first: they take longest
            jobs = jobs[len(inner):] + jobs[:len(inner)]
            for (i, j, _, _), e in zip(jobs, ex.map(_tile_edges, jobs)):
                if e is None:
                    wide.append((i, j))
                else:
                    out[(i, j)] = e
            failed = [t for t in inner if t not in out]
            if progress:
                progress(f"delaunay tiles: {len(inner) - len(failed)} of {len(inner)} inner tiles certified with a "
                         f"{margin:.0f} m margin; {tiles * tiles - len(inner)} border tiles with {wide_margin:.0f} m")
            for (i, j), e in zip(failed, ex.map(_tile_edges, [(i, j, wide_margin, False) for i, j in failed])):
                out[(i, j)] = e
    finally:
        _TILES.clear()
    return np.concatenate([out[(i, j)] for i in range(tiles) for j in range(tiles)])


def road_graph(n_points, seed, box=60000.0, clusters=200, sigma=2000.0, out_degree=2.8,
               oneway=0.05, hmax=800, alpha=ALPHA, slope=SLOPE, neg_share=0.10,
               name=None, tiles=0, progress=None) -> RoadGraph:
    """Road-like graph: a seeded clustered Poisson (Thomas) point process, its
    Delaunay triangulation thinned to a mean out-degree of ``out_degree`` (Euclidean
    MST plus random Delaunay edges, so the result stays connected), ``oneway`` of
    the kept streets one-way, largest SCC kept, integer fractal heights.  With
    ``tiles`` > 0 the triangulation is built tile by tile in parallel (for tens
    of millions of points); ``progress`` receives a line after each stage."""
    say = progress or (lambda msg: None)
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    centers = rng.uniform(0, box, (clusters, 2))
    pts = np.empty((0, 2))
    need = n_points
    while need > 0:
        c = centers[rng.integers(0, clusters, need)]
        p = c + rng.normal(0.0, sigma, (need, 2))
        p = p[(p >= 0).all(1) & (p <= box).all(1)]
        pts = np.concatenate([pts, p])
        need = n_points - pts.shape[0]
    pts = np.round(pts[:n_points], 3)
    # distinct points in lexicographic (x, y) order, through integer keys
    # (millimetre coordinates, at most 2^26 in the box) rather than row sorting
    kxy = (np.rint(pts[:, 0] * 1000).astype(np.int64) << 32) | np.rint(pts[:, 1] * 1000).astype(np.int64)
    first = np.empty(len(kxy), np.int64)
    first = first[:_gen().nwgen_unique_first(_ptr(kxy), len(kxy), _ptr(first))]
    pts = pts[first]
    pts = pts[rng.permutation(pts.shape[0])]
    n0 = pts.shape[0]
    say(f"points: {n0}")
    if tiles:
        e = tiled_delaunay_edges(pts, box, tiles, progress=say)
    else:
        tri = Delaunay(pts).simplices
        e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
        e.sort(axis=1)
    ke = np.ascontiguousarray(e[:, 0].astype(np.int64) * n0 + e[:, 1])
    ke = ke[:_gen().nwgen_unique_sorted(_ptr(ke), len(ke))]    # distinct edges, lexicographic
    e = np.stack([ke // n0, ke % n0], 1).astype(np.int32)
    say(f"delaunay edges: {len(e)}")
    ea, eb = np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1])
    px, py = np.ascontiguousarray(pts[:, 0]), np.ascontiguousarray(pts[:, 1])
    in_tree = np.empty(len(e), np.uint8)
    _gen().nwgen_mst(n0, _ptr(ea), _ptr(eb), len(e), _ptr(px), _ptr(py), _ptr(in_tree))
    in_tree = in_tree.view(bool)
    say("spanning tree done")
    mst_e = e[in_tree]
    rest = e[~in_tree]
    # target undirected edge count: arcs = E * (2 - oneway) = out_degree * n
    target = int(round(out_degree * n0 / (2.0 - oneway)))
    extra = max(0, min(rest.shape[0], target - mst_e.shape[0]))
    pick = rest[rng.choice(rest.shape[0], extra, replace=False)]
    und = np.concatenate([mst_e, pick])
    one = rng.random(und.shape[0]) < oneway
    flip = rng.random(und.shape[0]) < 0.5
    a, b = und[:, 0], und[:, 1]
    a1 = np.where(flip, b, a)
    b1 = np.where(flip, a, b)
    u = np.concatenate([a1, b1[~one]])
    v = np.concatenate([b1, a1[~one]])
    keepv, newid, sel = _largest_scc(n0, u, v)
    u, v = newid[u[sel]], newid[v[sel]]
    pts = pts[keepv]
    n = int(keepv.sum())
    off, tgt = _csr_from_pairs(n, u, v)
    say("largest strongly connected component and CSR done")
    tails = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    x, y = pts[:, 0].copy(), pts[:, 1].copy()
    ln = np.hypot(x[tgt] - x[tails], y[tgt] - y[tails])
    hf = sample_field(fractal_height(rng, box, hmax), box, x, y)
    dz = vertical_scale(x, y, hf, tails, tgt, alpha, slope, neg_share, ln=ln)
    h = np.rint(hf * dz).astype(np.int32)
    w = _energy(x, y, h, tails, tgt, alpha, slope, ln=ln)
    g = RoadGraph(n, off, tgt, w, x, y, h, name or f"road{n_points}_s{seed}")
    g.extra.update(dict(box=box, clusters=clusters, sigma=sigma, hmax=hmax, dz_scale=dz))
    return g


def vertical_scale(x, y, hf, tails, heads, alpha, slope, neg_share=0.10, ln=None):
    """Vertical exaggeration of a height field so that about ``neg_share`` of the arcs
    come out negative.  This is the reference's rule for its ``ev`` metric on road
    graphs (python/dimacs.py:155-160, python/bigcsr.py:155-161: the 80th percentile
    of downhill energy slopes is scaled to 1).  A literal 0-800 m field over a
    60 km box gives grades far below alpha/slope = 1.6 and no negative arc at all."""
    if ln is None:
        ln = np.hypot(x[heads] - x[tails], y[heads] - y[tails])
    drop = hf[tails] - hf[heads]
    sel = (drop > 0) & (ln > 0)
    if not sel.any():
        return 1.0
    s = slope * drop[sel] / (alpha * ln[sel])
    q = np.quantile(s, 1.0 - 2.0 * neg_share)
    return float(1.0 / q) if q > 0 else 1.0


def metric_family(g: RoadGraph, K: int, seed: int, hmax=1200, slope_range=(5, 40),
                  box=None, alpha=ALPHA, neg_share=0.10):
    """K metrics on one topology: metric k has its own slope coefficient drawn from
    ``slope_range`` and its own seeded fractal height field in [0, hmax].
    Returns int32 [K, m] weights and the slopes (vertical scale as in
    :func:`vertical_scale`)."""
    rng = np.random.default_rng(seed)
    box = box or g.extra.get("box", float(max(g.x.max(), g.y.max())))
    tails = g.tails()
    heads = g.tgt
    ln = np.hypot(g.x[heads] - g.x[tails], g.y[heads] - g.y[tails])
    W = np.empty((K, g.m), np.int32)
    slopes = rng.integers(slope_range[0], slope_range[1] + 1, K)
    fields = [fractal_height(rng, box, hmax) for _ in range(K)]   # the draws in metric order

    def one(k):
        hf = sample_field(fields[k], box, g.x, g.y)
        dz = vertical_scale(g.x, g.y, hf, tails, heads, alpha, int(slopes[k]), neg_share, ln=ln)
        h = np.rint(hf * dz).astype(np.int32)
        W[k] = _energy(g.x, g.y, h, tails, heads, alpha, int(slopes[k]), ln=ln)

    # metrics on threads (numpy's array loops release the GIL); at most 8 at
    # once bounds the temporaries
    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(max(1, min(8, K))) as ex:
        list(ex.map(one, range(K)))
    return W, slopes


def find_directed_triangle(g: RoadGraph, rng=None):
    """Arc indices (k1, k2, k3) of a directed 3-cycle u->v->w->u, or None."""
    rng = rng or np.random.default_rng(0)
    off, tgt = g.off, g.tgt
    start = int(rng.integers(0, g.n))
    for du in range(g.n):
        uu = (start + du) % g.n
        for k1 in range(off[uu], off[uu + 1]):
            v = int(tgt[k1])
            for k2 in range(off[v], off[v + 1]):
                ww = int(tgt[k2])
                if ww == uu:
                    continue
                for k3 in range(off[ww], off[ww + 1]):
                    if int(tgt[k3]) == uu:
                        return int(k1), int(k2), int(k3)
    return None


def inject_negative_cycle(w: np.ndarray, arcs, total=-1):
    """Set the weights of the given cycle arcs so that they sum to ``total``."""
    w = np.array(w, dtype=np.int32, copy=True)
    k1, k2, k3 = arcs
    w[k1], w[k2] = 0, 0
    w[k3] = total
    return w


CONFIGS = {
    0: dict(kind="grid", side=64, seed=0),
    1: dict(kind="road", n_points=264_000, seed=1, hmax=800),
    2: dict(kind="road", n_points=321_000, seed=2, hmax=800, out_degree=2.5),
    3: dict(kind="road", n_points=1_100_000, seed=3, hmax=800, out_degree=2.45),
    4: dict(kind="road", n_points=23_900_000, seed=4, hmax=3000, out_degree=2.44, tiles=8),
}


def make_config(idx: int, progress=None) -> RoadGraph:
    c = dict(CONFIGS[idx])
    kind = c.pop("kind")
    if kind == "grid":
        return grid_graph(**c)
    n_points = c.pop("n_points")
    return road_graph(n_points, name=f"c{idx}", progress=progress, **c)


def random_pairs(n, count, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, count).astype(np.int32), rng.integers(0, n, count).astype(np.int32)


# ------------------------------------------------------- small test families

def ev_grid(side=15, seed=0, alpha=10, up=10, down=6):
    """side x side 8-connected grid over random bump terrain, integer EV energy
    ``alpha*len + (up if climbing else down)*dz`` (the shape of the reference's
    python/graphs.py:63-79 generator; regenerative braking recovers less than
    climbing costs, so every cycle is positive while steep descents are negative)."""
    rng = np.random.default_rng(seed)
    n = side * side
    x = (np.arange(n) % side).astype(np.float64)
    y = (np.arange(n) // side).astype(np.float64)
    z = np.zeros(n)
    for _ in range(12):
        cx, cy = rng.uniform(0, side, 2)
        amp = rng.uniform(-0.6, 0.6) * side
        rad = rng.uniform(0.05, 0.2) * side
        z += amp * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * rad * rad))
    z = np.rint(z * 4).astype(np.int64)
    arcs = []
    for v in range(n):
        i, j = v % side, v // side
        for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (1, -1), (-1, 1), (-1, -1)):
            a, b = i + dx, j + dy
            if 0 <= a < side and 0 <= b < side:
                u = b * side + a
                dz = int(z[u] - z[v])
                ln = int(round(alpha * 4 * np.hypot(dx, dy)))
                arcs.append((v, u, ln + (up if dz > 0 else down) * dz))
    g = from_arcs(n, arcs, xy=list(zip(x, y)), name=f"evgrid{side}_s{seed}")
    g.h = z.astype(np.int32)
    return g


def ev_geometric(n=300, seed=0, k=3, alpha=100, up=10, down=6):
    """Random points in the unit square joined to their k nearest neighbours in
    both directions, integer EV energy (cf. python/graphs.py:82-132)."""
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    z = np.zeros(n)
    for _ in range(8):
        c = rng.random(2)
        z += rng.uniform(-1, 1) * np.exp(-((pts - c) ** 2).sum(1) / (2 * rng.uniform(0.05, 0.2) ** 2))
    z = np.rint(z * 60).astype(np.int64)
    d = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d, np.inf)
    nn = np.argsort(d, axis=1)[:, :k]
    pairs = set()
    for v in range(n):
        for u in nn[v]:
            pairs.add((v, int(u)))
            pairs.add((int(u), v))
    arcs = []
    for v, u in sorted(pairs):
        dz = int(z[u] - z[v])
        ln = int(round(alpha * np.sqrt(d[v, u])))
        arcs.append((v, u, ln + (up if dz > 0 else down) * dz))
    g = from_arcs(n, arcs, xy=[tuple(p) for p in pts], name=f"geo{n}_s{seed}")
    g.h = z.astype(np.int32)
    return g


def shifted_random(n=200, seed=0, deg=3, cmax=100, spread=300):
    """Random sparse digraph with weights c + phi(u) - phi(v), c > 0: every cycle is
    positive, many arcs are negative (python/graphs.py:135-151).  No geometry:
    coordinates are random, as the reference's test does (test_cch.py:35)."""
    rng = np.random.default_rng(seed)
    phi = rng.integers(0, spread, n)
    arcs = []
    for v in range(n):
        for _ in range(deg):
            u = int(rng.integers(0, n))
            if u != v:
                arcs.append((v, u, int(rng.integers(1, cmax)) + int(phi[v] - phi[u])))
    pts = rng.random((n, 2))
    return from_arcs(n, arcs, xy=[tuple(p) for p in pts], name=f"shift{n}_s{seed}")
