"""Drop-in for the object API of the reference's ``python/cch.py``: ``CCH(g, ...)``
with ``customize``, ``update_edges``, ``has_negative_cycle``,
``update_created_negative_cycle`` and ``query``, on the GPU.

Differences a caller sees (everything else keeps the reference's meaning):

* weights are integers (int32, ``2**31 - 1`` = no arc); ``query`` returns an
  ``int`` distance, or ``float('inf')`` when t is unreachable, and the number
  of arcs scanned;
* ``up`` / ``down`` are the two columns of one [m+, 2] int32 device tensor
  (``ud``), ``edge_w`` is an int32 device tensor;
* ``customize(g, stop_on_negative_cycle=True)`` returns the reference's witness
  but always leaves the complete customization in ``up``/``down`` (the
  reference leaves the partly customized arrays, python/cch.py:270-278);
* ``order`` is an input like the reference's; without one the library's
  inertial-flow nested dissection computes it on the host.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from ._lib import INF
from .device import DeviceTopology


def _as_csr(g):
    """(n, off, tgt, w, xy) of a RoadGraph or of a reference-style graph whose
    ``adj[u]`` lists ``(v, w)`` (python/graphs.py:21-39); arc k is the k-th
    entry in adjacency order (python/cch.py:257)."""
    if hasattr(g, "off") and hasattr(g, "tgt"):
        return g.n, np.asarray(g.off, np.int64), np.asarray(g.tgt, np.int32), np.asarray(g.w), (g.x, g.y)
    n = g.n
    deg = np.array([len(a) for a in g.adj], np.int64)
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    tgt = np.array([v for a in g.adj for v, _ in a], np.int32).reshape(-1)
    w = np.array([wt for a in g.adj for _, wt in a], np.float64).reshape(-1)
    xy = None
    if getattr(g, "coords", None):
        xy = (np.array([c[0] for c in g.coords], np.float64), np.array([c[1] for c in g.coords], np.float64))
    return n, off, tgt, w, xy


def _int_weights(w) -> np.ndarray:
    w = np.asarray(w)
    if w.dtype.kind == "f":
        fin = np.isfinite(w)
        if not np.array_equal(w[fin], np.round(w[fin])):
            raise ValueError("weights must be integers (the path computes in int32)")
        out = np.where(fin, w, INF).astype(np.int64)
    else:
        out = w.astype(np.int64)
    if out.size and (out.min() <= -INF or out.max() > INF):
        raise ValueError("weights must lie in (-2^31+1, 2^31-1]")
    return out.astype(np.int32)


class CCH:
    def __init__(self, g, xy=None, leaf=64, method="flow", log=print, order=None, device=None,
                 gpu_index=False):
        """The reference's constructor (python/cch.py:112); device = the GPU
        ordinal (default: the current one); gpu_index=True builds everything
        after the order -- the chordal completion, the arc map, the triangle
        index and the update side lists -- on that GPU
        (NWCCH_TOPO_DEVICE_SYMBOLIC | NWCCH_TOPO_DEVICE_INDEX) instead of the
        host: the same arrays, faster on big graphs."""
        t0 = time.perf_counter()
        n, off, tgt, _, gxy = _as_csr(g)
        self.n = n
        self._off, self._tgt = off, tgt
        if order is not None:
            rank = np.asarray(order, np.int64)
            if not np.array_equal(np.sort(rank), np.arange(n)):
                raise AssertionError("order must be a permutation")
            rank = rank.astype(np.int32)
        else:
            if xy is not None:
                x = np.array([p[0] for p in xy], np.float64)
                y = np.array([p[1] for p in xy], np.float64)
            elif gxy is not None:
                x, y = np.asarray(gxy[0], np.float64), np.asarray(gxy[1], np.float64)
            else:
                raise ValueError("an order or coordinates are needed")
            rank = _lib.order(n, off, tgt, x, y, leaf)
        self.rank = rank
        self.t_order = time.perf_counter() - t0
        t0 = time.perf_counter()
        if gpu_index:
            import torch
            dev_i = torch.cuda.current_device() if device is None else int(device)
            self.topo = _lib.Topology.from_graph(n, off, tgt, rank, device_index=dev_i, device_symbolic=dev_i)
        else:
            self.topo = _lib.Topology.from_graph(n, off, tgt, rank)
        ex = self.topo.export()
        self.t_contract = time.perf_counter() - t0
        self.first, self.head, self.parent = ex["first"], ex["head"], ex["parent"]
        self.edge_arc, self.edge_dir = ex["e_arc"], ex["e_up"]
        self.m = self.topo.mp
        self.depth = self.topo.depth
        self.loop_edges = np.nonzero(self.edge_arc < 0)[0].tolist()
        self.dev = DeviceTopology(self.topo, device)
        self._loop_w = None
        self.edge_w = None
        self.ud = None
        self._last = None
        log(f"  CCH: n={n:,} arcs={self.m:,} triangles={self.topo.triangles:,} "
            f"etree depth={self.depth} | order {self.t_order:.1f}s, contract {self.t_contract:.1f}s")

    # ----------------------------------------------------------- customization
    @property
    def up(self):
        return None if self.ud is None else self.ud[:, 0]

    @property
    def down(self):
        return None if self.ud is None else self.ud[:, 1]

    def base_weights(self):
        """(base_up, base_down): arc values from the original edges only (min over
        parallel edges), python/cch.py:234-247.  A host-side inspection helper;
        the device path folds the base into every arc evaluation."""
        w = self.edge_w.cpu().numpy()
        bu = np.full(self.m, INF, np.int64)
        bd = np.full(self.m, INF, np.int64)
        ea, ed = self.edge_arc, self.edge_dir
        sel = ea >= 0
        np.minimum.at(bu, ea[sel & (ed == 1)], w[sel & (ed == 1)])
        np.minimum.at(bd, ea[sel & (ed == 0)], w[sel & (ed == 0)])
        return bu.astype(np.int32), bd.astype(np.int32)

    @property
    def base_up(self):
        return self.base_weights()[0]

    @property
    def base_down(self):
        return self.base_weights()[1]

    def customize(self, g, stop_on_negative_cycle=False):
        """Full customization from the weights in g; the arcs witnessing a negative
        cycle (ascending), [] if none, [-1] for a negative self-loop
        (python/cch.py:249-287)."""
        _, _, _, w, _ = _as_csr(g)
        w = _int_weights(w)
        self.edge_w = self.dev.weights(w)
        self._loop_w = w[self.loop_edges].copy()
        self.ud = self.dev.customize(self.edge_w, out=self.ud)
        self._last = None
        if (self._loop_w < 0).any():
            # the reference returns before any triangle, leaving the base values
            self.ud = self.dev.stop_state(self.edge_w, self.ud, -1)
            self.dev.check_status(negative_cycle=True)
            self.cycle_arc = -1
            return [-1]
        lst = self.dev.negcycle_list(self.ud)
        self.dev.check_status(negative_cycle=bool(lst))
        if stop_on_negative_cycle and lst:
            # the reference checks the arcs of vertex x once a triangle with a
            # later "a" arc is reached (python/cch.py:270-278): vertices whose
            # row ends at or before the last such arc
            d = np.diff(self.first)
            rows = np.nonzero(d >= 2)[0]
            if len(rows):
                a_max = int((self.first[rows + 1] - 2).max())
                x = int(np.searchsorted(self.first, lst[0], side="right") - 1)
                if self.first[x + 1] <= a_max:
                    # the values the reference leaves when it stops at row x
                    self.ud = self.dev.stop_state(self.edge_w, self.ud, x)
                    self.dev.check_status(negative_cycle=True)
                    self.cycle_arc = lst[0]
                    return [lst[0]]
        return lst

    def has_negative_cycle(self):
        """python/cch.py:662-664."""
        if self._loop_w is not None and (self._loop_w < 0).any():
            return True
        return int(self.dev.negcycle_scan(self.ud)[0].item()) > 0

    # ----------------------------------------------------------- partial update
    def update_edges(self, changes):
        """Partial re-customization; changes = [(edge_index, new_weight)] in
        adjacency order.  Returns the number of arcs re-evaluated
        (python/cch.py:349-402)."""
        changes = list(changes)
        idx = np.array([k for k, _ in changes], np.int64)
        neww = _int_weights(np.array([v for _, v in changes], np.float64) if changes else np.zeros(0))
        res = self.dev.update(self.edge_w, self.ud, idx, neww)
        r = [int(v) for v in res.cpu()]
        if self.loop_edges:
            pos = {e: i for i, e in enumerate(self.loop_edges)}
            for k, v in zip(idx.tolist(), neww.tolist()):
                if k in pos:
                    self._loop_w[pos[k]] = v
        self._last = r
        self._last_changed = None
        self.dev.check_status(negative_cycle=self.update_created_negative_cycle())
        return r[0]

    @property
    def last_changed(self):
        """Arcs whose value changed in the last update, ascending."""
        if self._last is None:
            return []
        if self._last_changed is None:
            self._last_changed = self.dev.last_changed(self._last[1]).cpu().tolist()
        return self._last_changed

    def update_created_negative_cycle(self):
        """Negative-cycle test on the changed arcs only (python/cch.py:666-671)."""
        if self._loop_w is not None and (self._loop_w < 0).any():
            return True
        return self._last is not None and self._last[2] > 0

    def query_path(self, s, t, max_len: int = 1 << 16):
        """(distance, [original vertex ids]) of the shortest s-t path, shortcuts
        unpacked as python/cch.py:406-507 does; (inf, []) if unreachable."""
        d, plen, path = self.dev.query_path(self.ud, [int(self.rank[s])], [int(self.rank[t])], max_len=max_len)
        self.dev.check_status()
        L = int(plen[0].item())
        if L < 0:
            return float("inf"), []
        if not hasattr(self, "_inv"):
            self._inv = np.argsort(np.asarray(self.rank)).astype(np.int64)
        return int(d[0].item()), self._inv[path[0, :L].cpu().numpy()].tolist()

    def potential(self):
        """A feasible potential of the customized metric, by original vertex id:
        p(v) = min(0, min_u dist(u, v)) (python/cch.py:734-755); ints."""
        pr = self.dev.potential(self.ud).cpu().numpy()
        self.dev.check_status()
        return pr[np.asarray(self.rank)].tolist()

    # ---------------------------------------------------------------- queries
    def query(self, s, t):
        """(distance, arcs scanned) between original vertices s and t
        (python/cch.py:675-732)."""
        d, sc = self.dev.query(self.ud, [int(self.rank[s])], [int(self.rank[t])])
        v = int(d[0].item())
        return (float("inf") if v == INF else v), int(sc[0].item())

    def query_many(self, s, t, want_scanned=False):
        """Batch of queries between original vertices -> int32 device tensors."""
        rs = torch.as_tensor(self.rank[np.asarray(s)])
        rt = torch.as_tensor(self.rank[np.asarray(t)])
        return self.dev.query(self.ud, rs, rt, want_scanned=want_scanned)
