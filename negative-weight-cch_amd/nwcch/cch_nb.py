"""Drop-in for the flat-array API of the reference's ``python/cch_nb.py``.

Same function names, argument order and meaning; the metric-dependent functions
run on the GPU through libnwcch.  Differences a caller sees:

* weights are integers: ``w`` may be int32 or integral float64 (``inf`` = no
  arc); results are int32 device tensors with ``INF = 2**31 - 1`` for
  unreachable (``to_float64`` gives the reference's float64 with ``inf``);
* ``customize`` returns ``(upw, dnw)`` as the two columns of one [m+, 2] device
  tensor, which ``negative_cycle_arcs`` and ``query`` accept without a copy;
* ``query`` keeps the reference signature; ``df``, ``dr``, ``mark`` and ``q``
  are accepted and ignored (labels live in GPU LDS).  ``query_many`` runs a
  whole batch in one launch.

Host prerequisites (undirected CSR, ordering, symbolic contraction, arc map)
are the library's native host code.
"""
from __future__ import annotations

import weakref

import numpy as np
import torch
import xxhash

from . import _lib
from ._lib import INF
from .device import DeviceTopology, as_pairs

_cache: dict = {}
_CACHE_MAX = 4


# ------------------------------------------------------------ host prerequisites

def undirected_csr(n, off, tgt):
    """Symmetric, de-duplicated neighbour lists without loops, and reverse slots
    (python/cch_nb.py:22-71)."""
    off = np.asarray(off, np.int64)
    tgt = np.asarray(tgt, np.int64)
    u = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    keep = u != tgt
    a = np.concatenate([u[keep], tgt[keep]])
    b = np.concatenate([tgt[keep], u[keep]])
    key = np.unique(a * n + b)
    a, b = key // n, key % n
    noff = np.zeros(n + 1, np.int64)
    np.add.at(noff, a + 1, 1)
    noff = np.cumsum(noff)
    nb = b.astype(np.int32)
    rev = np.searchsorted(key, b * n + a).astype(np.int64)
    return noff, nb, rev


def nested_dissection(n, noff, nb, rev, x, y, leaf):
    """rank[v] by nested dissection with inertial-flow bisection
    (python/cch_nb.py:336-426)."""
    return _lib.order(n, noff, nb, x, y, leaf).astype(np.int64)


def symbolic(n, noff, nb, rank):
    """(first, head, parent) of the chordal completion in rank space
    (python/cch_nb.py:431-481)."""
    t = _lib.Topology.from_graph(n, noff, nb, np.asarray(rank, np.int32))
    ex = t.export()
    return ex["first"], ex["head"], ex["parent"].astype(np.int64)


def edge_arcs(n, off, tgt, rank, first, head):
    """(e_arc int64, e_up int8) for every original arc (python/cch_nb.py:496-515)."""
    t = _lib.Topology.from_graph(n, off, tgt, np.asarray(rank, np.int32))
    ex = t.export()
    if not (np.array_equal(ex["first"], np.asarray(first)) and np.array_equal(ex["head"], np.asarray(head))):
        raise ValueError("first/head do not belong to this graph and order")
    return ex["e_arc"], ex["e_up"]


def count_triangles(n, first, head):
    d = np.diff(np.asarray(first, np.int64))
    return int((d * (d - 1) // 2).sum())


# ------------------------------------------------------------ device topology

def _digest(*arrays) -> int:
    h = xxhash.xxh64()
    for a in arrays:
        if a is None:
            h.update(b"-")
            continue
        a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
        h.update(str(a.dtype).encode())
        h.update(np.ascontiguousarray(a).tobytes())
    return h.intdigest()


def _sample(a):
    """Shape, dtype and 16 evenly spaced entries of an array: an O(1) check
    that a cached topology's array was not rewritten in place."""
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        k = a.numel()
        idx = torch.linspace(0, max(k - 1, 0), min(k, 16)).long().to(a.device)
        return (tuple(a.shape), str(a.dtype), tuple(a.reshape(-1)[idx].tolist()) if k else ())
    a = np.asarray(a)
    k = a.size
    idx = np.linspace(0, max(k - 1, 0), min(k, 16)).astype(np.int64)
    return (a.shape, str(a.dtype), tuple(a.reshape(-1)[idx].tolist()) if k else ())


class _Seen:
    """A topology reached through these very array objects (held weakly)."""

    def __init__(self, dt, arrays):
        self.dt = dt
        self.refs = tuple(None if a is None else weakref.ref(a) for a in arrays)
        self.samples = tuple(_sample(a) for a in arrays)

    def matches(self, arrays) -> bool:
        for r, a in zip(self.refs, arrays):
            if (r is None) != (a is None) or (r is not None and r() is not a):
                return False
        return self.samples == tuple(_sample(a) for a in arrays)


_seen: dict = {}
_SEEN_MAX = 16


def device_topology(n, first, head, e_arc=None, e_up=None) -> DeviceTopology:
    """The cached device copy of a topology.  Calls that pass the same array
    objects again find it by identity (plus a 16-entry spot check) in O(1);
    other calls hash the arrays' content once to find or build it."""
    arrays = (first, head, e_arc, e_up)
    ikey = (int(n),) + tuple(id(a) for a in arrays)
    hit = _seen.get(ikey)
    if hit is not None and hit.matches(arrays):
        return hit.dt
    dt = _by_content(n, first, head, e_arc, e_up)
    try:
        seen = _Seen(dt, arrays)
    except TypeError:      # objects without weak references (lists): hashed on every call
        return dt
    if len(_seen) >= _SEEN_MAX:
        _seen.pop(next(iter(_seen)))
    _seen[ikey] = seen
    return dt


def _by_content(n, first, head, e_arc, e_up) -> DeviceTopology:
    skey = (int(n), int(len(head)), _digest(first, head))
    full = (skey, None if e_arc is None else _digest(e_arc, e_up))
    if full in _cache:
        return _cache[full]
    if e_arc is None:
        for k, v in _cache.items():
            if k[0] == skey:
                return v
    # the triangle index is built on the GPU that will hold the topology (the
    # same arrays as a host build, without the host's O(T) passes)
    topo = _lib.Topology.from_csr(int(n), np.asarray(first), np.asarray(head),
                                  None if e_arc is None else np.asarray(e_arc),
                                  None if e_up is None else np.asarray(e_up),
                                  device_index=torch.cuda.current_device())
    dt = DeviceTopology(topo)
    if len(_cache) >= _CACHE_MAX:
        _cache.pop(next(iter(_cache)))
    _cache[full] = dt
    return dt


# ------------------------------------------------------------ metric-dependent

def customize(n, first, head, e_arc, e_up, w):
    """(upw, dnw): customized shortcut weights (python/cch_nb.py:529-563)."""
    dt = device_topology(n, first, head, e_arc, e_up)
    ud = dt.customize(w)
    st = dt.status(clear=False)
    neg = bool(st & 1) and int(dt.negcycle_scan(ud)[0].item()) > 0
    dt.check_status(negative_cycle=neg)
    return ud[:, 0], ud[:, 1]


def negative_cycle_arcs(upw, dnw):
    """Number of shortcut arcs with upw + dnw < 0 (python/cch_nb.py:566-572)."""
    ud = as_pairs(upw, dnw)
    out = torch.empty(2, dtype=torch.int64, device=ud.device)
    from ._lib import check, lib, ptr
    check(lib().nwcch_negcycle_scan(ptr(ud), ud.shape[0], ptr(out), torch.cuda.current_stream().cuda_stream),
          "nwcch_negcycle_scan")
    return int(out[0].item())


def query(s_rank, t_rank, first, head, parent, upw, dnw, df=None, dr=None, mark=None, q=None):
    """ET query between ranks (python/cch_nb.py:577-615); float, inf if unreachable."""
    dt = device_topology(len(first) - 1, first, head)
    dist, _ = dt.query(as_pairs(upw, dnw), [int(s_rank)], [int(t_rank)], want_scanned=False)
    v = int(dist[0].item())
    return float("inf") if v == INF else float(v)


def query_many(s_ranks, t_ranks, first, head, upw, dnw, want_scanned=False):
    """Batch of ET queries in one launch -> int32 device tensors (dist, scanned)."""
    dt = device_topology(len(first) - 1, first, head)
    return dt.query(as_pairs(upw, dnw), s_ranks, t_ranks, want_scanned=want_scanned)


def to_float64(v) -> np.ndarray:
    """int32 results with the INF sentinel -> the reference's float64 with inf."""
    a = (v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)).astype(np.int64)
    out = a.astype(np.float64)
    out[a == INF] = np.inf
    return out
