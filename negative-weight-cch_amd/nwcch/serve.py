"""Continuous-batching server of ET queries and road-weight updates over
customized metrics kept resident in HBM.

The reference answers one query at a time on the host and applies an update
with ``CCH.update_edges`` followed by ``update_created_negative_cycle``
(python/cch.py:349-402, 666-671; python/cch_benchmark.py:100-156), rejecting a
cycle-creating update by applying the inverse update (docs/technical_note.md:147).
Here every admitted request becomes GPU work that overlaps with the rest:

* metrics are admitted while their device footprint fits a memory budget;
* the queries of all metrics that have arrived are packed into one launch
  (each query names its metric through a device table of metric addresses),
  with at most `max_inflight` launches on the device and launch buffers sized
  by a memory budget -- requests arriving while others run join the next
  launch (continuous batching);
* an update of metric k runs on the update stream after the query launches
  already issued.  In snapshot mode (the default) it works on a second copy of
  the metric while queries keep reading the committed version; an accepted
  update is committed by swapping the copies, a cycle-creating one is simply
  dropped.  In strict mode it works in place, metric k takes no new queries
  until the changed-arc negative-cycle verdict is known, and a cycle-creating
  update is undone by the inverse update;
* completion times come from HIP events (device clock), so latency is
  arrival -> result on the device.

Every answer carries the metric version it was computed on, and
``verify_alone`` recomputes each version from scratch and re-answers the same
queries alone, for the "same alone and in the mix" check.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np
import torch

from ._lib import INF
from .device import DeviceTopology


@dataclasses.dataclass
class Request:
    kind: str                      # "query" or "update"
    metric: int
    arrival: float                 # seconds after the start of run()
    s: int = 0                     # query: original vertex ids
    t: int = 0
    edges: np.ndarray | None = None   # update: original-arc indices
    weights: np.ndarray | None = None  # update: new weights
    # filled in by the server
    done: float | None = None
    dist: int | None = None
    version: int | None = None
    rejected: bool = False
    evaluated: int = 0
    changed: int = 0


@dataclasses.dataclass
class _Metric:
    mid: int
    slot: int                      # index in the device table of metric addresses
    w: torch.Tensor                # device int32 [m]
    ud: torch.Tensor               # device int32 [mp, 2]
    w_host: np.ndarray             # host mirror of w
    w_spare: torch.Tensor | None = None    # snapshot mode: the copy an update works on
    ud_spare: torch.Tensor | None = None
    version: int = 0
    history: list = dataclasses.field(default_factory=list)  # accepted updates, in order
    busy_update: object = None     # in-flight update state
    res_host: torch.Tensor | None = None   # pinned int64[4]: the verdict of the running update


class _Launch:
    """Buffers of one in-flight query launch (queries of any metrics) and the
    metric address table it was issued with."""

    def __init__(self, cap: int, dev, n_table: int):
        self.cap = cap
        self.h_q = torch.empty((3, cap), dtype=torch.int32, pin_memory=True)   # s, t, metric slot
        self.d_q = torch.empty((3, cap), dtype=torch.int32, device=dev)
        self.h_tab = torch.empty(n_table, dtype=torch.int64, pin_memory=True)  # metric addresses
        self.d_tab = torch.empty(n_table, dtype=torch.int64, device=dev)
        self.d_dist = torch.empty(cap, dtype=torch.int32, device=dev)
        self.h_dist = torch.empty(cap, dtype=torch.int32, pin_memory=True)
        self.event = torch.cuda.Event(enable_timing=True)
        self.ids = np.zeros(0, np.int64)   # request indices of the launch

    @staticmethod
    def nbytes(cap: int, n_table: int = 64) -> int:
        return cap * (12 + 4) * 2 + 16 * n_table


class Server:
    """metric_budget: device bytes for resident metrics (weights + customized
    pairs, twice in snapshot mode); slot_budget: bytes for query launch buffers
    (device + pinned host), which bounds the launches and queries in flight.
    Query launches alternate over `n_streams` streams, updates run on their own
    stream (all streams at the default priority)."""

    def __init__(self, dt: DeviceTopology, rank: np.ndarray, metric_budget: int, slot_budget: int,
                 max_batch: int = 65536, n_streams: int = 2, max_metrics: int = 64, snapshot: bool = True):
        self.dt = dt
        self.rank = np.asarray(rank, np.int32)
        self.metric_budget = int(metric_budget)
        self.snapshot = bool(snapshot)
        self.metrics: dict[int, _Metric] = {}
        self.used = 0
        dt.prepare_update()
        n_launch = max(1, int(slot_budget) // _Launch.nbytes(max_batch, max_metrics))
        self.launch_bufs = [_Launch(max_batch, dt.dev, max_metrics) for _ in range(n_launch)]
        self.free_bufs = list(range(n_launch))
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self.upd_stream = torch.cuda.Stream()
        # committed address of every metric's customization, copied into each
        # launch with its queries (a commit never races with a running launch)
        self.h_table = np.zeros(max_metrics, np.int64)
        self.free_table = list(range(max_metrics))
        self._rr = 0
        self.last_query_events: list = []
        self._updates_warm = False

    # ----------------------------------------------------------- metrics
    def metric_bytes(self) -> int:
        return (2 if self.snapshot else 1) * (4 * self.dt.m + 8 * self.dt.mp)

    def add_metric(self, mid: int, w) -> bool:
        """Customize and keep metric `mid` resident; False if it does not fit the
        budget or the table, or has a negative cycle (it is then not admitted)."""
        need = self.metric_bytes()
        if mid in self.metrics or self.used + need > self.metric_budget or not self.free_table:
            return False
        wd = self.dt.weights(w)
        ud = self.dt.customize(wd)
        neg = int(self.dt.negcycle_scan(ud)[0].item())
        st = self.dt.status(clear=True)
        if neg or st:
            return False
        slot = self.free_table.pop(0)
        self.h_table[slot] = ud.data_ptr()
        m = _Metric(mid, slot, wd, ud, np.asarray(w, np.int32).copy())
        # allocated once: a pinned allocation per update would stall the serving loop
        m.res_host = torch.empty(4, dtype=torch.int64, pin_memory=True)
        if self.snapshot:
            m.w_spare, m.ud_spare = torch.empty_like(wd), torch.empty_like(ud)
        self.metrics[mid] = m
        self.used += need
        if not self._updates_warm:
            self._warm_up_updates(m)
        return True

    def _warm_up_updates(self, m: _Metric):
        """One no-op update of three arcs on scratch copies, on the update
        stream: the library captures that stream's launch sequences on first
        use, which belongs to setup, not to the first served update."""
        k = min(3, self.dt.m)
        if k == 0:
            return
        idx = np.arange(k, dtype=np.int64)
        self.upd_stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self.upd_stream):
            w2, ud2 = m.w.clone(), m.ud.clone()
            self.dt.update(w2, ud2, idx, m.w_host[idx], stream=self.upd_stream)
        self.upd_stream.synchronize()
        self.dt.status(clear=True)
        self._updates_warm = True

    def drop_metric(self, mid: int):
        m = self.metrics.pop(mid, None)
        if m is not None:
            torch.cuda.synchronize()
            self.h_table[m.slot] = 0
            self.free_table.append(m.slot)
            self.used -= self.metric_bytes()

    # ------------------------------------------------------------ serving
    def run(self, requests: list[Request], max_seconds: float = 60.0, max_inflight: int = 3) -> dict:
        """Serve `requests` in real time (arrival times in seconds from the start
        of the call); returns statistics and writes the results into the
        Request objects.  At most `max_inflight` query launches are on the
        device; queries of all metrics that arrive meanwhile wait and go into
        the next launch together, so launches grow with the load (continuous
        batching).  A metric runs one update at a time; in strict mode it also
        takes no new queries until that update's verdict is known, in snapshot
        mode its queries go on at the committed version meanwhile."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        unknown = {r.metric for r in reqs} - set(self.metrics)
        if unknown:
            raise KeyError(f"requests for metrics that are not resident: {sorted(unknown)}")
        N = len(reqs)
        arrival = np.fromiter((r.arrival for r in reqs), np.float64, N)
        metric = np.fromiter((r.metric for r in reqs), np.int64, N)
        is_q = np.fromiter((r.kind == "query" for r in reqs), bool, N)
        s_rank = self.rank[np.fromiter((r.s for r in reqs), np.int64, N)]
        t_rank = self.rank[np.fromiter((r.t for r in reqs), np.int64, N)]
        slot_of = np.zeros(N, np.int32)
        for k, m in self.metrics.items():
            slot_of[metric == k] = m.slot
        dist = np.full(N, INF, np.int64)
        done = np.full(N, np.nan)
        version = np.zeros(N, np.int64)
        # per metric: its requests in arrival order, their arrivals, the positions
        # of its updates (a metric's queries never cross one of its updates), head
        queues = {}
        for k in self.metrics:
            ids = np.nonzero(metric == k)[0]
            queues[k] = [ids, arrival[ids], np.nonzero(~is_q[ids])[0], 0]
        torch.cuda.synchronize()
        # device-clock origin of the run: completion times are measured from it
        start = torch.cuda.Event(enable_timing=True)
        start.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        inflight: list = []
        launches = 0
        launched = 0
        while True:
            now = time.perf_counter() - t0
            if now > max_seconds:
                raise RuntimeError("serving run exceeded its time limit")
            # retire finished query launches
            still = []
            for b in inflight:
                L = self.launch_bufs[b]
                if L.event.query():
                    n = len(L.ids)
                    dist[L.ids] = L.h_dist.numpy()[:n]
                    done[L.ids] = start.elapsed_time(L.event) / 1000.0
                    self.free_bufs.append(b)
                else:
                    still.append(b)
            inflight = still
            busy = False
            parts = []
            room = self.launch_bufs[self.free_bufs[-1]].cap if self.free_bufs else 0
            can_launch = len(inflight) < max_inflight and room > 0
            for k, m in self.metrics.items():
                if m.busy_update is not None:
                    self._poll_update(m, start, done)
                    if m.busy_update is not None:
                        busy = True
                        if not self.snapshot:
                            continue
                ids, arr, upd_pos, head = queues[k]
                if head >= len(ids):
                    continue
                busy = True
                adm = int(np.searchsorted(arr, now, side="right"))
                if head >= adm:
                    continue
                if not is_q[ids[head]]:
                    if m.busy_update is not None:
                        continue
                    r = reqs[ids[head]]
                    r._idx = int(ids[head])
                    queues[k][3] = head + 1
                    self._launch_update(m, r)
                    continue
                if not can_launch or room == 0:
                    continue
                nu = int(np.searchsorted(upd_pos, head, side="left"))
                end = min(adm, head + room, int(upd_pos[nu]) if nu < len(upd_pos) else len(ids))
                parts.append(ids[head:end])
                version[ids[head:end]] = m.version
                room -= end - head
                queues[k][3] = end
            if parts:
                b = self.free_bufs.pop()
                batch = np.concatenate(parts)
                self._launch_queries(self.launch_bufs[b], batch, s_rank, t_rank, slot_of)
                inflight.append(b)
                launches += 1
                launched += len(batch)
            if not busy and not inflight:
                break
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        for i, r in enumerate(reqs):
            if is_q[i]:
                r.dist, r.done, r.version = int(dist[i]), float(done[i]), int(version[i])
        u = [r for r in reqs if r.kind == "update"]
        # overflow is expected only while a rejected update held a negative cycle
        self.dt.check_status(negative_cycle=any(r.rejected for r in u))
        lat_q = (done[is_q] - arrival[is_q]) * 1000.0 if is_q.any() else np.zeros(1)
        lat_u = np.array([r.done - r.arrival for r in u]) * 1000.0 if u else np.zeros(1)
        span = float(np.nanmax(done)) if N else 0.0
        nq = int(is_q.sum())
        return {
            "queries": nq, "updates": len(u), "rejected_updates": sum(r.rejected for r in u),
            "wall_s": wall, "queries_per_s": nq / span if span > 0 else 0.0,
            "query_p50_ms": float(np.percentile(lat_q, 50)), "query_p99_ms": float(np.percentile(lat_q, 99)),
            "update_p50_ms": float(np.percentile(lat_u, 50)), "update_p99_ms": float(np.percentile(lat_u, 99)),
            "query_launches": launches, "mean_batch": launched / max(launches, 1),
            "launch_buffers": len(self.launch_bufs), "max_inflight": max_inflight,
        }

    def _launch_queries(self, L: _Launch, batch: np.ndarray, s_rank, t_rank, slot_of):
        n = len(batch)
        hq = L.h_q.numpy()
        hq[0, :n] = s_rank[batch]
        hq[1, :n] = t_rank[batch]
        hq[2, :n] = slot_of[batch]
        L.h_tab.numpy()[:] = self.h_table
        stream = self.streams[self._rr % len(self.streams)]
        self._rr += 1
        with torch.cuda.stream(stream):
            L.d_q[:, :n].copy_(L.h_q[:, :n], non_blocking=True)
            L.d_tab.copy_(L.h_tab, non_blocking=True)
            self.dt.query_multi(L.d_tab, L.d_q[2], L.d_q[0], L.d_q[1], n, L.d_dist, stream)
            L.h_dist[:n].copy_(L.d_dist[:n], non_blocking=True)
            L.event.record(stream)
        L.ids = batch
        self.last_query_events = [e for e in self.last_query_events[-(len(self.streams) - 1):]] + [L.event] \
            if len(self.streams) > 1 else [L.event]

    def _launch_update(self, m: _Metric, r: Request, restore_of: Request | None = None):
        s = self.upd_stream
        # after every query launch already issued (the last one of each stream)
        for ev in self.last_query_events:
            s.wait_event(ev)
        old = m.w_host[r.edges].copy()
        with torch.cuda.stream(s):
            if self.snapshot:
                # the spare copy is free: every launch that read it as the
                # committed version was issued before the waits above
                m.w_spare.copy_(m.w)
                m.ud_spare.copy_(m.ud)
                res = self.dt.update(m.w_spare, m.ud_spare, r.edges, r.weights, stream=s)
            else:
                res = self.dt.update(m.w, m.ud, r.edges, r.weights, stream=s)
            host = m.res_host      # one update of a metric at a time
            host.copy_(res, non_blocking=True)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(s)
        m.busy_update = (r, ev, host, old, restore_of)

    def _poll_update(self, m: _Metric, start, done=None):
        r, ev, host, old, restore_of = m.busy_update
        if not ev.query():
            return
        evaluated, changed, neg, _ = (int(v) for v in host.tolist())
        m.busy_update = None
        if restore_of is not None:
            # the inverse update finished: the metric is back to its version
            restore_of.done = start.elapsed_time(ev) / 1000.0
            if done is not None and hasattr(restore_of, "_idx"):
                done[restore_of._idx] = restore_of.done
            return
        r.evaluated, r.changed = evaluated, changed
        if neg > 0:
            r.rejected = True
            if not self.snapshot:
                # apply the inverse update in place (Theorem 4 restores bitwise)
                inv = Request("update", m.mid, r.arrival, edges=r.edges, weights=old)
                self._launch_update(m, inv, restore_of=r)
                return
            # the committed copy was never touched: drop the spare
            r.done = start.elapsed_time(ev) / 1000.0
            if done is not None and hasattr(r, "_idx"):
                done[r._idx] = r.done
            return
        m.w_host[r.edges] = r.weights
        if self.snapshot:
            m.w, m.w_spare = m.w_spare, m.w
            m.ud, m.ud_spare = m.ud_spare, m.ud
            self.h_table[m.slot] = m.ud.data_ptr()
        m.version += 1
        m.history.append((r.edges.copy(), np.asarray(r.weights, np.int32).copy()))
        r.version = m.version
        r.done = start.elapsed_time(ev) / 1000.0
        if done is not None and hasattr(r, "_idx"):
            done[r._idx] = r.done

    # ------------------------------------------------------- verification
    def weights_at(self, mid: int, version: int, w0: np.ndarray) -> np.ndarray:
        w = np.asarray(w0, np.int32).copy()
        for e, v in self.metrics[mid].history[:version]:
            w[e] = v
        return w

    def verify_alone(self, requests: list[Request], w0: dict) -> dict:
        """Recompute every (metric, version) seen by a query from scratch (full
        customization of its weights) and answer its queries alone, one launch
        per (metric, version); count the answers that differ."""
        groups = collections.defaultdict(list)
        for r in requests:
            if r.kind == "query":
                groups[(r.metric, r.version)].append(r)
        mismatches = 0
        for (mid, ver), rs in groups.items():
            w = self.weights_at(mid, ver, w0[mid])
            ud = self.dt.customize(w)
            s = torch.as_tensor(self.rank[np.array([r.s for r in rs])])
            t = torch.as_tensor(self.rank[np.array([r.t for r in rs])])
            d, _ = self.dt.query(ud, s, t, want_scanned=False)
            got = np.array([r.dist for r in rs])
            mismatches += int((d.cpu().numpy() != got).sum())
        self.dt.check_status()
        return {"groups": len(groups), "mismatches": mismatches}


# ------------------------------------------------------------- workloads

def poisson_workload(g, n_metrics: int, rate: float, duration: float, seed: int, local_frac=0.2,
                     local_radius=5000.0, update_rate=50.0, update_sizes=(1, 10, 1000), delta=500):
    """Query requests as a Poisson stream at `rate`/s: uniform random s-t pairs,
    `local_frac` of them with t within `local_radius` metres of s; interleaved
    update events at `update_rate`/s changing 1, 10 or 1000 random arcs by
    +-U(0, delta) (the config-3 traffic).  Returns the request list."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    nq = rng.poisson(rate * duration)
    arr = np.sort(rng.uniform(0.0, duration, nq))
    s = rng.integers(0, g.n, nq)
    t = rng.integers(0, g.n, nq)
    local = rng.random(nq) < local_frac
    if local.any():
        tree = cKDTree(np.stack([g.x, g.y], 1))
        idx = np.nonzero(local)[0]
        for i in idx:
            nb = tree.query_ball_point([g.x[s[i]], g.y[s[i]]], local_radius)
            t[i] = nb[rng.integers(0, len(nb))]
    mids = rng.integers(0, n_metrics, nq)
    reqs = [Request("query", int(mids[i]), float(arr[i]), s=int(s[i]), t=int(t[i])) for i in range(nq)]
    nu = rng.poisson(update_rate * duration)
    for ua in np.sort(rng.uniform(0.0, duration, nu)):
        k = int(rng.choice(update_sizes))
        edges = rng.choice(g.m, k, replace=False).astype(np.int64)
        reqs.append(Request("update", int(rng.integers(0, n_metrics)), float(ua), edges=edges,
                            weights=None))
        reqs[-1].weights = rng.integers(-delta, delta + 1, k).astype(np.int64)  # a delta, made absolute later
    return reqs


def absolute_updates(reqs: list[Request], w0: dict):
    """Turn the update deltas of a workload into absolute weights, applied in
    arrival order per metric (as the server will apply them), keeping the
    weights inside int32."""
    cur = {k: np.asarray(v, np.int64).copy() for k, v in w0.items()}
    for r in sorted(reqs, key=lambda q: q.arrival):
        if r.kind != "update":
            continue
        w = cur[r.metric]
        new = np.clip(w[r.edges] + r.weights, -(2 ** 30), 2 ** 30)
        w[r.edges] = new
        r.weights = new.astype(np.int32)
    return reqs


def inf_to_float(v: int) -> float:
    return float("inf") if v == INF else float(v)
