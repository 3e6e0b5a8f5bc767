"""Autotuner of the launch shapes of the level kernels and of the query packing.

Every candidate is timed on the caller's own topology, metric(s) and queries
with HIP events, and its output is compared with the first candidate's: a shape
changes only how work is spread over lanes and wavefronts, never a result.
The winning values are set on the handle (``DeviceTopology.set_tuning``); the
slab width of a batched layout is returned for the caller to pass on.
"""
from __future__ import annotations

import statistics

import torch

from .device import DeviceTopology


def _ms(fn, reps: int) -> float:
    fn()
    evs = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        evs.append((a, b))
    torch.cuda.synchronize()
    return statistics.median(a.elapsed_time(b) for a, b in evs)


def _graph_ms(fn, reps: int) -> float:
    """Median time of `fn` replayed as a HIP graph: the sweep as deployed, with
    no host launch cost in the measurement (the level launches of a small graph
    can be shorter than the host needs to issue them)."""
    from .device import StepGraph
    g = StepGraph(fn)
    return _ms(g.replay, reps)


FLAT_OFF = 1 << 30


def tune_customize(dt: DeviceTopology, w, candidates=(1, 2, 4, 8, 16),
                   flat=(0, 64, 128, 256, FLAT_OFF), reps: int = 5, margin: float = 0.02,
                   graph: bool = True) -> dict:
    """Single-metric level sweep: triangles per lane of the team kernel, then
    the heaviest-arc size from which a level runs triangle-parallel instead
    (FLAT_OFF = never, 0 = every level), then the store mode and XCD slicing.
    A later candidate replaces an earlier one only if it is faster by more
    than `margin` (timing noise).  graph=True times each candidate as a
    captured HIP graph, False as eager launches."""
    w = dt.weights(w) if not (isinstance(w, torch.Tensor) and w.is_cuda) else w
    ref = None
    out = torch.empty((dt.mp, 2), dtype=torch.int32, device=dt.dev)

    def run(key, value):
        nonlocal ref
        dt.set_tuning(key, value)
        out.fill_(0)
        t = (_graph_ms if graph else _ms)(lambda: dt.customize(w, out=out), reps)
        if ref is None:
            ref = out.clone()
        elif not torch.equal(ref, out):
            raise AssertionError(f"{key}={value} changed the customization")
        return t

    dt.set_tuning("flat_min_tri", FLAT_OFF)
    times = {r: run("tri_per_lane", r) for r in candidates}
    best = min(times, key=times.get)
    dt.set_tuning("tri_per_lane", best)
    ftimes = {f: run("flat_min_tri", f) for f in flat}
    fbest = flat[0]
    for f in flat[1:]:
        if ftimes[f] < ftimes[fbest] * (1.0 - margin):
            fbest = f
    dt.set_tuning("flat_min_tri", fbest)
    layouts = {}
    if fbest != FLAT_OFF:
        # store mode x XCD slicing of the triangle-parallel levels
        best_l = None
        for store in (0, 1, 2):
            for xcd in (0, 1):
                dt.set_tuning("flat_store", store)
                t = run("flat_xcd", xcd)
                layouts[f"store{store}/xcd{xcd}"] = t
                if best_l is None or t < layouts[best_l[0]] * (1.0 - margin):
                    best_l = (f"store{store}/xcd{xcd}", store, xcd)
        _, store, xcd = best_l
        dt.set_tuning("flat_store", store)
        dt.set_tuning("flat_xcd", xcd)
    dt.check_status()
    return {"tri_per_lane": best, "flat_min_tri": fbest,
            "flat_store": dt.get_tuning("flat_store"), "flat_xcd": dt.get_tuning("flat_xcd"), "ms": times,
            "flat_ms": {("off" if f == FLAT_OFF else str(f)): v for f, v in ftimes.items()},
            "layout_ms": layouts}


def tune_batch(dt: DeviceTopology, W, slabs=(64, 16, 8), tris=(4, 8, 16), reps: int = 3) -> dict:
    """Slab width and triangles per lane of the batched sweep."""
    W = dt.weights_batch(W)
    K = W.shape[0]
    times = {}
    ref = None
    for sw in slabs:
        for r in tris:
            dt.set_tuning("batch_tri_per_lane", r)
            out = torch.empty(dt.batch_shape(K, sw), dtype=torch.int32, device=dt.dev)
            times[(sw, r)] = _ms(lambda: dt.customize_batch(W, out=out, slab=sw), reps)
            check = [DeviceTopology.metric_view(out, k) for k in (0, K - 1)]
            if ref is None:
                ref = [c.clone() for c in check]
            elif not all(torch.equal(a, b) for a, b in zip(ref, check)):
                raise AssertionError(f"slab={sw} batch_tri_per_lane={r} changed the customization")
    (sw, r) = min(times, key=times.get)
    dt.set_tuning("batch_tri_per_lane", r)
    # arc order inside wave classes x XCD slices, at the chosen shape; a
    # variant must beat the current choice by 2 %
    out = torch.empty(dt.batch_shape(K, sw), dtype=torch.int32, device=dt.dev)
    layouts, best = {}, None
    for order in (0, 1):
        for xcd in (0, 1):
            dt.set_tuning("batch_order", order)
            dt.set_tuning("batch_xcd", xcd)
            t = _ms(lambda: dt.customize_batch(W, out=out, slab=sw), reps)
            check = [DeviceTopology.metric_view(out, k) for k in (0, K - 1)]
            if not all(torch.equal(a, b) for a, b in zip(ref, check)):
                raise AssertionError(f"batch_order={order} batch_xcd={xcd} changed the customization")
            layouts[f"order{order}/xcd{xcd}"] = t
            if best is None or t < layouts[best[0]] * 0.98:
                best = (f"order{order}/xcd{xcd}", order, xcd)
    _, order, xcd = best
    dt.set_tuning("batch_order", order)
    dt.set_tuning("batch_xcd", xcd)
    return {"slab": sw, "batch_tri_per_lane": r, "batch_order": order, "batch_xcd": xcd,
            "ms": {f"{a}/{b}": v for (a, b), v in times.items()}, "layout_ms": layouts}


def tune_queries(dt: DeviceTopology, ud, s_rank, t_rank, reps: int = 3) -> dict:
    """Query packing: launch order as given, or longest ancestor chains first."""
    qs = torch.as_tensor(s_rank).to(dt.dev, torch.int32)
    qt = torch.as_tensor(t_rank).to(dt.dev, torch.int32)
    dist = torch.empty(qs.shape[0], dtype=torch.int32, device=dt.dev)
    order = dt.chain_order(qs, qt)
    t_plain = _ms(lambda: dt.query(ud, qs, qt, want_scanned=False, dist=dist), reps)
    ref = dist.clone()
    t_pack = _ms(lambda: dt.query(ud, qs, qt, want_scanned=False, dist=dist, order=order), reps)
    if not torch.equal(ref, dist):
        raise AssertionError("query packing changed a distance")
    best = "longest_first" if t_pack < t_plain else "none"
    return {"packing": best, "ms": {"none": t_plain, "longest_first": t_pack}}
