"""Benchmark of the metric-dependent CCH path on one MI355X.

Headline workload (`value`): the road-like graph of 264,000 clustered points
(seed 1; Delaunay thinned to mean out-degree 2.8, 5% one-way, largest SCC),
int32 energy weights, inertial-flow order (host, untimed).  One step = full
customization of one metric + the negative-cycle scan (the plain one-metric
form), inputs resident in HBM; value = lower triangles / step time.

The same run also reports, as extra fields:
  * ET queries on that metric: 100,000 random s-t pairs in one launch (with
    their algorithmic bytes against the HBM peak), and single-query launches
    (p50/p99 latency);
  * partial updates of 1, 10 and 1,000 arcs, the CCH potential, 2,000 paths
    with their vertices, and the metric-independent index built on the host
    and on the GPU;
  * the serving loop on the 1.1M-point graph (16 resident metrics, Poisson
    queries at 0.2-2 M/s interleaved with updates), p50/p99;
  * the batched form: 64 metrics (slopes 5-40, own height fields, one with an
    injected 3-arc cycle of total -1) customized in one level sweep on the
    321,000-point graph (seed 2), against 64 plain single-metric customizations;
  * parity of every GPU result checked bit-for-bit against the CPU oracle;
  * the CPU baseline: the reference's own C++ customization and ET queries
    (cpp/cch.cpp, compiled from its sources) on the host, or the oracle port
    where the reference was not built.

Other modes: `--full-scale` runs the 23.4 M-vertex road graph (customization,
batched sweep, queries, checks against the reference's C++); `--prep` times
everything after the order for `--config` on the host and on the GPU.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "negative-weight-cch_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cust_bytes(m, mp, T, mapped, K=1):
    """Compulsory HBM bytes of one customization sweep in the pull layout: the
    (a, b) triangle pairs (8 B each) and the 16-byte schedule record of every
    arc once; the arc->edge lists (4 B per mapped arc + 4 B offsets); per metric
    the weights (4 B per arc) and every (up, down) pair written once and read
    once (16 B per arc)."""
    return 8 * T + 16 * mp + 4 * mapped + 4 * (mp + 1) + K * (4 * m + 16 * mp)


def level_bytes(mp, T):
    """Compulsory HBM bytes of all level launches of one customization: every
    lower triangle's two operand arc ids (8 B), and per arc its base (up, down)
    pair read and its final pair written (16 B).  Schedule records and per-
    triangle arc ids are layout overhead and not counted."""
    return 8 * T + 16 * mp


def read_traffic(kernel, algorithmic):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy of known size);
    None when there is no summary for it."""
    p = ROOT / "profiles" / f"{kernel}_traffic.md"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            v = float(line.split(":", 1)[1])
            return {"bytes_per_launch": v, "vs_algorithmic": v / algorithmic, "source": str(p.relative_to(ROOT))}
    return None


def timed(fn, reps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def event_ms(fn, reps):
    """Per-call device time of fn (events on the current stream, fn enqueues only)."""
    import torch
    evs = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        evs.append((a, b))
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in evs]


def prepare(cfg, progress=None, device_index=None):
    from nwcch import _lib, graphs
    t0 = time.perf_counter()
    g = graphs.make_config(cfg, progress=progress)
    t_gen = time.perf_counter() - t0
    log(f"config {cfg}: graph n={g.n} m={g.m} generated in {t_gen:.1f}s")
    t0 = time.perf_counter()
    rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    t_order = time.perf_counter() - t0
    log(f"config {cfg}: nested dissection {t_order:.1f}s")
    t0 = time.perf_counter()
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=device_index)
    t_topo = time.perf_counter() - t0
    log(f"config {cfg}: n={g.n} m={g.m} neg={g.negative_fraction():.3f} | m+={topo.mp} T={topo.triangles} "
        f"depth={topo.depth} levels={topo.levels} | gen {t_gen:.1f}s order {t_order:.1f}s topo {t_topo:.1f}s")
    return g, rank, topo, dict(gen_s=t_gen, order_s=t_order, topology_s=t_topo)


def plain(args, out):
    import numpy as np
    import torch

    from nwcch import graphs
    from nwcch.device import DeviceTopology, StepGraph

    g, rank, topo, host = prepare(args.config)
    dt = DeviceTopology(topo, 0)
    w = dt.weights(g.w)
    ud = torch.empty((topo.mp, 2), dtype=torch.int32, device=dt.dev)
    scan = torch.empty(2, dtype=torch.int64, device=dt.dev)

    def step():
        dt.customize(w, out=ud)
        dt.negcycle_scan(ud, out=scan)

    from nwcch import tune
    tuning = tune.tune_customize(dt, w, reps=3) if not args.no_tune else None
    for _ in range(args.warmup):
        step()
    dt.check_status()
    # timed region: K steps between two synchronisations
    t_step = timed(step, args.steps)
    # the same step replayed from one HIP graph
    if args.no_graph:
        t_graph = float("inf")
    else:
        graph = StepGraph(step)
        timed(graph.replay, args.warmup)
        t_graph = timed(graph.replay, args.steps)
    dt.check_status()
    # device time of the customization alone, per call: replayed from a HIP
    # graph (eager launches of a sweep this fast are bound by the host's issue
    # rate), or back-to-back eager launches without graphs
    if args.no_graph:
        cust_ms = event_ms(lambda: dt.customize(w, out=ud), args.steps)
    else:
        cgraph = StepGraph(lambda: dt.customize(w, out=ud))
        cgraph.replay()
        cust_ms = event_ms(cgraph.replay, args.steps)
        del cgraph
    scan_ms = event_ms(lambda: dt.negcycle_scan(ud, out=scan), args.steps)
    # per-launch durations of the level kernel: HIP events recorded by the
    # library around each launch, on the stream the launch goes to
    dt.profile_begin(args.steps * (topo.levels + 2))
    for _ in range(args.steps):
        dt.customize(w, out=ud)
    kid, kms = dt.profile_end()
    n_lv = int((kid == 0).sum()) // args.steps   # level launches per customization
    base_ms = float(np.median(kms[kid == 5]))
    launches = 1 + n_lv          # the base pass + one launch per level with triangles
    best = min(t_step, t_graph)
    B = cust_bytes(g.m, topo.mp, topo.triangles, topo.n_mapped)
    B_lv = level_bytes(topo.mp, topo.triangles)
    cmed = statistics.median(cust_ms)
    # average level launch: the customization minus its base pass, over the
    # level launches (events around every launch would add their own gaps)
    lv_avg_us = 1000.0 * (cmed - base_ms) / n_lv
    achieved = (B_lv / n_lv) / (lv_avg_us * 1e-6) / 1e9
    flat = dt.get_tuning("flat_min_tri")
    kernel = "k_customize_flat" if flat == 0 else ("k_customize_level" if flat >= 1 << 30 else
                                                   "k_customize_flat+k_customize_level")
    traffic = read_traffic(kernel, B_lv / n_lv)
    if args.copy_mib:
        # a plain device copy of known size, for scaling the HBM byte counters
        src = torch.ones(args.copy_mib << 18, dtype=torch.int32, device=dt.dev)
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    # the drop-in path on host buffers: weights uploaded, customized, scanned,
    # the (up, down) pairs copied back (PCIe included)
    w_host = g.w.astype(np.int32)

    def host_step():
        u = dt.customize(w_host)
        dt.negcycle_scan(u, out=scan)
        return u.cpu()
    host_step()
    t_host = timed(host_step, max(1, args.steps // 2))

    # ET queries
    qs_h, qt_h = graphs.random_pairs(g.n, args.queries, seed=7)
    qs = torch.as_tensor(rank[qs_h]).to(dt.dev)
    qt = torch.as_tensor(rank[qt_h]).to(dt.dev)
    dist = torch.empty(args.queries, dtype=torch.int32, device=dt.dev)
    qtune = tune.tune_queries(dt, ud, qs, qt, reps=2) if not args.no_tune else {"packing": "none"}
    order = dt.chain_order(qs, qt) if qtune["packing"] == "longest_first" else None
    qfn = lambda: dt.query(ud, qs, qt, want_scanned=False, dist=dist, order=order)  # noqa: E731
    timed(qfn, 2)
    qms = statistics.median(event_ms(qfn, 5))
    # algorithmic bytes of the batch: 8 per scanned arc (head + value), 12 per
    # chain vertex (parent + first), 12 per query (s, t, answer)
    _, qsc = dt.query(ud, qs, qt, order=order)
    vdep = topo.export()["depth"].astype(np.int64)
    chain = vdep[rank[qs_h]] + vdep[rank[qt_h]] + 2
    qbytes = 8 * int(qsc.to(torch.int64).sum().item()) + 12 * int(chain.sum()) + 12 * args.queries
    q_gbs = qbytes / (qms / 1000.0) / 1e9
    single = []
    for i in range(args.single):
        single.append(event_ms(lambda i=i: dt.query(ud, qs[i:i + 1], qt[i:i + 1], want_scanned=False), 1)[0])
    lat = np.array(single)
    t1 = timed(lambda: [dt.query(ud, qs[i:i + 1], qt[i:i + 1], want_scanned=False) for i in range(100)], 3) / 100
    dt.check_status()

    updates = update_timing(args, g, rank, dt, w, ud)

    # the steps after the path: the CCH potential, and queries with unpacked paths
    pot = torch.empty(g.n, dtype=torch.int32, device=dt.dev)
    dt.potential(ud, out=pot)
    pot_ms = statistics.median(event_ms(lambda: dt.potential(ud, out=pot), 5))
    npq = min(args.queries, 2000)
    pd, plen, _ = dt.query_path(ud, qs[:npq], qt[:npq], max_len=8192)
    pth_ms = statistics.median(event_ms(lambda: dt.query_path(ud, qs[:npq], qt[:npq], max_len=8192), 3))
    dt.check_status()
    plen_h = plen.cpu().numpy()
    paths = {"queries": npq, "ms": pth_ms, "queries_per_s": npq / (pth_ms / 1e3),
             "mean_vertices": float(plen_h[plen_h > 0].mean()) if (plen_h > 0).any() else 0.0}

    parity = None
    if not args.no_check:
        from oracle import oracle as O
        st = O.Structure(g.n, g.off, g.tgt, rank)
        up, dn, _ = st.customize(g.w)
        got = ud.cpu().numpy()
        ok = np.array_equal(got[:, 0], up) and np.array_equal(got[:, 1], dn)
        ok &= tuple(int(v) for v in scan.cpu()) == st.negcycle(up, dn)
        nchk = min(args.queries, 20000)
        d0, _, _ = st.query(up, dn, qs_h[:nchk], qt_h[:nchk], threads=os.cpu_count() or 1)
        ok &= np.array_equal(dist[:nchk].cpu().numpy(), d0)
        ok &= np.array_equal(pd.cpu().numpy(), d0[:npq] if nchk >= npq else pd.cpu().numpy())
        pj, _ = O.johnson(g.n, g.off, g.tgt, g.w)
        ok &= np.array_equal(pot.cpu().numpy()[rank].astype(np.int64), pj)
        parity = bool(ok)
        log(f"plain parity vs oracle: {parity}")

    cpu = None
    if not args.no_cpu:
        # a sample of about 10 s of host work: 15 customizations each way and
        # 3 x 50,000 reference queries
        ncq = min(args.queries, 50000)
        cpu = cpu_baseline(g, rank, topo, reps=15, queries=(qs_h[:ncq], qt_h[:ncq], dist[:ncq].cpu().numpy()))
    prep = preprocessing(g, rank, topo, host, dt, updates["prepare_s"])

    out.update({
        "value": topo.triangles / best,
        "ms_per_step": 1000.0 * best,
        "config": {"workload": f"road-like graph, config {args.config}: n={g.n} m={g.m} m+={topo.mp} "
                               f"T={topo.triangles} ET depth={topo.depth}; one metric: customize + "
                               f"negative-cycle scan per step",
                   "n": g.n, "m": g.m, "m_plus": topo.mp, "triangles": topo.triangles, "depth": topo.depth,
                   "levels": topo.levels, "metrics": 1, "negative_arc_fraction": round(g.negative_fraction(), 4)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_vs_algorithmic": traffic["vs_algorithmic"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None, "kernel": kernel,
                     "algorithmic_bytes_per_launch": B_lv / n_lv, "avg_launch_us": lv_avg_us,
                     "avg_launch_us_isolated": 1000.0 * float(kms[kid == 0].astype(np.float64).mean()),
                     "base_pass_us": 1000.0 * base_ms,
                     "level_launches": n_lv, "launches_per_customization": launches,
                     "customization_algorithmic_bytes": B, "customize_ms": cmed,
                     "customization_GB_s": B / (cmed / 1000.0) / 1e9},
        "cpu_baseline": cpu,
        "host_buffers": {"ms": 1000.0 * t_host, "triangles_per_s": topo.triangles / t_host,
                         "what": "customize + scan with int32 weights from host memory and the (up, down) "
                                 "pairs copied back to host memory (PCIe included)"},
        "step_ms": {"eager": 1000.0 * t_step, "hip_graph": None if args.no_graph else 1000.0 * t_graph},
        "scan": {"ms": statistics.median(scan_ms),
                 "GB_s": 8 * topo.mp / (statistics.median(scan_ms) / 1000.0) / 1e9},
        "queries": {"batch": args.queries, "batch_ms": qms, "queries_per_s": args.queries / (qms / 1000.0),
                    "roofline": {"bound": "hbm", "achieved": q_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": q_gbs / HBM_PEAK_GBS, "algorithmic_bytes": qbytes,
                                 "bytes_per_query": qbytes / args.queries},
                    "batch_latency_ms": qms,
                    "single_launch_queries_per_s": 1.0 / t1,
                    "single_p50_ms": float(np.percentile(lat, 50)),
                    "single_p99_ms": float(np.percentile(lat, 99))},
        "updates": updates,
        "potential_ms": pot_ms,
        "path_queries": paths,
        "parity_vs_oracle": parity,
        "tuning": {"customize": tuning, "queries": qtune},
        "host": dict(host, nproc=os.cpu_count()),
        "preprocessing": prep,
    })
    del dt


def update_timing(args, g, rank, dt, w, ud, reps=5):
    """Partial re-customization (python/cch.py:349-402) of k = 1, 10, 1000
    random original arcs changed by +-U(0, 500), each on a fresh copy of the
    customized metric: GPU time per update (HIP events, host validation and
    staging included), arcs evaluated, and beside it the oracle's C
    restatement of update_edges on the same changes (the reference has no
    compiled update)."""
    import numpy as np
    import torch

    from oracle import oracle as O
    rng = np.random.default_rng(17)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dt.prepare_update()          # the side lists, once
    torch.cuda.synchronize()
    prep_s = time.perf_counter() - t0
    upd = None
    if not args.no_cpu:
        st = O.Structure(g.n, g.off, g.tgt, rank)
        base_up, base_dn, _ = st.customize(g.w)
        upd = O.Updater(st, g.w, base_up, base_dn)   # restored after every change set
    res = {}
    for k in (1, 10, 1000):
        gpu_ms, cpu_ms, evals, same = [], [], [], True
        for _ in range(reps):
            idx = rng.choice(g.m, size=k, replace=False).astype(np.int64)
            neww = np.clip(g.w[idx].astype(np.int64) + rng.integers(-500, 501, size=k), -(1 << 20), 1 << 20)
            neww = neww.astype(np.int32)
            w2, ud2 = w.clone(), ud.clone()
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            r = dt.update(w2, ud2, idx, neww)
            b.record()
            torch.cuda.synchronize()
            gpu_ms.append(a.elapsed_time(b))
            ev = int(r[0].item())
            evals.append(ev)
            if upd is not None:
                t0 = time.perf_counter()
                ev_c, _, _ = upd.update(idx, neww)
                cpu_ms.append(1000 * (time.perf_counter() - t0))
                got = ud2.cpu().numpy()
                same &= ev_c == ev and np.array_equal(got[:, 0], upd.up) and np.array_equal(got[:, 1], upd.dn)
                upd.update(idx, g.w[idx])
        dt.status(clear=True)
        res[str(k)] = {"gpu_ms_median": statistics.median(gpu_ms), "gpu_ms_max": max(gpu_ms),
                       "evaluated_median": int(statistics.median(evals)),
                       "cpu_port_ms_median": statistics.median(cpu_ms) if cpu_ms else None,
                       "equal_to_oracle": same if upd is not None else None}
        log(f"update k={k}: GPU {statistics.median(gpu_ms):.3f} ms, evaluated {int(statistics.median(evals))}"
            + (f", oracle (1 core) {statistics.median(cpu_ms):.2f} ms, equal {same}" if cpu_ms else ""))
    res["prepare_s"] = prep_s
    res["what"] = ("k random original arcs changed by +-U(0, 500); GPU: nwcch_update on a fresh copy, HIP events; "
                   "CPU: the oracle's C restatement of update_edges, one thread (cores 1, kind port)")
    return res


def preprocessing(g, rank, topo, host, dt, t_upd_h):
    """The metric-independent index built on the host and on the GPU
    (NWCCH_TOPO_DEVICE_INDEX): wall-clock of each step, and whether the
    device-built triangle and side lists equal the host's."""
    import numpy as np
    import torch

    from nwcch import _lib
    from nwcch.device import DeviceTopology

    def wall(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    td, t_topo_d = wall(lambda: _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0))
    dd, t_dev_d = wall(lambda: DeviceTopology(td, 0))
    _, t_upd_d = wall(dd.prepare_update)
    same = bool(np.array_equal(dd.export_triangles(), topo.export_triangles()[2]))
    so_h, st_h = dt.export_sides()
    so_d, st_d = dd.export_sides()
    same &= bool(np.array_equal(so_h, so_d) and np.array_equal(st_h, st_d))
    del dd, td
    # everything after the order on the GPU: chordal completion and arc map too
    ts, t_topo_s = wall(lambda: _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0,
                                                         device_symbolic=0))
    same_s = same_structure(topo, ts)
    del ts
    res = {"host_index": {"topology_s": host["topology_s"], "update_index_s": t_upd_h},
           "device_index": {"topology_s": t_topo_d, "dev_create_s": t_dev_d, "update_index_s": t_upd_d},
           "device_symbolic_and_index": {"topology_s": t_topo_s, "identical": same_s},
           "identical": same,
           "what": "topology = symbolic contraction + arc map + triangle index + schedule (device index: "
                   "triangle counts on the GPU, lists filled by dev_create; device symbolic: chordal "
                   "completion and arc map on the GPU); update_index = side lists"}
    log(f"preprocessing: host topology {host['topology_s']:.3f} s + side lists {t_upd_h:.3f} s; device-index "
        f"topology {t_topo_d:.3f} s + dev_create {t_dev_d:.3f} s + side lists {t_upd_d:.3f} s; identical {same}; "
        f"device symbolic + index topology {t_topo_s:.3f} s, identical {same_s}")
    return res


def same_structure(a, b):
    """Two topologies hold the same contraction, arc map, levels and schedule."""
    import numpy as np
    ea, eb = a.export(), b.export()
    if (a.n, a.m, a.mp, a.triangles, a.levels) != (b.n, b.m, b.mp, b.triangles, b.levels):
        return False
    if any(not np.array_equal(ea[k], eb[k]) for k in ea):
        return False
    sa, oa, _ = a.export_triangles(with_list=False)
    sb, ob, _ = b.export_triangles(with_list=False)
    return bool(np.array_equal(sa, sb) and np.array_equal(oa, ob))


def prep(args, out):
    """Everything after the order for one configuration, timed three ways: on
    the host; with the triangle index on the GPU; with the chordal completion,
    the arc map and the triangle index on the GPU (one line per phase on
    stderr).  All three must give the same structure."""
    import numpy as np
    import torch

    from nwcch import _lib, graphs

    os.environ["NWCCH_PROGRESS"] = "1"
    cfg = args.config
    t0 = time.perf_counter()
    g = graphs.make_config(cfg, progress=lambda msg: log(f"config {cfg}: {msg}"))
    log(f"config {cfg}: graph n={g.n} m={g.m} generated in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    log(f"config {cfg}: nested dissection {time.perf_counter() - t0:.1f}s")
    torch.cuda.init()
    torch.empty(1, device="cuda:0")
    sm = graphs.ev_grid(6, 0)     # first use of the GPU by the library, outside the timings
    _lib.Topology.from_graph(sm.n, sm.off, sm.tgt, np.arange(sm.n, dtype=np.int32), device_index=0,
                             device_symbolic=0)
    res, ref = {}, None
    for name, kw in (("host", {}), ("device_index", {"device_index": 0}),
                     ("device_symbolic_and_index", {"device_index": 0, "device_symbolic": 0})):
        log(f"config {cfg}: topology, {name}")
        t0 = time.perf_counter()
        t = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, **kw)
        dt = time.perf_counter() - t0
        res[name] = {"topology_s": dt}
        if ref is None:
            ref = t
            res["structure"] = {"n": g.n, "m": g.m, "mp": t.mp, "triangles": t.triangles, "depth": t.depth,
                                "levels": t.levels}
        else:
            res[name]["identical"] = same_structure(ref, t)
            del t
        log(f"config {cfg}: {name} {dt:.2f} s {res[name]}")
    out["prep"] = res


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(g, rank, topo, reps=5, queries=None):
    """The reference's C++ customization (serial and OpenMP) timed on the host;
    with queries = (s, t, gpu_dist), also its ET query (Query::run, one thread)
    on that sample, whose distances must equal the GPU's."""
    import numpy as np

    from oracle import oracle as O
    if O.have_ref():
        ref = O.Reference(g.n, g.off, g.tgt, rank)
        run = lambda: ref.customize(g.w)  # noqa: E731
        kind = "reference"
        what = "cpp/cch.cpp CCH::customize (serial, float64)"
    else:
        st = O.Structure(g.n, g.off, g.tgt, rank)
        run = lambda: st.customize(g.w)  # noqa: E731
        kind = "port"
        what = "oracle restatement of cch_nb.customize (serial, int32)"
    run()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    tmed = statistics.median(ts)
    log(f"cpu baseline ({kind}): {1000 * tmed:.2f} ms per customization")
    res = {"value": topo.triangles / tmed, "unit": "triangles/s", "cores": 1, "kind": kind,
           "sample": f"{what} on the same graph and order, median of {reps} full customizations "
                     f"(range {1000 * min(ts):.1f}-{1000 * max(ts):.1f} ms)",
           "ms": 1000 * tmed}
    if kind == "reference":
        # the reference's own multi-threaded form beside it, all host threads
        thr = os.cpu_count() or 1
        ref.customize(g.w, threads=thr)
        tp = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ref.customize(g.w, threads=thr)
            tp.append(time.perf_counter() - t0)
        tpm = statistics.median(tp)
        log(f"cpu baseline (reference, {thr} threads): {1000 * tpm:.2f} ms per customization")
        res["parallel"] = {"value": topo.triangles / tpm, "cores": thr, "ms": 1000 * tpm,
                           "sample": "cpp/cch.cpp CCH::customize_parallel (OpenMP, level buckets), "
                                     f"median of {reps}"}
        if queries is not None:
            qs, qt, gd = queries
            up, dn = ref.customize(g.w)
            ref.query(up, dn, qs[:100], qt[:100])   # warm-up, not timed
            tq = []
            for _ in range(3):
                t0 = time.perf_counter()
                d = ref.query(up, dn, qs, qt)
                tq.append(time.perf_counter() - t0)
            tqm = statistics.median(tq)
            d = np.where(np.isinf(d), O.INF, d).astype(np.int64)
            same = bool(np.array_equal(d, gd.astype(np.int64)))
            log(f"cpu baseline (reference queries): {len(qs)} in {1000 * tqm:.1f} ms; equal to the GPU's: {same}")
            res["queries"] = {"value": len(qs) / tqm, "unit": "queries/s", "cores": 1, "kind": "reference",
                              "sample": f"cpp/cch.cpp Query::run, {len(qs)} random s-t pairs of the GPU batch, "
                                        "median of 3",
                              "ms": 1000 * tqm, "equal_to_gpu": same}
    res["cpu_model"] = cpu_model()
    return res


def batched(args, out):
    """64 metrics of config 2 in one level sweep vs 64 single-metric sweeps."""
    import numpy as np
    import torch

    from nwcch import graphs
    from nwcch.device import DeviceTopology

    g, rank, topo, host = prepare(2)
    K = args.metrics
    W, slopes = graphs.metric_family(g, K, seed=2, hmax=1200)
    tri = graphs.find_directed_triangle(g, np.random.default_rng(2))
    bad = K // 3
    if tri is not None:
        W[bad] = graphs.inject_negative_cycle(W[bad], tri)
    dt = DeviceTopology(topo, 0)
    Wd = dt.weights_batch(W)
    from nwcch import tune
    btune = tune.tune_batch(dt, Wd, reps=2) if not args.no_tune else {"slab": 0}
    slab = btune["slab"]
    UD = torch.empty(dt.batch_shape(K, slab), dtype=torch.int32, device=dt.dev)
    sc = torch.empty((2, K), dtype=torch.int64, device=dt.dev)

    def step():
        dt.customize_batch(Wd, out=UD, slab=slab)
        dt.negcycle_scan_batch(UD, K, out=sc)

    timed(step, 2)
    dt.check_status(negative_cycle=True)
    t_b = timed(step, args.batch_steps)
    t_bg = None
    if not args.no_graph:
        from nwcch.device import StepGraph
        bgraph = StepGraph(step)
        timed(bgraph.replay, 1)
        t_bg = timed(bgraph.replay, args.batch_steps)
        dt.check_status(negative_cycle=True)
    t_b_eager = t_b
    if t_bg is not None:
        t_b = min(t_b, t_bg)
    ud1 = torch.empty((topo.mp, 2), dtype=torch.int32, device=dt.dev)
    s1 = torch.empty(2, dtype=torch.int64, device=dt.dev)
    rows = [Wd[k].contiguous() for k in range(K)]

    def single_all():
        for k in range(K):
            dt.customize(rows[k], out=ud1)
            dt.negcycle_scan(ud1, out=s1)

    timed(single_all, 1)
    t_s = timed(single_all, 2)
    bms = statistics.median(event_ms(lambda: dt.customize_batch(Wd, out=UD, slab=slab), args.batch_steps))
    B = cust_bytes(g.m, topo.mp, topo.triangles, topo.n_mapped, K)
    dt.profile_begin(args.batch_steps * (topo.levels + 4))
    for _ in range(args.batch_steps):
        dt.customize_batch(Wd, out=UD, slab=slab)
    kid, kms = dt.profile_end()
    n_lv = int((kid == 4).sum()) // args.batch_steps
    bbase_ms = float(np.median(kms[kid == 5]))
    blv_us = 1000.0 * (bms - bbase_ms) / n_lv
    B_blv = 8 * topo.triangles + K * 16 * topo.mp   # operand ids shared, 16 B per arc per metric
    b_ach = (B_blv / n_lv) / (blv_us * 1e-6) / 1e9
    flagged = [int(k) for k in np.nonzero(sc[0].cpu().numpy() > 0)[0]]
    parity = None
    if not args.no_check:
        from oracle import oracle as O
        st = O.Structure(g.n, g.off, g.tgt, rank)
        ok = True
        for k in sorted({0, bad, K - 1}):
            up, dn, _ = st.customize(W[k])
            got = dt.metric_view(UD, k).cpu().numpy()
            ok &= np.array_equal(got[:, 0], up) and np.array_equal(got[:, 1], dn)
            ok &= (int(sc[0, k]), int(sc[1, k])) == st.negcycle(up, dn)
        parity = bool(ok)
        log(f"batched parity vs oracle (metrics 0, {bad}, {K - 1}): {parity}; flagged {flagged}")
    out["batched"] = {
        "workload": f"config 2: n={g.n} m={g.m} m+={topo.mp} T={topo.triangles} depth={topo.depth}; "
                    f"{K} metrics in one level sweep + per-metric scan",
        "metrics": K, "ms_per_step": 1000 * t_b, "triangles_per_s": K * topo.triangles / t_b,
        "step_ms": {"eager": 1000 * t_b_eager, "hip_graph": None if t_bg is None else 1000 * t_bg},
        "plain_ms_for_all_metrics": 1000 * t_s, "plain_triangles_per_s": K * topo.triangles / t_s,
        "speedup_vs_plain": t_s / t_b,
        "roofline": {"kernel": "k_customize_batch_level", "bound": "hbm", "achieved": b_ach,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b_ach / HBM_PEAK_GBS,
                     "algorithmic_bytes_per_launch": B_blv / n_lv, "avg_launch_us": blv_us, "level_launches": n_lv,
                     "avg_launch_us_isolated": 1000.0 * float(kms[kid == 4].astype(np.float64).mean()),
                     "base_pass_us": 1000.0 * bbase_ms,
                     "customization_algorithmic_bytes": B, "customize_ms": bms,
                     "customization_GB_s": B / (bms / 1000.0) / 1e9},
        "injected_cycle_metric": bad if tri is not None else None, "flagged_metrics": flagged,
        "parity_vs_oracle": parity, "tuning": btune, "host": host,
    }
    del dt


def serving(args, out):
    """Config 3: 16 resident metrics on the 1.1M-point graph (seed 3), Poisson
    query streams with interleaved 1/10/1000-arc updates; p50/p99 latency and
    throughput per arrival rate, every answer re-checked alone."""
    import numpy as np

    from nwcch import graphs, serve
    from nwcch.device import DeviceTopology

    g, rank, topo, host = prepare(3)
    K = 16
    W, _ = graphs.metric_family(g, K, seed=3, hmax=800)
    dt = DeviceTopology(topo, 0)
    srv = serve.Server(dt, rank, metric_budget=K * 2 * (4 * topo.m + 8 * topo.mp), slot_budget=64 << 20,
                       max_batch=65536, n_streams=args.serve_streams, snapshot=True)
    w0 = {}
    for k in range(K):
        if not srv.add_metric(k, W[k]):
            raise RuntimeError(f"metric {k} was not admitted")
        w0[k] = W[k].copy()
    # warm-up: a short query-only run (first launches, lazy allocations); not reported
    warm = serve.poisson_workload(g, K, rate=args.serve_rates[0], duration=0.05, seed=29, update_rate=0)
    srv.run(warm, max_seconds=60, max_inflight=args.serve_inflight)
    runs = []
    all_reqs = []
    for i, rate in enumerate(args.serve_rates):
        reqs = serve.poisson_workload(g, K, rate=rate, duration=args.serve_seconds, seed=30 + i,
                                      update_rate=args.serve_update_rate)
        reqs = serve.absolute_updates(reqs, {k: srv.weights_at(k, srv.metrics[k].version, w0[k])
                                             for k in range(K)})
        st = srv.run(reqs, max_seconds=60 + 10 * args.serve_seconds, max_inflight=args.serve_inflight)
        st["arrival_rate"] = rate
        runs.append(st)
        all_reqs += reqs
        log(f"serving {rate:.0f}/s: {st['queries_per_s']:.0f} q/s, p50 {st['query_p50_ms']:.3f} ms, "
            f"p99 {st['query_p99_ms']:.3f} ms, batch {st['mean_batch']:.0f}, updates {st['updates']} "
            f"(rejected {st['rejected_updates']}), update p99 {st['update_p99_ms']:.2f} ms")
    ver = srv.verify_alone(all_reqs, w0) if not args.no_check else None
    if ver is not None:
        log(f"serving: every answer vs alone recompute: {ver}")
    out["serving"] = {
        "workload": f"config 3: n={g.n} m={g.m} m+={topo.mp} T={topo.triangles} depth={topo.depth}; {K} "
                    f"resident metrics; Poisson queries (20% within 5 km) + 1/10/1000-arc updates",
        "update_mode": "snapshot (updates on a second copy, committed by swap)",
        "runs": runs, "verify_alone": ver, "host": host,
    }
    del srv, dt


def full_scale(args, out):
    """Config 4: the 23.9M-point road graph (seed 4, continental fractal heights
    0-3,000 m).  One metric and K = 8 metrics customized on the GPU, the
    negative-cycle scans, 1M random ET queries; every shortcut value checked
    against the reference's own C++ customization (cpp/cch.cpp, compiled from
    its sources) and a query sample against Johnson + Dijkstra; the reference's
    serial and OpenMP customization timed on the host beside it."""
    import concurrent.futures as cf

    import numpy as np
    import torch

    from nwcch import graphs
    from nwcch.device import DeviceTopology

    os.environ["NWCCH_PROGRESS"] = "1"   # a line per big nested-dissection cell and preprocessing phase
    # the triangle index is counted and filled on the GPU
    g, rank, topo, host = prepare(4, progress=lambda msg: log(f"config 4: {msg}"), device_index=0)
    t0 = time.perf_counter()
    dt = DeviceTopology(topo, 0)
    torch.cuda.synchronize()
    host["dev_create_s"] = time.perf_counter() - t0
    log(f"config 4: device topology created in {host['dev_create_s']:.1f} s (triangle lists filled on the GPU)")
    free_b, total_b = torch.cuda.mem_get_info()
    log(f"device topology uploaded ({(total_b - free_b) / 1e9:.1f} GB of HBM in use)")
    w = dt.weights(g.w)
    ud = torch.empty((topo.mp, 2), dtype=torch.int32, device=dt.dev)
    scan = torch.empty(2, dtype=torch.int64, device=dt.dev)
    dt.customize(w, out=ud)
    dt.check_status()
    cust_ms = event_ms(lambda: dt.customize(w, out=ud), args.steps)
    scan_ms = event_ms(lambda: dt.negcycle_scan(ud, out=scan), args.steps)
    dt.profile_begin(topo.levels + 2)
    dt.customize(w, out=ud)
    kid, kms = dt.profile_end()
    n_lv = int((kid == 0).sum())
    base_ms = float(np.median(kms[kid == 5]))
    cmed = statistics.median(cust_ms)
    B_lv = level_bytes(topo.mp, topo.triangles)
    lv_avg_us = 1000.0 * (cmed - base_ms) / n_lv
    achieved = (B_lv / n_lv) / (lv_avg_us * 1e-6) / 1e9
    log(f"config 4 customize {cmed:.1f} ms ({topo.triangles / (cmed / 1e3):.3g} tri/s), level launches "
        f"{n_lv}, {achieved:.0f} GB/s algorithmic; scan {statistics.median(scan_ms):.2f} ms")
    # K = 8 metrics in one sweep
    K = 8
    W, _ = graphs.metric_family(g, K, seed=4, hmax=3000)
    Wd = dt.weights_batch(W)
    del W
    UD = torch.empty(dt.batch_shape(K, 0), dtype=torch.int32, device=dt.dev)
    sc = torch.empty((2, K), dtype=torch.int64, device=dt.dev)
    dt.customize_batch(Wd, out=UD)
    dt.check_status()
    bms = statistics.median(event_ms(lambda: dt.customize_batch(Wd, out=UD), max(2, args.steps // 2)))
    dt.negcycle_scan_batch(UD, K, out=sc)
    flagged = [int(k) for k in np.nonzero(sc[0].cpu().numpy() > 0)[0]]
    log(f"config 4 K={K}: {bms:.1f} ms ({K * topo.triangles / (bms / 1e3):.3g} tri/s), flagged {flagged}")
    del UD, Wd
    # 1M random ET queries on metric 0
    nq = args.full_queries
    qs_h, qt_h = graphs.random_pairs(g.n, nq, seed=41)
    qs = torch.as_tensor(rank[qs_h]).to(dt.dev)
    qt = torch.as_tensor(rank[qt_h]).to(dt.dev)
    dist = torch.empty(nq, dtype=torch.int32, device=dt.dev)
    order = dt.chain_order(qs, qt)
    # the query kernels on this very deep hierarchy: the throughput kernel for
    # the big launch, then each latency kernel for the big launch and alone
    forms = {}
    for name, chunk, deep in (("chain_walk_throughput", 0, 0), ("chunked_latency", 1, 1 << 30),
                              ("chain_walk_latency", 0, 1 << 30)):
        dt.set_tuning("query_chunk", chunk)
        dt.set_tuning("query_deep_max", deep)
        dt.query(ud, qs, qt, want_scanned=False, dist=dist, order=order)
        bms_q = statistics.median(event_ms(lambda: dt.query(ud, qs, qt, want_scanned=False, dist=dist, order=order),
                                           2))
        one = [event_ms(lambda i=i: dt.query(ud, qs[i:i + 1], qt[i:i + 1], want_scanned=False), 1)[0]
               for i in range(30)]
        forms[name] = {"batch_ms": bms_q, "queries_per_s": nq / (bms_q / 1e3),
                       "single_p50_ms": float(np.percentile(one, 50))}
        log(f"config 4 queries, {name}: {nq} in {bms_q:.1f} ms; single p50 {np.percentile(one, 50):.2f} ms")
    best_b = min(forms, key=lambda k: forms[k]["batch_ms"])
    best_s = min(forms, key=lambda k: forms[k]["single_p50_ms"])
    dt.set_tuning("query_chunk", 1 if best_b == "chunked_latency" else 0)
    dt.set_tuning("query_deep_max", 0 if best_b == "chain_walk_throughput" else 1 << 30)
    dt.query(ud, qs, qt, want_scanned=False, dist=dist, order=order)
    qms = forms[best_b]["batch_ms"]
    dt.set_tuning("query_chunk", 1 if best_s == "chunked_latency" else 0)
    dt.set_tuning("query_deep_max", 0 if best_s == "chain_walk_throughput" else 1 << 30)
    single = [event_ms(lambda i=i: dt.query(ud, qs[i:i + 1], qt[i:i + 1], want_scanned=False), 1)[0]
              for i in range(50)]
    dt.check_status()
    log(f"config 4 queries: {nq} in {qms:.1f} ms ({nq / (qms / 1e3):.3g} q/s); single p50 "
        f"{np.percentile(single, 50):.2f} ms")
    res = {
        "workload": f"config 4: n={g.n} m={g.m} m+={topo.mp} T={topo.triangles} ET depth={topo.depth}",
        "n": g.n, "m": g.m, "m_plus": topo.mp, "triangles": topo.triangles, "depth": topo.depth,
        "levels": topo.levels, "negative_arc_fraction": round(g.negative_fraction(), 4), "host": host,
        "customize_ms": cmed, "triangles_per_s": topo.triangles / (cmed / 1e3),
        "roofline": {"kernel": "level sweep", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": B_lv / n_lv,
                     "avg_launch_us": lv_avg_us, "level_launches": n_lv},
        "scan_ms": statistics.median(scan_ms),
        "batched": {"metrics": K, "ms": bms, "triangles_per_s": K * topo.triangles / (bms / 1e3),
                    "flagged_metrics": flagged},
        "queries": {"count": nq, "ms": qms, "queries_per_s": nq / (qms / 1e3), "batch_kernel": best_b,
                    "single_kernel": best_s, "single_p50_ms": float(np.percentile(single, 50)),
                    "single_p99_ms": float(np.percentile(single, 99)), "forms": forms},
    }
    if not args.no_check or not args.no_cpu:
        from oracle import oracle as O
        got = ud.cpu().numpy()
        d_gpu = dist[:args.full_check].cpu().numpy()
        del ud
        torch.cuda.empty_cache()
        if O.have_ref():
            t0 = time.perf_counter()
            ref = O.Reference(g.n, g.off, g.tgt, rank)
            t_build = time.perf_counter() - t0
            t0 = time.perf_counter()
            up, dn = ref.customize(g.w)
            t_ser = time.perf_counter() - t0
            ok = True
            for col, r in ((0, up), (1, dn)):
                r = np.where(np.isinf(r), O.INF, r).astype(np.int64)
                ok &= bool(np.array_equal(got[:, col].astype(np.int64), r))
            del up, dn
            thr = os.cpu_count() or 1
            t0 = time.perf_counter()
            ref.customize(g.w, threads=thr)
            t_par = time.perf_counter() - t0
            res["parity_vs_reference_customization"] = ok
            res["cpu_baseline"] = [
                {"value": topo.triangles / t_ser, "unit": "triangles/s", "cores": 1, "kind": "reference",
                 "sample": "cpp/cch.cpp CCH::customize (serial, float64), one full customization", "ms": 1e3 * t_ser},
                {"value": topo.triangles / t_par, "unit": "triangles/s", "cores": thr, "kind": "reference",
                 "sample": "cpp/cch.cpp CCH::customize_parallel (OpenMP, level buckets), one full customization",
                 "ms": 1e3 * t_par}]
            res["reference_build_s"] = t_build
            log(f"config 4 reference: build {t_build:.1f}s, serial {t_ser:.1f}s, {thr} threads {t_par:.1f}s; "
                f"parity {ok}")
            del ref
        if not args.no_check:
            t0 = time.perf_counter()
            p, neg = O.johnson(g.n, g.off, g.tgt, g.w)
            t_j = time.perf_counter() - t0
            nchk = args.full_check
            workers = min(8, os.cpu_count() or 1)
            import queue
            free = queue.Queue()   # one searcher (with its own scratch arrays) per running task
            for _ in range(workers):
                free.put(O.Dijkstra(g.n, g.off, g.tgt, g.w, p))

            def run(i):
                dj = free.get()
                try:
                    return dj(int(qs_h[i]), int(qt_h[i]))
                finally:
                    free.put(dj)
            t0 = time.perf_counter()
            with cf.ThreadPoolExecutor(workers) as ex:
                want = np.array(list(ex.map(run, range(nchk), chunksize=1)), np.int64)
            t_d = time.perf_counter() - t0
            want = np.where(want >= O.INF, O.INF, want)
            okq = bool(np.array_equal(d_gpu.astype(np.int64), want))
            if not okq:
                bad = np.nonzero(d_gpu.astype(np.int64) != want)[0]
                log(f"config 4 query mismatches at {bad[:10].tolist()}: gpu {d_gpu[bad[:5]].tolist()} "
                    f"dijkstra {want[bad[:5]].tolist()}")
            res["query_check"] = {"pairs": nchk, "agree_with_johnson_dijkstra": okq, "johnson_s": t_j,
                                  "dijkstra_s": t_d, "negative_cycle": neg}
            log(f"config 4 queries vs Johnson+Dijkstra on {nchk}: {okq} (Johnson {t_j:.1f}s, Dijkstra {t_d:.1f}s)")
    # the metric-independent index: the side lists on the GPU, then the same
    # topology built with the host's triangle index for the time beside it and
    # an entry-for-entry comparison of the triangle lists
    t0 = time.perf_counter()
    dt.prepare_update()
    torch.cuda.synchronize()
    t_side = time.perf_counter() - t0
    prep = {"device_index": {"topology_s": host["topology_s"], "dev_create_s": host["dev_create_s"],
                             "update_index_s": t_side}}
    log(f"config 4: update side lists on the GPU in {t_side:.1f} s")
    if not args.no_check:
        from nwcch import _lib
        t0 = time.perf_counter()
        th = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
        t_host = time.perf_counter() - t0
        sh, oh, tab = th.export_triangles()
        sd, od, _ = topo.export_triangles(with_list=False)
        same = bool(np.array_equal(sh, sd) and np.array_equal(oh, od))
        del sh, oh, sd, od
        th.close()
        same &= bool(np.array_equal(dt.export_triangles(), tab))
        del tab
        prep["host_index"] = {"topology_s": t_host}
        prep["identical"] = same
        log(f"config 4: host-built index {t_host:.1f} s; device-built lists identical: {same}")
    res["preprocessing"] = prep
    out["full_scale"] = res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--queries", type=int, default=100_000)
    ap.add_argument("--single", type=int, default=500, help="single-query launches for latency")
    ap.add_argument("--metrics", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (scales HBM byte counters)")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP graph replay of the step")
    ap.add_argument("--full-scale", action="store_true",
                    help="run the 23.9M-point road graph (config 4) instead of the default workloads")
    ap.add_argument("--prep", action="store_true",
                    help="time everything after the order for --config on the host and on the GPU")
    ap.add_argument("--full-queries", type=int, default=1_000_000)
    ap.add_argument("--full-check", type=int, default=100, help="queries checked against Johnson + Dijkstra")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library defaults)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[200_000, 1_000_000, 2_000_000])
    ap.add_argument("--serve-seconds", type=float, default=0.25)
    ap.add_argument("--serve-inflight", type=int, default=3, help="query launches on the device at once")
    ap.add_argument("--serve-streams", type=int, default=2, help="streams the query launches rotate over")
    ap.add_argument("--serve-update-rate", type=float, default=40.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); no path sharded over "
              "several GPUs is built", file=sys.stderr)
        sys.exit(2)
    out = {
        "metric": "triangles/s customize; ET queries/s + p99 latency, batched/served vs plain GPU",
        "value": None, "unit": "triangles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
    }
    if args.full_scale:
        full_scale(args, out)
        print(json.dumps(out), flush=True)
        return
    if args.prep:
        prep(args, out)
        print(json.dumps(out), flush=True)
        return
    plain(args, out)
    if not args.no_batched:
        batched(args, out)
    if not args.no_serving:
        serving(args, out)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
