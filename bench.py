"""Benchmark of the metric-dependent CCH path on one MI355X.

Headline workload (`value`, the default run): config 4, the full-scale
road-like graph -- 23.9 M clustered points (seed 4), Delaunay thinned to a mean
out-degree of 2.44, 5% one-way, largest SCC, continental fractal heights
0-3,000 m, int32 energy weights, inertial-flow order (host, untimed).  One step
= K = 8 metrics customized in one level sweep plus the per-metric
negative-cycle scans, inputs resident in HBM, replayed from one HIP graph;
value = K x lower triangles / step time.

The same run also reports, as extra fields: the plain one-metric
customization and scan; 1 M random ET queries in one launch (three kernel
forms) and single-query launches (p50/p99); per-launch roofline figures of
the level kernels; the CPU baseline -- the reference's own C++ (cpp/cch.cpp,
compiled from its sources) on the host: serial and OpenMP customization,
Query::run on one core and on all cores; and parity: every arc of one metric
against the reference's customization, every batched metric against its
single-metric run, 100 queries against Johnson + Dijkstra, a query sample
against the reference's Query::run.

Other modes: `--config 1` (plain one-metric workload on the 264,000-point
graph, with updates, potential, paths, and the batched 64-metric run on
config 2 and the serving run on config 3 after it); `--prep` times everything
after the order for `--config` on the host and on the GPU.

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


def cust_bytes_8d(n, m, mp, T, K):
    """Algorithmic bytes of one customization of K metrics: a 4-byte c-index
    per lower triangle, first and head, per metric every (up, down) pair read
    and written once and the weight vector, the arc map (4 + 1 bytes per
    original arc)."""
    return 4 * T + 4 * (n + 1) + 4 * mp + K * (16 * mp + 4 * m) + 5 * m


def level_bytes_8d(n, mp, T, K):
    """The part of cust_bytes_8d the level launches move: the c-index stream,
    first/head and the K (up, down) pairs of every arc read and written."""
    return 4 * T + 4 * (n + 1) + 4 * mp + 16 * K * mp


def layout_level_bytes(mp, T, K):
    """The same in this build's layout: an (a, b) operand pair (8 B) per
    triangle and a 16-byte schedule record per arc instead of the c-index."""
    return 8 * T + 16 * mp + 16 * K * mp


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
    # the first updates capture the stream's launch graphs
    idx = rng.choice(g.m, size=1000, replace=False).astype(np.int64)
    for mode in (0, 2):
        w2, ud2 = w.clone(), ud.clone()
        dt.update(w2, ud2, idx, g.w[idx] + 1, mode=mode)
    torch.cuda.synchronize()
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
    # beside the default (auto: a full recustomization diffed against the old
    # values from 3 changes on), the frontier climbing the heights one launch
    # per height (mode 2)
    for k in (10, 1000):
        gpu_ms = []
        for _ in range(reps):
            idx = rng.choice(g.m, size=k, replace=False).astype(np.int64)
            neww = np.clip(g.w[idx].astype(np.int64) + rng.integers(-500, 501, size=k), -(1 << 20), 1 << 20)
            w2, ud2 = w.clone(), ud.clone()
            gpu_ms.append(event_ms(lambda: dt.update(w2, ud2, idx, neww.astype(np.int32), mode=2), 1)[0])
        dt.status(clear=True)
        res[str(k)]["gpu_ms_median_per_height_driver"] = statistics.median(gpu_ms)
        log(f"update k={k}: per-height driver {statistics.median(gpu_ms):.3f} ms")
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
    # the plain form beside it on the same graph and metric 0: one query per
    # launch (latency) and one large launch (throughput), no updates
    import torch
    ud0 = srv.metrics[0].ud
    qs_h, qt_h = graphs.random_pairs(g.n, 100_000, seed=33)
    qs = torch.as_tensor(rank[qs_h]).to(dt.dev)
    qt = torch.as_tensor(rank[qt_h]).to(dt.dev)
    dt.query(ud0, qs[:1], qt[:1], want_scanned=False)
    one = [event_ms(lambda i=i: dt.query(ud0, qs[i:i + 1], qt[i:i + 1], want_scanned=False), 1)[0]
           for i in range(args.single)]
    dist = torch.empty(len(qs_h), dtype=torch.int32, device=dt.dev)
    dt.query(ud0, qs, qt, want_scanned=False, dist=dist)
    bms = statistics.median(event_ms(lambda: dt.query(ud0, qs, qt, want_scanned=False, dist=dist), 3))
    plain = {"single_launches": len(one), "single_p50_ms": float(np.percentile(one, 50)),
             "single_p99_ms": float(np.percentile(one, 99)),
             "single_launch_queries_per_s": 1e3 / float(np.mean(one)),
             "batch": len(qs_h), "batch_ms": bms, "batch_queries_per_s": len(qs_h) / (bms / 1e3)}
    log(f"config 3 plain: one query per launch p50 {plain['single_p50_ms']:.3f} ms p99 "
        f"{plain['single_p99_ms']:.3f} ms; {len(qs_h)} in one launch {bms:.2f} ms "
        f"({plain['batch_queries_per_s']:.3g} q/s)")
    out["serving"] = {
        "workload": f"config 3: n={g.n} m={g.m} m+={topo.mp} T={topo.triangles} depth={topo.depth}; {K} "
                    f"resident metrics; Poisson queries (20% within 5 km) + 1/10/1000-arc updates",
        "update_mode": "snapshot (updates on a second copy, committed by swap)",
        "runs": runs, "verify_alone": ver, "plain_one_metric": plain, "host": host,
    }
    del srv, dt


def config4_inputs(args, progress):
    """The config-4 graph, its order and the K metrics.  While the host orders
    the graph (native, threads), the K weight vectors are drawn and, as the
    checker of the query sample, Johnson's potentials and Dijkstra's answers
    are computed on the original graph (neither needs the order).  With
    --cache FILE everything is read from / written to one .npz."""
    import concurrent.futures as cf

    import numpy as np

    from nwcch import _lib, graphs
    if args.cache and Path(args.cache).exists():
        t0 = time.perf_counter()
        z = np.load(args.cache)
        g = graphs.RoadGraph(int(z["n"]), z["off"], z["tgt"], z["w"], z["x"], z["y"], None, "c4")
        g.extra["box"] = float(z["box"])
        log(f"config {args.graph}: graph, order and metrics read from {args.cache} in {time.perf_counter() - t0:.1f}s")
        chk = (z["chk_s"], z["chk_t"], z["chk_d"]) if "chk_d" in z else None
        return g, z["rank"], z["W"], chk, {"cached": True}
    host = {}
    t0 = time.perf_counter()
    g = graphs.make_config(args.graph, progress=progress)
    host["gen_s"] = time.perf_counter() - t0
    log(f"config {args.graph}: graph n={g.n} m={g.m} generated in {host['gen_s']:.1f}s")
    qs_h, qt_h = graphs.random_pairs(g.n, args.full_check, seed=43)

    def checker():
        # Johnson's virtual-source potentials, then one Dijkstra per pair
        from oracle import oracle as O
        import queue
        t1 = time.perf_counter()
        p, neg = O.johnson(g.n, g.off, g.tgt, g.w)
        free = queue.Queue()        # one searcher (own scratch arrays) per running task
        for _ in range(4):
            free.put(O.Dijkstra(g.n, g.off, g.tgt, g.w, p))

        def run(i):
            dj = free.get()
            try:
                return dj(int(qs_h[i]), int(qt_h[i]))
            finally:
                free.put(dj)
        with cf.ThreadPoolExecutor(4) as ex:
            want = np.array(list(ex.map(run, range(len(qs_h)))), np.int64)
        want = np.where(want >= O.INF, O.INF, want)
        log(f"config {args.graph}: Johnson + Dijkstra answers for {len(qs_h)} pairs in {time.perf_counter() - t1:.1f}s")
        return want, bool(neg)

    with cf.ThreadPoolExecutor(2) as side:
        fut_w = side.submit(graphs.metric_family, g, args.metrics, 4, graphs.CONFIGS[args.graph].get("hmax", 3000))
        fut_c = side.submit(checker) if not args.no_check else None
        t0 = time.perf_counter()
        rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
        host["order_s"] = time.perf_counter() - t0
        log(f"config {args.graph}: nested dissection {host['order_s']:.1f}s")
        t0 = time.perf_counter()
        W, _ = fut_w.result()
        chk = None
        if fut_c is not None:
            want, _neg = fut_c.result()
            chk = (qs_h, qt_h, want)
        host["side_wait_s"] = time.perf_counter() - t0
    if args.cache:
        extra = {} if chk is None else {"chk_s": chk[0], "chk_t": chk[1], "chk_d": chk[2]}
        np.savez(args.cache, n=g.n, off=g.off, tgt=g.tgt, w=g.w, x=g.x, y=g.y, box=g.extra["box"], rank=rank,
                 W=W, **extra)
        log(f"config {args.graph}: inputs written to {args.cache}")
    return g, rank, W, chk, host


def counter_run(args, out):
    """A short config-4 run for hardware-counter passes (rocprofv3 --pmc): the
    graph, its order and K = 8 metrics, then 2 batched customizations (and 2
    more with the arcs of each wave class in id order and one slice per XCD), 2
    single-metric ones, 2 launches of 20,000 queries on each query form, and
    one device copy of --copy-mib MiB to scale the byte counters.  No CPU
    baseline, no parity check, no timing loop."""
    import torch

    from nwcch import _lib, graphs
    from nwcch.device import DeviceTopology

    os.environ["NWCCH_PROGRESS"] = "1"
    t0 = time.perf_counter()
    g = graphs.make_config(args.graph, progress=lambda msg: log(f"config {args.graph}: {msg}"))
    rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    host = {"gen_and_order_s": time.perf_counter() - t0}
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0, device_symbolic=0)
    dt = DeviceTopology(topo, 0)
    W, _ = graphs.metric_family(g, args.metrics, 4, graphs.CONFIGS[args.graph].get("hmax", 3000))
    K = W.shape[0]
    Wd = dt.weights_batch(W)
    UD = torch.empty(dt.batch_shape(K), dtype=torch.int32, device=dt.dev)
    for _ in range(2):
        dt.customize_batch(Wd, out=UD)
    # the same sweep with the arcs of each wave class in id order, one
    # contiguous slice of each class per XCD (rows of one tail on one XCD)
    dt.set_tuning("batch_order", 1)
    dt.set_tuning("batch_xcd", 1)
    for _ in range(2):
        dt.customize_batch(Wd, out=UD)
    dt.set_tuning("batch_order", 0)
    dt.set_tuning("batch_xcd", 0)
    w0 = dt.weights(W[0])
    ud = torch.empty((topo.mp, 2), dtype=torch.int32, device=dt.dev)
    for _ in range(2):
        dt.customize(w0, out=ud)
    nq = 20000
    qs_h, qt_h = graphs.random_pairs(g.n, nq, seed=41)
    qs = torch.as_tensor(rank[qs_h]).to(dt.dev)
    qt = torch.as_tensor(rank[qt_h]).to(dt.dev)
    up, dn = dt.split(ud)
    dt.set_tuning("query_blk_waves", 2)
    for _ in range(2):
        dt.query(ud, qs, qt, want_scanned=False)
        dt.query_split(up, dn, qs, qt, want_scanned=False)
    dt.set_tuning("query_blk_waves", 0)
    if args.copy_mib:
        src = torch.ones(args.copy_mib << 18, dtype=torch.int32, device=dt.dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
    dt.check_status()
    torch.cuda.synchronize()
    out.update({"config": {"workload": "counter run: config 4, 2 batched (K=8) + 2 single customizations, "
                                       "2 x 20,000 queries per form"}, "host": host})
    del dt


def headline(args, out):
    """Config 4, the full-scale road graph (23.9 M points, seed 4): K = 8
    metrics customized in one level sweep plus their negative-cycle scans per
    step (the timed work), the plain one-metric customization beside it,
    1 M random ET queries, the reference's own C++ on the host as the CPU
    baseline, and parity: every arc of one metric against the reference's
    customization, every batched metric against its single-metric run, and
    100 queries against Johnson + Dijkstra."""
    import numpy as np
    import torch

    from nwcch import _lib, graphs
    from nwcch.device import DeviceTopology, StepGraph

    os.environ["NWCCH_PROGRESS"] = "1"   # a line per big nested-dissection cell and preprocessing phase
    g, rank, W, chk, host = config4_inputs(args, lambda msg: log(f"config {args.graph}: {msg}"))
    K = W.shape[0]
    t0 = time.perf_counter()
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank, device_index=0, device_symbolic=0)
    host["topology_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    dt = DeviceTopology(topo, 0)
    torch.cuda.synchronize()
    host["dev_create_s"] = time.perf_counter() - t0
    free_b, total_b = torch.cuda.mem_get_info()
    log(f"config {args.graph}: n={g.n} m={g.m} m+={topo.mp} T={topo.triangles} depth={topo.depth} levels={topo.levels}; "
        f"topology {host['topology_s']:.1f}s, device {host['dev_create_s']:.1f}s, "
        f"{(total_b - free_b) / 1e9:.1f} GB of HBM in use")

    # ---- the timed step: K metrics in one sweep + per-metric scans, one HIP graph
    Wd = dt.weights_batch(W)
    UD = torch.empty(dt.batch_shape(K), dtype=torch.int32, device=dt.dev)
    sc = torch.empty((2, K), dtype=torch.int64, device=dt.dev)

    def step():
        dt.customize_batch(Wd, out=UD)
        dt.negcycle_scan_batch(UD, K, out=sc)

    step()
    dt.check_status()
    graph = StepGraph(step)
    for _ in range(args.warmup):
        graph.replay()
    dt.check_status()
    t_step = timed(graph.replay, args.steps)
    t_eager = timed(step, 2)
    dt.check_status()
    flagged = [int(k) for k in np.nonzero(sc[0].cpu().numpy() > 0)[0]]
    # device time of the batched customization alone (graph replay, HIP events)
    cgraph = StepGraph(lambda: dt.customize_batch(Wd, out=UD))
    cgraph.replay()
    bms = statistics.median(event_ms(cgraph.replay, max(3, args.steps // 4)))
    del cgraph
    bscan_ms = statistics.median(event_ms(lambda: dt.negcycle_scan_batch(UD, K, out=sc), 5))
    # per-launch durations: HIP events the library records around each launch
    dt.profile_begin(topo.levels + 8)
    dt.customize_batch(Wd, out=UD)
    kid, kms = dt.profile_end()
    n_lv = int((kid == 4).sum())
    bbase_ms = float(kms[kid == 5].astype(np.float64).sum())
    lv_avg_us = 1000.0 * (bms - bbase_ms) / n_lv
    B8d_lv = level_bytes_8d(g.n, topo.mp, topo.triangles, K)
    B8d = cust_bytes_8d(g.n, g.m, topo.mp, topo.triangles, K)
    achieved = (B8d_lv / n_lv) / (lv_avg_us * 1e-6) / 1e9
    traffic = read_traffic("k_customize_batch_level_road23.4m", B8d_lv / n_lv)
    log(f"config {args.graph} K={K}: step {1e3 * t_step:.2f} ms ({K * topo.triangles / t_step:.3g} tri/s); customization "
        f"{bms:.2f} ms = base {bbase_ms:.2f} ms + {n_lv} level launches of {lv_avg_us:.1f} us; "
        f"{achieved:.0f} GB/s by the algorithmic bytes; flagged {flagged}")

    # ---- the plain one-metric form beside it
    w0 = dt.weights(W[0])
    ud = torch.empty((topo.mp, 2), dtype=torch.int32, device=dt.dev)
    scan = torch.empty(2, dtype=torch.int64, device=dt.dev)

    def step1():
        dt.customize(w0, out=ud)
        dt.negcycle_scan(ud, out=scan)

    step1()
    g1 = StepGraph(step1)
    g1.replay()
    t_plain = timed(g1.replay, max(3, args.steps // 2))
    cust1_ms = statistics.median(event_ms(lambda: dt.customize(w0, out=ud), max(3, args.steps // 4)))
    scan1_ms = statistics.median(event_ms(lambda: dt.negcycle_scan(ud, out=scan), 5))
    dt.profile_begin(topo.levels + 8)
    dt.customize(w0, out=ud)
    kid1, kms1 = dt.profile_end()
    n_lv1 = int((kid1 == 0).sum())
    base1_ms = float(kms1[kid1 == 5].astype(np.float64).sum())
    lv1_us = 1000.0 * (cust1_ms - base1_ms) / n_lv1
    B8d_lv1 = level_bytes_8d(g.n, topo.mp, topo.triangles, 1)
    ach1 = (B8d_lv1 / n_lv1) / (lv1_us * 1e-6) / 1e9
    dt.check_status()
    log(f"config {args.graph} one metric: step {1e3 * t_plain:.2f} ms, customization {cust1_ms:.2f} ms "
        f"({topo.triangles / (cust1_ms / 1e3):.3g} tri/s), scan {scan1_ms:.3f} ms")

    # ---- ET queries on metric 0: a batch of 1 M in one launch, and one query per launch
    queries = None
    if args.full_queries:
        nq = args.full_queries
        qs_h, qt_h = graphs.random_pairs(g.n, nq, seed=41)
        qs = torch.as_tensor(rank[qs_h]).to(dt.dev)
        qt = torch.as_tensor(rank[qt_h]).to(dt.dev)
        dist = torch.empty(nq, dtype=torch.int32, device=dt.dev)
        order = dt.chain_order(qs, qt)
        # query kernel forms: (query_blk, wavefronts per side, query_chunk, query_deep_max)
        kinds = {"blocked_4": (1, 4, 0, 1 << 30), "blocked_2": (1, 2, 0, 1 << 30), "blocked_1": (1, 1, 0, 1 << 30),
                 "chunked_latency": (0, 0, 1, 1 << 30), "chain_walk_throughput": (0, 0, 0, 0)}

        def use(name):
            blk, waves, chunk, deep = kinds[name]
            dt.set_tuning("query_blk", blk)
            dt.set_tuning("query_blk_waves", waves)
            dt.set_tuning("query_chunk", chunk)
            dt.set_tuning("query_deep_max", deep)
        forms = {}
        for name in kinds:
            use(name)
            dt.query(ud, qs, qt, want_scanned=False, dist=dist, order=order)
            ms = statistics.median(event_ms(lambda: dt.query(ud, qs, qt, want_scanned=False, dist=dist,
                                                             order=order), 2))
            forms[name] = {"batch_ms": ms, "queries_per_s": nq / (ms / 1e3)}
            log(f"config {args.graph} queries, {name}: {nq} in {ms:.1f} ms")
        # the metric split once into the reference's separate up / down arrays
        # (timed apart), then the blocked kernel on it with the fastest
        # wavefront count of the pair form
        best_blk = min(("blocked_1", "blocked_2", "blocked_4"), key=lambda k: forms[k]["batch_ms"])
        use(best_blk)
        up_s, dn_s = dt.split(ud)
        split_ms = statistics.median(event_ms(lambda: dt.split(ud, out=(up_s, dn_s)), 3))
        dt.query_split(up_s, dn_s, qs, qt, want_scanned=False, dist=dist, order=order)
        ms = statistics.median(event_ms(lambda: dt.query_split(up_s, dn_s, qs, qt, want_scanned=False, dist=dist,
                                                               order=order), 2))
        forms["split"] = {"batch_ms": ms, "queries_per_s": nq / (ms / 1e3), "split_ms": split_ms,
                          "wavefronts_per_side": kinds[best_blk][1]}
        log(f"config {args.graph} queries, split ({best_blk}): {nq} in {ms:.1f} ms (metric split {split_ms:.3f} ms)")
        # one query per launch: the blocked form with 8 wavefronts per side (its
        # default for launches of up to 256 queries) with 8 or 16 64-arc chunks
        # per wavefront in flight, the same on the split metric, and the older
        # chunked form: (query_blk, query_chunk, query_blk_unroll, split)
        single_forms = {"blocked_8": (1, 0, 8, False), "blocked_8_u16": (1, 0, 16, False),
                        "split_8_u16": (1, 0, 16, True), "chunked_latency": (0, 1, 8, False)}

        def use_single(name):
            blk, chunk, unroll, split = single_forms[name]
            dt.set_tuning("query_blk", blk)
            dt.set_tuning("query_blk_waves", 0)
            dt.set_tuning("query_chunk", chunk)
            dt.set_tuning("query_deep_max", 1 << 30)
            dt.set_tuning("query_blk_unroll", unroll)
            if split:
                return lambda i: dt.query_split(up_s, dn_s, qs[i:i + 1], qt[i:i + 1], want_scanned=False)
            return lambda i: dt.query(ud, qs[i:i + 1], qt[i:i + 1], want_scanned=False)
        singles = {}
        for name in single_forms:
            f1 = use_single(name)
            f1(0)
            one = [event_ms(lambda i=i: f1(i), 1)[0] for i in range(30)]
            singles[name] = float(np.median(one))
            log(f"config {args.graph} single query, {name}: p50 {np.median(one):.2f} ms")
        best_b = min(forms, key=lambda k: forms[k]["batch_ms"])
        best_s = min(singles, key=lambda k: singles[k])
        use(best_blk if best_b == "split" else best_b)
        if best_b == "split":
            dist_full, scanned = dt.query_split(up_s, dn_s, qs, qt, order=order)
        else:
            dist_full, scanned = dt.query(ud, qs, qt, order=order)
        qms = forms[best_b]["batch_ms"]
        vdep = topo.export()["depth"].astype(np.int64)
        chain = vdep[rank[qs_h]] + vdep[rank[qt_h]] + 2
        qbytes = 8 * int(scanned.to(torch.int64).sum().item()) + 12 * int(chain.sum()) + 12 * nq
        f1 = use_single(best_s)
        single = [event_ms(lambda i=i: f1(i), 1)[0] for i in range(args.single_full)]
        dt.set_tuning("query_blk_unroll", 8)
        dt.check_status()
        queries = {"count": nq, "batch_ms": qms, "queries_per_s": nq / (qms / 1e3), "batch_kernel": best_b,
                   "roofline": {"bound": "hbm", "achieved": qbytes / (qms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": qbytes / (qms / 1e3) / 1e9 / HBM_PEAK_GBS,
                                "bytes_per_query": qbytes / nq,
                                "what": "8 B per scanned arc + 12 B per chain vertex + 12 B per query"},
                   "single_kernel": best_s, "single_launches": len(single),
                   "single_p50_ms": float(np.percentile(single, 50)),
                   "single_p99_ms": float(np.percentile(single, 99)), "forms": forms,
                   "single_p50_by_form_ms": singles}
        log(f"config {args.graph} queries: {nq} in {qms:.1f} ms ({nq / (qms / 1e3):.3g} q/s); single p50 "
            f"{np.percentile(single, 50):.2f} ms p99 {np.percentile(single, 99):.2f} ms")

    # ---- parity (the checker runs outside every timed region)
    parity = {}
    if not args.no_check:
        # every batched metric equals its single-metric customization
        same_b = True
        for k in range(K):
            dt.customize(dt.weights(W[k]), out=ud)
            same_b &= bool(torch.equal(dt.metric_view(UD, k), ud))
        parity["batched_equals_single"] = same_b
        if chk is not None:
            # the checker's answers are on the graph's own metric
            cs, ct, want = chk
            dt.customize(dt.weights(g.w), out=ud)
            d, _ = dt.query(ud, rank[cs], rank[ct], want_scanned=False)
            parity["queries_vs_johnson_dijkstra"] = {
                "pairs": len(cs), "equal": bool(np.array_equal(d.cpu().numpy().astype(np.int64), want))}
        dt.customize(w0, out=ud)
        dt.check_status()
    got = ud.cpu().numpy()
    cpu = None
    from oracle import oracle as O
    if (not args.no_cpu or not args.no_check) and O.have_ref():
        t0 = time.perf_counter()
        ref = O.Reference(g.n, g.off, g.tgt, rank)
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        up, dn = ref.customize(W[0])
        t_ser = time.perf_counter() - t0
        ok = True
        for col, r in ((0, up), (1, dn)):
            ok &= bool(np.array_equal(got[:, col].astype(np.int64), O.float_to_i32(r).astype(np.int64)))
        parity["customization_vs_reference_cpp"] = {"arcs": topo.mp, "equal": ok}
        log(f"config {args.graph} reference C++: structure {t_build:.1f}s, serial customization {t_ser:.2f}s; "
            f"all {topo.mp} arcs equal: {ok}")
        if not args.no_cpu:
            thr = min(16, os.cpu_count() or 1)
            t0 = time.perf_counter()
            ref.customize(W[0], threads=thr)
            t_par = time.perf_counter() - t0
            nqc = min(args.cpu_queries, args.full_queries) if args.full_queries else 0
            qres = None
            if nqc:
                ref.query(up, dn, qs_h[:50], qt_h[:50])   # warm-up
                t0 = time.perf_counter()
                dq = ref.query(up, dn, qs_h[:nqc], qt_h[:nqc])
                t_q1 = time.perf_counter() - t0
                t0 = time.perf_counter()
                dqp = ref.query_par(up, dn, qs_h[:nqc * 4], qt_h[:nqc * 4], thr)
                t_qp = time.perf_counter() - t0
                eq = bool(np.array_equal(O.float_to_i32(dq), dist_full[:nqc].cpu().numpy()))
                eqp = bool(np.array_equal(O.float_to_i32(dqp), dist_full[:nqc * 4].cpu().numpy()))
                parity["queries_vs_reference_cpp"] = {"pairs": nqc * 4, "equal": eq and eqp}
                qres = {"serial": {"value": nqc / t_q1, "unit": "queries/s", "cores": 1, "kind": "reference",
                                   "sample": f"cpp/cch.cpp Query::run, {nqc} random pairs of the GPU batch"},
                        "parallel": {"value": 4 * nqc / t_qp, "unit": "queries/s", "cores": thr,
                                     "kind": "reference",
                                     "sample": f"Query::run on {thr} OpenMP threads (one Query each), "
                                               f"{4 * nqc} pairs"}}
                log(f"config {args.graph} reference queries: {nqc / t_q1:.0f} q/s on 1 core, {4 * nqc / t_qp:.0f} q/s on "
                    f"{thr}; equal to the GPU's: {eq and eqp}")
            cpu = {"value": topo.triangles / t_ser, "unit": "triangles/s", "cores": 1, "kind": "reference",
                   "sample": f"cpp/cch.cpp CCH::customize (serial, float64) of metric 0 of the same graph and "
                             f"order, one full customization ({t_ser:.2f} s)",
                   "ms": 1e3 * t_ser, "cpu_model": cpu_model(),
                   "parallel": {"value": topo.triangles / t_par, "cores": thr, "ms": 1e3 * t_par,
                                "sample": "cpp/cch.cpp CCH::customize_parallel (OpenMP, level buckets), one "
                                          "full customization"},
                   "queries": qres, "reference_structure_s": t_build}
            log(f"config {args.graph} reference C++ on {thr} threads: {t_par:.2f}s")
        del ref, up, dn
    if parity:
        log(f"config {args.graph} parity: {parity}")

    out.update({
        "value": K * topo.triangles / t_step,
        "ms_per_step": 1000.0 * t_step,
        "config": {"workload": f"road-like graph, config {args.graph}: {graphs.CONFIGS[args.graph].get('n_points', 0):,} "
                               f"clustered points, seed {graphs.CONFIGS[args.graph].get('seed')}, n={g.n} m={g.m} "
                               f"m+={topo.mp} T={topo.triangles} ET depth={topo.depth}; {K} metrics customized "
                               f"in one level sweep + per-metric negative-cycle scans per step",
                   "n": g.n, "m": g.m, "m_plus": topo.mp, "triangles": topo.triangles, "depth": topo.depth,
                   "levels": topo.levels, "metrics": K, "negative_arc_fraction": round(g.negative_fraction(), 4)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_vs_algorithmic": traffic["vs_algorithmic"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "kernel": "k_customize_batch_level", "level_launches": n_lv,
                     "level_kernels": "k_customize_batch_level (triangle list) and k_rows_batch (row-merge "
                                      "form, levels of 8-40 rows); traffic counted on the triangle-list kernel",
                     "algorithmic_bytes_per_launch": B8d_lv / n_lv, "avg_launch_us": lv_avg_us,
                     "what": "algorithmic bytes 4 T + 4 (n+1) + 4 m+ + 16 K m+ over the level launches "
                             "(c-index stream, first/head, every (up, down) pair read and written once); "
                             "average launch = (customization - base pass) / level launches, HIP events",
                     "layout_bytes_per_launch": layout_level_bytes(topo.mp, topo.triangles, K) / n_lv,
                     "customization_ms": bms, "base_pass_ms": bbase_ms,
                     "customization_algorithmic_bytes": B8d,
                     "customization_frac": B8d / (bms / 1e3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "speedup_vs_cpu_baseline": (K * topo.triangles / t_step) / cpu["value"] if cpu else None,
        "step_ms": {"hip_graph": 1e3 * t_step, "eager": 1e3 * t_eager, "batched_scan_ms": bscan_ms},
        "flagged_metrics": flagged,
        "plain_one_metric": {"step_ms": 1e3 * t_plain, "customize_ms": cust1_ms, "scan_ms": scan1_ms,
                             "triangles_per_s": topo.triangles / (cust1_ms / 1e3),
                             "batched_speedup_per_metric": (K * cust1_ms) / bms,
                             "roofline": {"kernel": "k_customize_flat", "achieved": ach1, "frac": ach1 / HBM_PEAK_GBS,
                                          "avg_launch_us": lv1_us, "level_launches": n_lv1,
                                          "base_pass_ms": base1_ms}},
        "queries": queries,
        "parity": parity,
        "host": dict(host, nproc=os.cpu_count()),
    })
    del dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=4,
                    help="4 (default): the full-scale headline; 1-3: the plain, batched and serving workloads")
    ap.add_argument("--cache", default=None, help="config 4: read/write graph, order and metrics to this .npz")
    ap.add_argument("--graph", type=int, default=4, help="config 4 mode on another config's graph (a quick run)")
    ap.add_argument("--single-full", type=int, default=200, help="config 4: single-query launches for p50/p99")
    ap.add_argument("--cpu-queries", type=int, default=2000, help="config 4: reference queries timed on 1 core")
    ap.add_argument("--queries", type=int, default=100_000)
    ap.add_argument("--single", type=int, default=500, help="single-query launches for latency")
    ap.add_argument("--metrics", type=int, default=None, help="metrics per batch (8 at config 4, 64 at config 2)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (scales HBM byte counters)")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP graph replay of the step")
    ap.add_argument("--counter-run", action="store_true",
                    help="config 4: a short run of the sweep and query kernels for rocprofv3 --pmc passes")
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
    if args.metrics is None:
        args.metrics = 8 if args.config == 4 else 64
    if args.prep:
        prep(args, out)
        print(json.dumps(out), flush=True)
        return
    if args.counter_run:
        counter_run(args, out)
        print(json.dumps(out), flush=True)
        return
    if args.config == 4:
        headline(args, out)
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
