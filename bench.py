"""Benchmark of the metric-dependent CCH path on one MI355X.

Workload: the road-like graph of 264,000 clustered points (seed 1, Delaunay
thinned to mean out-degree 2.8, 5% one-way, largest SCC), int32 energy
weights, inertial-flow order (host, untimed).  One step = full customization of
one metric + the negative-cycle scan (the plain one-metric form), with inputs
resident in HBM.  The same run times a batch of 100,000 random s-t ET queries
(one launch) and 1,000 single-query launches (p50/p99 latency), checks the
results bit-for-bit against the CPU oracle, and times the reference's own
C++ customization (or the oracle port) on the host cores as the CPU baseline.

Prints ONE JSON line (rank 0).
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


def cust_bytes(n, m, mp, T, mapped):
    """Compulsory HBM bytes of one customization in the pull layout: the (a, b)
    triangle pairs (8 B each), per-arc schedule/offset records (8 B tri_off +
    4 B sched_arc + 4 B ae_off), the arc->edge list (4 B per mapped arc), the
    weights (4 B per arc), and each (up, dn) pair read once and written once."""
    return 8 * T + 16 * (mp + 1) + 4 * mapped + 4 * m + 16 * mp


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--queries", type=int, default=100_000)
    ap.add_argument("--single", type=int, default=1000, help="single-query launches for latency")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design is "
              "described in DESIGN.md and is not built", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from nwcch import _lib, graphs
    from nwcch.device import DeviceTopology

    t0 = time.perf_counter()
    g = graphs.make_config(args.config)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    t_order = time.perf_counter() - t0
    t0 = time.perf_counter()
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    t_topo = time.perf_counter() - t0
    log(f"graph n={g.n} m={g.m} neg={g.negative_fraction():.3f} | m+={topo.mp} T={topo.triangles} "
        f"depth={topo.depth} levels={topo.levels} | gen {t_gen:.1f}s order {t_order:.1f}s topo {t_topo:.1f}s")

    dt = DeviceTopology(topo, 0)
    w = dt.weights(g.w)
    ud = torch.empty((topo.mp, 2), dtype=torch.int32, device=dt.dev)
    scan = torch.empty(2, dtype=torch.int64, device=dt.dev)
    qs_h, qt_h = graphs.random_pairs(g.n, args.queries, seed=7)
    qs = torch.as_tensor(rank[qs_h]).to(dt.dev)
    qt = torch.as_tensor(rank[qt_h]).to(dt.dev)
    dist = torch.empty(args.queries, dtype=torch.int32, device=dt.dev)

    def step():
        dt.customize(w, out=ud)
        dt.negcycle_scan(ud, out=scan)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    dt.check_status()

    # ---- timed region: K customization + scan steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    ms_per_step = 1000.0 * t_steps / args.steps
    tri_per_s = topo.triangles / (t_steps / args.steps)

    # ---- the same step replayed as one HIP graph (level launches captured once)
    from nwcch.device import StepGraph
    graph = StepGraph(step)
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    ms_graph = 1000.0 * (time.perf_counter() - t0) / args.steps
    dt.check_status()
    eager_ms = ms_per_step
    if ms_graph < ms_per_step:
        ms_per_step = ms_graph
        tri_per_s = topo.triangles / (ms_graph / 1000.0)

    # ---- per-launch kernel durations with HIP events over the same K steps
    dt.profile_begin(args.steps * (topo.levels + 4))
    for _ in range(args.steps):
        step()
    kid, kms = dt.profile_end()
    cust_ms = kms[kid == 0]
    launches_per_cust = len(cust_ms) / max(args.steps, 1)
    cust_kernel_ms = float(cust_ms.sum()) / args.steps          # sum of level-kernel durations
    avg_launch_us = 1000.0 * float(cust_ms.mean()) if len(cust_ms) else float("nan")
    sev0, sev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    sev0.record()
    for _ in range(args.steps):
        dt.negcycle_scan(ud, out=scan)
    sev1.record()
    torch.cuda.synchronize()
    scan_ms = sev0.elapsed_time(sev1) / args.steps

    B = cust_bytes(g.n, g.m, topo.mp, topo.triangles, topo.n_mapped)
    achieved = B / (cust_kernel_ms / 1000.0) / 1e9
    scan_gbs = 8 * topo.mp / (scan_ms / 1000.0) / 1e9

    # ---- ET queries: one batch launch, and single-query launches
    for _ in range(2):
        dt.query(ud, qs, qt, want_scanned=False)
    torch.cuda.synchronize()
    qreps = 5
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(qreps):
        d_b, _ = dt.query(ud, qs, qt, want_scanned=False)
    e1.record()
    torch.cuda.synchronize()
    batch_ms = e0.elapsed_time(e1) / qreps
    qps_batch = args.queries / (batch_ms / 1000.0)
    lat = []
    for i in range(args.single):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        dt.query(ud, qs[i:i + 1], qt[i:i + 1], want_scanned=False)
        b.record()
        lat.append((a, b))
    torch.cuda.synchronize()
    lat_ms = np.array([a.elapsed_time(b) for a, b in lat])
    t0 = time.perf_counter()
    for i in range(args.single):
        dt.query(ud, qs[i:i + 1], qt[i:i + 1], want_scanned=False)
    torch.cuda.synchronize()
    qps_single = args.single / (time.perf_counter() - t0)
    dt.check_status()

    # ---- parity against the CPU oracle (untimed)
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
        ok &= np.array_equal(d_b[:nchk].cpu().numpy(), d0)
        parity = bool(ok)
        log(f"parity vs oracle: {parity}")

    # ---- CPU baseline on the host cores (bounded sample: full customization x 5)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        reps = 5
        if O.have_ref():
            ref = O.Reference(g.n, g.off, g.tgt, rank)
            ref.customize(g.w)
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                ref.customize(g.w)
                ts.append(time.perf_counter() - t0)
            kind = "reference"
            sample = (f"cpp/cch.cpp CCH::customize (serial, float64), same graph and order, "
                      f"median of {reps} full customizations")
        else:
            st = O.Structure(g.n, g.off, g.tgt, rank)
            st.customize(g.w)
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                st.customize(g.w)
                ts.append(time.perf_counter() - t0)
            kind = "port"
            sample = (f"oracle restatement of cch_nb.customize (serial, int32), same graph and order, "
                      f"median of {reps} full customizations")
        tmed = statistics.median(ts)
        cpu = {"value": topo.triangles / tmed, "unit": "triangles/s", "cores": 1, "kind": kind,
               "sample": sample, "ms": 1000 * tmed}
        log(f"cpu baseline ({kind}): {1000 * tmed:.2f} ms per customization")

    out = {
        "metric": "triangles/s customize; ET queries/s + p99 latency, batched/served vs plain GPU",
        "value": tri_per_s,
        "unit": "triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"road-like graph config {args.config}: n={g.n} m={g.m} m+={topo.mp} "
                               f"T={topo.triangles} depth={topo.depth}; 1 metric customize + "
                               f"negative-cycle scan; {args.queries} ET queries",
                   "n": g.n, "m": g.m, "m_plus": topo.mp, "triangles": topo.triangles,
                   "depth": topo.depth, "levels": topo.levels, "metrics": 1,
                   "negative_arc_fraction": round(g.negative_fraction(), 4)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_customize_level", "algorithmic_bytes_per_customization": B,
                     "launches_per_customization": launches_per_cust,
                     "avg_launch_us": avg_launch_us, "kernel_ms_per_customization": cust_kernel_ms},
        "cpu_baseline": cpu,
        "queries": {"batch": args.queries, "batch_ms": batch_ms, "queries_per_s": qps_batch,
                    "p99_ms_batch": batch_ms, "single_launch_queries_per_s": qps_single,
                    "single_p50_ms": float(np.percentile(lat_ms, 50)),
                    "single_p99_ms": float(np.percentile(lat_ms, 99))},
        "step_ms": {"eager": eager_ms, "hip_graph": ms_graph},
        "scan": {"ms": scan_ms, "GB_s": scan_gbs, "frac": scan_gbs / HBM_PEAK_GBS},
        "parity_vs_oracle": parity,
        "host": {"order_s": t_order, "topology_s": t_topo, "nproc": os.cpu_count()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
