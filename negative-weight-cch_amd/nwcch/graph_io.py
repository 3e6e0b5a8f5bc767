"""The reference's on-disk graph format, read and written.

A graph directory holds (python/export_graph.py:22-52, read by
cpp/cch.cpp:72-95):

* ``meta.txt``: n on the first line, m on the second, the metric names
  separated by spaces on the third;
* ``off.i64`` (n + 1 int64), ``tgt.i32`` (m int32): the directed CSR;
* ``x.f64``, ``y.f64`` (n float64 each): coordinates, for the ordering;
* ``z.f64`` (optional, n float64): elevations;
* ``w_<metric>.f64`` (m float64 each): the weights of each metric.

An order file is the raw int32 array rank[v] (cpp/cch.cpp:766-777).

The reference's weights are float64.  The path computes in int32, so
:func:`int_weights` converts them: integer values pass unchanged, ``inf`` becomes
the INF sentinel, and fractional values are refused unless a fixed-point scale
is given.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from ._lib import INF
from .graphs import RoadGraph


def read_meta(d) -> tuple[int, int, list[str]]:
    lines = (Path(d) / "meta.txt").read_text().split("\n")
    n, m = int(lines[0]), int(lines[1])
    names = lines[2].split() if len(lines) > 2 else []
    return n, m, names


def _read(path, dtype, count):
    a = np.fromfile(path, dtype=dtype)
    if a.shape[0] != count:
        raise ValueError(f"{path}: {a.shape[0]} values, expected {count}")
    return a


def read_weights(d, metric: str) -> np.ndarray:
    """The float64 weights of one metric, in CSR arc order."""
    n, m, _ = read_meta(d)
    return _read(Path(d) / f"w_{metric}.f64", np.float64, m)


def int_weights(w, scale: float | None = None) -> np.ndarray:
    """Reference float64 weights -> int32 with the INF sentinel.  Without a
    scale every finite weight must be an integer; with one, weights are
    multiplied by it and rounded to the nearest integer."""
    w = np.asarray(w, np.float64)
    fin = np.isfinite(w)
    if np.isnan(w).any() or (np.isinf(w) & (w < 0)).any():
        raise ValueError("weights must be finite numbers or +inf")
    v = w[fin] * scale if scale is not None else w[fin]
    if scale is None and not np.array_equal(v, np.round(v)):
        raise ValueError("fractional weights: pass a fixed-point scale")
    v = np.round(v)
    if v.size and (v.min() <= -INF or v.max() >= INF):
        raise ValueError("weights do not fit in int32")
    out = np.full(w.shape, INF, np.int64)
    out[fin] = v.astype(np.int64)
    return out.astype(np.int32)


def read_graph_dir(d, metric: str | None = None, scale: float | None = None) -> RoadGraph:
    """A RoadGraph from a graph directory, with the int32 weights of `metric`
    (the first metric listed if None; zero weights if the directory has none)."""
    d = Path(d)
    n, m, names = read_meta(d)
    off = _read(d / "off.i64", np.int64, n + 1)
    tgt = _read(d / "tgt.i32", np.int32, m)
    if off[0] != 0 or off[-1] != m or (np.diff(off) < 0).any():
        raise ValueError(f"{d}: off.i64 is not a CSR offset array for m = {m}")
    if m and (tgt.min() < 0 or tgt.max() >= n):
        raise ValueError(f"{d}: tgt.i32 holds vertex ids outside [0, n)")
    x = _read(d / "x.f64", np.float64, n)
    y = _read(d / "y.f64", np.float64, n)
    z = np.fromfile(d / "z.f64", np.float64) if (d / "z.f64").exists() else None
    if metric is None and names:
        metric = names[0]
    w = int_weights(read_weights(d, metric), scale) if metric else np.zeros(m, np.int32)
    g = RoadGraph(n, off, tgt, w, x, y, None, name=d.name)
    g.extra.update(dict(metrics=names, metric=metric, z=z))
    return g


def write_graph_dir(d, g: RoadGraph, metrics: dict | None = None, z=None) -> Path:
    """Write a graph directory; `metrics` maps names to weight arrays (int32
    with the INF sentinel, or float64), default {"w": g.w}."""
    d = Path(d)
    d.mkdir(parents=True, exist_ok=True)
    metrics = {"w": g.w} if metrics is None else metrics
    np.asarray(g.off, np.int64).tofile(d / "off.i64")
    np.asarray(g.tgt, np.int32).tofile(d / "tgt.i32")
    np.asarray(g.x, np.float64).tofile(d / "x.f64")
    np.asarray(g.y, np.float64).tofile(d / "y.f64")
    if z is not None:
        np.asarray(z, np.float64).tofile(d / "z.f64")
    for name, w in metrics.items():
        w = np.asarray(w)
        if w.shape != (g.m,):
            raise ValueError(f"metric {name}: {w.shape[0]} weights, expected {g.m}")
        wf = w.astype(np.float64)
        if np.issubdtype(w.dtype, np.integer):
            wf[w == INF] = np.inf
        wf.tofile(d / f"w_{name}.f64")
    (d / "meta.txt").write_text(f"{g.n}\n{g.m}\n{' '.join(metrics)}\n")
    return d


def read_order(path, n: int | None = None) -> np.ndarray:
    rank = np.fromfile(path, np.int32)
    if n is not None and rank.shape[0] != n:
        raise ValueError(f"{path}: {rank.shape[0]} ranks, expected {n}")
    if not np.array_equal(np.sort(rank), np.arange(rank.shape[0], dtype=np.int32)):
        raise ValueError(f"{path}: not a permutation")
    return rank


def write_order(path, rank) -> None:
    np.asarray(rank, np.int32).tofile(path)
