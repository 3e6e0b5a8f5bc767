"""Device side of the path: a topology resident in HBM and the HIP entry points
(customize, negative-cycle scan, ET queries) called through the C ABI.

PyTorch provides the device memory and the stream; every computation is one of
libnwcch's kernels.  Nothing here falls back to the CPU: without a GPU or
without the built library the calls raise.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import INF, NwcchError, check, lib, ptr


def _stream(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def require_gpu():
    if not torch.cuda.is_available():
        raise NwcchError("no GPU: the MI355X path needs a HIP device (there is no CPU fallback)")


def as_pairs(up, dn) -> torch.Tensor:
    """The [mp, 2] int32 tensor behind (up, dn): the storage itself when up/dn
    are the two columns of one such tensor (what customize returns), else an
    interleaved device copy."""
    if (isinstance(up, torch.Tensor) and isinstance(dn, torch.Tensor) and up.is_cuda
            and up.dtype == torch.int32 and dn.dtype == torch.int32 and up.dim() == 1
            and up.stride() == (2,) and dn.stride() == (2,)
            and dn.data_ptr() == up.data_ptr() + 4 and up.untyped_storage().data_ptr() == dn.untyped_storage().data_ptr()):
        base = up.as_strided((up.shape[0], 2), (2, 1))
        return base
    dev = torch.device("cuda")
    u = torch.as_tensor(up).to(dev, torch.int32)
    d = torch.as_tensor(dn).to(dev, torch.int32)
    return torch.stack([u, d], 1).contiguous()


class StepGraph:
    """A HIP graph of a fixed sequence of library calls on fixed device buffers
    (e.g. the level launches of one customization plus the scan), captured once
    and replayed with one launch.  `fn` must only enqueue work on the current
    stream; the library records no memset nodes."""

    def __init__(self, fn):
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            fn()  # warm-up outside the capture
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            fn()

    def replay(self):
        self.graph.replay()


class DeviceTopology:
    """One topology uploaded to one GPU (owns the device copy)."""

    def __init__(self, topo: _lib.Topology, device: int | None = None):
        require_gpu()
        self.topo = topo
        self.device = torch.cuda.current_device() if device is None else int(device)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().nwcch_dev_create(topo._h, self.device, C.byref(h)), "nwcch_dev_create")
        self._h = h
        self.n, self.m, self.mp = topo.n, topo.m, topo.mp
        self.triangles, self.depth, self.levels = topo.triangles, topo.depth, topo.levels
        self._scan_out = torch.empty(2, dtype=torch.int64, device=self.dev)

    @property
    def dev(self):
        return torch.device("cuda", self.device)

    # ---------------------------------------------------------------- metric
    def weights(self, w) -> torch.Tensor:
        w = torch.as_tensor(w)
        if w.dtype.is_floating_point:
            wf = w.double()
            fin = torch.isfinite(wf)
            if not torch.equal(wf[fin], torch.round(wf[fin])):
                raise ValueError("weights must be integers (the path computes in int32)")
            w = torch.where(fin, wf, torch.full_like(wf, INF)).to(torch.int64)
        w = w.to(torch.int64)
        if w.numel() and (int(w.min()) <= -INF or int(w.max()) > INF):
            raise ValueError("weights must lie in (-2^31+1, 2^31-1]; 2^31-1 means no arc")
        if w.shape != (self.m,):
            raise ValueError(f"expected {self.m} original arc weights, got {tuple(w.shape)}")
        return w.to(self.dev, torch.int32).contiguous()

    def customize(self, w, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """Full customization -> ud [mp, 2] int32 (up, down)."""
        w = self.weights(w) if not (isinstance(w, torch.Tensor) and w.is_cuda and w.dtype == torch.int32) else w
        if out is None:
            out = torch.empty((self.mp, 2), dtype=torch.int32, device=self.dev)
        assert out.shape == (self.mp, 2) and out.dtype == torch.int32 and out.is_contiguous()
        check(lib().nwcch_customize(self._h, ptr(w), ptr(out), _stream(stream)), "nwcch_customize")
        return out

    def stop_state(self, w: torch.Tensor, ud_full: torch.Tensor, x_stop: int, out: torch.Tensor | None = None,
                   stream=None) -> torch.Tensor:
        """The values python/cch.py's early-stopping customize leaves when it
        stops at the row of rank x_stop (-1: base values), from the complete
        customization ud_full of the device weights w."""
        if out is None:
            out = torch.empty_like(ud_full)
        assert out.shape == (self.mp, 2) and ud_full.shape == (self.mp, 2) and out.data_ptr() != ud_full.data_ptr()
        check(lib().nwcch_customize_stop_state(self._h, ptr(w), ptr(ud_full), int(x_stop), ptr(out),
                                               _stream(stream)), "nwcch_customize_stop_state")
        return out

    def negcycle_scan(self, ud: torch.Tensor, out: torch.Tensor | None = None, stream=None):
        """Device int64[2] = (count, first witness or -1); not synchronised."""
        out = self._scan_out if out is None else out
        check(lib().nwcch_negcycle_scan(ptr(ud), ud.shape[0], ptr(out), _stream(stream)), "nwcch_negcycle_scan")
        return out

    def query_path(self, ud: torch.Tensor, s_rank, t_rank, max_len: int = 4096, stream=None):
        """Distances and unpacked paths (ranks) -> (dist int32 [nq], path_len int32
        [nq], path int32 [nq, max_len]); device tensors, not synchronised."""
        assert ud.dtype == torch.int32 and ud.shape == (self.mp, 2) and ud.is_contiguous()
        qs = torch.as_tensor(s_rank).to(self.dev, torch.int32).contiguous()
        qt = torch.as_tensor(t_rank).to(self.dev, torch.int32).contiguous()
        nq = qs.shape[0]
        dist = torch.empty(nq, dtype=torch.int32, device=self.dev)
        plen = torch.empty(nq, dtype=torch.int32, device=self.dev)
        path = torch.empty((nq, max_len), dtype=torch.int32, device=self.dev)
        check(lib().nwcch_query_path(self._h, ptr(ud), ptr(qs), ptr(qt), nq, int(max_len), ptr(dist), ptr(plen),
                                     ptr(path), _stream(stream)), "nwcch_query_path")
        return dist, plen, path

    def potential(self, ud: torch.Tensor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """CCH potential indexed by rank: device int32 [n] (cpp/cch.cpp:635-655)."""
        assert ud.dtype == torch.int32 and ud.shape == (self.mp, 2) and ud.is_contiguous()
        if out is None:
            out = torch.empty(self.n, dtype=torch.int32, device=self.dev)
        check(lib().nwcch_potential(self._h, ptr(ud), ptr(out), _stream(stream)), "nwcch_potential")
        return out

    def negcycle_list(self, ud: torch.Tensor, stream=None) -> list[int]:
        """All witness arcs, ascending (python/cch.py:287)."""
        cnt, _ = (int(v) for v in self.negcycle_scan(ud, stream=stream).cpu())
        if cnt == 0:
            return []
        lst = torch.empty(cnt, dtype=torch.int64, device=self.dev)
        num = torch.empty(1, dtype=torch.int64, device=self.dev)
        check(lib().nwcch_negcycle_list(ptr(ud), ud.shape[0], ptr(lst), cnt, ptr(num), _stream(stream)),
              "nwcch_negcycle_list")
        return sorted(lst.cpu().tolist())

    def query(self, ud: torch.Tensor, s_rank, t_rank, want_scanned=True, stream=None, metric: int = 0,
              dist: torch.Tensor | None = None, order: torch.Tensor | None = None):
        """ET queries between ranks -> (dist int32 [nq], scanned int32 [nq] or None).
        ud is one metric [mp, 2] or a batched result [slabs, mp, SW, 2] (then
        `metric` picks one).  `order` (device int32 permutation) packs the
        queries: wavefront i serves query order[i]."""
        qs = torch.as_tensor(s_rank).to(self.dev, torch.int32).contiguous()
        qt = torch.as_tensor(t_rank).to(self.dev, torch.int32).contiguous()
        if qs.shape != qt.shape or qs.dim() != 1:
            raise ValueError("s_rank and t_rank must be 1-d of equal length")
        if order is not None:
            assert order.dtype == torch.int32 and order.is_cuda and order.shape == qs.shape
        if ud.dim() == 4:
            slabs, _, SW, _ = ud.shape
            if not 0 <= metric < slabs * SW:
                raise ValueError("metric out of range")
            assert ud.shape == (slabs, self.mp, SW, 2) and ud.is_contiguous()
            base = ud.data_ptr() + 8 * ((metric // SW) * self.mp * SW + metric % SW)
            stride = SW
        else:
            assert ud.shape == (self.mp, 2) and ud.is_contiguous()
            base, stride = ud.data_ptr(), 1
        nq = qs.shape[0]
        if dist is None:
            dist = torch.empty(nq, dtype=torch.int32, device=self.dev)
        sc = torch.empty(nq, dtype=torch.int32, device=self.dev) if want_scanned else None
        check(lib().nwcch_query_strided(self._h, base, stride, ptr(qs), ptr(qt), nq, ptr(order), ptr(dist),
                                        ptr(sc), _stream(stream)), "nwcch_query_strided")
        return dist, sc

    def split(self, ud: torch.Tensor, out: tuple[torch.Tensor, torch.Tensor] | None = None, stream=None):
        """One metric's (up, down) pairs [mp, 2] as separate (up [mp], dn [mp])
        arrays, the reference's layout (nwcch_metric_split)."""
        assert ud.shape == (self.mp, 2) and ud.is_contiguous()
        up, dn = out if out is not None else (torch.empty(self.mp, dtype=torch.int32, device=self.dev),
                                               torch.empty(self.mp, dtype=torch.int32, device=self.dev))
        check(lib().nwcch_metric_split(ud.data_ptr(), self.mp, ptr(up), ptr(dn), _stream(stream)),
              "nwcch_metric_split")
        return up, dn

    def query_split(self, up: torch.Tensor, dn: torch.Tensor, s_rank, t_rank, want_scanned=True, stream=None,
                    dist: torch.Tensor | None = None, order: torch.Tensor | None = None):
        """ET queries on a metric given as separate up / down arrays
        (nwcch_query_split) -> (dist, scanned or None), as `query`."""
        qs = torch.as_tensor(s_rank).to(self.dev, torch.int32).contiguous()
        qt = torch.as_tensor(t_rank).to(self.dev, torch.int32).contiguous()
        if qs.shape != qt.shape or qs.dim() != 1:
            raise ValueError("s_rank and t_rank must be 1-d of equal length")
        assert up.shape == (self.mp,) and dn.shape == (self.mp,) and up.dtype == dn.dtype == torch.int32
        if order is not None:
            assert order.dtype == torch.int32 and order.is_cuda and order.shape == qs.shape
        nq = qs.shape[0]
        if dist is None:
            dist = torch.empty(nq, dtype=torch.int32, device=self.dev)
        sc = torch.empty(nq, dtype=torch.int32, device=self.dev) if want_scanned else None
        check(lib().nwcch_query_split(self._h, ptr(up), ptr(dn), ptr(qs), ptr(qt), nq, ptr(order), ptr(dist),
                                      ptr(sc), _stream(stream)), "nwcch_query_split")
        return dist, sc

    def query_raw(self, ud: torch.Tensor, qs: torch.Tensor, qt: torch.Tensor, n: int, dist: torch.Tensor,
                  stream) -> None:
        """The first n queries of device int32 rank arrays qs/qt on a one-metric ud,
        into dist, with no argument conversion (the serving loop's launch path)."""
        check(lib().nwcch_query_strided(self._h, ud.data_ptr(), 1, qs.data_ptr(), qt.data_ptr(), n, None,
                                        dist.data_ptr(), None, stream.cuda_stream), "nwcch_query_strided")

    def query_multi(self, table: torch.Tensor, qmetric: torch.Tensor, qs: torch.Tensor, qt: torch.Tensor, n: int,
                    dist: torch.Tensor, stream=None) -> None:
        """The first n queries, query i on the metric whose ud address is
        table[qmetric[i]] (device int64 table, int32 metric indices)."""
        assert table.dtype == torch.int64 and qmetric.dtype == torch.int32
        check(lib().nwcch_query_multi(self._h, table.data_ptr(), table.numel(), qmetric.data_ptr(), qs.data_ptr(),
                                      qt.data_ptr(), n, dist.data_ptr(), None, _stream(stream)),
              "nwcch_query_multi")

    def chain_order(self, s_rank, t_rank) -> torch.Tensor:
        """Query packing: the permutation serving the longest ancestor chains
        (depth(s) + depth(t)) first, so the longest queries start earliest."""
        if not hasattr(self, "_depth"):
            self._depth = torch.as_tensor(self.topo.export()["depth"]).to(self.dev)
        qs = torch.as_tensor(s_rank).to(self.dev, torch.long)
        qt = torch.as_tensor(t_rank).to(self.dev, torch.long)
        key = self._depth[qs] + self._depth[qt]
        return torch.argsort(key, descending=True).to(torch.int32)

    def set_tuning(self, key: str, value: int):
        check(lib().nwcch_set_tuning(self._h, key.encode(), int(value)), "nwcch_set_tuning")

    def get_tuning(self, key: str) -> int:
        v = np.zeros(1, np.int32)
        check(lib().nwcch_get_tuning(self._h, key.encode(), ptr(v)), "nwcch_get_tuning")
        return int(v[0])

    # ---------------------------------------------------------------- batched
    def weights_batch(self, W) -> torch.Tensor:
        W = torch.as_tensor(W)
        if W.dim() != 2 or W.shape[1] != self.m:
            raise ValueError(f"expected [K, {self.m}] weights")
        if W.dtype != torch.int32 or not W.is_cuda:
            W64 = W.to(torch.int64)
            if W64.numel() and (int(W64.min()) <= -INF or int(W64.max()) > INF):
                raise ValueError("weights must lie in (-2^31+1, 2^31-1]")
            W = W64.to(self.dev, torch.int32)
        return W.contiguous()

    def batch_shape(self, K: int, slab: int = 0):
        """Shape of the batched result: int32 [slabs, mp, SW, 2] with slab width
        SW = min(slab or 64, K rounded up to a power of two), slabs = ceil(K/SW)."""
        sw = lib().nwcch_batch_slab_width(int(K), int(slab))
        if sw < 0:
            raise ValueError("slab must be 0 or a power of two <= 64")
        return ((K + sw - 1) // sw, self.mp, sw, 2)

    @staticmethod
    def metric_view(UD: torch.Tensor, k: int) -> torch.Tensor:
        """The [mp, 2] (up, down) values of metric k of a batched result."""
        sw = UD.shape[2]
        return UD[k // sw, :, k % sw, :]

    def customize_batch(self, W, out: torch.Tensor | None = None, slab: int = 0, stream=None) -> torch.Tensor:
        """K metrics in one level sweep: W [K, m] -> UD [slabs, mp, SW, 2]."""
        W = self.weights_batch(W)
        K = W.shape[0]
        shape = self.batch_shape(K, slab)
        if out is None:
            out = torch.empty(shape, dtype=torch.int32, device=self.dev)
        assert tuple(out.shape) == shape and out.dtype == torch.int32 and out.is_contiguous()
        check(lib().nwcch_customize_batch(self._h, ptr(W), K, int(slab), ptr(out), _stream(stream)),
              "nwcch_customize_batch")
        return out

    def negcycle_scan_batch(self, UD: torch.Tensor, K: int, out: torch.Tensor | None = None, stream=None):
        """Device int64 [2, K]: per-metric witness counts and smallest witness (or
        -1).  K is the number of metrics (the last slab may be partly unused)."""
        slab = UD.shape[2]
        assert tuple(UD.shape) == self.batch_shape(K, slab)
        if out is None:
            out = torch.empty((2, K), dtype=torch.int64, device=self.dev)
        check(lib().nwcch_negcycle_scan_batch(ptr(UD), self.mp, K, int(slab), ptr(out), _stream(stream)),
              "nwcch_negcycle_scan_batch")
        return out

    # ------------------------------------------------------------- read-back
    def export_triangles(self) -> np.ndarray:
        """The device's triangle list in schedule order, [T, 2] int32 on the host."""
        tab = np.empty(2 * self.triangles, np.int32)
        with torch.cuda.device(self.device):
            check(lib().nwcch_dev_export_triangles(self._h, ptr(tab)), "nwcch_dev_export_triangles")
        return tab.reshape(-1, 2)

    def export_sides(self):
        """(side_off [mp+1] int64, side_top [2T] int32) of the update index."""
        self.prepare_update()
        so = np.empty(self.mp + 1, np.int64)
        with torch.cuda.device(self.device):
            check(lib().nwcch_dev_export_sides(self._h, ptr(so), None), "nwcch_dev_export_sides")
            st = np.empty(int(so[-1]), np.int32)
            check(lib().nwcch_dev_export_sides(self._h, None, ptr(st)), "nwcch_dev_export_sides")
        return so, st

    # ---------------------------------------------------------- partial update
    def prepare_update(self):
        if not getattr(self, "_upd_ready", False):
            check(lib().nwcch_update_prepare(self._h, self.topo._h), "nwcch_update_prepare")
            self._upd_ready = True
            self._upd_res = torch.empty(4, dtype=torch.int64, device=self.dev)

    def update(self, w: torch.Tensor, ud: torch.Tensor, edge_idx, new_w, mode: int = 0,
               changed: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """Partial re-customization in place (w and ud are modified).  Changes
        with a repeated edge index keep the last weight (sequential assignment,
        python/cch.py:359-360).  Returns the device int64[4] result
        (evaluated, changed, negative changed, first negative changed or -1);
        not synchronised."""
        self.prepare_update()
        e = np.asarray(edge_idx, np.int64).reshape(-1)
        v = np.asarray(new_w, np.int64).reshape(-1)
        if e.shape != v.shape:
            raise ValueError("edge_idx and new_w must have the same length")
        if len(e) and (v.min() <= -INF or v.max() > INF):
            raise ValueError("weights must lie in (-2^31+1, 2^31-1]")
        if len(e):
            last = len(e) - 1 - np.unique(e[::-1], return_index=True)[1]
            last.sort()
            e, v = e[last], v[last]
        e = np.ascontiguousarray(e)
        v = np.ascontiguousarray(v.astype(np.int32))
        assert w.dtype == torch.int32 and w.is_cuda and w.shape == (self.m,)
        assert ud.dtype == torch.int32 and ud.shape == (self.mp, 2) and ud.is_contiguous()
        cap = 0 if changed is None else changed.numel()
        check(lib().nwcch_update(self._h, ptr(w), ptr(ud), ptr(e), ptr(v), len(e), int(mode),
                                 ptr(changed), cap, ptr(self._upd_res), _stream(stream)), "nwcch_update")
        return self._upd_res

    def last_changed(self, count: int, stream=None) -> torch.Tensor:
        """The changed arcs of the last update, ascending (python/cch.py:380, 396)."""
        out = torch.empty(max(count, 1), dtype=torch.int32, device=self.dev)
        if count:
            check(lib().nwcch_update_changed(self._h, ptr(out), count, _stream(stream)), "nwcch_update_changed")
        return torch.sort(out[:count]).values

    # -------------------------------------------------------------- profiling
    def profile_begin(self, max_launches: int):
        self._prof_cap = int(max_launches)
        check(lib().nwcch_profile_begin(self._h, self._prof_cap), "nwcch_profile_begin")

    def profile_end(self):
        """(kernel ids int32[], durations ms float32[]) of the launches since
        profile_begin (synchronises)."""
        cap = getattr(self, "_prof_cap", 0)
        kid = np.zeros(max(cap, 1), np.int32)
        ms = np.zeros(max(cap, 1), np.float32)
        cnt = np.zeros(1, np.int64)
        check(lib().nwcch_profile_end(self._h, ptr(kid), ptr(ms), cap, ptr(cnt)), "nwcch_profile_end")
        k = min(int(cnt[0]), cap)
        return kid[:k], ms[:k]

    # ---------------------------------------------------------------- status
    def status(self, clear=True, stream=None) -> int:
        st = np.zeros(1, np.uint32)
        check(lib().nwcch_status(self._h, ptr(st), int(clear), _stream(stream)), "nwcch_status")
        return int(st[0])

    def check_status(self, stream=None, negative_cycle: bool = False):
        """Raise on any sticky status bit.  With negative_cycle=True an int32
        overflow is accepted: under a negative cycle the customized values run
        away (the reference's float64 values would just keep falling), they are
        undefined as distances, and the verdict -- computed with clamped values
        that stay negative -- is what the caller gets."""
        st = self.status(clear=True, stream=stream)
        if negative_cycle:
            st &= ~1
        if st:
            bits = [v for k, v in _lib.STATUS_BITS.items() if st & k]
            raise NwcchError("device status: " + ", ".join(bits))

    def close(self):
        if getattr(self, "_h", None):
            lib().nwcch_dev_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
