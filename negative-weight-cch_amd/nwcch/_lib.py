"""ctypes binding of libnwcch.so (the C ABI declared in include/nwcch.h).

The product path is this library and nothing else: if it is missing, importing
the GPU entry points raises instead of falling back to any CPU code.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libnwcch.so"
INF = 2147483647

NWCCH_OK = 0
NWCCH_TOPO_DEVICE_INDEX = 1
NWCCH_TOPO_DEVICE_SYMBOLIC = 2
STATUS_BITS = {1: "int32 overflow", 2: "INT32_MIN weight", 4: "query vertex out of range",
               8: "update frontier full or path longer than max_len",
               16: "path unpacking passed 4n + 16 vertices (negative cycle)"}

_lib = None


class NwcchError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NwcchError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        p, i32, i64, vp = C.c_void_p, C.c_int32, C.c_int64, C.c_void_p
        sigs = {
            "nwcch_last_error": ([], C.c_char_p),
            "nwcch_order": ([i64, p, p, p, p, i32, p], C.c_int),
            "nwcch_topo_create": ([i64, p, p, p, C.POINTER(vp)], C.c_int),
            "nwcch_topo_create_ex": ([i64, p, p, p, i32, C.c_int, C.POINTER(vp)], C.c_int),
            "nwcch_topo_from_csr": ([i64, p, p, i64, p, p, C.POINTER(vp)], C.c_int),
            "nwcch_topo_from_csr_ex": ([i64, p, p, i64, p, p, i32, C.c_int, C.POINTER(vp)], C.c_int),
            "nwcch_topo_destroy": ([vp], None),
            "nwcch_topo_info": ([vp, p], C.c_int),
            "nwcch_topo_export": ([vp, p, p, p, p, p, p, p], C.c_int),
            "nwcch_topo_export_triangles": ([vp, p, p, p], C.c_int),
            "nwcch_dev_create": ([vp, C.c_int, C.POINTER(vp)], C.c_int),
            "nwcch_dev_destroy": ([vp], None),
            "nwcch_dev_info": ([vp, p], C.c_int),
            "nwcch_dev_export_triangles": ([vp, p], C.c_int),
            "nwcch_dev_export_sides": ([vp, p, p], C.c_int),
            "nwcch_customize": ([vp, p, p, vp], C.c_int),
            "nwcch_customize_stop_state": ([vp, p, p, i32, p, vp], C.c_int),
            "nwcch_negcycle_scan": ([p, i64, p, vp], C.c_int),
            "nwcch_negcycle_list": ([p, i64, p, i64, p, vp], C.c_int),
            "nwcch_query": ([vp, p, p, p, i64, p, p, vp], C.c_int),
            "nwcch_status": ([vp, p, C.c_int, vp], C.c_int),
            "nwcch_profile_begin": ([vp, i64], C.c_int),
            "nwcch_update_prepare": ([vp, vp], C.c_int),
            "nwcch_update": ([vp, p, p, p, p, i64, i32, p, i64, p, vp], C.c_int),
            "nwcch_update_changed": ([vp, p, i64, vp], C.c_int),
            "nwcch_update_noalloc": ([vp, p, p, p, p, i64, i32, p, i64, p, vp], C.c_int),
            "nwcch_profile_end": ([vp, p, p, i64, p], C.c_int),
            "nwcch_query_strided": ([vp, p, i64, p, p, i64, p, p, p, vp], C.c_int),
            "nwcch_query_multi": ([vp, p, i32, p, p, p, i64, p, p, vp], C.c_int),
            "nwcch_metric_split": ([p, i64, p, p, vp], C.c_int),
            "nwcch_query_split": ([vp, p, p, p, p, i64, p, p, p, vp], C.c_int),
            "nwcch_customize_batch": ([vp, p, i32, i32, p, vp], C.c_int),
            "nwcch_negcycle_scan_batch": ([p, i64, i32, i32, p, vp], C.c_int),
            "nwcch_batch_slab_width": ([i32, i32], C.c_int),
            "nwcch_set_tuning": ([vp, C.c_char_p, i32], C.c_int),
            "nwcch_get_tuning": ([vp, C.c_char_p, p], C.c_int),
            "nwcch_potential": ([vp, p, p, vp], C.c_int),
            "nwcch_query_path": ([vp, p, p, p, i64, i32, p, p, p, vp], C.c_int),
        }
        optional = {}
        for name, (args, res) in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        for name, (args, res) in optional.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != NWCCH_OK:
        msg = lib().nwcch_last_error().decode(errors="replace")
        raise NwcchError(f"{what} failed ({rc}): {msg}")


def ptr(a) -> int | None:
    """Address of a numpy array or torch tensor (None for None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "arrays passed to libnwcch must be contiguous"
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        assert a.is_contiguous(), "tensors passed to libnwcch must be contiguous"
        return a.data_ptr()
    raise TypeError(type(a))


def exported_symbols():
    """Names declared in include/nwcch.h (for the symbol-export test)."""
    return ["nwcch_last_error", "nwcch_order", "nwcch_topo_create", "nwcch_topo_from_csr",
            "nwcch_topo_destroy", "nwcch_topo_info", "nwcch_topo_export",
            "nwcch_topo_export_triangles", "nwcch_dev_create", "nwcch_dev_destroy",
            "nwcch_dev_info", "nwcch_customize", "nwcch_negcycle_scan", "nwcch_negcycle_list",
            "nwcch_query", "nwcch_status", "nwcch_profile_begin", "nwcch_profile_end",
            "nwcch_update_prepare", "nwcch_update", "nwcch_update_changed", "nwcch_query_strided",
            "nwcch_customize_batch", "nwcch_negcycle_scan_batch", "nwcch_batch_slab_width",
            "nwcch_set_tuning", "nwcch_get_tuning", "nwcch_query_multi", "nwcch_potential",
            "nwcch_query_path", "nwcch_topo_create_ex", "nwcch_dev_export_triangles",
            "nwcch_dev_export_sides", "nwcch_topo_from_csr_ex", "nwcch_update_noalloc",
            "nwcch_customize_stop_state"]


# ------------------------------------------------------------ host topology

class Topology:
    """Metric-independent CCH structure built by the native host code."""

    def __init__(self, handle):
        self._h = handle
        info = np.zeros(8, np.int64)
        check(lib().nwcch_topo_info(self._h, ptr(info)), "nwcch_topo_info")
        (self.n, self.m, self.mp, self.triangles, self.depth, self.levels,
         self.n_loops, self.n_mapped) = (int(v) for v in info)

    @classmethod
    def from_graph(cls, n, off, tgt, rank, device_index=None, device_symbolic=None):
        """device_index = a GPU ordinal: that GPU builds the lower-triangle index
        (NWCCH_TOPO_DEVICE_INDEX); the host then keeps no triangle list.
        device_symbolic = a GPU ordinal: that GPU builds the chordal completion
        and the arc map (NWCCH_TOPO_DEVICE_SYMBOLIC)."""
        off = np.ascontiguousarray(off, np.int64)
        tgt = np.ascontiguousarray(tgt, np.int32)
        rank = np.ascontiguousarray(rank, np.int32)
        if off.shape != (n + 1,) or rank.shape != (n,):
            raise ValueError("off must be [n+1] and rank [n]")
        if (device_index is not None and device_symbolic is not None
                and int(device_index) != int(device_symbolic)):
            raise ValueError("device_index and device_symbolic name different GPUs")
        flags = ((0 if device_index is None else NWCCH_TOPO_DEVICE_INDEX)
                 | (0 if device_symbolic is None else NWCCH_TOPO_DEVICE_SYMBOLIC))
        dev = device_index if device_index is not None else device_symbolic
        h = C.c_void_p()
        if flags == 0:
            check(lib().nwcch_topo_create(n, ptr(off), ptr(tgt), ptr(rank), C.byref(h)), "nwcch_topo_create")
        else:
            check(lib().nwcch_topo_create_ex(n, ptr(off), ptr(tgt), ptr(rank), flags, int(dev), C.byref(h)),
                  "nwcch_topo_create_ex")
        t = cls(h)
        t.rank = rank.copy()
        t.device_index = device_index
        return t

    @classmethod
    def from_csr(cls, n, first, head, e_arc=None, e_up=None, device_index=None):
        """device_index = a GPU ordinal: that GPU builds the triangle index."""
        first = np.ascontiguousarray(first, np.int64)
        head = np.ascontiguousarray(head, np.int32)
        m = 0 if e_arc is None else int(len(e_arc))
        ea = None if e_arc is None else np.ascontiguousarray(e_arc, np.int64)
        eu = None if e_up is None else np.ascontiguousarray(e_up, np.int8)
        h = C.c_void_p()
        if device_index is None:
            check(lib().nwcch_topo_from_csr(n, ptr(first), ptr(head), m, ptr(ea), ptr(eu), C.byref(h)),
                  "nwcch_topo_from_csr")
        else:
            check(lib().nwcch_topo_from_csr_ex(n, ptr(first), ptr(head), m, ptr(ea), ptr(eu),
                                               NWCCH_TOPO_DEVICE_INDEX, int(device_index), C.byref(h)),
                  "nwcch_topo_from_csr_ex")
        t = cls(h)
        t.rank = None
        return t

    def export(self):
        first = np.empty(self.n + 1, np.int64)
        head = np.empty(self.mp, np.int32)
        parent = np.empty(self.n, np.int32)
        e_arc = np.empty(self.m, np.int64)
        e_up = np.empty(self.m, np.int8)
        level = np.empty(self.n, np.int32)
        depth = np.empty(self.n, np.int32)
        check(lib().nwcch_topo_export(self._h, ptr(first), ptr(head), ptr(parent), ptr(e_arc),
                                      ptr(e_up), ptr(level), ptr(depth)), "nwcch_topo_export")
        return dict(first=first, head=head, parent=parent, e_arc=e_arc, e_up=e_up,
                    level=level, depth=depth)

    def export_triangles(self, with_list=True):
        """(sched_arc, tri_off, tri_ab[T, 2]); tri_ab is None without with_list
        (a topology whose list lives on the device has only the first two)."""
        sched = np.empty(self.mp, np.int32)
        toff = np.empty(self.mp + 1, np.int64)
        tab = np.empty(2 * self.triangles, np.int32) if with_list else None
        check(lib().nwcch_topo_export_triangles(self._h, ptr(sched), ptr(toff), ptr(tab)),
              "nwcch_topo_export_triangles")
        return sched, toff, (tab.reshape(-1, 2) if with_list else None)

    def close(self):
        if getattr(self, "_h", None):
            lib().nwcch_topo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def order(n, off, tgt, x, y, leaf=64) -> np.ndarray:
    """Nested-dissection order (inertial flow), rank[v]."""
    off = np.ascontiguousarray(off, np.int64)
    tgt = np.ascontiguousarray(tgt, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    rank = np.empty(n, np.int32)
    check(lib().nwcch_order(n, ptr(off), ptr(tgt), ptr(x), ptr(y), int(leaf), ptr(rank)), "nwcch_order")
    return rank
