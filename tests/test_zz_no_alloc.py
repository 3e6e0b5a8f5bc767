"""nwcch_update_noalloc (the update that allocates nothing: its staging is a
static device array) against nwcch_update through the C ABI: the same weights,
customization, result and changed set on consecutive calls, the later ones
smaller; no device memory taken by a call; the size limit rejected."""
import numpy as np
import pytest
import torch

from nwcch import _lib
from nwcch.device import DeviceTopology
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MAX_CHANGES = 65536


def _call(fn, dt, w, ud, idx, new, mode, changed, res):
    e = np.ascontiguousarray(idx, np.int64)
    v = np.ascontiguousarray(new, np.int32)
    rc = fn(dt._h, _lib.ptr(w), _lib.ptr(ud), _lib.ptr(e), _lib.ptr(v), len(e), mode, _lib.ptr(changed),
            changed.numel(), _lib.ptr(res), torch.cuda.current_stream().cuda_stream)
    return rc


def test_update_noalloc_equals_update(c0):
    g, rank = c0
    print("case: config 0, nwcch_update_noalloc vs nwcch_update, k = 1000, 5, 2", flush=True)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    dt.prepare_update()
    L = _lib.lib()
    w_a = dt.weights(g.w)
    ud_a = dt.customize(w_a)
    w_b, ud_b = w_a.clone(), ud_a.clone()
    mp = topo.mp
    chg_a = torch.empty(mp, dtype=torch.int32, device=dt.dev)
    chg_b = torch.empty(mp, dtype=torch.int32, device=dt.dev)
    res_a = torch.empty(4, dtype=torch.int64, device=dt.dev)
    res_b = torch.empty(4, dtype=torch.int64, device=dt.dev)
    st = O.Structure(g.n, g.off, g.tgt, rank)
    up, dn, _ = st.customize(g.w)
    upd = O.Updater(st, g.w, up, dn)
    rng = np.random.default_rng(5)
    for k, mode in ((1000, 0), (5, 0), (2, 1)):
        idx = rng.choice(g.m, k, replace=False)
        new = (upd.w[idx] + rng.integers(-500, 501, k)).astype(np.int32)
        assert _call(L.nwcch_update, dt, w_a, ud_a, idx, new, mode, chg_a, res_a) == 0
        assert _call(L.nwcch_update_noalloc, dt, w_b, ud_b, idx, new, mode, chg_b, res_b) == 0
        torch.cuda.synchronize()
        ra, rb = res_a.cpu().numpy(), res_b.cpu().numpy()
        assert np.array_equal(ra, rb), (k, ra, rb)
        assert torch.equal(w_a, w_b) and torch.equal(ud_a, ud_b), k
        nch = int(ra[1])
        assert np.array_equal(np.sort(chg_a[:nch].cpu().numpy()), np.sort(chg_b[:nch].cpu().numpy())), k
        # and both equal the oracle's restatement of update_edges
        ev0, ch0, _ = upd.update(idx, new)
        got = ud_b.cpu().numpy()
        assert int(rb[0]) == ev0 and nch == len(ch0), k
        assert np.array_equal(got[:, 0], upd.up) and np.array_equal(got[:, 1], upd.dn), k
    dt.status(clear=True)


def test_update_noalloc_takes_no_device_memory(c0):
    """A call of 60,000 changes (past the staging nwcch_update starts with)
    leaves the device's free memory unchanged."""
    g, rank = c0
    print("case: config 0, nwcch_update_noalloc k = 60000, free device memory", flush=True)
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    dt.prepare_update()
    L = _lib.lib()
    w = dt.weights(g.w)
    ud = dt.customize(w)
    chg = torch.empty(topo.mp, dtype=torch.int32, device=dt.dev)
    res = torch.empty(4, dtype=torch.int64, device=dt.dev)
    rng = np.random.default_rng(6)
    idx = rng.integers(0, g.m, 60000)
    new = g.w[idx].astype(np.int32)
    assert _call(L.nwcch_update_noalloc, dt, w, ud, idx[:10], new[:10], 0, chg, res) == 0   # first use
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    assert _call(L.nwcch_update_noalloc, dt, w, ud, idx, new, 0, chg, res) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)
    dt.status(clear=True)


def test_update_noalloc_rejects_too_many_changes(c0):
    g, rank = c0
    topo = _lib.Topology.from_graph(g.n, g.off, g.tgt, rank)
    dt = DeviceTopology(topo)
    dt.prepare_update()
    w = dt.weights(g.w)
    ud = dt.customize(w)
    chg = torch.empty(1, dtype=torch.int32, device=dt.dev)
    res = torch.empty(4, dtype=torch.int64, device=dt.dev)
    idx = np.zeros(MAX_CHANGES + 1, np.int64)
    rc = _call(_lib.lib().nwcch_update_noalloc, dt, w, ud, idx, np.zeros(len(idx), np.int32), 0, chg, res)
    assert rc == -5     # NWCCH_ERR_CAPACITY
    assert "65536" in _lib.lib().nwcch_last_error().decode()
