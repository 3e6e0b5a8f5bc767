"""The oracle against golden vectors written by the reference's own python/cch.py
(tests/golden/make_cch_py_vectors.py): full customization, the partial update's
evaluated count and changed list (python/cch.py:349-402), and the witnesses of
customize with and without the early stop (python/cch.py:249-287)."""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden" / "cch_py_vectors.npz"


def cases():
    with np.load(GOLD) as d:
        return [str(c) for c in d["cases"]]


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("name", cases())
def test_customization_equals_reference_python(gold, name):
    off, tgt, w, rank = (gold[f"{name}/{k}"] for k in ("off", "tgt", "w", "rank"))
    st = O.Structure(len(rank), off, tgt, rank)
    up, dn, ovf = st.customize(w)
    assert not ovf
    assert np.array_equal(up, gold[f"{name}/up0"]) and np.array_equal(dn, gold[f"{name}/dn0"])


@pytest.mark.parametrize("name", cases())
def test_update_counts_equal_reference_python(gold, name):
    off, tgt, w, rank = (gold[f"{name}/{k}"] for k in ("off", "tgt", "w", "rank"))
    st = O.Structure(len(rank), off, tgt, rank)
    up, dn, _ = st.customize(w)
    upd = O.Updater(st, w, up, dn)
    for u in range(int(gold[f"{name}/n_updates"])):
        ev, changed, ovf = upd.update(gold[f"{name}/u{u}/edges"], gold[f"{name}/u{u}/new_w"])
        assert not ovf
        assert ev == int(gold[f"{name}/u{u}/evaluated"]), (name, u)
        assert changed.tolist() == gold[f"{name}/u{u}/last_changed"].tolist(), (name, u)
        neg = st.negcycle(upd.up[changed], upd.dn[changed])[0] > 0
        assert neg == bool(gold[f"{name}/u{u}/negative"])
    assert np.array_equal(upd.up, gold[f"{name}/up_end"]) and np.array_equal(upd.dn, gold[f"{name}/dn_end"])


@pytest.mark.parametrize("name", cases())
def test_witnesses_equal_reference_python(gold, name):
    off, tgt, rank = (gold[f"{name}/{k}"] for k in ("off", "tgt", "rank"))
    wn = gold[f"{name}/neg_w"]
    st = O.Structure(len(rank), off, tgt, rank)
    up, dn, _ = st.customize(wn)
    neg = np.nonzero((up != O.INF) & (dn != O.INF) & (up.astype(np.int64) + dn < 0))[0]
    assert neg.tolist() == gold[f"{name}/neg_witnesses"].tolist()
    assert [st.customize_stop(wn)] == gold[f"{name}/neg_stop_witness"].tolist()


@pytest.mark.parametrize("name", cases())
def test_query_path_equals_reference_python(gold, name):
    """Unpacked shortest paths (python/cch.py:406-507): the same distance and the
    same vertex sequence, ties resolved as the reference resolves them."""
    off, tgt, w, rank = (gold[f"{name}/{k}"] for k in ("off", "tgt", "w", "rank"))
    st = O.Structure(len(rank), off, tgt, rank)
    up, dn, _ = st.customize(w)
    po, pv = gold[f"{name}/path_off"], gold[f"{name}/path_v"]
    for i, (s, t) in enumerate(zip(gold[f"{name}/path_s"], gold[f"{name}/path_t"])):
        d, path = st.query_path(up, dn, int(s), int(t))
        assert d == int(gold[f"{name}/path_dist"][i]), (name, i)
        assert path == pv[po[i]:po[i + 1]].tolist(), (name, i)


@pytest.mark.parametrize("name", cases())
def test_query_scanned_equals_reference_python(gold, name):
    """ET queries (python/cch.py:675-732): the distance and the number of arcs
    scanned, 300 pairs per graph, unreachable and negative pairs among them."""
    off, tgt, w, rank = (gold[f"{name}/{k}"] for k in ("off", "tgt", "w", "rank"))
    st = O.Structure(len(rank), off, tgt, rank)
    up, dn, _ = st.customize(w)
    d, sc, _ = st.query(up, dn, gold[f"{name}/q_s"], gold[f"{name}/q_t"])
    assert np.array_equal(d.astype(np.int64), gold[f"{name}/q_dist"])
    assert np.array_equal(sc.astype(np.int64), gold[f"{name}/q_scanned"])


@pytest.mark.parametrize("name", cases())
def test_stop_state_equals_reference_python(gold, name):
    """The values customize(g, stop_on_negative_cycle=True) leaves when it stops
    (python/cch.py:270-278)."""
    off, tgt, rank = (gold[f"{name}/{k}"] for k in ("off", "tgt", "rank"))
    st = O.Structure(len(rank), off, tgt, rank)
    wit, up, dn = st.customize_stop(gold[f"{name}/neg_w"], with_state=True)
    assert [wit] == gold[f"{name}/neg_stop_witness"].tolist()
    assert np.array_equal(up, gold[f"{name}/neg_stop_up"]) and np.array_equal(dn, gold[f"{name}/neg_stop_dn"])


@pytest.mark.parametrize("name", cases())
def test_order_reproduces_the_fixture(gold, name):
    """The nested-dissection order is deterministic: it reproduces the ranks the
    fixture was written with (any maximum flow gives the same source-side
    minimum cut, so the flow algorithm does not change a separator)."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "negative-weight-cch_amd"))
    from nwcch import graphs, _lib
    g = graphs.make_config(0) if name == "grid" else \
        [graphs.ev_grid(12, seed=1), graphs.ev_geometric(200, seed=2), graphs.shifted_random(150, seed=3)][int(name[-1])]
    assert np.array_equal(g.off, gold[f"{name}/off"]) and np.array_equal(g.tgt, gold[f"{name}/tgt"])
    rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 8 if g.n < 1000 else 64)
    assert np.array_equal(rank, gold[f"{name}/rank"])
