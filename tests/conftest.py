import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "negative-weight-cch_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def c0():
    from nwcch import _lib, graphs
    g = graphs.make_config(0)
    rank = _lib.order(g.n, g.off, g.tgt, g.x, g.y, 64)
    return g, rank


def small_graphs():
    """The reference's test family (python/test_cch.py:30-32), integer versions:
    4 grids of 15x15, 4 geometric graphs of 300, 4 shifted random graphs of 200."""
    from nwcch import graphs
    return ([graphs.ev_grid(15, seed=i) for i in range(4)]
            + [graphs.ev_geometric(300, seed=i) for i in range(4)]
            + [graphs.shifted_random(200, seed=i) for i in range(4)])


def order_of(g, leaf=8):
    from nwcch import _lib
    return _lib.order(g.n, g.off, g.tgt, g.x, g.y, leaf)
