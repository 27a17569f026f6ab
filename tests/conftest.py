import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpgbsc.so on the device)")


@pytest.fixture(scope="session")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def golden_graph(g):
    """(n, colptr, rowidx) of a fixture graph exactly as the reference loads it."""
    import subgraph_amd as sg
    if g["kind"] == "edges":
        n, s, d = sg.remove_holes(g["n"], np.array(g["src"], np.int32), np.array(g["dst"], np.int32))
    else:
        s, d = sg.rmat_edges(g["scale"], g["edge_factor"], seed=g["seed"])
        assert len(s) == g["m"], "R-MAT generator changed: fixture no longer matches"
        from golden_util import edge_checksum
        assert edge_checksum(s, d) == g["checksum"], "R-MAT generator changed: fixture no longer matches"
        n, s, d = sg.remove_holes(1 << g["scale"], s, d)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    return n, colptr, rowidx


def golden_colors(c, n, k):
    import subgraph_amd as sg
    if c["kind"] == "explicit":
        return np.array(c["data"], np.uint8).reshape(-1, n)
    return sg.colors_fill(c["seed"], c["first"], c["count"], n, k)


@pytest.fixture(scope="session")
def gpu_ctx():
    import subgraph_amd as sg
    return sg.Context(0)
