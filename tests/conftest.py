import json
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


def load_lattice(name):
    """Hopping matrix (dense nx x nx) of a built-in lattice stored in tests/golden/lattices.json."""
    d = json.loads((GOLDEN / "lattices.json").read_text())[name]
    nx = d["nx"]
    h = np.zeros((nx, nx))
    for (i, j), hop in zip(d["adjacency"], d["hopping"]):
        if hop != 0:
            h[i, j] = h[j, i] = hop
    return h


@pytest.fixture
def lattice():
    return load_lattice
