import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ["knot3_d4", "knot3_d5", "knot3_d4_params", "clifford4_d3"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def case_args(g):
    """Reference command-line arguments of a fixture; fixtures without aRadius are dense-grid runs."""
    args = {k[4:]: (g[k].item()) for k in g if k.startswith("arg_")}
    args.setdefault("aRadius", -1)
    return args


def read_grid_values(path):
    """Values of a RegularGrid file (Misha/RegularGrid.inl:440-471): "G<dim>",
    "<n> DOUBLE", the resolutions, dim+1 transform rows, then the binary values
    (an array of shape (count,) or (count, n))."""
    b = open(path, "rb").read()
    o = 0
    lines = []
    while len(lines) < 3:
        e = b.index(b"\n", o); lines.append(b[o:e].decode()); o = e + 1
    dim = int(lines[0][1:])
    nv = int(lines[1].split()[0])
    for _ in range(dim + 1):
        o = b.index(b"\n", o) + 1
    v = np.frombuffer(b, dtype="<f8", offset=o)
    return v if nv == 1 else v.reshape(-1, nv)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    scale = max(np.abs(b).max(), 1e-300) if b.size else 1.0
    return float(np.abs(a - b).max() / scale) if b.size else 0.0


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle
