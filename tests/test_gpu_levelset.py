"""Parity of the GPU simplicial extraction (epr_subgrid_triangulation,
epr_level_set through the C ABI) with the reference's outputs (fixtures) and
with the oracle on larger inputs: vertex numbering, simplex order and vertex
data identical bit for bit."""
import glob
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "ls_*.npz")))


def vertex_values(values, res, V):
    idx = np.zeros(len(V), dtype=np.int64)
    for d in range(V.shape[1]):
        idx = idx * (res + 1) + np.floor(V[:, d] + 0.5).astype(np.int64)
    return values[idx]


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_gpu_matches_reference(path):
    from exteriorpoissonrecon_amd import level_set, subgrid_triangulation
    g = np.load(path)
    dim, res = int(g["dim"]), int(g["res"])
    print(f"[gpu] level set {os.path.basename(path)}", flush=True)
    V, S = subgrid_triangulation(dim, g["cells"])
    np.testing.assert_array_equal(V.ravel(), g["tri_vertices"])
    np.testing.assert_array_equal(S.ravel(), g["tri_simplices"].astype(np.uint32))
    vals = vertex_values(g["node_values"], res, V)
    LV, LS = level_set(V, S, vals)
    np.testing.assert_array_equal(LV.ravel(), g["ls_vertices"])
    np.testing.assert_array_equal(LS.ravel(), g["ls_simplices"].astype(np.uint32))
    for c in range(2):                       # the codimension-1 sets of --levelSets
        LV1, LS1 = level_set(V, S, vals[:, c])
        np.testing.assert_array_equal(LV1.ravel(), g[f"ls1_{c}_vertices"])
        np.testing.assert_array_equal(LS1.ravel(), g[f"ls1_{c}_simplices"].astype(np.uint32))


@pytest.mark.parametrize("dim,res,kind", [(3, 40, "sphere"), (3, 12, "noise"), (4, 12, "sphere"), (4, 5, "noise")])
def test_gpu_matches_oracle_larger(dim, res, kind, oracle_lib):
    from exteriorpoissonrecon_amd import level_set, subgrid_triangulation
    rng = np.random.default_rng(res)
    nodes = np.indices((res + 1,) * dim).reshape(dim, -1).T.astype(np.float64)
    if kind == "sphere":
        c = np.full(dim, res / 2.0) + rng.normal(size=dim) * 0.3
        n = rng.normal(size=dim)
        n /= np.linalg.norm(n)
        q = nodes - c
        values = np.stack([(q * q).sum(1) - (0.35 * res) ** 2, q @ n], 1)
    else:
        values = rng.uniform(-1, 1, size=(len(nodes), 2))
    cells = np.array([e[::-1] for e in np.ndindex(*((res,) * dim))], dtype=np.uint32)
    cells = cells[rng.permutation(len(cells))]
    print(f"[gpu] level set {dim}D res {res} {kind}: {len(cells)} cells", flush=True)
    V, S = subgrid_triangulation(dim, cells)
    oV, oS = oracle_lib.subgrid_triangulation(dim, cells)
    np.testing.assert_array_equal(V, oV)
    np.testing.assert_array_equal(S, oS)
    vals = vertex_values(values, res, V)
    LV, LS = level_set(V, S, vals)
    oLV, oLS = oracle_lib.level_set(V, S, vals)
    assert len(LV) > 0
    np.testing.assert_array_equal(LV, oLV)
    np.testing.assert_array_equal(LS, oLS)
    LV1, LS1 = level_set(V, S, vals[:, 0])
    oLV1, oLS1 = oracle_lib.level_set(V, S, vals[:, 0])
    np.testing.assert_array_equal(LV1, oLV1)
    np.testing.assert_array_equal(LS1, oLS1)


def test_gpu_empty_inputs():
    from exteriorpoissonrecon_amd import level_set, subgrid_triangulation
    V, S = subgrid_triangulation(3, np.zeros((0, 3), dtype=np.uint32))
    assert V.shape == (0, 3) and S.shape == (0, 4)
    V, S = subgrid_triangulation(3, np.array([[0, 0, 0]], dtype=np.uint32))
    LV, LS = level_set(V, S, np.ones((len(V), 2)))          # no zero crossing
    assert LV.shape == (0, 3) and LS.shape == (0, 2)
