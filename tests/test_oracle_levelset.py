"""The simplicial-extraction oracle against the reference's own outputs and
against what the geometry fixes.

Fixtures tests/golden/ls_*.npz hold the reference build's
RegularSubGridTriangulation and LevelSet outputs (codimension 2 and both
codimension-1 sets) on circle/2-sphere fields, grid-aligned fields with exact
zeros on nodes and edges, and random fields (make_levelset_golden.py).
Topology and vertex data must agree bit for bit.
"""
import glob
import math
import os

import numpy as np
import pytest

from conftest import ROOT

FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "ls_*.npz")))


def vertex_values(g, V):
    dim, res = int(g["dim"]), int(g["res"])
    idx = np.zeros(len(V), dtype=np.int64)
    for d in range(dim):
        idx = idx * (res + 1) + np.floor(V[:, d] + 0.5).astype(np.int64)
    return g["node_values"][idx]


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_matches_reference(path, oracle_lib):
    g = np.load(path)
    dim = int(g["dim"])
    V, S = oracle_lib.subgrid_triangulation(dim, g["cells"])
    np.testing.assert_array_equal(V.ravel(), g["tri_vertices"])
    np.testing.assert_array_equal(S.ravel(), g["tri_simplices"].astype(np.uint32))
    vals = vertex_values(g, V)
    LV, LS = oracle_lib.level_set(V, S, vals)
    np.testing.assert_array_equal(LV.ravel(), g["ls_vertices"])
    np.testing.assert_array_equal(LS.ravel(), g["ls_simplices"].astype(np.uint32))
    for c in range(2):
        LV1, LS1 = oracle_lib.level_set(V, S, vals[:, c])
        np.testing.assert_array_equal(LV1.ravel(), g[f"ls1_{c}_vertices"])
        np.testing.assert_array_equal(LS1.ravel(), g[f"ls1_{c}_simplices"].astype(np.uint32))


@pytest.mark.parametrize("dim", [3, 4])
def test_cube_simplices_tile_the_cube(dim, oracle_lib):
    """dim! distinct simplices of volume 1/dim! (edge-matrix determinant +1:
    positively oriented once oriented), so they tile the unit cube."""
    V, S = oracle_lib.subgrid_triangulation(dim, np.zeros((1, dim), dtype=np.uint32))
    assert len(S) == math.factorial(dim) and len(V) == 2 ** dim
    vols = []
    for s in S:
        M = V[s[1:]] - V[s[0]]
        vols.append(np.linalg.det(M))
    assert np.allclose(vols, 1.0)
    keys = {tuple(sorted(map(int, s))) for s in S}
    assert len(keys) == len(S)


def test_triangulation_shares_vertices(oracle_lib):
    """A 2x2x2 block of cells has 27 distinct vertices; numbering is by first appearance."""
    cells = np.array([[i, j, k] for k in range(2) for j in range(2) for i in range(2)], dtype=np.uint32)
    V, S = oracle_lib.subgrid_triangulation(3, cells)
    assert len(V) == 27 and len(S) == 8 * 6
    first = []
    for v in S.ravel():
        if v not in first:
            first.append(int(v))
    assert first == list(range(27))


@pytest.mark.parametrize("dim", [3, 4])
def test_level_set_of_linear_fields(dim, oracle_lib):
    """Linear values: the barycentric interpolation is exact, so every output
    vertex lies on both planes (to rounding), and the 3D set is a polyline."""
    rng = np.random.default_rng(dim)
    res = 5
    cells = np.array([e[::-1] for e in np.ndindex(*((res,) * dim))], dtype=np.uint32)
    V, S = oracle_lib.subgrid_triangulation(dim, cells)
    A = rng.normal(size=(2, dim))
    b = -A @ np.full(dim, res / 2.0) + rng.normal(size=2) * 0.1
    vals = V @ A.T + b
    LV, LS = oracle_lib.level_set(V, S, vals)
    assert len(LV) > 0
    np.testing.assert_allclose(LV @ A.T + b, 0, atol=1e-12)
    if dim == 3:
        deg = np.bincount(LS.ravel(), minlength=len(LV))
        assert deg.max() <= 2 and (deg == 1).sum() == 2     # one open polyline through the box
