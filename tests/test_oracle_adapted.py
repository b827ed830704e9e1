"""The oracle's adapted mode against the reference's own --aRadius runs
(tests/golden/adapted_*.npz, written by the reference build).

The hierarchy is decoded from the reference's tree (hierarchy_from_tree), so
this pins the oracle's restatement of what an adapted hierarchy changes in the
solve (ExteriorPoisson.h:126-409 over HierarchicalAdaptedIndexer): zero
coefficients outside it, relaxation of active functions only, B over indexed
elements, c over indexed elements, prolongation and Galerkin products with
indexed rows and columns accumulated in the hierarchy's numbering, the random
initial guess in that numbering.  Every hierarchical coefficient of every depth
must agree to 1e-11 relative (measured: <= 1.1e-13)."""
import glob
import os

import numpy as np
import pytest

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "adapted_*.npz")))


def fixture_args(g):
    return {k[4:]: g[k].item() for k in g.files if k.startswith("arg_") and k != "arg_aRadius"}


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_oracle_adapted_solve_matches_reference(path, oracle_lib):
    from oracle.pyoracle import Oracle, hierarchy_from_tree
    g = np.load(path)
    dim, depth = g["in_pos"].shape[1], int(g["arg_depth"])
    h = hierarchy_from_tree(g, dim, depth)
    o = Oracle(g["in_pos"], g["in_skew"], hierarchy=h, **fixture_args(g))
    r = o.solve()
    for d in range(depth + 1):
        order = h[d][2]
        for c, key in (("x", "hx"), ("y", "hy")):
            ref = g[f"{c}{d}"]
            mine = r[f"{key}{d}"]
            assert np.abs(mine[order] - ref).max() <= 1e-11 * max(np.abs(ref).max(), 1e-300), (d, c)
            # nothing outside the hierarchy
            outside = np.ones(len(mine), bool); outside[order] = False
            assert not np.any(mine[outside])
    # the energy the reference prints (6 significant digits)
    assert abs(r["final_E"].sum() - g["printed_E"][0]) <= 1e-5 * g["printed_E"][0]


def test_oracle_adapted_without_hierarchy_is_dense(oracle_lib):
    """A hierarchy that indexes every function and element, all active,
    reproduces the dense oracle exactly (same numbering)."""
    from oracle.pyoracle import Oracle
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "adapted_knot3_d4.npz"))
    depth = 3
    dim = g["in_pos"].shape[1]
    h = []
    for l in range(depth + 1):
        n1 = (1 << l) + 1
        h.append((np.ones(n1 ** dim, np.uint8), np.ones((1 << l) ** dim, np.uint8), np.arange(n1 ** dim)))
    a = Oracle(g["in_pos"], g["in_skew"], hierarchy=h, depth=depth).solve()
    b = Oracle(g["in_pos"], g["in_skew"], depth=depth).solve()
    for d in range(depth + 1):
        assert np.array_equal(a[f"hx{d}"], b[f"hx{d}"]) and np.array_equal(a[f"hy{d}"], b[f"hy{d}"])
