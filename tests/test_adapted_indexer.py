"""The adapted hierarchy built on the host (epr_adapted_indexer_write) against
the reference's serialised HierarchicalAdaptedIndexer (tests/golden/adapted_*.npz,
written by the reference build with --aRadius; Include/Hat.adaptive.inl:349-729).

Exact: every tree node (depth, offset, children flag, function and element
index, in file order), the active counts and the function/element counts."""
import ctypes
import glob
import os
import tempfile

import numpy as np
import pytest

from exteriorpoissonrecon_amd._lib import load
from exteriorpoissonrecon_amd.treefile import read_adapted_indexer

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "adapted_*.npz")))


def build_indexer(pos, depth, radius, scale=1.1):
    L = load()
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "ix.bin")
        rc = L.epr_adapted_indexer_write(pos.shape[1], depth, radius, scale, pos.shape[0],
                                         pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         path.encode())
        assert rc == 0, L.epr_last_error(None)
        b = open(path, "rb").read()
    ix, o = read_adapted_indexer(b, 0, pos.shape[1])
    assert o == len(b)
    return ix


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_adapted_indexer_matches_reference(path):
    g = np.load(path)
    ix = build_indexer(g["in_pos"], int(g["arg_depth"]), int(g["arg_aRadius"]))
    assert np.array_equal(ix["active_sizes"], g["active_sizes"])
    assert np.array_equal(ix["num_functions"], g["num_functions"])
    assert np.array_equal(ix["num_elements"], g["num_elements"])
    assert len(ix["node_depth"]) == len(g["node_depth"])
    assert np.array_equal(ix["node_depth"], g["node_depth"])
    assert np.array_equal(ix["node_offset"], g["node_offset"])
    assert np.array_equal(ix["node_has_children"], g["node_has_children"])
    assert np.array_equal(ix["node_function"], g["node_function"])
    assert np.array_equal(ix["node_element"], g["node_element"])


def test_adapted_radius_out_of_range():
    L = load()
    pos = np.random.default_rng(0).random((64, 3))
    for r in (6, -2):
        rc = L.epr_adapted_indexer_write(3, 3, r, 1.1, 64, pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         b"/dev/null")
        assert rc != 0
        assert b"kernel radius" in L.epr_last_error(None)


def test_adapted_hierarchy_invariants():
    """The indexer's own invariants (BaseIndexer, Hat.adaptive.h:37-45): the
    elements of active functions exist, and the functions of indexed elements
    exist; every active function's one-ring is indexed."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "adapted_knot3_d5_r0.npz"))
    dim = g["in_pos"].shape[1]
    ix = build_indexer(g["in_pos"], int(g["arg_depth"]), int(g["arg_aRadius"]))
    depth = int(g["arg_depth"])
    for d in range(depth + 1):
        sel = ix["node_depth"] == d + 1
        F = {tuple(o): f for o, f in zip(ix["node_offset"][sel], ix["node_function"][sel]) if f >= 0}
        E = {tuple(o) for o, e in zip(ix["node_offset"][sel], ix["node_element"][sel]) if e >= 0}
        res = 1 << d
        active = [o for o, f in F.items() if f < ix["active_sizes"][d]]
        for o in active:
            for w in np.ndindex(*(3,) * dim):
                G = tuple(o[k] + w[k] - 1 for k in range(dim))
                if all(0 <= c <= res for c in G):
                    assert G in F
                Eo = tuple(o[k] + w[k] - 1 for k in range(dim))
                if all(w[k] <= 1 for k in range(dim)) and all(0 <= c < res for c in Eo):
                    assert Eo in E
        for e in E:
            for c in np.ndindex(*(2,) * dim):
                assert tuple(e[k] + c[k] for k in range(dim)) in F
