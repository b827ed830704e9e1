"""The oracle's --separate and --jitter against the reference's own runs
(tests/golden/variant_*.npz, dense-grid cases; the adapted ones are checked on
the GPU).

--separate replaces the paired 2x2 node step by SeparateGaussSeidelSolve's x-
then-y step (ExteriorPoissonSolvers.h:198-250, whose guards test
fabs( H(k,k)>1e-20 ), i.e. H(k,k)>1e-20); --jitter translates the unit-cube
samples by a draw from the unseeded rand() sequence (ExteriorPoissonRecon.cpp:
174-180).  Every hierarchical coefficient must agree to 1e-9 relative per
depth (measured ~1e-13) and the final energy to the 6 digits printed."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN

GOLD = [p for p in sorted(glob.glob(os.path.join(GOLDEN, "variant_*.npz")))
        if int(np.load(p)["arg_aRadius"]) == -1]


def args_of(g):
    a = {k[4:]: g[k].item() for k in g.files if k.startswith("arg_") and k != "arg_aRadius"}
    if "separate" in a:
        a["separate"] = int(a["separate"])
    return a


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_oracle_variant_matches_reference(path, oracle_lib):
    from oracle.pyoracle import Oracle
    g = np.load(path)
    o = Oracle(g["in_pos"], g["in_skew"], **args_of(g)).solve()
    depth = int(g["arg_depth"])
    for d in range(depth + 1):
        for key, ref in ((f"hx{d}", g[f"x{d}"]), (f"hy{d}", g[f"y{d}"])):
            assert np.abs(o[key] - ref).max() <= 1e-9 * max(np.abs(ref).max(), 1e-300), (key, d)
    assert abs(o["final_E"].sum() - g["printed_E"][0]) <= 1e-5 * g["printed_E"][0]


def test_oracle_separate_changes_the_solve(oracle_lib):
    """The two steps are different algorithms: the same case without --separate ends elsewhere."""
    from oracle.pyoracle import Oracle
    g = np.load(os.path.join(GOLDEN, "variant_knot3_d4_separate.npz"))
    a = args_of(g)
    paired = Oracle(g["in_pos"], g["in_skew"], **{k: v for k, v in a.items() if k != "separate"}).solve()
    assert np.abs(paired["hx4"] - g["x4"]).max() > 1e-6 * np.abs(g["x4"]).max()
