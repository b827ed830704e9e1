"""The multigrid oracle (oracle/pymg.py) against the reference's own outputs
(tests/golden/mg_*.npz from Include/MG.h, make_mg_golden.py) and against what
the algebra fixes."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_mg_golden as mk  # noqa: E402
from oracle import pymg  # noqa: E402

CASES = sorted(mk.CASES)


def load(name):
    g = np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"))
    M, Ps, b, cols = mk.build(mk.CASES[name])
    np.testing.assert_array_equal(b, g["b"])
    mc = {l + 1: [c[1][c[0][k]:c[0][k + 1]] for k in range(len(c[0]) - 1)] for l, c in enumerate(cols)}
    return g, M, Ps, b, mc


@pytest.mark.parametrize("name", CASES)
def test_relaxation_bit_exact(name):
    g, M, Ps, b, mc = load(name)
    x = pymg.GSRelaxer(M)(b, np.zeros(len(b)), mc[len(Ps)])
    np.testing.assert_array_equal(x, g["gs_x"])


@pytest.mark.parametrize("name", CASES)
def test_vcycles_match_reference(name):
    g, M, Ps, b, mc = load(name)
    s = pymg.MGSolver(M, Ps)
    x = np.zeros(len(b))
    for c in range(int(g["cycles"])):
        x = s.vcycle(b, x, int(g["iters"]), lambda l: mc[l])
        ref = g[f"x{c}"]
        assert np.abs(x - ref).max() <= 1e-11 * np.abs(ref).max(), c
        r = np.linalg.norm(b - M @ x)
        assert abs(r - g[f"res{c}"][0]) <= 1e-6 * g[f"res{c}"][0] + 1e-14


def test_galerkin_products_and_direct_solve():
    M, Ps, b, _ = mk.build(mk.CASES["mg_hat3"])
    s = pymg.MGSolver(M, Ps)
    ref = (Ps[-1].T @ M @ Ps[-1]).toarray()
    np.testing.assert_allclose(s.Ms[-1].toarray(), ref, rtol=1e-13, atol=1e-15)
    # no prolongation: the cycle is the direct solve
    d = pymg.MGSolver(M, [])
    x = d.vcycle(b, np.zeros(len(b)), 1, lambda l: None)
    np.testing.assert_allclose(M @ x, b, rtol=1e-10, atol=1e-12)


def test_relaxation_decreases_energy():
    """Each Gauss-Seidel step exactly minimises E(x) = x^T M x / 2 - b^T x along e_i."""
    M, Ps, b, cols = mk.build(mk.CASES["mg_hat3"])
    st, idx = cols[-1]
    mc = [idx[st[k]:st[k + 1]] for k in range(len(st) - 1)]
    x = np.zeros(len(b))
    E = lambda v: 0.5 * v @ (M @ v) - b @ v  # noqa: E731
    e0 = E(x)
    for _ in range(3):
        x = pymg.GSRelaxer(M)(b, x, mc)
        e1 = E(x)
        assert e1 < e0
        e0 = e1
