"""The CPU oracle against vectors written by the reference itself.

Fixtures (tests/golden/*.npz) come from the reference compiled from its own
sources (oracle/ref_dump.cpp driving ExteriorPoisson::Reconstructor, serial),
see tests/golden/make_golden.py.  Tolerances: the oracle reproduces the
reference's summation order, so most quantities agree to the last bit; the
noise estimate uses a different eigen-solver (Jacobi vs Eigen) and differs at
~1e-12, which then enters R and everything after it at ~1e-15 relative.
"""
import numpy as np
import pytest

from conftest import CASES, case_args, load_golden, rel_err


@pytest.fixture(scope="module", params=CASES)
def case(request, oracle_lib):
    g = load_golden(request.param)
    o = oracle_lib.Oracle(g["in_pos"], g["in_skew"], **case_args(g))
    return request.param, g, o


def _strided(g, key, arr, width=1):
    s = g.get(key + "_stride")
    if s is None:
        return arr
    s = int(s)
    return arr.reshape(-1, width)[::s].ravel() if width > 1 else arr[::s]


def test_binning_exact(case):
    name, g, o = case
    np.testing.assert_array_equal(o.get("unit_pos"), g["unit_pos"])
    np.testing.assert_array_equal(o.get("cells"), g["cells"])
    np.testing.assert_array_equal(o.get("cell_counts"), g["cell_counts"])


def test_estimator(case):
    name, g, o = case
    assert rel_err(o.get("measure"), g["measure"]) < 1e-13
    assert rel_err(o.get("sample_depth"), g["sample_depth"]) < 1e-13
    assert rel_err(o.get("noise"), g["noise"]) < 1e-9


def test_target_field(case):
    name, g, o = case
    dim = o.dim
    cd = 3 if dim == 3 else 6
    S = _strided(g, "S_elem", o.get("S_elem"), cd)
    assert rel_err(S, g["S_elem"]) < 1e-12
    assert rel_err(o.get("c"), g["c"]) < 1e-12


def test_operators(case, oracle_lib):
    name, g, o = case
    D = o.depth
    for d in range(D + 1):
        if f"R{d}_v" not in g:
            continue
        v = oracle_lib.probe(o.nfun(d), 0.25 + d)
        assert rel_err(_strided(g, f"R{d}_v", o.matvec(d, "R", v)), g[f"R{d}_v"]) < 1e-11, d
        assert rel_err(_strided(g, f"B{d}_v", o.matvec(d, "B", v)), g[f"B{d}_v"]) < 1e-12, d
        if f"P{d}_v" in g:
            np.testing.assert_array_equal(_strided(g, f"P{d}_v", o.matvec(d, "P", v)), g[f"P{d}_v"])


def test_energies_fits_sweeps(case, oracle_lib):
    name, g, o = case
    D = o.depth
    for d in range(D + 1):
        if f"E{d}" not in g:
            continue
        N = o.nfun(d)
        x = oracle_lib.probe(N, 1.5 + d) * 0.01
        y = oracle_lib.probe(N, 2.75 + 0.5 * d) * 0.01
        assert rel_err(o.energies(d, x, y), g[f"E{d}"]) < 1e-12, d
        assert rel_err([o.wedge(d, x, y)], g[f"W{d}"]) < 1e-13, d
        if f"Q{d}" in g:
            assert rel_err(_strided(g, f"Q{d}", o.qfit(d, x, y).ravel(), 5), g[f"Q{d}"]) < 1e-10, d
            gx, gy = o.gs(d, x, y, 1)
            assert rel_err(_strided(g, f"gs{d}_x", gx), g[f"gs{d}_x"]) < 1e-10, d
            assert rel_err(_strided(g, f"gs{d}_y", gy), g[f"gs{d}_y"]) < 1e-10, d


def test_init_and_solve(case):
    name, g, o = case
    x, y = o.init()
    np.testing.assert_array_equal(x, g["init_x"])
    np.testing.assert_array_equal(y, g["init_y"])
    s = o.solve()
    assert rel_err(s["final_E"], g["final_E"]) < 1e-10
    assert rel_err(_strided(g, "final_x", s["final_x"]), g["final_x"]) < 1e-8
    assert rel_err(_strided(g, "final_y", s["final_y"]), g["final_y"]) < 1e-8
    for d in range(o.depth + 1):
        if f"hx{d}" in g:
            assert rel_err(_strided(g, f"hx{d}", s[f"hx{d}"]), g[f"hx{d}"]) < 1e-8
