"""Many reconstructions at once (epr_build_batch / epr_solve_batch).

Bar: every reconstruction of a batch ends with the coefficients (every level,
hierarchical) and energies its own solve() gives on the same build, bit for
bit -- the batch runs the same kernels on the same data in the same per-node
order.  Builds splat with FP64 atomics, so two builds of one input differ in
the last bits: a batch build is held to a separate build of the same input
within 1e-9 relative in the final energy.  One batch member also matches the
oracle's solve within the solve tolerance of the single-reconstruction tests
(1e-6 relative, the energy bar)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL_SOLVE = 1e-6     # the energy and (phi1, phi2) bar of the single-reconstruction solve tests


def _inputs3():
    from exteriorpoissonrecon_amd import samples
    return [
        (samples.torus_knot(2048, 2, 3, seed=1), dict(depth=4, aRadius=-1)),
        (samples.spiral(3000, 3, seed=2), dict(depth=5, aRadius=-1, seed=7)),
        (samples.torus_knot(4096, 3, 5, seed=3), dict(depth=6, aRadius=-1, separate=1)),
        (samples.spiral(2500, 5, seed=4), dict(depth=3, aRadius=-1)),
        (samples.torus_knot(3000, 2, 5, seed=5), dict(depth=5, aRadius=-1, minSolveDepth=2, gsIters=6)),
        (samples.torus_knot(2048, 3, 4, seed=6), dict(depth=5, aRadius=-1, maxSolveDepth=3)),
        (samples.spiral(2048, 2, seed=7), dict(depth=4, aRadius=-1, singleLevel=1)),
        (samples.torus_knot(4096, 2, 7, seed=8), dict(depth=5, aRadius=3)),
        (samples.spiral(4096, 4, seed=9), dict(depth=6, aRadius=2)),
        (samples.torus_knot(2048, 5, 3, seed=10), dict(depth=4, aRadius=3, iMult=1.5, dWeight=0.01)),
    ]


def _solve_each(recs):
    """solve() of every (built) reconstruction, one after another."""
    out = []
    for r in recs:
        E = r.solve()
        out.append((E, [r.coefficients(d) for d in range(r.depth + 1)]))
    return out


def _separate_builds(inputs):
    from exteriorpoissonrecon_amd import Reconstructor
    out = []
    for (pos, skew), prm in inputs:
        r = Reconstructor(pos, skew, **prm).build()
        out.append(r.solve())
        r.close()
    return out


def _check_same(recs, E, ref):
    for k, (r, (Ek, coeffs)) in enumerate(zip(recs, ref)):
        assert np.array_equal(E[k], Ek), (k, E[k], Ek)
        for d in range(r.depth + 1):
            x, y = r.coefficients(d)
            assert np.array_equal(x, coeffs[d][0]) and np.array_equal(y, coeffs[d][1]), (k, d)


def test_gpu_batch_3d_matches_single_solves():
    """Dense and adapted, mixed depths 3-6, --separate, --singleLevel, min/max
    solve depths, iteration multipliers and Dirichlet weights in one batch."""
    from exteriorpoissonrecon_amd import Reconstructor, build_batch, solve_batch
    inputs = _inputs3()
    print("case batch3d", len(inputs), flush=True)
    recs = [Reconstructor(pos, skew, **prm) for (pos, skew), prm in inputs]
    build_batch(recs, threads=4)
    ref = _solve_each(recs)
    E = solve_batch(recs)
    _check_same(recs, E, ref)
    # solving the same batch again starts from the same random guess
    E2 = solve_batch(recs)
    assert np.array_equal(E, E2)
    for Ek, Es in zip(E, _separate_builds(inputs)):
        assert rel_err(Ek, Es) < 1e-9, (Ek, Es)
    for r in recs:
        r.close()


def test_gpu_batch_4d_matches_single_solves():
    from exteriorpoissonrecon_amd import Reconstructor, build_batch, samples, solve_batch
    inputs = [
        (samples.clifford_torus(side=24, seed=1), dict(depth=3, aRadius=-1)),
        (samples.clifford_torus(side=32, seed=2), dict(depth=4, aRadius=-1)),
        (samples.clifford_torus(side=20, seed=3), dict(depth=2, aRadius=-1, seed=3)),
        (samples.clifford_torus(side=32, seed=4), dict(depth=4, aRadius=2)),
    ]
    print("case batch4d", len(inputs), flush=True)
    recs = build_batch([Reconstructor(pos, skew, **prm) for (pos, skew), prm in inputs])
    ref = _solve_each(recs)
    E = solve_batch(recs)
    _check_same(recs, E, ref)
    for Ek, Es in zip(E, _separate_builds(inputs)):
        assert rel_err(Ek, Es) < 1e-9, (Ek, Es)
    for r in recs:
        r.close()


def test_gpu_batch_member_matches_oracle(oracle_lib):
    from exteriorpoissonrecon_amd import Reconstructor, build_batch, samples, solve_batch
    pos, skew = samples.spiral(4096, 3, seed=11)
    others = [samples.torus_knot(3000, 2, 3, seed=12), samples.spiral(2048, 6, seed=13)]
    print("case batch-oracle", flush=True)
    recs = [Reconstructor(pos, skew, depth=5, aRadius=-1)] + [Reconstructor(p, s, depth=d, aRadius=-1) for (p, s), d in zip(others, (4, 6))]
    build_batch(recs)
    E = solve_batch(recs)
    so = oracle_lib.Oracle(pos, skew, depth=5).solve()
    assert rel_err(E[0], so["final_E"]) < TOL_SOLVE
    scale = np.abs(so["final_x"]).max() + np.abs(so["final_y"]).max()
    for d in range(6):
        hx, hy = recs[0].coefficients(d)
        assert np.abs(hx - so[f"hx{d}"]).max() < TOL_SOLVE * scale
        assert np.abs(hy - so[f"hy{d}"]).max() < TOL_SOLVE * scale
    for r in recs:
        r.close()


def test_gpu_batch_errors():
    from exteriorpoissonrecon_amd import EprError, Reconstructor, build_batch, samples, solve_batch
    a = Reconstructor(*samples.torus_knot(1024, seed=1), depth=3, aRadius=-1)
    b = Reconstructor(*samples.clifford_torus(side=16, seed=1), depth=2, aRadius=-1)
    with pytest.raises(EprError, match="epr_build has not run"):
        solve_batch([a])
    with pytest.raises(EprError, match="appears twice"):
        build_batch([a, a])
    build_batch([a, b])                      # builds of different dimensions may share a batch
    with pytest.raises(EprError, match="different dimensions"):
        solve_batch([a, b])
    with pytest.raises(EprError, match="appears twice"):
        solve_batch([a, a])
    assert solve_batch([]).shape == (0, 2)
    a.close(), b.close()


def test_gpu_batch_fp32_matches_single_solves():
    """Contexts with --fp32 (their float iterations in shared float grids, the
    last iteration of each colour-launch level in FP64) next to FP64 and
    adapted ones: each ends bit for bit where its own solve() ends, and the
    replayed graph of a second call gives the same bits."""
    from exteriorpoissonrecon_amd import Reconstructor, build_batch, samples, solve_batch
    inputs = [
        (samples.torus_knot(4096, 2, 3, seed=21), dict(depth=6, aRadius=-1, fp32Sweeps=1)),
        (samples.spiral(3000, 4, seed=22), dict(depth=5, aRadius=-1, fp32Sweeps=1, seed=5)),
        (samples.torus_knot(4096, 3, 5, seed=23), dict(depth=6, aRadius=-1)),
        (samples.spiral(2048, 3, seed=24), dict(depth=4, aRadius=-1, fp32Sweeps=1, separate=1)),
        (samples.torus_knot(3000, 2, 5, seed=25), dict(depth=5, aRadius=3, fp32Sweeps=1)),
        (samples.torus_knot(2048, 3, 4, seed=26), dict(depth=6, aRadius=-1, fp32Sweeps=1, iMult=1.5)),
    ]
    print("case batch3d-fp32", len(inputs), flush=True)
    recs = build_batch([Reconstructor(pos, skew, **prm) for (pos, skew), prm in inputs])
    ref = _solve_each(recs)
    E = solve_batch(recs)
    _check_same(recs, E, ref)
    E2 = solve_batch(recs)
    assert np.array_equal(E, E2)
    _check_same(recs, E2, ref)
    for r in recs:
        r.close()


def test_gpu_batch_follows_rebuilds():
    """New samples of the same depths rebuilt into the same contexts keep their
    buffers, so the recorded graph is replayed; it must follow the new
    operators, target constants and random starts: equal to single solves
    after the rebuild."""
    from exteriorpoissonrecon_amd import Reconstructor, build_batch, samples, solve_batch
    first = [samples.torus_knot(2048, 2, 3, seed=31), samples.spiral(3000, 3, seed=32), samples.torus_knot(3000, 3, 4, seed=33)]
    second = [samples.spiral(2500, 5, seed=41), samples.torus_knot(4096, 2, 5, seed=42), samples.spiral(2048, 2, seed=43)]
    prms = [dict(depth=4, aRadius=-1), dict(depth=5, aRadius=-1, fp32Sweeps=1), dict(depth=5, aRadius=3)]
    print("case batch-rebuild", flush=True)
    recs = build_batch([Reconstructor(p, s, **prm) for (p, s), prm in zip(first, prms)])
    E1 = solve_batch(recs)
    for r, (p, s) in zip(recs, second):
        r.set_samples(p, s)
    build_batch(recs)
    ref = _solve_each(recs)
    E = solve_batch(recs)
    _check_same(recs, E, ref)
    assert not np.array_equal(E, E1)
    for r in recs:
        r.close()


def test_gpu_autotune_plans():
    """The autotuners time every candidate plan of every level and keep one per
    level; solves under the tuned plans stay within the solve bar of the
    built-in plans' (the kernels differ only in their summation order), the
    table's text form round-trips, and clearing it restores the built-in
    plans bit for bit (the batch records a new graph when a plan changes)."""
    from exteriorpoissonrecon_amd import (Reconstructor, autotune_batch, build_batch, samples, set_tune_table, solve_batch,
                                          tune_table)
    set_tune_table(None)
    assert tune_table() == ""
    print("case autotune", flush=True)
    recs = build_batch([Reconstructor(*samples.torus_knot(4096, 2, 3, seed=51), depth=6, aRadius=-1, fp32Sweeps=1),
                        Reconstructor(*samples.spiral(3000, 3, seed=52), depth=5, aRadius=-1),
                        Reconstructor(*samples.torus_knot(3000, 3, 5, seed=53), depth=4, aRadius=3)])
    E0 = solve_batch(recs)
    S0 = recs[0].solve()
    try:
        trials = recs[0].autotune(fp32=True)
        levels = sorted({t[0] for t in trials})
        assert levels == list(range(7))
        for l in levels:
            assert sum(1 for t in trials if t[0] == l and t[3] == "fp64" and t[-1]) == 1
            assert all(t[6] > 0 for t in trials if t[0] == l)
        assert any(t[3] == "fp32" and t[-1] for t in trials)
        S1 = recs[0].solve()
        assert rel_err(S1, S0) < TOL_SOLVE, (S1, S0)
        bt = autotune_batch(recs, fp32=True)
        assert sorted({t[0] for t in bt}) == list(range(7))
        for l in range(7):
            assert sum(1 for t in bt if t[0] == l and t[4] == "fp64" and t[-1]) == 1
        E1 = solve_batch(recs)
        assert rel_err(E1, E0) < TOL_SOLVE, (E1, E0)
        text = tune_table()
        lines = text.splitlines()
        assert len(lines) >= 14 and all(len(l.split()) == 7 for l in lines)
        set_tune_table(text)
        assert tune_table() == text
        assert np.array_equal(solve_batch(recs), E1)
        with pytest.raises(Exception, match="bad tuning line"):
            set_tune_table("3 4 0 0 3 96 1.0\n")
        assert tune_table() == text
    finally:
        set_tune_table(None)
    assert np.array_equal(solve_batch(recs), E0)
    assert np.array_equal(recs[0].solve(), S0)
    for r in recs:
        r.close()


def test_gpu_measurement_hooks_leave_data():
    """epr_time_op and epr_time_batch_sweeps time kernels on scratch outputs
    or restore what they sweep: the solved coefficients are unchanged, the
    byte counts are the documented algorithmic ones, and bad requests are
    errors."""
    from exteriorpoissonrecon_amd import EprError, Reconstructor, build_batch, samples, solve_batch, time_batch_sweeps
    print("case hooks", flush=True)
    recs = build_batch([Reconstructor(*samples.torus_knot(3000, 2, 3, seed=61), depth=5, aRadius=-1, fp32Sweeps=1),
                        Reconstructor(*samples.spiral(3000, 4, seed=62), depth=5, aRadius=-1)])
    solve_batch(recs)
    before = [[r.coefficients(d) for d in range(6)] for r in recs]
    r = recs[0]
    N = [r.num_functions(d) for d in range(6)]
    for op, level, expect in (("prolong", 4, 16 * (N[4] + 2 * N[5])), ("galerkin_t", 5, 16 * N[5] * 27),
                              ("galerkin_c", 5, 8 * (N[5] + N[4]) * 27), ("energy", 5, 8 * (2 * N[5] + 2 * N[5] * 27))):
        ms, b = r.time_op(op, level, reps=2)
        assert ms > 0 and b == expect, (op, ms, b, expect)
    for fp32 in (False, True):
        ms, launches, nodes = time_batch_sweeps(recs, 5, fp32=fp32, iters=2)
        assert ms > 0 and launches == 16 and nodes == 2 * N[5]
    for rr, bb in zip(recs, before):
        for d in range(6):
            x, y = rr.coefficients(d)
            assert np.array_equal(x, bb[d][0]) and np.array_equal(y, bb[d][1]), d
    with pytest.raises(EprError, match="unknown operation"):
        r.time_op("bogus", 3)
    with pytest.raises(EprError, match="below the depth"):
        r.time_op("prolong", 5)
    with pytest.raises(EprError, match="no dense context"):
        time_batch_sweeps(recs, 6)
    for rr in recs:
        rr.close()
