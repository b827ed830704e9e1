"""Inputs written by the reference's own Sample program
(tests/golden/sample_plys.npz) through the whole GPU path: the PLY as the
command line reads it, and the solve against the oracle's on the same samples
within the 1e-6 solve bar (energy and coefficients)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu
TOL_SOLVE = 1e-6     # the energy and (phi1, phi2) bar of the single-reconstruction solve tests
CLI = os.path.join(ROOT, "exteriorpoissonrecon_amd", "bin", "ExteriorPoissonRecon")


@pytest.mark.parametrize("name,depth", [("knot3_noisy", 5), ("borromean3", 5), ("clifford4_noisy", 3), ("hopf4", 3)])
def test_gpu_solves_reference_sample_inputs(name, depth, tmp_path, oracle_lib):
    from exteriorpoissonrecon_amd import Reconstructor, samples
    print("case", name, depth, flush=True)
    g = np.load(os.path.join(ROOT, "tests", "golden", "sample_plys.npz"))
    ply = str(tmp_path / (name + ".ply"))
    open(ply, "wb").write(g[name].tobytes())
    pos, skew = samples.read_ply(ply)
    r = Reconstructor(pos, skew, depth=depth, aRadius=-1).build()
    E = r.solve()
    so = oracle_lib.Oracle(pos, skew, depth=depth).solve()
    assert rel_err(E, so["final_E"]) < TOL_SOLVE
    scale = np.abs(so["final_x"]).max() + np.abs(so["final_y"]).max()
    hx, hy = r.coefficients(depth)
    assert np.abs(hx - so[f"hx{depth}"]).max() < TOL_SOLVE * scale
    assert np.abs(hy - so[f"hy{depth}"]).max() < TOL_SOLVE * scale
    r.close()
    # the command line reads the reference's file itself
    res = subprocess.run([CLI, "--in", ply, "--depth", str(depth), "--aRadius", "-1"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0 and "[ERROR]" not in res.stderr, res.stderr
    m = re.search(r"Error: \S+ -> (\S+) = (\S+) \+ (\S+)", res.stdout)
    assert m, res.stdout
    assert rel_err([float(m.group(2)), float(m.group(3))], so["final_E"]) < TOL_SOLVE
