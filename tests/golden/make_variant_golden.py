"""Writes the solver-variant fixtures (tests/golden/variant_*.npz) with the reference build.

Runs oracle/_ref/ExteriorPoissonRecon (the reference command line compiled by
`make -C oracle ref`) with --separate (SeparateGaussSeidelSolve,
ExteriorPoissonSolvers.h:198-250) or --jitter (ExteriorPoissonRecon.cpp:174-180)
on synthetic inputs and stores the inputs, the arguments, the final energy it
prints and what its .tree holds (voxel-to-world, hierarchical x and y per
depth, scattered into the dense layout for adapted runs).  Run here only:

    python tests/golden/make_variant_golden.py
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from exteriorpoissonrecon_amd import samples  # noqa: E402
from exteriorpoissonrecon_amd.treefile import TreeFile  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ExteriorPoissonRecon")

CASES = {
    "variant_knot3_d4_separate": dict(gen=lambda: samples.torus_knot(1024), args=dict(depth=4, aRadius=-1, separate=True)),
    "variant_knot3_d5_separate_adapted": dict(gen=lambda: samples.torus_knot(1024, p=3, q=5, seed=3),
                                              args=dict(depth=5, aRadius=2, separate=True, gsIters=6)),
    "variant_clifford4_d3_separate": dict(gen=lambda: samples.clifford_torus(side=24, ang_deg=2.0, seed=1),
                                          args=dict(depth=3, aRadius=-1, separate=True)),
    "variant_knot3_d4_jitter": dict(gen=lambda: samples.torus_knot(1024, seed=2), args=dict(depth=4, aRadius=-1, jitter=0.05)),
    "variant_clifford4_d3_jitter": dict(gen=lambda: samples.clifford_torus(side=24, ang_deg=2.0, seed=1),
                                        args=dict(depth=3, aRadius=1, jitter=0.02)),
}


def run(name, case):
    pos, skew = case["gen"]()
    with tempfile.TemporaryDirectory() as td:
        ply = os.path.join(td, "in.ply")
        samples.write_ply(ply, pos, skew)
        cmd = [REF, "--in", ply, "--out", os.path.join(td, "o")]
        for k, v in case["args"].items():
            cmd += ["--" + k] if v is True else ["--" + k, str(v)]
        txt = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
        t = TreeFile(os.path.join(td, "o.tree"))
    m = re.search(r"->\s*(\S+)\s*=\s*(\S+)\s*\+\s*(\S+)", txt)
    out = {"in_pos": pos, "in_skew": skew, "voxel_to_world": t.voxel_to_world,
           "printed_E": np.array([float(m.group(1)), float(m.group(2)), float(m.group(3))])}
    for d in range(len(t.x)):
        out[f"x{d}"], out[f"y{d}"] = t.dense(d, t.x[d]), t.dense(d, t.y[d])
    for k, v in case["args"].items():
        out["arg_" + k] = np.array(v)
    path = os.path.join(os.path.dirname(__file__), name + ".npz")
    np.savez_compressed(path, **out)
    print(name, out["printed_E"], os.path.getsize(path), "bytes")


if __name__ == "__main__":
    for name, case in CASES.items():
        if len(sys.argv) == 1 or name in sys.argv[1:]:
            run(name, case)
