"""Writes the adapted-hierarchy fixtures (tests/golden/adapted_*.npz) with the reference build.

Runs oracle/_ref/ExteriorPoissonRecon (the reference command line compiled by
`make -C oracle ref`) with --aRadius K on synthetic inputs from
exteriorpoissonrecon_amd.samples and stores the inputs, the arguments and what
its .tree holds: the hierarchical (x, y) per depth in the hierarchy's
numbering, and the serialised HierarchicalAdaptedIndexer (every tree node's
depth, offset, children flag, function and element index; the active counts;
the per-depth function and element counts).  Run here only:

    python tests/golden/make_adapted_golden.py
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
    # the reference's default radius on the 1024-sample torus knot at depths 4 and 5
    "adapted_knot3_d4": dict(gen=lambda: samples.torus_knot(1024), args=dict(depth=4, aRadius=3)),
    "adapted_knot3_d5": dict(gen=lambda: samples.torus_knot(1024), args=dict(depth=5, aRadius=3)),
    # small radii: padding and sparse coarse levels
    "adapted_knot3_d5_r0": dict(gen=lambda: samples.torus_knot(1024, p=3, q=5, seed=3),
                                args=dict(depth=5, aRadius=0, gsIters=4, seed=5)),
    "adapted_knot3_d5_r1": dict(gen=lambda: samples.torus_knot(512, seed=2), args=dict(depth=5, aRadius=1, maxSolveDepth=4, gsIters=6, iMult=1.5)),
    # 4D Clifford torus
    "adapted_clifford4_d3": dict(gen=lambda: samples.clifford_torus(side=24, ang_deg=2.0, seed=1), args=dict(depth=3, aRadius=1)),
    "adapted_clifford4_d4": dict(gen=lambda: samples.clifford_torus(side=32, ang_deg=2.0, seed=2), args=dict(depth=4, aRadius=0)),
}


def run(name, case):
    pos, skew = case["gen"]()
    with tempfile.TemporaryDirectory() as td:
        ply = os.path.join(td, "in.ply")
        samples.write_ply(ply, pos, skew)
        cmd = [REF, "--in", ply, "--out", os.path.join(td, "o")]
        for k, v in case["args"].items():
            cmd += ["--" + k, str(v)]
        txt = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
        t = TreeFile(os.path.join(td, "o.tree"), adapted=True)
    m = re.search(r"->\s*(\S+)\s*=\s*(\S+)\s*\+\s*(\S+)", txt)
    out = {"in_pos": pos, "in_skew": skew, "hash": np.array(t.hash, dtype=np.uint64),
           "voxel_to_world": t.voxel_to_world, "node_depth": t.node_depth.astype(np.int16),
           "node_offset": t.node_offset.astype(np.int16), "node_has_children": t.node_has_children.astype(np.int8),
           "node_function": t.node_function.astype(np.int32), "node_element": t.node_element.astype(np.int32),
           "active_sizes": t.active_sizes, "num_functions": t.num_functions, "num_elements": t.num_elements,
           "printed_E": np.array([float(m.group(1)), float(m.group(2)), float(m.group(3))]) if m else np.zeros(3)}
    for k, v in case["args"].items():
        out["arg_" + k] = np.array(v)
    for d in range(len(t.x)):
        out[f"x{d}"] = t.x[d]
        out[f"y{d}"] = t.y[d]
    np.savez_compressed(os.path.join(os.path.dirname(__file__), name + ".npz"), **out)
    print(name, len(t.node_depth), "nodes;", "functions", list(t.num_functions), "active", list(t.active_sizes),
          "elements", list(t.num_elements), "E", out["printed_E"], sum(v.nbytes for v in out.values()) // 1024, "KiB")


if __name__ == "__main__":
    if not os.path.exists(REF):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    for name, case in CASES.items():
        if len(sys.argv) > 1 and name not in sys.argv[1:]:
            continue
        run(name, case)
