"""Writes the golden fixtures of tests/golden/ with the reference build.

Runs oracle/_ref/ref_dump (the reference compiled from /root/reference by
`make -C oracle ref`) on synthetic inputs from exteriorpoissonrecon_amd.samples
and stores inputs and reference outputs as .npz files.  Run here only (the
reference does not exist on GPU machines):

    python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from exteriorpoissonrecon_amd import samples  # noqa: E402
from oracle.pyoracle import load_dump  # noqa: E402

REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")

CASES = {
    # the 1024-sample (2,3) torus knot at reduced depth (full dump)
    "knot3_d4": dict(gen=lambda: samples.torus_knot(1024), args=dict(depth=4)),
    # the same knot at depth 5 (finest arrays subsampled)
    "knot3_d5": dict(gen=lambda: samples.torus_knot(1024), args=dict(depth=5), stride=7,
                     keep=("unit_pos", "cells", "cell_counts", "measure", "noise", "sample_depth", "S_elem", "c",
                           "init_W", "init_x", "init_y", "final_E", "final_x", "final_y", "E5", "W5", "Q5",
                           "gs5_x", "gs5_y", "R5_v", "B5_v", "E0", "W0", "Q0", "gs0_x", "gs0_y")),
    # non-default parameters
    "knot3_d4_params": dict(gen=lambda: samples.torus_knot(1024, p=3, q=5, seed=3),
                            args=dict(depth=4, tRadius=2, dRadius=0, sWeight=50, dWeight=0.01,
                                      gsIters=3, iMult=1.5, seed=7, nSigma=0.3, scale=1.25)),
    # 4D Clifford torus, 576 samples
    "clifford4_d3": dict(gen=lambda: samples.clifford_torus(side=24, ang_deg=2.0, seed=1), args=dict(depth=3)),
}


def run(name, case):
    pos, skew = case["gen"]()
    with tempfile.TemporaryDirectory() as td:
        ply = os.path.join(td, "in.ply")
        samples.write_ply(ply, pos, skew)
        dump = os.path.join(td, "out.bin")
        cmd = [REF_DUMP, "--in", ply, "--dump", dump, "--out", os.path.join(td, "tree")]
        for k, v in case["args"].items():
            cmd += ["--" + k, str(v)]
        subprocess.check_call(cmd, stdout=subprocess.DEVNULL)
        d = load_dump(dump)
        tree = np.frombuffer(open(os.path.join(td, "tree.tree"), "rb").read(), dtype=np.uint8)
    out = {"in_pos": pos, "in_skew": skew, "tree_head": tree[:12].copy()}
    for k, v in case["args"].items():
        out["arg_" + k] = np.array(v)
    stride = case.get("stride")
    depth = case["args"]["depth"]
    keep = case.get("keep")
    for k, v in d.items():
        if keep and k not in keep:
            continue
        if stride and (k in ("final_x", "final_y", f"hx{depth}", f"hy{depth}", f"Q{depth}", f"gs{depth}_x", f"gs{depth}_y",
                             f"R{depth}_v", f"B{depth}_v", f"P{depth - 1}_v", "S_elem")):
            if k.startswith("Q"):
                v = v.reshape(-1, 5)[::stride].ravel()
            elif k == "S_elem":
                v = v.reshape(-1, 3 if pos.shape[1] == 3 else 6)[::stride].ravel()
            else:
                v = v[::stride]
            out[k + "_stride"] = np.array(stride)
        out[k] = v
    np.savez_compressed(os.path.join(os.path.dirname(__file__), name + ".npz"), **out)
    print(name, sum(v.nbytes for v in out.values()) // 1024, "KiB")


if __name__ == "__main__":
    if not os.path.exists(REF_DUMP):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    for name, case in CASES.items():
        if len(sys.argv) > 1 and name not in sys.argv[1:]:
            continue
        run(name, case)
