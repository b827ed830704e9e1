"""Writes the density-file fixtures (tests/golden/density_*.npz) with the reference build.

oracle/_ref/ExteriorPoissonRecon (the reference command line compiled by
`make -C oracle ref`) writes <out>.density.tree beside the .tree
(ExteriorPoissonRecon.cpp:272-277).  A fixture holds the input samples, the
arguments, and that file decoded: the voxel-to-world matrix, every node's
depth, offsets and "has children" flag in file order, its DensityAndNoiseInfo
statistics (weightSum, positionSum, covarianceSum) and the target samples per
cell.  Run here only:

    python tests/golden/make_density_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from exteriorpoissonrecon_amd import samples  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ExteriorPoissonRecon")

CASES = {
    "density_knot3_d4": dict(gen=lambda: samples.torus_knot(2048, seed=5), args=dict(depth=4)),
    "density_knot3_d4_r2": dict(gen=lambda: samples.torus_knot(1024, p=3, q=5, seed=3), args=dict(depth=4, dRadius=2, aRadius=-1)),
    "density_clifford4_d2": dict(gen=lambda: samples.clifford_torus(side=16, ang_deg=2.0, seed=1), args=dict(depth=2)),
}


def decode(path, dim):
    """Decodes a density file: (voxel_to_world, depth, offsets, has_children, stats, target)."""
    b = open(path, "rb").read()
    o = 8
    v2w = np.frombuffer(b, "<f8", (dim + 1) ** 2, o).reshape(dim + 1, dim + 1).copy()
    o += 8 * (dim + 1) ** 2
    head = ((2 * (dim + 1) + 7) // 8) * 8
    rec = head + 16 + 8 * (2 + dim + dim * dim)
    n = (len(b) - o - 8) // rec
    assert o + n * rec + 8 == len(b)
    raw = np.frombuffer(b, np.uint8, n * rec, o).reshape(n, rec)
    h = raw[:, :2 * (dim + 1)].copy().view("<u2").reshape(n, dim + 1)
    ch = raw[:, head + 8:head + 16].copy().view("<u8").ravel()
    data = raw[:, head + 16:].copy().view("<f8").reshape(n, 2 + dim + dim * dim)
    stats = np.concatenate([data[:, :1 + dim], data[:, 2 + dim:]], axis=1)    # the matrix's vtable slot dropped
    target = np.frombuffer(b, "<f8", 1, len(b) - 8)[0]
    return v2w, h[:, 0].astype(np.int16), h[:, 1:].astype(np.int16), (ch != 0).astype(np.int8), stats, target


def run(name, case):
    pos, skew = case["gen"]()
    dim = pos.shape[1]
    with tempfile.TemporaryDirectory() as td:
        ply = os.path.join(td, "in.ply")
        samples.write_ply(ply, pos, skew)
        cmd = [REF, "--in", ply, "--out", os.path.join(td, "o")]
        for k, v in case["args"].items():
            cmd += ["--" + k, str(v)]
        subprocess.run(cmd, check=True, capture_output=True)
        v2w, depth, off, kids, stats, target = decode(os.path.join(td, "o.density.tree"), dim)
    out = {"in_pos": pos, "in_skew": skew, "voxel_to_world": v2w, "node_depth": depth, "node_offset": off,
           "node_has_children": kids, "stats": stats, "target": np.array(target)}
    for k, v in case["args"].items():
        out["arg_" + k] = np.array(v)
    path = os.path.join(os.path.dirname(__file__), name + ".npz")
    np.savez_compressed(path, **out)
    print(name, len(depth), "nodes", os.path.getsize(path), "bytes")


if __name__ == "__main__":
    for name, case in CASES.items():
        if len(sys.argv) == 1 or name in sys.argv[1:]:
            run(name, case)
