"""Writes the level-set fixtures (tests/golden/ls_*.npz) with the reference build.

Runs oracle/_ref/ref_levelset (MarchingSimplices::RegularSubGridTriangulation
and MarchingSimplices::LevelSet of the reference, compiled by
`make -C oracle ref`) on node-value grids and cell lists generated here, and
stores inputs and reference outputs.  Run here only:

    python tests/golden/make_levelset_golden.py

Cases: analytic fields whose joint zero set is a circle (3D) or a 2-sphere (4D)
on the cells whose corner values straddle zero in both components (in grid
order and shuffled, since vertex numbering follows the cell order); fields with
exact zeros on grid nodes and edges (the degenerate paths of the barycentric
snapping); uniform random fields on every cell.
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from oracle.pyoracle import load_dump  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_levelset")


def node_coords(dim, res):
    """(res+1)^dim x dim node coordinates, coordinate 0 most significant."""
    g = np.indices((res + 1,) * dim).reshape(dim, -1).T
    return g.astype(np.float64)


def straddling_cells(dim, res, vals):
    """Cells whose corner values contain zero in both components, coordinate 0 fastest."""
    V = vals.reshape((res + 1,) * dim + (2,))
    cells = []
    for e in np.ndindex(*((res,) * dim)):
        E = e[::-1]                                   # coordinate 0 fastest
        sl = tuple(slice(E[d], E[d] + 2) for d in range(dim))
        c = V[sl].reshape(-1, 2)
        if (c.min(0) <= 0).all() and (c.max(0) >= 0).all():
            cells.append(E)
    return np.array(cells, dtype=np.uint32).reshape(-1, dim)


def all_cells(dim, res):
    return np.array([e[::-1] for e in np.ndindex(*((res,) * dim))], dtype=np.uint32).reshape(-1, dim)


def sphere_field(dim, res, c, r, n):
    p = node_coords(dim, res)
    n = np.array(n, dtype=np.float64)
    n /= np.linalg.norm(n)
    q = p - np.array(c)
    return np.stack([(q * q).sum(1) - r * r, q @ n], 1)


def cases():
    rng = np.random.default_rng(2024)
    out = {}
    v = sphere_field(3, 16, (7.3, 8.1, 7.7), 5.2, (0.3, 0.5, 0.8))
    cells = straddling_cells(3, 16, v)
    out["ls_sphere3"] = (3, 16, cells, v)
    out["ls_sphere3_shuffled"] = (3, 16, cells[rng.permutation(len(cells))], v)
    p = node_coords(3, 8)
    out["ls_grid3"] = (3, 8, all_cells(3, 8), np.stack([p[:, 0] - 3, p[:, 1] - 5], 1))
    p6 = node_coords(3, 6)
    out["ls_diag3"] = (3, 6, all_cells(3, 6), np.stack([p6[:, 0] + p6[:, 1] - 3, p6[:, 2] - p6[:, 1]], 1))
    out["ls_noise3"] = (3, 6, all_cells(3, 6), rng.uniform(-1, 1, size=(7 ** 3, 2)))
    v4 = sphere_field(4, 8, (3.6, 4.2, 3.9, 4.1), 2.9, (0.2, 0.4, 0.1, 0.9))
    out["ls_sphere4"] = (4, 8, straddling_cells(4, 8, v4), v4)
    p4 = node_coords(4, 4)
    out["ls_grid4"] = (4, 4, all_cells(4, 4), np.stack([p4[:, 0] - 2, p4[:, 1] + p4[:, 2] - 3], 1))
    out["ls_noise4"] = (4, 3, all_cells(4, 3), rng.uniform(-1, 1, size=(4 ** 4, 2)))
    return out


def run(name, dim, res, cells, vals):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<IIQ", dim, res, len(cells)))
            f.write(np.ascontiguousarray(cells, dtype=np.uint32).tobytes())
            f.write(np.ascontiguousarray(vals, dtype=np.float64).tobytes())
        subprocess.check_call([REF, fin, fout])
        d = load_dump(fout)
    res_ = {"dim": np.array(dim), "res": np.array(res), "cells": cells, "node_values": vals}
    res_.update(d)
    np.savez_compressed(os.path.join(os.path.dirname(__file__), name + ".npz"), **res_)
    print(name, len(cells), "cells,", len(d["tri_vertices"]) // dim, "mesh vertices,",
          len(d["ls_vertices"]) // dim, "level-set vertices,", len(d["ls_simplices"]) // (dim - 1), "simplices")


if __name__ == "__main__":
    if not os.path.exists(REF):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    for name, args in cases().items():
        if len(sys.argv) > 1 and name not in sys.argv[1:]:
            continue
        run(name, *args)
