"""Writes the multigrid fixtures (tests/golden/mg_*.npz) with the reference build.

Runs oracle/_ref/ref_mg (the reference's Include/MG.h MGSolver and GSRelaxer,
compiled by `make -C oracle ref`) on hat-basis systems built here: mass plus
(weighted) stiffness of the multilinear hat basis on a regular grid, the
hat prolongations between grids, parity colour lists per level.  Run here only:

    python tests/golden/make_mg_golden.py
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np
import scipy.sparse as sp

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from oracle.pyoracle import load_dump  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_mg")


def hat_1d(res):
    h = 1.0 / res
    n = res + 1
    m = sp.diags([np.full(n - 1, h / 6), np.r_[h / 3, np.full(n - 2, 2 * h / 3), h / 3], np.full(n - 1, h / 6)], [-1, 0, 1])
    k = sp.diags([np.full(n - 1, -1 / h), np.r_[1 / h, np.full(n - 2, 2 / h), 1 / h], np.full(n - 1, -1 / h)], [-1, 0, 1])
    return m.tocsc(), k.tocsc()


def prolong_1d(res):
    """(res+1) x (res/2+1): coarse node i -> fine 2i (1) and 2i +- 1 (1/2)."""
    rows, cols, vals = [], [], []
    for i in range(res // 2 + 1):
        for o, w in ((-1, 0.5), (0, 1.0), (1, 0.5)):
            f = 2 * i + o
            if 0 <= f <= res:
                rows.append(f), cols.append(i), vals.append(w)
    return sp.csc_matrix((vals, (rows, cols)), shape=(res + 1, res // 2 + 1))


def kron_all(mats):
    out = mats[0]
    for m in mats[1:]:
        out = sp.kron(out, m, format="csc")
    return sp.csc_matrix(out)


def system(dim, res, weights):
    m, k = hat_1d(res)
    M = kron_all([m] * dim)
    for d in range(dim):
        M = M + weights[d] * kron_all([k if e == d else m for e in range(dim)])
    M = sp.csc_matrix(M)
    M.sort_indices()
    return M


def colours(dim, res):
    F = np.indices((res + 1,) * dim).reshape(dim, -1).T
    c = np.zeros(len(F), dtype=np.int64)
    for d in range(dim):
        c |= (F[:, d] & 1) << d
    lists = [np.nonzero(c == k)[0] for k in range(1 << dim)]
    starts = np.cumsum([0] + [len(x) for x in lists])
    return starts.astype(np.uint32), np.concatenate(lists).astype(np.uint32)


CASES = {
    "mg_hat3": dict(dim=3, res=8, levels=2, weights=(1.0, 1.0, 1.0), iters=2, cycles=4),
    "mg_aniso3": dict(dim=3, res=16, levels=3, weights=(1.0, 10.0, 0.1), iters=3, cycles=3),
    "mg_hat4": dict(dim=4, res=4, levels=2, weights=(1.0, 1.0, 1.0, 1.0), iters=2, cycles=3),
}


def csc_bytes(A):
    A = sp.csc_matrix(A)
    A.sort_indices()
    return (struct.pack("<I", A.nnz) + A.indptr.astype(np.uint32).tobytes() + A.indices.astype(np.uint32).tobytes()
            + A.data.astype(np.float64).tobytes())


def build(case, seed=0):
    dim, res, L = case["dim"], case["res"], case["levels"]
    M = system(dim, res, case["weights"])
    Ps = []
    for i in range(L):                      # Ps[i]: level i -> level i+1, Ps[L-1] ends at M's grid
        r = res >> (L - 1 - i)
        Ps.append(kron_all([prolong_1d(r)] * dim))
    b = np.random.default_rng(seed).normal(size=M.shape[0])
    cols = [colours(dim, res >> (L - l)) for l in range(1, L + 1)]   # level l = 1..L
    return M, Ps, b, cols


def run(name, case):
    M, Ps, b, cols = build(case)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<IIII", M.shape[0], len(Ps), case["iters"], case["cycles"]))
            f.write(csc_bytes(M))
            for P in Ps:
                f.write(struct.pack("<II", *P.shape))
                f.write(csc_bytes(P))
            f.write(b.astype(np.float64).tobytes())
            for st, idx in cols:
                f.write(struct.pack("<II", len(idx), len(st) - 1))
                f.write(st.tobytes() + idx.tobytes())
        subprocess.check_call([REF, fin, fout])
        d = load_dump(fout)
    out = {"dim": np.array(case["dim"]), "res": np.array(case["res"]), "levels": np.array(case["levels"]),
           "iters": np.array(case["iters"]), "cycles": np.array(case["cycles"]), "b": b}
    out.update(d)
    np.savez_compressed(os.path.join(os.path.dirname(__file__), name + ".npz"), **out)
    print(name, M.shape[0], "unknowns;", "residuals", [float(d[f"res{c}"][0]) for c in range(case["cycles"])])


if __name__ == "__main__":
    if not os.path.exists(REF):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    for name, case in CASES.items():
        if len(sys.argv) > 1 and name not in sys.argv[1:]:
            continue
        run(name, case)
