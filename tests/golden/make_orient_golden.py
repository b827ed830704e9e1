"""Writes the orientation fixtures (tests/golden/orient_cases.npz) with the reference build.

Runs oracle/_ref/ref_orient (MarchingSimplices::Orient compiled from the
reference's header by `make -C oracle ref`) on simplicial complexes with known
topology, their simplices shuffled and some flipped at random:

  torus      triangulated torus grid (closed, orientable)
  mobius     triangulated Moebius strip (open, not orientable)
  klein      triangulated Klein bottle (closed, not orientable)
  mixed      a torus, a Moebius strip and a disc, interleaved
  curves     closed and open polylines (segments)

Stores every input and the reference's output.  Run here only:

    python tests/golden/make_orient_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_orient")


def grid(n, m, wrap_i, twist_i, wrap_j, off=0):
    """Triangles of an n x m vertex grid; i wraps (optionally reversing j), j wraps optionally."""
    def v(i, j):
        if j == m:
            j = 0
        if i == n:
            i, j = 0, (m - 1 - j if twist_i else j)
        return off + i * m + j
    T = []
    for i in range(n if wrap_i else n - 1):
        for j in range(m if wrap_j else m - 1):
            a, b, c, d = v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)
            T += [(a, b, c), (a, c, d)]
    return np.array(T, dtype=np.uint32)


def disc(k, off):
    c = off + k
    return np.array([(off + i, off + (i + 1) % k, c) for i in range(k)], dtype=np.uint32)


def scramble(S, rng):
    S = S.copy()
    flip = rng.random(len(S)) < 0.4
    S[flip, 0], S[flip, 1] = S[flip, 1], S[flip, 0].copy()
    return S[rng.permutation(len(S))]


def curves(rng):
    segs, off = [], 0
    for n, closed in ((7, True), (12, False), (5, True), (9, False)):
        for i in range(n if closed else n - 1):
            segs.append((off + i, off + (i + 1) % n))
        off += n
    return np.array(segs, dtype=np.uint32)


def run(S):
    with tempfile.TemporaryDirectory() as td:
        a, b = os.path.join(td, "in"), os.path.join(td, "out")
        with open(a, "wb") as f:
            f.write(np.uint32(S.shape[1]).tobytes() + np.uint64(len(S)).tobytes() + S.astype("<u4").tobytes())
        subprocess.run([REF, a, b], check=True)
        return np.fromfile(b, "<u4").reshape(S.shape)


def main():
    rng = np.random.default_rng(7)
    cases = {
        "torus": grid(9, 7, True, False, True),
        "mobius": grid(10, 4, True, True, False),
        "klein": grid(8, 6, True, True, True),
        "mixed": np.concatenate([grid(6, 5, True, False, True), grid(8, 3, True, True, False, off=30), disc(9, 60)]),
        "curves": curves(rng),
    }
    out = {}
    for name, S in cases.items():
        Si = scramble(S, rng)
        out[name + "_in"] = Si
        out[name + "_out"] = run(Si)
    path = os.path.join(os.path.dirname(__file__), "orient_cases.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
