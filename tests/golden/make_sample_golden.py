"""Writes tests/golden/sample_plys.npz: PLY files written by the reference's own
Sample program (oracle/_ref/Sample, compiled from the reference sources by
`make -C oracle ref`), for the PLY reader/writer parity tests.

Every case stores the exact bytes the reference wrote and its command line.
The positional, angular and spot noise of Sample.cpp:137-200 use
std::default_random_engine with its default seed, so the files are
reproducible.  Run here only:

    python tests/golden/make_sample_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
SAMPLE = os.path.join(ROOT, "oracle", "_ref", "Sample")

CASES = {
    "knot3_noisy": ["--type", "torus_knot:2:3", "--res", "512", "--aNoise", "5", "--pNoise", "0.005", "--sNoise", "0.1"],
    "borromean3": ["--type", "borromean_rings", "--res", "600"],
    "spiral3_regular": ["--type", "spiral:3", "--res", "256", "--regular"],
    "clifford4_noisy": ["--type", "clifford_torus", "--res", "24", "--aNoise", "2"],
    "hopf4": ["--type", "hopf_torus:3:0.3", "--res", "20"],
}


def main():
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, args in CASES.items():
            path = os.path.join(td, name + ".ply")
            subprocess.run([SAMPLE, *args, "--out", path], check=True, capture_output=True)
            out[name] = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
            out[name + "_args"] = np.array(" ".join(args))
            print(name, len(out[name]), "bytes")
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "sample_plys.npz"), **out)


if __name__ == "__main__":
    sys.exit(main())
