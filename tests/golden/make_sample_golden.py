"""Writes tests/golden/sample_plys.npz: PLY files written by the reference's own
Sample program (oracle/_ref/Sample, compiled from the reference sources by
`make -C oracle ref`), for the PLY reader/writer parity tests.

Every case stores the exact bytes the reference wrote and its command line.
The positional, angular and spot noise of Sample.cpp:137-200 use
std::default_random_engine with its default seed, so the files are
reproducible.  Run here only:

    python tests/golden/make_sample_golden.py          # sample_plys.npz
    python tests/golden/make_sample_golden.py extra    # sample_extra.npz
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


# further cases for the generator's byte parity only (tests/test_host.py):
# other parameters, every noise kind, a knot whose gcd is 2, the spherical curve
EXTRA = {
    "knot35_regular_pnoise_snoise": ["--type", "torus_knot:3:5", "--res", "300", "--regular", "--pNoise", "0.01", "--sNoise", "0.05"],
    "knot24_gcd2": ["--type", "torus_knot:2:4", "--res", "256"],
    "spiral4_anoise": ["--type", "spiral:4", "--res", "400", "--aNoise", "10"],
    "borromean_all_noise": ["--type", "borromean_rings", "--res", "700", "--pNoise", "0.002", "--aNoise", "3", "--sNoise", "0.2"],
    "spherical_curve": ["--type", "spherical:1:3:0.5", "--res", "200"],
    "clifford_regular_pnoise": ["--type", "clifford_torus", "--res", "16", "--regular", "--pNoise", "0.01"],
    "hopf2_anoise_snoise": ["--type", "hopf_torus:2:0.4", "--res", "18", "--aNoise", "4", "--sNoise", "0.1"],
}


def write(cases, fname):
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, args in cases.items():
            path = os.path.join(td, name + ".ply")
            subprocess.run([SAMPLE, *args, "--out", path], check=True, capture_output=True)
            out[name] = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
            out[name + "_args"] = np.array(" ".join(args))
            print(name, len(out[name]), "bytes")
    np.savez_compressed(os.path.join(os.path.dirname(__file__), fname), **out)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "extra":
        write(EXTRA, "sample_extra.npz")
    else:
        write(CASES, "sample_plys.npz")


if __name__ == "__main__":
    sys.exit(main())
