"""Writes the pinned energies and coefficients of the benchmark workloads.

Outputs (tests/golden/):
  bench_reference.json   final (E_fit, E_reg) per workload, read by bench.py
  bench_<workload>.npz   the inputs' seed-free description is the generator
                         call in bench.py; this holds the final energy and a
                         strided subsample of the finest level's regular and
                         hierarchical coefficients (x, y), read by
                         tests/test_gpu_headline.py

Sources:
  * borromean3d_d7, knot3d_d5, clifford4d_d5: the reference itself
    (oracle/_ref/ref_dump, compiled from the reference sources by
    `make -C oracle ref`) on the same samples, default parameters, dense grid,
    its thread pool on (--threads; the energies' parallel sums differ from a
    serial run in the last bits only);
  * clifford4d_d6: the CPU oracle (oracle/epr_oracle.cpp, pinned against the
    reference by tests/test_oracle_golden.py), because the reference's Eigen
    matrices at 4D depth 6 do not fit this machine's memory.

Run here only (the reference is not on GPU machines):

    python tests/golden/make_bench_reference.py [workload ...]
"""
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from exteriorpoissonrecon_amd import samples  # noqa: E402
from oracle.pyoracle import Oracle, load_dump  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
OUT_JSON = os.path.join(HERE, "bench_reference.json")

# the generators of bench.py's workloads (same calls, same samples)
WORKLOADS = {
    "borromean3d_d7": (lambda: samples.borromean(16384), 7, "reference"),
    "knot3d_d5": (lambda: samples.torus_knot(1024), 5, "reference"),
    "clifford4d_d5": (lambda: samples.clifford_torus(256, ang_deg=2.0), 5, "reference"),
    "clifford4d_d6": (lambda: samples.clifford_torus(256, ang_deg=2.0), 6, "oracle"),
}
SUBSAMPLE = 4099   # coefficients kept per vector (a prime stride over the finest level)


def run_reference(pos, skew, depth):
    with tempfile.TemporaryDirectory() as td:
        ply = os.path.join(td, "in.ply")
        samples.write_ply(ply, pos, skew)
        dump = os.path.join(td, "out.bin")
        subprocess.check_call([REF_DUMP, "--in", ply, "--dump", dump, "--depth", str(depth), "--threads"],
                              stdout=subprocess.DEVNULL)
        d = load_dump(dump)
    return {"final_E": d["final_E"], "final_x": d["final_x"], "final_y": d["final_y"],
            "hx": d[f"hx{depth}"], "hy": d[f"hy{depth}"]}


def run_oracle(pos, skew, depth):
    r = Oracle(pos, skew, depth=depth).solve()
    return {"final_E": r["final_E"], "final_x": r["final_x"], "final_y": r["final_y"],
            "hx": r[f"hx{depth}"], "hy": r[f"hy{depth}"]}


def main(names):
    pinned = json.load(open(OUT_JSON)) if os.path.exists(OUT_JSON) else {}
    for name in names:
        gen, depth, source = WORKLOADS[name]
        pos, skew = gen()
        t0 = time.time()
        r = (run_reference if source == "reference" else run_oracle)(pos, skew, depth)
        dt = time.time() - t0
        n = len(r["final_x"])
        stride = max(1, n // SUBSAMPLE) | 1
        out = {"final_E": r["final_E"], "stride": np.array(stride), "depth": np.array(depth),
               "source": np.array(source), "seconds": np.array(dt)}
        for k in ("final_x", "final_y", "hx", "hy"):
            out[k] = r[k][::stride].copy()
        np.savez_compressed(os.path.join(HERE, f"bench_{name}.npz"), **out)
        pinned[name] = [float(v) for v in r["final_E"]]
        json.dump(pinned, open(OUT_JSON, "w"), indent=1)
        print(f"{name}: E={r['final_E'].tolist()} ({source}, {dt:.1f} s, {n} nodes, stride {stride})", flush=True)


if __name__ == "__main__":
    if not os.path.exists(REF_DUMP):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    main(sys.argv[1:] or list(WORKLOADS))
