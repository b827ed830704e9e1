"""A small program for hardware-counter runs of one kernel family (no
reference counterpart): the batched workload's dominant kernel, the colour
grids of the finest level of its depth-6 inputs, on one host thread and one
stream, next to the KCopy calibration kernel that the counter summaries
(profsum traffic) scale by.

    python -m exteriorpoissonrecon_amd.kernel_probe batch3d_256 [--fp64] [--iters K]

Builds the 256 inputs of samples.batch_workload (one build thread), launches
epr_time_copy (1 GiB read + 1 GiB written, 5 times after a warm-up), then K
iterations of the finest level's colour grids over every depth-6 input
(epr_time_batch_sweeps; FP32 on the float copies unless --fp64).  Prints the
per-launch time.
"""
import argparse
import ctypes
import sys
import time

from . import Reconstructor, build_batch, samples, time_batch_sweeps
from ._lib import load

COPY_BYTES = 1 << 30


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["batch3d_256"])
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--iters", type=int, default=2)
    a = ap.parse_args(argv)
    fp32 = 0 if a.fp64 else 1
    inputs = samples.batch_workload(256)
    t0 = time.perf_counter()
    recs = build_batch([Reconstructor(p, s, depth=d, aRadius=-1, fp32Sweeps=fp32) for p, s, d, _ in inputs], 1)
    print(f"[probe] built {len(recs)} inputs in {time.perf_counter() - t0:.2f} s", flush=True)
    ms = ctypes.c_double()
    if load().epr_time_copy(COPY_BYTES, 5, ctypes.byref(ms)) != 0:
        sys.exit("epr_time_copy failed")
    print(f"[probe] copy kernel {2 * COPY_BYTES / (ms.value * 1e-3) / 1e9:.1f} GB/s", flush=True)
    top = max(d for _, _, d, _ in inputs)
    deep = [r for r, (_, _, d, _) in zip(recs, inputs) if d == top]
    t, launches, nodes = time_batch_sweeps(deep, top, fp32=bool(fp32), iters=a.iters)
    print(f"[probe] KGsPhaseBatch32<3, {'float' if fp32 else 'double'}>: {launches} launches over {len(deep)} inputs, "
          f"{t * 1e3 / launches:.1f} us per launch ({nodes / 8:.0f} nodes)", flush=True)
    for r in recs:
        r.close()


if __name__ == "__main__":
    main()
