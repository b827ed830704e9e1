"""--fp32 (north-star pattern 1): FP32 Gauss-Seidel iterations with an FP64
last iteration per level, FP64 energies.

Bar: the final energy within 1e-6 relative of the reference's (3D Borromean
depth 7, 4D Clifford depth 5: the reference's own runs; 4D depth 6: the CPU
oracle, when its fixture exists) -- the metric's accuracy bar -- and the
finest regular and hierarchical coefficients within 2e-4 relative of the
pinned run.  The coefficients carry the FP32 iterate's rounding (~1e-7
relative per step) along the non-convex trajectory, and most of it lies along
the energy's flat directions (a rotation or SL(2) drift of (x, y) that no
FP64 iteration corrects); measured 2e-5 (3D depth 7) against the reference
and 8e-5 against the FP64 path, while the energy stays within 1.2e-7.  The
FP64 path keeps the 1e-6 coefficient bar (tests/test_gpu_headline.py).  The
per-result error against the FP64 path on the same build is printed.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
TOL_E = 1e-6
TOL_X = 2e-4


def _fixture(name):
    path = os.path.join(GOLDEN, f"bench_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"no pinned run for {name}")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", ["borromean3d_d7", "clifford4d_d5", "clifford4d_d6"])
def test_fp32_sweeps_reach_the_pinned_energy(name):
    from exteriorpoissonrecon_amd import Reconstructor, samples
    from exteriorpoissonrecon_amd.learned import regular_coefficients
    g = _fixture(name)
    if name.startswith("borromean"):
        (pos, skew), depth = samples.borromean(16384), 7
    else:
        (pos, skew), depth = samples.clifford_torus(256, ang_deg=2.0), int(name[-1])
    print(f"[gpu] {name} --fp32", flush=True)
    out = {}
    for fp32 in (0, 1):
        r = Reconstructor(pos, skew, depth=depth, aRadius=-1, fp32Sweeps=fp32).build()
        E = r.solve()
        fx, fy = regular_coefficients(r, depth)
        hx, hy = r.coefficients(depth)
        out[fp32] = (E, fx, fy, hx, hy)
        r.close()
    E, fx, fy, hx, hy = out[1]
    s = int(g["stride"])
    relE = abs(E.sum() - g["final_E"].sum()) / abs(g["final_E"].sum())
    scale = max(np.abs(g["final_x"]).max(), np.abs(g["final_y"]).max())
    errX = max(np.abs(a[::s] - b).max() for a, b in ((fx, g["final_x"]), (fy, g["final_y"]), (hx, g["hx"]), (hy, g["hy"]))) / scale
    E64, fx64, fy64 = out[0][:3]
    print(f"[gpu]   E {E.sum():.12g} pinned {g['final_E'].sum():.12g} rel {relE:.2e}; coefficients {errX:.2e}; "
          f"against FP64 on the same build: E {abs(E.sum() - E64.sum()) / E64.sum():.2e}, "
          f"coefficients {max(np.abs(fx - fx64).max(), np.abs(fy - fy64).max()) / scale:.2e}", flush=True)
    assert relE < TOL_E
    assert errX < TOL_X
