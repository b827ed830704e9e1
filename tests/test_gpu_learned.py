"""What the trained network gains: time-to-energy on the 4D Hopf torus at depth 6.

Workload: the Hopf torus of the spherical curve theta(s) = pi/2 + 0.3 sin(3 s),
phi = s (samples.hopf_torus(256, lobes=3, amp=0.3), 65 536 samples), dense
grid, depth 6 (17.9 M nodes), default parameters.  E_ref is the plain
solver's final FP64 energy (the reference's cascade from its random start at
level 0); the target is E_ref (1 + 1e-6).

Both forms run the same harness (learned.time_to_energy): the reference's
schedule on the levels below the finest, then finest-level iterations; a
first run checks the energy after each to find the count k that reaches the
target, a second run with exactly k (one energy evaluation at the end, which
must give the same bits) is timed.  The plain form starts at level 0 from the
random guess, the learned form at level 4 from the shipped network's
prediction (exteriorpoissonrecon_amd/weights/unet4d_l4.npz, trained on the
device by `python -m exteriorpoissonrecon_amd.learned` on 2 048 solver pairs
of random Hopf and two-circle tori -- this Hopf torus is not among them).

Bars (deterministic): the learned form reaches the target, with fewer node
updates than the plain form; epr_solve_learned (the reference's schedule from
level 4) ends within the target.  Measured and printed: the wall times of
both (best of three), their ratio, and, for contrast, the plain solver
started from its random guess at level 4 instead (minSolveDepth 4).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _updates(rec, lo, finest_iters):
    D = rec.depth
    n = sum(rec.num_functions(d) * int(np.ceil(2.0 ** (D - d) * 10)) for d in range(lo, D))
    return n + rec.num_functions(D) * finest_iters


def test_learned_start_reaches_the_reference_energy_sooner():
    from exteriorpoissonrecon_amd import Reconstructor, learned, samples
    assert os.path.exists(learned.WEIGHTS_4D)
    net = learned.UNet.load(learned.WEIGHTS_4D)
    pos, skew = samples.hopf_torus(256, lobes=3, amp=0.3, latitude="sin")
    print(f"[gpu] Hopf torus, {len(pos)} samples, depth 6", flush=True)
    rec = Reconstructor(pos, skew, depth=6, aRadius=-1).build()
    E_ref = rec.solve().sum()
    target = E_ref * (1 + 1e-6)
    plain = min((learned.time_to_energy(rec, target, None) for _ in range(3)), key=lambda r: r[0] or 1e9)
    lrn = min((learned.time_to_energy(rec, target, net, max_finest=40) for _ in range(3)), key=lambda r: r[0] or 1e9)
    print(f"[gpu] E_ref {E_ref:.12g}; plain {plain[0]:.3f} s ({plain[1]} finest iterations); "
          f"learned {lrn[0]} s ({lrn[1]} finest iterations, E {lrn[2]:.12g})", flush=True)
    assert plain[0] is not None and plain[2] <= target
    assert lrn[0] is not None and lrn[2] <= target
    u_plain, u_lrn = _updates(rec, 0, plain[1]), _updates(rec, 4, lrn[1])
    print(f"[gpu] node updates: plain {u_plain:.4e} learned {u_lrn:.4e}; wall-time ratio {plain[0] / lrn[0]:.3f}", flush=True)
    assert u_lrn < u_plain
    E_l = learned.solve_learned(rec, net).sum()
    print(f"[gpu] epr_solve_learned E {E_l:.12g} (target {target:.12g})", flush=True)
    assert E_l <= target
    r4 = Reconstructor(pos, skew, depth=6, aRadius=-1, minSolveDepth=4).build()
    E4 = r4.solve().sum()
    print(f"[gpu] for contrast, the random start at level 4: E {E4:.12g} ({'reaches' if E4 <= target else 'misses'} the target)", flush=True)
    r4.close()
    rec.close()
    net.close()
