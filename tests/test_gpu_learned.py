"""What the trained network gains: time-to-energy on the 4D Hopf torus at depth 6.

Workload: the Hopf torus of the spherical curve theta(s) = pi/2 + 0.3 sin(3 s),
phi = s (samples.hopf_torus(256, lobes=3, amp=0.3), 65 536 samples), dense
grid, depth 6 (17.9 M nodes), default parameters.  E_ref is the plain
solver's final FP64 energy (the reference's cascade from its random start at
level 0); the target is E_ref (1 + 1e-6).

Both forms run the same harness (learned.time_to_energy): the iteration
counts of the levels below the finest, then finest-level iterations; a first
run checks the FP64 energy after each to find the count k that reaches the
target, then the whole schedule with k finest iterations runs as one
epr_solve_schedule call, which must reach the target.  The plain form starts
at level 0 from the random guess with the reference's schedule; the learned
form starts at level 4 from the shipped network's prediction
(exteriorpoissonrecon_amd/weights/unet4d_l4.npz, trained on the device by
`python -m exteriorpoissonrecon_amd.learned` on 2 048 solver pairs of random
Hopf and two-circle tori -- this Hopf torus is not among them), with the
counts of levels 4 and 5 searched (learned.search_schedule).

Bars (deterministic): the learned form reaches the target with fewer node
updates than the plain form, and with the same coarse counts the random
start at level 4 misses it (what the network gains: the coarse levels it
replaces cannot simply be skipped); epr_solve_learned (the reference's
schedule from level 4) ends within the target.  Measured and printed: the
wall times of both, their ratio.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _updates(rec, iters, lo):
    return sum(rec.num_functions(d) * iters[d] for d in range(lo, rec.depth + 1))


def test_learned_start_reaches_the_reference_energy_sooner():
    from exteriorpoissonrecon_amd import Reconstructor, learned, samples
    assert os.path.exists(learned.WEIGHTS_4D)
    net = learned.UNet.load(learned.WEIGHTS_4D)
    pos, skew = samples.hopf_torus(256, lobes=3, amp=0.3, latitude="sin")
    print(f"[gpu] Hopf torus, {len(pos)} samples, depth 6", flush=True)
    rec = Reconstructor(pos, skew, depth=6, aRadius=-1).build()
    E_ref = rec.solve().sum()
    target = E_ref * (1 + 1e-6)
    plain = learned.time_to_energy(rec, target, None)
    lrn, trials = learned.search_schedule(rec, net, target, grid=(5, 10, 20))
    assert plain[0] is not None and plain[2] <= target
    assert lrn is not None and lrn[2] <= target
    print(f"[gpu] E_ref {E_ref:.12g}; plain {plain[0]:.3f} s ({plain[1]} finest iterations); "
          f"learned {lrn[0]:.3f} s (schedule {lrn[3]}, E {lrn[2]:.12g}); ratio {plain[0] / lrn[0]:.3f}", flush=True)
    u_plain, u_lrn = _updates(rec, plain[3], 0), _updates(rec, lrn[3], 4)
    print(f"[gpu] node updates: plain {u_plain:.4e} learned {u_lrn:.4e}", flush=True)
    assert u_lrn < u_plain
    # the random start at level 4 with the learned form's counts
    k = learned._first_crossing(rec, None, lrn[3], 16, target, start_level=4)
    print(f"[gpu] random start at level 4, learned schedule: {'reaches' if k else 'misses'} the target in 16 finest iterations", flush=True)
    assert k is None
    E_l = learned.solve_learned(rec, net).sum()
    print(f"[gpu] epr_solve_learned E {E_l:.12g} (target {target:.12g})", flush=True)
    assert E_l <= target
    rec.close()
    net.close()
