"""epr_energies_noalloc: the energies with the partial sums in the device's
static scratch instead of the context's buffers.  Bar: the same result bits as
epr_energies (same kernels, same summation order), on two calls in a row --
the second on a smaller level, so stale partial sums of the first call would
show -- and on a second context; a level too large for the scratch is a clear
error, and the entry point keeps working afterwards."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(r, level):
    from exteriorpoissonrecon_amd._lib import _p, load
    out = np.zeros(2)
    rc = load().epr_energies_noalloc(r._h, level, _p(out))
    return rc, out


def _check_level(r, level):
    rc, E = _noalloc(r, level)
    assert rc == 0, rc
    Eref = r.energies(level)
    print(f"[gpu] dim {r.dim} level {level}: {E.tolist()} vs {Eref.tolist()}", flush=True)
    assert np.array_equal(E, Eref), (E, Eref)


def test_noalloc_energies_match_epr_energies():
    from exteriorpoissonrecon_amd import Reconstructor, samples
    from exteriorpoissonrecon_amd._lib import load
    print("[gpu] no-alloc energies: 3D Borromean depth 7", flush=True)
    r3 = Reconstructor(*samples.borromean(16384), depth=7, aRadius=-1).build()
    r3.solve()
    _check_level(r3, 7)        # the largest 3D level the scratch holds
    _check_level(r3, 4)        # then a smaller one: partial sums of the first call must not leak in
    r3.init(7)                 # new coefficients on the finest level
    _check_level(r3, 7)
    r3.close()
    print("[gpu] no-alloc energies: 4D Clifford depth 6", flush=True)
    r4 = Reconstructor(*samples.clifford_torus(256, ang_deg=2.0), depth=6, aRadius=-1).build()
    r4.solve()
    _check_level(r4, 5)
    _check_level(r4, 2)
    rc, _ = _noalloc(r4, 6)    # 17.9 M functions: beyond the scratch
    assert rc != 0
    msg = load().epr_last_error(r4._h).decode()
    assert "too large" in msg and "scratch" in msg, msg
    _check_level(r4, 3)        # still usable after the refusal
    r4.close()


def test_noalloc_energies_adapted():
    from exteriorpoissonrecon_amd import Reconstructor, samples
    print("[gpu] no-alloc energies: 3D knot depth 6, adapted", flush=True)
    r = Reconstructor(*samples.torus_knot(4096, 3, 5, seed=3), depth=6, aRadius=3).build()
    r.solve()
    for level in (6, 3, 5):
        _check_level(r, level)
    r.close()
