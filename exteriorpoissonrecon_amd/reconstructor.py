"""Host-side mirror of the reference's ExteriorPoisson::Reconstructor.

Reference interface: ExteriorPoissonRecon/ExteriorPoisson.h:56-94 (construct
from samples, solve(), x(depth)/y(depth) per-level coefficients) and the
energy methods of Include/SystemEnergy.h:60-125.  All work happens in
libepr_hip.so on the current HIP device.
"""
import ctypes

import numpy as np

from ._lib import EprError, _p, default_params, load


class Reconstructor:
    """One dense-grid reconstruction on the GPU.

    Parameters follow the reference command line (depth, seed, sWeight,
    dWeight, gsIters, iMult, nSigma, scale, tRadius, dRadius, minSolveDepth,
    maxSolveDepth, singleLevel, verbose)."""

    def __init__(self, positions, skews, **params):
        pos = np.asarray(positions)
        if pos.ndim != 2 or pos.shape[1] not in (3, 4):
            raise ValueError("positions must be n x 3 or n x 4")
        self.dim = pos.shape[1]
        self.params = default_params(**params)
        self.depth = self.params.depth
        L = load()
        h = ctypes.c_void_p()
        rc = L.epr_create(self.dim, ctypes.byref(self.params), ctypes.byref(h))
        self._h = h
        self._check(rc)
        self.set_samples(positions, skews)

    def set_samples(self, positions, skews):
        """Hands over (new) samples: the unit-cube map and the ordered sampler's
        cells (ExteriorPoisson.h:100-123, OrderedSamples.inl:32-85); build()
        must run again afterwards."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        sk = np.ascontiguousarray(skews, dtype=np.float64)
        if pos.ndim != 2 or pos.shape[1] != self.dim:
            raise ValueError(f"positions must be n x {self.dim}")
        n = pos.shape[0]
        if sk.shape != (n, self.dim, self.dim):
            raise ValueError("skews must be n x dim x dim")
        self._check(load().epr_set_samples(self._h, n, _p(pos), _p(sk)))
        return self

    def _check(self, rc):
        if rc != 0:
            msg = load().epr_last_error(self._h) if self._h else b"create failed"
            raise EprError(f"epr error {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None):
            load().epr_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # -- phases ---------------------------------------------------------------
    def build(self):
        self._check(load().epr_build(self._h))
        return self

    def solve(self):
        E = np.zeros(2)
        self._check(load().epr_solve(self._h, _p(E)))
        return E

    def init(self, level, x=None, y=None):
        if x is None:
            self._check(load().epr_init(self._h, level, None, None))
        else:
            x = np.ascontiguousarray(x, dtype=np.float64); y = np.ascontiguousarray(y, dtype=np.float64)
            self._check(load().epr_init(self._h, level, _p(x), _p(y)))

    def gauss_seidel(self, level, iters):
        self._check(load().epr_gs_sweeps(self._h, level, iters))

    def gauss_seidel_colour(self, level, colour):
        """One colour phase of a Gauss-Seidel iteration (levels swept by colour launches)."""
        self._check(load().epr_gs_colour(self._h, level, colour))

    def time_gauss_seidel(self, level, iters):
        """(device milliseconds, kernel launches) of iters iterations at a level (HIP events)."""
        ms = ctypes.c_double(); n = ctypes.c_int()
        self._check(load().epr_time_gs_sweeps(self._h, level, iters, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def time_gauss_seidel_fp32(self, level, iters):
        """The same iterations in FP32 (measurement): (device ms, launches); the
        level's coefficients are left at the FP32 iterate."""
        ms = ctypes.c_double(); n = ctypes.c_int()
        self._check(load().epr_time_gs_sweeps_fp32(self._h, level, iters, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def prolong_add(self, level):
        self._check(load().epr_prolong_add(self._h, level))

    # -- queries --------------------------------------------------------------
    def num_functions(self, level):
        return load().epr_num_functions(self._h, level)

    def coefficients(self, level):
        n = self.num_functions(level)
        x = np.zeros(n); y = np.zeros(n)
        self._check(load().epr_get_coeffs(self._h, level, _p(x), _p(y)))
        return x, y

    def x(self, level):
        return self.coefficients(level)[0]

    def y(self, level):
        return self.coefficients(level)[1]

    def set_coefficients(self, level, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64); y = np.ascontiguousarray(y, dtype=np.float64)
        self._check(load().epr_set_coeffs(self._h, level, _p(x), _p(y)))

    def energies(self, level):
        out = np.zeros(2)
        self._check(load().epr_energies(self._h, level, _p(out)))
        return out

    def quadratic_fits(self, level):
        q = np.zeros(5 * self.num_functions(level))
        self._check(load().epr_quadratic_fits(self._h, level, _p(q)))
        return q.reshape(-1, 5)

    def apply(self, level, which, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        n = self.num_functions(level + 1) if which == "P" else self.num_functions(level)
        out = np.zeros(n)
        self._check(load().epr_apply(self._h, level, which.encode(), _p(v), _p(out)))
        return out

    def array(self, name):
        cnt = ctypes.c_size_t()
        self._check(load().epr_get_array(self._h, name.encode(), None, 0, ctypes.byref(cnt)))
        out = np.zeros(cnt.value)
        self._check(load().epr_get_array(self._h, name.encode(), _p(out), cnt.value, ctypes.byref(cnt)))
        return out

    def write_tree(self, path, indexer_hash):
        self._check(load().epr_write_tree(self._h, path.encode(), indexer_hash))

    def write_grid(self, path):
        self._check(load().epr_write_grid(self._h, path.encode()))

    def write_density(self, path):
        """<out>.density.tree: the density estimator as ExteriorPoissonRecon.cpp:272-277 writes it."""
        self._check(load().epr_write_density(self._h, path.encode()))

    def seconds(self, which):
        return load().epr_last_seconds(self._h, 0 if which == "build" else 1)

    def time_op(self, op, level, reps=10):
        """(ms per launch, algorithmic HBM bytes) of one kernel at a level
        (epr_time_op: "prolong", "galerkin_t", "galerkin_c", "energy")."""
        ms, b = ctypes.c_double(), ctypes.c_double()
        self._check(load().epr_time_op(self._h, op.encode(), int(level), int(reps), ctypes.byref(ms), ctypes.byref(b)))
        return ms.value, b.value

    def autotune(self, fp32=False):
        """Times every sweep plan of each level of this (built) reconstruction's
        solve and keeps the fastest for single solves (epr_autotune).  Returns
        the trials as (level, res, nodes, precision, kind, threads,
        us_per_iteration, kept) tuples."""
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(load().epr_autotune(self._h, 1 if fp32 else 0, buf, len(buf)))
        return _trials(buf.value.decode(), 7)


def _handles(recs):
    hs = [r._h for r in recs]
    if any(not h for h in hs):
        raise EprError("a reconstruction in the batch is closed")
    return (ctypes.c_void_p * len(hs))(*hs)


def _batch_check(rc):
    if rc != 0:
        msg = load().epr_last_error(None)
        raise EprError(f"epr error {rc}: {msg.decode() if msg else ''}")


def build_batch(recs, threads=0):
    """Reconstructor.build() of every reconstruction, from up to `threads` host
    threads (0: four), each on its own stream."""
    recs = list(recs)
    _batch_check(load().epr_build_batch(len(recs), _handles(recs), int(threads)))
    return recs


def solve_batch(recs):
    """Reconstructor.solve() of every (built) reconstruction of one dimension,
    with each level's sweeps of all of them in shared launches.  Every
    reconstruction ends with the coefficients and energies its own solve()
    gives.  Returns an n x 2 array of (E_fit, E_reg)."""
    recs = list(recs)
    E = np.zeros((len(recs), 2))
    _batch_check(load().epr_solve_batch(len(recs), _handles(recs), _p(E)))
    return E


def _trials(text, nfields):
    """Report lines of the autotuners as tuples: integer fields, the kind and
    precision names, the time (the last field) and whether the trial was kept."""
    out = []
    for line in text.splitlines():
        f = line.split()
        kept = len(f) > nfields and f[nfields] == "*"
        row = []
        for k, v in enumerate(f[:nfields]):
            if k == nfields - 1:
                row.append(float(v))
            else:
                try:
                    row.append(int(v))
                except ValueError:
                    row.append(v)
        out.append(tuple(row) + (kept,))
    return out


def autotune_batch(recs, fp32=False):
    """Times every sweep plan of each level of solve_batch(recs) and keeps the
    fastest for batched solves (epr_autotune_batch).  Returns the trials as
    (level, res, nodes, contexts, precision, kind, threads, us, kept) tuples."""
    recs = list(recs)
    buf = ctypes.create_string_buffer(1 << 16)
    _batch_check(load().epr_autotune_batch(len(recs), _handles(recs), 1 if fp32 else 0, buf, len(buf)))
    return _trials(buf.value.decode(), 8)


def time_batch_sweeps(recs, level, fp32=False, iters=1):
    """(ms, launches, node updates per iteration) of `iters` iterations of a
    level's colour grids as solve_batch runs them (epr_time_batch_sweeps)."""
    recs = list(recs)
    ms, nl, nn = ctypes.c_double(), ctypes.c_int(), ctypes.c_longlong()
    _batch_check(load().epr_time_batch_sweeps(len(recs), _handles(recs), int(level), 1 if fp32 else 0, int(iters),
                                              ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(nn)))
    return ms.value, nl.value, nn.value


def tune_table():
    """The current device's tuning table as text (epr_tune_get): one line per
    entry, "dim res batched fp32 kind threads us"."""
    n = ctypes.c_size_t(0)
    _batch_check(load().epr_tune_get(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _batch_check(load().epr_tune_get(buf, len(buf), None))
    return buf.value.decode()


def set_tune_table(text):
    """Replaces the current device's tuning table (epr_tune_set); None or ""
    restores the built-in rules."""
    _batch_check(load().epr_tune_set(text.encode() if text else None))
