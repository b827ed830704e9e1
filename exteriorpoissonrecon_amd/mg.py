"""Multigrid on the GPU: the reference's Include/MG.h.

`MGSolver(M, Ps)` mirrors MGSolver( M , Ps , pNum ) (MG.h:83-108) for a
scipy-style CSC matrix M and prolongations Ps[0..pNum-1] (Ps[i] maps level i
to level i+1, Ps[-1] ends at M's size); `vcycle` is vCycle( b , x , iters ,
MCI ) (MG.h:146-186) with colour lists set per level; `relax` is the
multicolour GSRelaxer (MG.h:62-79).  All work runs in libepr_hip.so.
"""
import ctypes

import numpy as np

from ._lib import EprError, _p, load

_up = ctypes.POINTER(ctypes.c_uint)


def _csc(A):
    """(indptr, indices, data) of a CSC matrix with sorted row indices (scipy.sparse or a tuple)."""
    if isinstance(A, tuple):
        indptr, indices, data, shape = A
    else:
        A = A.tocsc()
        A.sort_indices()
        indptr, indices, data, shape = A.indptr, A.indices, A.data, A.shape
    return (np.ascontiguousarray(indptr, dtype=np.uint32), np.ascontiguousarray(indices, dtype=np.uint32),
            np.ascontiguousarray(data, dtype=np.float64), shape)


def _check(rc):
    if rc != 0:
        msg = load().epr_last_error(None)
        raise EprError(f"epr error {rc}: {msg.decode() if msg else ''}")


class MGSolver:
    def __init__(self, M, Ps=()):
        L = load()
        cp, ri, v, shape = _csc(M)
        n = shape[0]
        self._keep = [cp, ri, v]
        P = [_csc(p) for p in Ps]
        pn = len(P)
        rows = (ctypes.c_size_t * max(pn, 1))(*[p[3][0] for p in P])
        cols = (ctypes.c_size_t * max(pn, 1))(*[p[3][1] for p in P])
        pcp = (_up * max(pn, 1))(*[p[0].ctypes.data_as(_up) for p in P])
        pri = (_up * max(pn, 1))(*[p[1].ctypes.data_as(_up) for p in P])
        pv = (ctypes.POINTER(ctypes.c_double) * max(pn, 1))(*[_p(p[2]) for p in P])
        self._h = ctypes.c_void_p()
        _check(L.epr_mg_create(n, cp.ctypes.data_as(_up), ri.ctypes.data_as(_up), _p(v), pn, rows, cols,
                               pcp, pri, pv, ctypes.byref(self._h)))
        self.levels = L.epr_mg_num_levels(self._h)
        self.sizes = [L.epr_mg_level_size(self._h, i) for i in range(self.levels)]

    def set_colours(self, level, lists):
        """Colour lists (sequence of index arrays) of a level 1..pNum."""
        starts = np.cumsum([0] + [len(x) for x in lists]).astype(np.uint32)
        idx = np.ascontiguousarray(np.concatenate([np.asarray(x) for x in lists]), dtype=np.uint32)
        _check(load().epr_mg_set_colours(self._h, level, len(lists), starts.ctypes.data_as(_up), idx.ctypes.data_as(_up)))

    def vcycle(self, b, x, iters):
        """One V-cycle; returns the updated x (a new array)."""
        bb = np.ascontiguousarray(b, dtype=np.float64)
        xx = np.array(x, dtype=np.float64)
        _check(load().epr_mg_vcycle(self._h, _p(bb), _p(xx), iters))
        return xx

    def relax(self, level, b, x, iters=1):
        bb = np.ascontiguousarray(b, dtype=np.float64)
        xx = np.array(x, dtype=np.float64)
        _check(load().epr_mg_relax(self._h, level, _p(bb), _p(xx), iters))
        return xx

    def close(self):
        if getattr(self, "_h", None):
            load().epr_mg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
