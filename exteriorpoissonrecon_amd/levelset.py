"""Simplicial extraction on the GPU (the reference's Extract path).

Mirrors Misha/MarchingSimplices.h: RegularSubGridTriangulation (:622-706,
the dim! simplices of each listed grid cell) and LevelSet (:410-576, the
joint zero set of two values per vertex).  Vertex numbering and simplex order
are the reference's; both run in libepr_hip.so.
"""
import ctypes

import numpy as np

from ._lib import EprError, _p, load


def _mesh(h):
    L = load()
    nv, ns = ctypes.c_size_t(), ctypes.c_size_t()
    vd, k = ctypes.c_int(), ctypes.c_int()
    L.epr_mesh_info(h, ctypes.byref(nv), ctypes.byref(vd), ctypes.byref(ns), ctypes.byref(k))
    V = np.zeros((nv.value, vd.value))
    S = np.zeros((ns.value, k.value), dtype=np.uint32)
    L.epr_mesh_copy(h, _p(V), S.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    L.epr_mesh_destroy(h)
    return V, S


def _check(rc):
    if rc != 0:
        msg = load().epr_last_error(None)
        raise EprError(f"epr error {rc}: {msg.decode() if msg else ''}")


def subgrid_triangulation(dim, cells, orient=True):
    """(vertices nV x dim, simplices nS x (dim+1)) of the listed cells (nCells x dim offsets)."""
    c = np.ascontiguousarray(cells, dtype=np.uint32).reshape(-1, dim)
    h = ctypes.c_void_p()
    _check(load().epr_subgrid_triangulation(dim, len(c), c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                                            int(orient), ctypes.byref(h)))
    return _mesh(h)


def level_set(vertices, simplices, values, iso=None):
    """(vertices, simplices) of the zero set of values - iso over the mesh: codimension 2 for
    values nV x 2 (the joint zero set), codimension 1 for values of length nV; output vertices
    interpolate the input vertex rows."""
    V = np.ascontiguousarray(vertices, dtype=np.float64)
    S = np.ascontiguousarray(simplices, dtype=np.uint32)
    val = np.ascontiguousarray(values, dtype=np.float64)
    codim = 1 if val.ndim == 1 else val.shape[1]
    if len(val) != len(V) or codim not in (1, 2):
        raise ValueError("one value (codimension 1) or one pair (codimension 2) per vertex")
    iso_ = None if iso is None else np.ascontiguousarray(np.atleast_1d(iso), dtype=np.float64)
    h = ctypes.c_void_p()
    _check(load().epr_level_set(S.shape[1] - 1, codim, V.shape[1], len(V), _p(V), len(S),
                                S.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), _p(val),
                                None if iso_ is None else _p(iso_), ctypes.byref(h)))
    return _mesh(h)
