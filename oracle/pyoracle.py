"""ctypes binding of oracle/_build/liboracle.so (test infrastructure only).

The oracle is a CPU restatement of the reference algorithm; see epr_oracle.h.
"""
import ctypes
import os
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


class Params(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_int), ("seed", ctypes.c_int), ("tRadius", ctypes.c_int),
                ("dRadius", ctypes.c_int), ("minSolveDepth", ctypes.c_int),
                ("maxSolveDepth", ctypes.c_int), ("gsIters", ctypes.c_int),
                ("singleLevel", ctypes.c_int), ("nSigma", ctypes.c_double),
                ("scale", ctypes.c_double), ("sWeight", ctypes.c_double),
                ("dWeight", ctypes.c_double), ("iMult", ctypes.c_double),
                ("separate", ctypes.c_int), ("jitter", ctypes.c_double)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, sz = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t
        L.orc_default_params.argtypes = [ctypes.POINTER(Params)]
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_int, ctypes.POINTER(Params), sz, dp, dp]
        L.orc_create_adapted.restype = vp
        L.orc_create_adapted.argtypes = [ctypes.c_int, ctypes.POINTER(Params), sz, dp, dp,
                                         ctypes.POINTER(ctypes.c_ubyte), ctypes.POINTER(ctypes.c_ubyte),
                                         ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(sz)]
        L.orc_destroy.argtypes = [vp]
        L.orc_error.restype = ctypes.c_char_p
        L.orc_num_functions.restype = sz
        L.orc_num_functions.argtypes = [vp, ctypes.c_int]
        L.orc_num_elements.restype = sz
        L.orc_num_elements.argtypes = [vp, ctypes.c_int]
        L.orc_get.restype = sz
        L.orc_get.argtypes = [vp, ctypes.c_char_p, dp, sz]
        L.orc_matvec.argtypes = [vp, ctypes.c_int, ctypes.c_char, dp, dp]
        L.orc_energies.argtypes = [vp, ctypes.c_int, dp, dp, dp]
        L.orc_wedge.restype = ctypes.c_double
        L.orc_wedge.argtypes = [vp, ctypes.c_int, dp, dp]
        L.orc_qfit.argtypes = [vp, ctypes.c_int, dp, dp, dp]
        L.orc_gs.argtypes = [vp, ctypes.c_int, dp, dp, ctypes.c_int]
        L.orc_init.argtypes = [vp, dp, dp]
        L.orc_solve.argtypes = [vp, dp, dp, dp, dp, dp]
        L.orc_write_tree.argtypes = [vp, ctypes.c_char_p, dp, dp, ctypes.c_ulonglong]
        up = ctypes.POINTER(ctypes.c_uint)
        L.orc_cube_simplices.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_subgrid_triangulation.argtypes = [ctypes.c_int, sz, up, ctypes.c_int, ctypes.POINTER(sz),
                                                ctypes.POINTER(dp), ctypes.POINTER(sz), ctypes.POINTER(up)]
        L.orc_level_set.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, sz, dp, sz, up, dp, dp,
                                    ctypes.POINTER(sz), ctypes.POINTER(dp), ctypes.POINTER(sz), ctypes.POINTER(up)]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_extract.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(sz), ctypes.POINTER(up),
                                  ctypes.POINTER(dp), ctypes.POINTER(sz), ctypes.POINTER(dp), ctypes.POINTER(sz),
                                  ctypes.POINTER(up)]
        L.orc_extract_level_set.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(sz), ctypes.POINTER(dp),
                                            ctypes.POINTER(sz), ctypes.POINTER(up)]
        L.orc_extract_trimmed.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_double,
                                          ctypes.POINTER(sz), ctypes.POINTER(up), ctypes.POINTER(dp), ctypes.POINTER(sz),
                                          ctypes.POINTER(dp), ctypes.POINTER(sz), ctypes.POINTER(up), ctypes.POINTER(sz),
                                          ctypes.POINTER(dp)]
        L.orc_extract_error.restype = ctypes.c_char_p
        L.orc_indexer_hash.restype = ctypes.c_ulonglong
        L.orc_orient.argtypes = [ctypes.c_int, sz, up]
        L.orc_indexer_hash.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_density_at.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_size_t, dp, dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_params(**kw):
    p = Params()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def hierarchy_from_tree(fields, dim, depth):
    """Per-level (fmask, emask, forder) of an adapted hierarchy from the node
    records of a reference .tree (exteriorpoissonrecon_amd.treefile fields or a
    tests/golden/adapted_*.npz): function flags over the dense (2^l+1)^dim
    layout (coordinate 0 most significant; 1 active, 2 padding), element flags
    over (2^l)^dim (coordinate 0 least significant), and the dense index of
    every function in the hierarchy's numbering."""
    nd, off = np.asarray(fields["node_depth"]), np.asarray(fields["node_offset"]).astype(np.int64)
    nf, ne = np.asarray(fields["node_function"]), np.asarray(fields["node_element"])
    act = np.asarray(fields["active_sizes"])
    out = []
    for l in range(depth + 1):
        res = 1 << l
        sel = nd == l + 1
        fsel = sel & (nf >= 0)
        F = off[fsel]
        fl = np.zeros(len(F), dtype=np.int64)
        for k in range(dim):
            fl = fl * (res + 1) + F[:, k]
        order = np.zeros(len(F), dtype=np.int64)
        order[nf[fsel]] = fl
        fmask = np.zeros((res + 1) ** dim, dtype=np.uint8)
        fmask[order[:act[l]]] = 1
        fmask[order[act[l]:]] = 2
        E = off[sel & (ne >= 0)]
        el = np.zeros(len(E), dtype=np.int64)
        for k in range(dim - 1, -1, -1):
            el = el * res + E[:, k]
        emask = np.zeros(res ** dim, dtype=np.uint8)
        emask[el] = 1
        out.append((fmask, emask, order))
    return out


class Oracle:
    """One reconstruction built by the CPU oracle (hierarchy: the result of
    hierarchy_from_tree for an adapted reconstruction)."""

    def __init__(self, pos, skew, hierarchy=None, **params):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        skew = np.ascontiguousarray(skew, dtype=np.float64)
        self.dim = pos.shape[1]
        self.params = default_params(**params)
        if hierarchy is None:
            self.h = lib().orc_create(self.dim, ctypes.byref(self.params), pos.shape[0], _p(pos), _p(skew))
        else:
            fm = np.ascontiguousarray(np.concatenate([h[0] for h in hierarchy]), dtype=np.uint8)
            em = np.ascontiguousarray(np.concatenate([h[1] for h in hierarchy]), dtype=np.uint8)
            fo = np.ascontiguousarray(np.concatenate([h[2] for h in hierarchy]), dtype=np.int64)
            no = np.ascontiguousarray([len(h[2]) for h in hierarchy], dtype=np.uint64)
            ub = ctypes.POINTER(ctypes.c_ubyte)
            self.h = lib().orc_create_adapted(self.dim, ctypes.byref(self.params), pos.shape[0], _p(pos), _p(skew),
                                              fm.ctypes.data_as(ub), em.ctypes.data_as(ub),
                                              fo.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                              no.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)))
        if not self.h:
            raise RuntimeError(lib().orc_error().decode())
        self.depth = self.params.depth

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def nfun(self, level):
        return lib().orc_num_functions(self.h, level)

    def nelem(self, level):
        return lib().orc_num_elements(self.h, level)

    def get(self, name):
        n = lib().orc_get(self.h, name.encode(), None, 0)
        out = np.zeros(n)
        lib().orc_get(self.h, name.encode(), _p(out), n)
        return out

    def matvec(self, level, which, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n_out = self.nfun(level + 1) if which == "P" else self.nfun(level)
        y = np.zeros(n_out)
        lib().orc_matvec(self.h, level, which.encode(), _p(x), _p(y))
        return y

    def energies(self, level, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64); y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(2)
        lib().orc_energies(self.h, level, _p(x), _p(y), _p(out))
        return out

    def wedge(self, level, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64); y = np.ascontiguousarray(y, dtype=np.float64)
        return lib().orc_wedge(self.h, level, _p(x), _p(y))

    def qfit(self, level, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64); y = np.ascontiguousarray(y, dtype=np.float64)
        q = np.zeros(5 * self.nfun(level))
        lib().orc_qfit(self.h, level, _p(x), _p(y), _p(q))
        return q.reshape(-1, 5)

    def gs(self, level, x, y, iters=1):
        x = np.array(x, dtype=np.float64); y = np.array(y, dtype=np.float64)
        lib().orc_gs(self.h, level, _p(x), _p(y), iters)
        return x, y

    def init(self):
        n = self.nfun(min(self.params.minSolveDepth, self.depth))
        x = np.zeros(n); y = np.zeros(n)
        lib().orc_init(self.h, _p(x), _p(y))
        return x, y

    def solve(self):
        tot = sum(self.nfun(l) for l in range(self.depth + 1))
        hx = np.zeros(tot); hy = np.zeros(tot)
        n = self.nfun(self.depth)
        fx = np.zeros(n); fy = np.zeros(n); E = np.zeros(2)
        lib().orc_solve(self.h, _p(hx), _p(hy), _p(fx), _p(fy), _p(E))
        out = {"final_x": fx, "final_y": fy, "final_E": E}
        o = 0
        for l in range(self.depth + 1):
            m = self.nfun(l)
            out[f"hx{l}"] = hx[o:o + m]; out[f"hy{l}"] = hy[o:o + m]
            o += m
        out["_hx"] = hx; out["_hy"] = hy
        return out

    def write_tree(self, path, hx, hy, indexer_hash):
        hx = np.ascontiguousarray(hx); hy = np.ascontiguousarray(hy)
        return lib().orc_write_tree(self.h, path.encode(), _p(hx), _p(hy), indexer_hash)


def probe(n, salt):
    """The probe vectors of oracle/ref_dump.cpp: v[i] = cos(0.7*i + salt)."""
    return np.cos(0.7 * np.arange(n, dtype=np.float64) + salt)


def load_dump(path):
    """Reads a ref_dump file: {u32 nameLen, name, u64 count, count*f64}*."""
    d = {}
    b = open(path, "rb").read()
    o = 0
    while o < len(b):
        l = struct.unpack_from("<I", b, o)[0]; o += 4
        name = b[o:o + l].decode(); o += l
        c = struct.unpack_from("<Q", b, o)[0]; o += 8
        d[name] = np.frombuffer(b, dtype="<f8", count=c, offset=o).copy(); o += 8 * c
    return d


# ---- simplicial extraction ------------------------------------------------
def _take(ptr, n, dtype):
    out = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].astype(dtype, copy=True)
    lib().orc_free(ctypes.cast(ptr, ctypes.c_void_p))
    return out


def cube_simplices(dim):
    """The canonical cube's simplices as vertex bit-indices, (dim!, dim+1)."""
    out = (ctypes.c_int * (24 * 5))()
    n = lib().orc_cube_simplices(dim, out)
    return np.array(out[:n * (dim + 1)]).reshape(n, dim + 1)


def subgrid_triangulation(dim, cells, orient=True):
    """(vertices nV x dim, simplices nS x (dim+1)) of the reference's RegularSubGridTriangulation."""
    c = np.ascontiguousarray(cells, dtype=np.uint32)
    nV, nS = ctypes.c_size_t(), ctypes.c_size_t()
    V, S = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_uint)()
    rc = lib().orc_subgrid_triangulation(dim, len(c), c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), int(orient),
                                         ctypes.byref(nV), ctypes.byref(V), ctypes.byref(nS), ctypes.byref(S))
    if rc:
        raise ValueError("bad triangulation input")
    return _take(V, nV.value * dim, np.float64).reshape(-1, dim), _take(S, nS.value * (dim + 1), np.uint32).reshape(-1, dim + 1)


def level_set(vertices, simplices, values, iso=None):
    """(vertices, simplices) of the reference's MarchingSimplices::LevelSet; values nV x codim."""
    V = np.ascontiguousarray(vertices, dtype=np.float64)
    S = np.ascontiguousarray(simplices, dtype=np.uint32)
    val = np.ascontiguousarray(values, dtype=np.float64)
    codim = 1 if val.ndim == 1 else val.shape[1]
    dim, vdim = S.shape[1] - 1, V.shape[1]
    iso = np.zeros(codim) if iso is None else np.ascontiguousarray(iso, dtype=np.float64)
    nLV, nLS = ctypes.c_size_t(), ctypes.c_size_t()
    LV, LS = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_uint)()
    rc = lib().orc_level_set(dim, codim, vdim, len(V), _p(V), len(S), S.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                             _p(val), _p(iso), ctypes.byref(nLV), ctypes.byref(LV), ctypes.byref(nLS), ctypes.byref(LS))
    if rc:
        raise ValueError("bad level-set input")
    k = dim - codim + 1
    return _take(LV, nLV.value * vdim, np.float64).reshape(-1, vdim), _take(LS, nLS.value * k, np.uint32).reshape(-1, k)


# ---- the Extract pipeline ---------------------------------------------------
def tree_header(path):
    """(dim, depth) of a .tree: u32 dim, u64 hash, the XForm (8 + (dim+1)^2 * 8 bytes), u64 depth count."""
    with open(path, "rb") as f:
        b = f.read(4 + 8 + 8 + 25 * 8 + 8)
    dim = struct.unpack_from("<I", b, 0)[0]
    n = struct.unpack_from("<Q", b, 20 + 8 * (dim + 1) ** 2)[0]
    return dim, n - 1


def extract(tree, all=True, density=None, trim=0.0):
    """The oracle's Extract.cpp:109-259 on a .tree: dict with cells (level-set
    nodes, tree order), values ((2^D+1)^dim x 2 finest values), vertices (nV x
    3 dim: world position, n_0, n_1), simplices (nS x (dim-1), oriented) and,
    with a density file and trim > 0, the trimmed simplices and every
    component's largest samples-per-cell estimate (component_max)."""
    dim, D = tree_header(tree)
    sz, up, dp = ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_double)
    nC, nV, nS, nK = sz(), sz(), sz(), sz()
    C, S = up(), up()
    val, V, K = dp(), dp(), dp()
    rc = lib().orc_extract_trimmed(tree.encode(), int(all), density.encode() if density else None, float(trim),
                                   ctypes.byref(nC), ctypes.byref(C), ctypes.byref(val),
                                   ctypes.byref(nV), ctypes.byref(V), ctypes.byref(nS), ctypes.byref(S),
                                   ctypes.byref(nK), ctypes.byref(K))
    if rc:
        raise ValueError(lib().orc_extract_error().decode())
    nF = ((1 << D) + 1) ** dim
    return {"cells": _take(C, nC.value * dim, np.uint32).reshape(-1, dim),
            "values": _take(val, 2 * nF, np.float64).reshape(-1, 2),
            "vertices": _take(V, nV.value * 3 * dim, np.float64).reshape(-1, 3 * dim),
            "simplices": _take(S, nS.value * (dim - 1), np.uint32).reshape(-1, dim - 1),
            "component_max": _take(K, nK.value, np.float64)}


def extract_level_set(tree, component):
    """The oracle's Extract.cpp:296-310 for one component: (world vertices, simplices)."""
    dim, _ = tree_header(tree)
    nV, nS = ctypes.c_size_t(), ctypes.c_size_t()
    V, S = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_uint)()
    rc = lib().orc_extract_level_set(tree.encode(), component, ctypes.byref(nV), ctypes.byref(V),
                                     ctypes.byref(nS), ctypes.byref(S))
    if rc:
        raise ValueError(lib().orc_extract_error().decode())
    return _take(V, nV.value * dim, np.float64).reshape(-1, dim), _take(S, nS.value * dim, np.uint32).reshape(-1, dim)


def indexer_hash(dim, adapted):
    return int(lib().orc_indexer_hash(dim, int(adapted)))


def density_at(density_path, dim, points):
    """Samples per cell of a .density.tree at unit-cube points (n x dim)."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    out = np.zeros(len(pts))
    rc = lib().orc_density_at(density_path.encode(), int(dim), len(pts), _p(pts), _p(out))
    if rc != 0:
        raise RuntimeError(lib().orc_extract_error().decode())
    return out


def orient(simplices):
    """The oracle's MarchingSimplices::Orient of n x 2 or n x 3 simplices (a copy)."""
    S = np.ascontiguousarray(simplices, dtype=np.uint32).copy()
    if lib().orc_orient(S.shape[1], len(S), S.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))):
        raise ValueError(lib().orc_extract_error().decode())
    return S
