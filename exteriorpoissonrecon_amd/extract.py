"""Surface extraction from a .tree on the GPU: the reference's Extract
(Extract/Extract.cpp:96-330) over libepr_hip.so.

`IsoTree(path)` reads a dense or adapted hierarchy and holds its finest
regular values on the device (Include/IsoTree.inl:45-73, 314-338);
`level_set_nodes`, `extract` and `level_set` are the stages of Extract.cpp;
`mesh_ply_bytes` / `level_sets_ply_bytes` serialise results exactly as the
reference's PLY writer does (Misha/Ply.inl:673-737, binary little endian).
"""
import ctypes

import numpy as np

from ._lib import EprError, load
from .levelset import _mesh

AXES = "xyzw"


class IsoTree:
    def __init__(self, path):
        L = load()
        h = ctypes.c_void_p()
        if L.epr_iso_open(str(path).encode(), ctypes.byref(h)):
            msg = L.epr_iso_last_error(None)
            raise EprError(msg.decode() if msg else "epr_iso_open failed")
        self._h = h
        dim, depth, adapted = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        xf = (ctypes.c_double * 25)()
        L.epr_iso_info(h, ctypes.byref(dim), ctypes.byref(depth), ctypes.byref(adapted), xf)
        self.dim, self.depth, self.adapted = dim.value, depth.value, bool(adapted.value)
        self.voxel_to_world = np.array(xf[:(self.dim + 1) ** 2]).reshape(self.dim + 1, self.dim + 1)

    def _check(self, rc):
        if rc:
            msg = load().epr_iso_last_error(self._h)
            raise EprError(f"epr error {rc}: {msg.decode() if msg else ''}")

    def values(self):
        """Finest regular values, ((2^D+1)^dim, 2), last coordinate fastest."""
        out = np.zeros((((1 << self.depth) + 1) ** self.dim, 2))
        self._check(load().epr_iso_values(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def level_set_nodes(self, all=True):
        """IsoTree::levelSetNodes after refineZeroCrossing: (n, dim) finest element offsets in tree order."""
        n = ctypes.c_size_t()
        self._check(load().epr_iso_level_set_nodes(self._h, int(all), 0, None, ctypes.byref(n)))
        out = np.zeros((n.value, self.dim), dtype=np.uint32)
        self._check(load().epr_iso_level_set_nodes(self._h, int(all), n.value,
                                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), ctypes.byref(n)))
        return out

    def read_density(self, path):
        """TreeEstimator::read of a density file (ExteriorPoissonRecon's <out>.density.tree)."""
        self._check(load().epr_iso_read_density(self._h, str(path).encode()))

    def extract(self, all=True, trim_density=0.0):
        """The oriented joint zero set: (vertices (nV, 3 dim): world position, n_0, n_1; simplices (nS, dim-1)).
        With a density read and trim_density > 0, only the face-connected components whose
        largest samples-per-cell estimate exceeds trim_density (Extract.cpp:217-238)."""
        h = ctypes.c_void_p()
        self._check(load().epr_iso_extract_trimmed(self._h, int(all), float(trim_density), ctypes.byref(h)))
        return _mesh(h)

    def component_densities(self):
        """Per component of the last trimmed extraction: the largest samples-per-cell estimate."""
        n = ctypes.c_size_t()
        load().epr_iso_component_densities(self._h, 0, None, ctypes.byref(n))
        out = np.zeros(n.value)
        load().epr_iso_component_densities(self._h, n.value, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n))
        return out

    def level_set(self, component):
        """The codimension-1 zero set of one component: (world vertices (nV, dim), simplices (nS, dim))."""
        h = ctypes.c_void_p()
        self._check(load().epr_iso_level_set(self._h, int(component), ctypes.byref(h)))
        return _mesh(h)

    def write_grid(self, header):
        self._check(load().epr_iso_write_grid(self._h, str(header).encode()))

    def seconds(self, which):
        return load().epr_iso_seconds(self._h, which)

    def close(self):
        if self._h:
            load().epr_iso_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def orient(simplices):
    """MarchingSimplices::Orient of segments (n x 2) or triangles (n x 3): a copy, oriented."""
    from ._lib import EprError as _E
    S = np.ascontiguousarray(simplices, dtype=np.uint32).copy()
    if S.ndim != 2 or S.shape[1] not in (2, 3):
        raise ValueError("simplices must be n x 2 or n x 3")
    if load().epr_orient(S.shape[1], len(S), S.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))):
        msg = load().epr_iso_last_error(None)
        raise _E(msg.decode() if msg else "epr_orient failed")
    return S


def density_grid(density_path, dim, depth, grid_path):
    """Writes grid_path: the density estimator's samples per cell at every node
    of the (2^depth+1)^dim grid (a .grid with one DOUBLE per node), from a
    <out>.density.tree."""
    if load().epr_density_grid(density_path.encode(), int(dim), int(depth), grid_path.encode()):
        msg = load().epr_last_error(None)
        raise EprError(msg.decode() if msg else "epr_density_grid failed")


def _ply(props, nrows, rows, faces):
    k = faces.shape[1] if faces.ndim == 2 else 0
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {nrows}"]
    head += [f"property {p}" for p in props]
    head += [f"element face {len(faces)}", "property list int int vertex_indices", "end_header", ""]
    f = np.empty((len(faces), k + 1), dtype="<i4")
    f[:, 0] = k
    f[:, 1:] = faces
    return "\n".join(head).encode() + rows + f.tobytes()


def mesh_ply_bytes(dim, vertices, simplices):
    """Extract.cpp:263-282's --out file: positions x.., frames n_c_d as doubles."""
    props = [f"double {a}" for a in AXES[:dim]] + [f"double n_{c}_{d}" for c in range(2) for d in range(dim)]
    return _ply(props, len(vertices), np.ascontiguousarray(vertices, dtype="<f8").tobytes(),
                np.asarray(simplices).reshape(-1, dim - 1))


COLOURS = ((223, 41, 53), (55, 114, 255))


def level_sets_ply_bytes(dim, sets):
    """Extract.cpp:285-327's --levelSets file from [(vertices, simplices)] per component:
    positions as doubles and the component's colour as uchar red, green, blue."""
    props = [f"double {a}" for a in AXES[:dim]] + ["uchar red", "uchar green", "uchar blue"]
    row = np.dtype([("p", "<f8", (dim,)), ("c", "u1", (3,))])
    rows, faces, off = [], [], 0
    for c, (V, S) in enumerate(sets):
        r = np.zeros(len(V), dtype=row)
        r["p"] = V
        r["c"] = COLOURS[c]
        rows.append(r.tobytes())
        faces.append(np.asarray(S, dtype=np.int64).reshape(-1, dim) + off)
        off += len(V)
    return _ply(props, off, b"".join(rows), np.concatenate(faces).astype("<i4") if faces else np.zeros((0, dim), "<i4"))


def write_ply(path, data):
    with open(path, "wb") as f:
        f.write(data)
