"""Reader of the reference's .tree files (ExteriorPoissonRecon.cpp:259-270).

Layout: u32 dim, u64 indexer hash, the voxel-to-world XForm<double, dim+1>
written raw (the 8-byte vtable pointer of its Algebra base, Misha/Algebra.h:168,
then (dim+1)^2 f64), the
hierarchical x and y (each: u64 depth count, then per depth u64 size and the
f64 values), then the indexer:

* regular (HierarchicalRegularIndexer::write, Hat.adaptive.inl:262-263): u32
  maxDepth;
* adapted (HierarchicalAdaptedIndexer::write, Hat.adaptive.inl:408-414): the
  tree written raw and depth first (Misha/RegularTree.inl:616-641: the root,
  then every node that has children followed by its 2^dim children), the
  per-depth active counts, and the per-depth function and element node tables
  (u64 count, u64 depth sizes, 8-byte pointers).  A node is the struct
  RegularTreeNode<dim, NodeData, unsigned short> (Misha/RegularTree.h:95-105):
  u16 depth, u16 offset[dim], padding to 8, parent and children pointers,
  u64 functionIndex, u64 elementIndex.

The adapted tables index functions and elements in the hierarchy's numbering;
`functions(d)` / `elements(d)` map that numbering to multi-indices through the
tree nodes (internal depth d+1 holds external depth d).
"""
import struct

import numpy as np

NONE = 0xFFFFFFFFFFFFFFFF


def read_adapted_indexer(b, o, dim):
    """Parses HierarchicalAdaptedIndexer::write output starting at byte o of b.
    Returns (fields, end offset): node_depth / node_offset / node_has_children /
    node_function / node_element per tree node in file order, active_sizes,
    num_functions, num_elements."""
    head = ((2 * (dim + 1) + 7) // 8) * 8
    nb = head + 32
    recs = []

    def rec():
        nonlocal o
        h = struct.unpack_from("<%dH" % (dim + 1), b, o)
        _, children, fi, ei = struct.unpack_from("<4Q", b, o + head)
        o += nb
        recs.append((h[0],) + tuple(h[1:]) + (1 if children else 0, fi, ei))
        return children != 0

    # ReadChildren (RegularTree.inl:645-662): a node with children is followed
    # by its brood, then each child's subtree (recursion depth = tree depth)
    def read_children(has):
        if has:
            for k in [rec() for _ in range(1 << dim)]:
                read_children(k)

    read_children(rec())
    a = np.array(recs, dtype=np.uint64)
    fi, ei = a[:, 2 + dim], a[:, 3 + dim]
    out = {"node_depth": a[:, 0].astype(np.int64), "node_offset": a[:, 1:1 + dim].astype(np.int64),
           "node_has_children": a[:, 1 + dim].astype(np.int64),
           "node_function": np.where(fi == NONE, -1, fi.astype(np.int64)),
           "node_element": np.where(ei == NONE, -1, ei.astype(np.int64))}

    def u64():
        nonlocal o
        v = struct.unpack_from("<Q", b, o)[0]
        o += 8
        return v

    n = u64()
    out["active_sizes"] = np.array([u64() for _ in range(n)], dtype=np.int64)
    for key in ("num_functions", "num_elements"):
        n = u64()
        s = []
        for _ in range(n):
            c = u64()
            s.append(c)
            o += 8 * c
        out[key] = np.array(s, dtype=np.int64)
    return out, o


class TreeFile:
    def __init__(self, path, adapted=None):
        b = open(path, "rb").read()
        self.raw = b
        o = 0

        def u32():
            nonlocal o
            v = struct.unpack_from("<I", b, o)[0]
            o += 4
            return v

        def u64():
            nonlocal o
            v = struct.unpack_from("<Q", b, o)[0]
            o += 8
            return v

        self.dim = dim = u32()
        self.hash = u64()
        o += 8   # vtable pointer
        self.voxel_to_world = np.frombuffer(b, np.float64, (dim + 1) ** 2, o).reshape(dim + 1, dim + 1).copy()
        o += 8 * (dim + 1) ** 2
        coeffs = []
        for _ in range(2):
            n = u64()
            lv = []
            for _ in range(n):
                s = u64()
                lv.append(np.frombuffer(b, np.float64, s, o).copy())
                o += 8 * s
            coeffs.append(lv)
        self.x, self.y = coeffs
        self.depth = len(self.x) - 1
        rest = len(b) - o
        if adapted is None:
            adapted = rest != 4
        self.adapted = adapted
        if not adapted:
            self.max_depth = u32()
            return
        ix, o = read_adapted_indexer(b, o, dim)
        for k, v in ix.items():
            setattr(self, k, v)
        if o != len(b):
            raise ValueError(f"{len(b) - o} trailing bytes in {path}")

    def _table(self, d, which):
        sel = (self.node_depth == d + 1) & ((self.node_function if which == "f" else self.node_element) >= 0)
        idx = (self.node_function if which == "f" else self.node_element)[sel]
        off = self.node_offset[sel]
        n = self.num_functions[d] if which == "f" else self.num_elements[d]
        out = np.full((n, self.dim), -1, dtype=np.int64)
        out[idx] = off
        return out

    def functions(self, d):
        """Multi-index of every function of depth d, in the hierarchy's numbering."""
        return self._table(d, "f")

    def elements(self, d):
        """Multi-index of every element of depth d, in the hierarchy's numbering."""
        return self._table(d, "e")

    def dense(self, d, values):
        """Scatters per-function values of depth d into the dense (2^d+1)^dim layout
        (coordinate 0 most significant)."""
        n1 = (1 << d) + 1
        out = np.zeros(n1 ** self.dim)
        if not self.adapted:
            out[:] = values
            return out
        F = self.functions(d)
        flat = np.zeros(len(F), dtype=np.int64)
        for k in range(self.dim):
            flat = flat * n1 + F[:, k]
        out[flat] = values
        return out
