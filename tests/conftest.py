import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ["knot3_d4", "knot3_d5", "knot3_d4_params", "clifford4_d3"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def case_args(g):
    """Reference command-line arguments of a fixture; fixtures without aRadius are dense-grid runs."""
    args = {k[4:]: (g[k].item()) for k in g if k.startswith("arg_")}
    args.setdefault("aRadius", -1)
    return args


def read_grid_values(path):
    """Values of a RegularGrid file (Misha/RegularGrid.inl:440-471): "G<dim>",
    "<n> DOUBLE", the resolutions, dim+1 transform rows, then the binary values
    (an array of shape (count,) or (count, n))."""
    b = open(path, "rb").read()
    o = 0
    lines = []
    while len(lines) < 3:
        e = b.index(b"\n", o); lines.append(b[o:e].decode()); o = e + 1
    dim = int(lines[0][1:])
    nv = int(lines[1].split()[0])
    for _ in range(dim + 1):
        o = b.index(b"\n", o) + 1
    v = np.frombuffer(b, dtype="<f8", offset=o)
    return v if nv == 1 else v.reshape(-1, nv)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    scale = max(np.abs(b).max(), 1e-300) if b.size else 1.0
    return float(np.abs(a - b).max() / scale) if b.size else 0.0


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


# ---- the reference's Extract PLY outputs (Misha/Ply.inl:673-737, PlyFile.inl:314-356:
# binary little endian, faces as "list int int vertex_indices") -------------------
def ply_bytes(props, nrows, row_bytes, faces, k):
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {nrows}"]
    head += [f"property {p}" for p in props]
    head += [f"element face {len(faces)}", "property list int int vertex_indices", "end_header", ""]
    f = np.empty((len(faces), k + 1), dtype="<i4")
    f[:, 0] = k
    f[:, 1:] = np.asarray(faces).reshape(-1, k)
    return "\n".join(head).encode() + row_bytes + f.tobytes()


def mesh_ply(dim, V, S):
    """Extract.cpp --out: x.. and n_c_d doubles per vertex, (dim-1)-vertex faces."""
    props = [f"double {a}" for a in "xyzw"[:dim]] + [f"double n_{c}_{d}" for c in range(2) for d in range(dim)]
    return ply_bytes(props, len(V), np.ascontiguousarray(V, "<f8").tobytes(), S, dim - 1)


def level_sets_ply(dim, sets):
    """Extract.cpp --levelSets: both components' sets, colours (223,41,53) and (55,114,255) as uchar."""
    props = [f"double {a}" for a in "xyzw"[:dim]] + ["uchar red", "uchar green", "uchar blue"]
    colours = ((223, 41, 53), (55, 114, 255))
    row = np.dtype([("p", "<f8", (dim,)), ("c", "u1", (3,))])
    rows, faces, off = [], [], 0
    for c, (V, S) in enumerate(sets):
        r = np.zeros(len(V), dtype=row)
        r["p"], r["c"] = np.asarray(V).reshape(-1, dim), colours[c]
        rows.append(r.tobytes())
        faces.append(np.asarray(S, np.int64).reshape(-1, dim) + off)
        off += len(V)
    return ply_bytes(props, off, b"".join(rows), np.concatenate(faces) if faces else np.zeros((0, dim)), dim)


def sha256(b):
    import hashlib
    return hashlib.sha256(b).hexdigest()


EXTRACT_FIXTURES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.startswith("extract_") and f.endswith(".npz"))


def fixture_tree(name, tmp_path):
    """Writes the .tree bytes of an Extract fixture to tmp_path and returns (path, fixture);
    a density file, when the fixture has one, goes to tmp_path/<name>.density.tree."""
    g = load_golden(name)
    p = str(tmp_path / (name + ".tree"))
    with open(p, "wb") as f:
        f.write(g["tree"].tobytes())
    if "density" in g:
        with open(str(tmp_path / (name + ".density.tree")), "wb") as f:
            f.write(g["density"].tobytes())
    return p, g


TRIM_FIXTURES = [n for n in EXTRACT_FIXTURES if "density" in np.load(os.path.join(GOLDEN, n + ".npz")).files]


def decode_density(path, dim):
    """A density file (ExteriorPoissonRecon's <out>.density.tree): the voxel-to-world
    matrix, then per node in file order depth, offsets, has-children flag and the
    DensityAndNoiseInfo statistics (weightSum, positionSum, covarianceSum; the
    matrix's vtable slot dropped), then the target samples per cell."""
    b = open(path, "rb").read()
    o = 8
    v2w = np.frombuffer(b, "<f8", (dim + 1) ** 2, o).reshape(dim + 1, dim + 1).copy()
    o += 8 * (dim + 1) ** 2
    head = ((2 * (dim + 1) + 7) // 8) * 8
    rec = head + 16 + 8 * (2 + dim + dim * dim)
    n = (len(b) - o - 8) // rec
    assert o + n * rec + 8 == len(b)
    raw = np.frombuffer(b, np.uint8, n * rec, o).reshape(n, rec)
    h = raw[:, :2 * (dim + 1)].copy().view("<u2").reshape(n, dim + 1)
    ch = raw[:, head + 8:head + 16].copy().view("<u8").ravel()
    data = raw[:, head + 16:].copy().view("<f8").reshape(n, 2 + dim + dim * dim)
    stats = np.concatenate([data[:, :1 + dim], data[:, 2 + dim:]], axis=1)
    return {"voxel_to_world": v2w, "node_depth": h[:, 0].astype(np.int16), "node_offset": h[:, 1:].astype(np.int16),
            "node_has_children": (ch != 0).astype(np.int8), "stats": stats,
            "target": np.frombuffer(b, "<f8", 1, len(b) - 8)[0]}
