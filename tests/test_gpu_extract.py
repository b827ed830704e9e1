"""The GPU Extract (epr_iso_*, bin/Extract) against the reference's own Extract
runs (tests/golden/extract_*.npz) and against the oracle on larger trees.

Bar: every PLY byte-identical to the reference's -- the oriented joint zero
set with frames (all = true and all = false) and both codimension-1 sets --
and the finest values, level-set nodes and meshes identical bit for bit to the
oracle's; with a density file, the trimmed mesh byte-identical to the
reference's and the per-component maxima identical to the oracle's.  The
--grid values agree with the reference's to 1e-12 relative (its prolongation
sums with parallel atomics) and their headers byte for byte."""
import os
import subprocess

import numpy as np
import pytest

from conftest import EXTRACT_FIXTURES, ROOT, TRIM_FIXTURES, fixture_tree, level_sets_ply, mesh_ply, read_grid_values, sha256

pytestmark = pytest.mark.gpu
EXTRACT = os.path.join(ROOT, "exteriorpoissonrecon_amd", "bin", "Extract")
REF_EXTRACT = os.path.join(ROOT, "oracle", "_ref", "Extract")


@pytest.mark.parametrize("name", EXTRACT_FIXTURES)
def test_gpu_extract_matches_reference(name, tmp_path, oracle_lib):
    from exteriorpoissonrecon_amd.extract import IsoTree, level_sets_ply_bytes, mesh_ply_bytes
    print("case", name, flush=True)
    tree, g = fixture_tree(name, tmp_path)
    iso = IsoTree(tree)
    dim = iso.dim
    for key, all_ in (("mesh", True), ("mesh_any", False)):
        V, S = iso.extract(all_)
        assert [len(V), len(S)] == g[key + "_counts"].tolist(), key
        b = mesh_ply_bytes(dim, V, S)
        assert b == mesh_ply(dim, V, S)
        assert sha256(b) == str(g[key + "_sha256"]), key
    sets = [iso.level_set(c) for c in range(2)]
    b = level_sets_ply_bytes(dim, sets)
    assert b == level_sets_ply(dim, sets)
    assert sha256(b) == str(g["level_sets_sha256"])
    # the stages against the oracle, bit for bit
    o = oracle_lib.extract(tree, True)
    assert np.array_equal(iso.values(), o["values"])
    assert np.array_equal(iso.level_set_nodes(True), o["cells"])
    if "grid0" in g:
        iso.write_grid(str(tmp_path / "g"))
        for c in range(2):
            mine = read_grid_values(str(tmp_path / f"g.{c}.grid"))
            ref = g[f"grid{c}"]
            assert np.abs(mine - ref).max() <= 1e-12 * np.abs(ref).max()
            head = open(str(tmp_path / f"g.{c}.grid"), "rb").read()[:-8 * len(mine)]
            assert head == g["grid_header"].tobytes()
    iso.close()


@pytest.mark.parametrize("name", TRIM_FIXTURES)
def test_gpu_density_trimming_matches_reference(name, tmp_path, oracle_lib):
    from exteriorpoissonrecon_amd.extract import IsoTree, mesh_ply_bytes
    print("case", name, flush=True)
    tree, g = fixture_tree(name, tmp_path)
    dens = str(tmp_path / (name + ".density.tree"))
    iso = IsoTree(tree)
    iso.read_density(dens)
    V, S = iso.extract(True, float(g["trim"]))
    assert [len(V), len(S)] == g["trimmed_counts"].tolist()
    assert sha256(mesh_ply_bytes(iso.dim, V, S)) == str(g["trimmed_sha256"])
    o = oracle_lib.extract(tree, True, dens, float(g["trim"]))
    assert np.array_equal(iso.component_densities(), o["component_max"])
    # no trimming without a threshold, even with a density read
    V0, S0 = iso.extract(True, 0.0)
    assert sha256(mesh_ply_bytes(iso.dim, V0, S0)) == str(g["mesh_sha256"])
    iso.close()
    res = subprocess.run([EXTRACT, "--in", tree, "--out", str(tmp_path / "t.ply"), "--density", dens,
                          "--trimDensity", str(float(g["trim"])), "--verbose"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0 and "[ERROR]" not in res.stderr, res.stderr
    assert sha256(open(str(tmp_path / "t.ply"), "rb").read()) == str(g["trimmed_sha256"])
    printed = [float(ln.split()[-1]) for ln in res.stdout.splitlines() if ln.startswith("Max Density")]
    assert printed == g["printed_component_max"].tolist()


@pytest.mark.parametrize("name", [n for n in EXTRACT_FIXTURES if "knot3_d4" in n or "clifford4_d3" == n[8:]])
def test_gpu_extract_command_line(name, tmp_path):
    """bin/Extract writes the reference's files from the same flags."""
    tree, g = fixture_tree(name, tmp_path)
    m, a, l = (str(tmp_path / f) for f in ("m.ply", "a.ply", "l.ply"))
    for cmd in ([EXTRACT, "--in", tree, "--out", m], [EXTRACT, "--in", tree, "--out", a, "--levelSets", l]):
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert res.returncode == 0 and "[ERROR]" not in res.stderr, res.stderr
    assert sha256(open(m, "rb").read()) == str(g["mesh_sha256"])
    assert sha256(open(a, "rb").read()) == str(g["mesh_any_sha256"])
    assert sha256(open(l, "rb").read()) == str(g["level_sets_sha256"])


@pytest.mark.parametrize("case", [
    dict(dim=3, depth=6, aRadius=3, n=4096),
    dict(dim=3, depth=6, aRadius=-1, n=4096),
    dict(dim=4, depth=4, aRadius=2, side=40),
], ids=["knot3_d6_r3", "knot3_d6_dense", "clifford4_d4_r2"])
def test_gpu_extract_product_trees(case, tmp_path, oracle_lib):
    """Trees written by the GPU reconstructor: the GPU Extract equals the
    oracle's bit for bit and, where the reference build is present, the
    reference Extract's PLY byte for byte."""
    from exteriorpoissonrecon_amd import Reconstructor, samples
    from exteriorpoissonrecon_amd.extract import IsoTree, mesh_ply_bytes
    case = dict(case)
    print("case", case, flush=True)
    dim = case.pop("dim")
    if dim == 3:
        pos, skew = samples.torus_knot(case.pop("n"), seed=4)
    else:
        pos, skew = samples.clifford_torus(side=case.pop("side"), ang_deg=2.0, seed=4)
    r = Reconstructor(pos, skew, **case).build()
    r.solve()
    tree = str(tmp_path / "o.tree")
    r.write_tree(tree, oracle_lib.indexer_hash(dim, case["aRadius"] != -1))
    r.close()
    iso = IsoTree(tree)
    V, S = iso.extract(True)
    o = oracle_lib.extract(tree, True)
    assert np.array_equal(iso.values(), o["values"])
    assert np.array_equal(iso.level_set_nodes(True), o["cells"])
    assert np.array_equal(V, o["vertices"]) and np.array_equal(S, o["simplices"])
    for c in range(2):
        LV, LS = iso.level_set(c)
        OV, OS = oracle_lib.extract_level_set(tree, c)
        assert np.array_equal(LV, OV) and np.array_equal(LS, OS)
    if os.path.exists(REF_EXTRACT):
        ref = str(tmp_path / "ref.ply")
        res = subprocess.run([REF_EXTRACT, "--in", tree, "--out", ref], capture_output=True, text=True, timeout=600)
        assert res.returncode == 0, res.stderr
        assert mesh_ply_bytes(dim, V, S) == open(ref, "rb").read()
    iso.close()


def test_gpu_extract_empty_level_set(tmp_path, oracle_lib):
    """A hierarchy whose functions never vanish together: no level-set nodes, empty outputs."""
    from exteriorpoissonrecon_amd.extract import IsoTree, mesh_ply_bytes
    name = [n for n in EXTRACT_FIXTURES if n == "extract_knot3_d4"][0]
    tree, g = fixture_tree(name, tmp_path)
    b = bytearray(g["tree"].tobytes())
    # shift phi1 up by a constant at depth 0 (regular values of every depth shift with it)
    dim = 3
    o = 4 + 8 + 8 + 8 * (dim + 1) ** 2
    n0 = int(np.frombuffer(b, "<u8", 1, o + 8)[0])
    x0 = np.frombuffer(b, "<f8", n0, o + 16).copy() + 1e6
    b[o + 16:o + 16 + 8 * n0] = x0.tobytes()
    p = str(tmp_path / "shifted.tree")
    open(p, "wb").write(bytes(b))
    iso = IsoTree(p)
    assert len(iso.level_set_nodes(True)) == 0
    V, S = iso.extract(True)
    assert V.shape == (0, 9) and S.shape == (0, 2)
    o = oracle_lib.extract(p, True)
    assert len(o["vertices"]) == 0 and len(o["cells"]) == 0
    assert mesh_ply_bytes(3, V, S).endswith(b"end_header\n")
    iso.close()


@pytest.mark.parametrize("case", ["torus", "mobius", "klein", "mixed", "curves"])
def test_gpu_orient_matches_reference(case):
    """epr_orient: the device parity solve on orientable components, the host
    walk on the Moebius strip and the Klein bottle; the reference's own output
    (tests/golden/orient_cases.npz) in every case."""
    from conftest import GOLDEN
    from exteriorpoissonrecon_amd.extract import orient
    g = np.load(os.path.join(GOLDEN, "orient_cases.npz"))
    assert np.array_equal(orient(g[case + "_in"]), g[case + "_out"])


def test_gpu_orient_rejects_three_incidences():
    from exteriorpoissonrecon_amd import EprError
    from exteriorpoissonrecon_amd.extract import orient
    with pytest.raises(EprError, match="Both faces occupied"):
        orient(np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], dtype=np.uint32))
