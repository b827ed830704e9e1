"""The oracle's Extract pipeline (oracle/epr_extract.cpp) against the reference's
own Extract runs (tests/golden/extract_*.npz, written by the reference build).

Every PLY the reference wrote -- the oriented joint zero set with frames for
all = true (--out alone) and all = false (--out with --levelSets), and the two
codimension-1 sets of --levelSets -- must be reproduced byte for byte from the
same .tree: cumulative values, zero-crossing refinement, tree order,
triangulation, level set, orientation, gradient frames and the voxel-to-world
map all take part; with a density file and --trimDensity, the samples-per-cell
estimates, the face-adjacent components and their traversal order too.  The
--grid values agree to 1e-12 relative (the reference
sums the prolongation with parallel atomics, so its last bits vary).  The
orientation is also checked for consistency on its own: every face shared by
two simplices appears with opposite orientations."""
import numpy as np
import pytest

from conftest import EXTRACT_FIXTURES, TRIM_FIXTURES, fixture_tree, level_sets_ply, mesh_ply, sha256


@pytest.mark.parametrize("name", EXTRACT_FIXTURES)
def test_oracle_extract_matches_reference(name, tmp_path, oracle_lib):
    tree, g = fixture_tree(name, tmp_path)
    dim = oracle_lib.tree_header(tree)[0]
    for key, all_ in (("mesh", True), ("mesh_any", False)):
        r = oracle_lib.extract(tree, all_)
        b = mesh_ply(dim, r["vertices"], r["simplices"])
        assert [len(r["vertices"]), len(r["simplices"])] == g[key + "_counts"].tolist(), key
        assert sha256(b) == str(g[key + "_sha256"]), key
    sets = [oracle_lib.extract_level_set(tree, c) for c in range(2)]
    assert sha256(level_sets_ply(dim, sets)) == str(g["level_sets_sha256"])
    if "grid0" in g:
        vals = r["values"]
        for c in range(2):
            ref = g[f"grid{c}"]
            assert np.abs(vals[:, c] - ref).max() <= 1e-12 * np.abs(ref).max()


def oriented(S):
    """MarchingSimplices::IsOriented (MarchingSimplices.h:249-286) restated on index arrays."""
    K = S.shape[1] - 1
    seen = {}
    for s in S:
        for v in range(K + 1):
            if K == 1:
                face, even = (int(s[1]),) if v == 0 else (int(s[0]),), v == 0
            else:
                f = [int(s[(v + d + 1) % (K + 1)]) for d in range(K)]
                if (K & 1) and (v & 1):
                    f[0], f[1] = f[1], f[0]
                inv = sum(1 for i in range(len(f)) for j in range(i + 1, len(f)) if f[i] > f[j])
                face, even = tuple(sorted(f)), inv % 2 == 0
            if face in seen:
                if seen[face] == even:
                    return False
            else:
                seen[face] = even
    return True


@pytest.mark.parametrize("name", [n for n in EXTRACT_FIXTURES if "knot3" in n])
def test_oracle_orientation_is_consistent(name, tmp_path, oracle_lib):
    tree, _ = fixture_tree(name, tmp_path)
    r = oracle_lib.extract(tree, True)
    assert oriented(r["simplices"])
    # planting a flip breaks it
    S = r["simplices"].copy()
    S[len(S) // 2, [0, 1]] = S[len(S) // 2, [1, 0]]
    assert not oriented(S)


@pytest.mark.parametrize("name", TRIM_FIXTURES)
def test_oracle_density_trimming_matches_reference(name, tmp_path, oracle_lib):
    tree, g = fixture_tree(name, tmp_path)
    dim = oracle_lib.tree_header(tree)[0]
    r = oracle_lib.extract(tree, True, str(tmp_path / (name + ".density.tree")), float(g["trim"]))
    assert [len(r["vertices"]), len(r["simplices"])] == g["trimmed_counts"].tolist()
    assert sha256(mesh_ply(dim, r["vertices"], r["simplices"])) == str(g["trimmed_sha256"])
    # the maxima the reference prints (6 significant digits)
    printed = g["printed_component_max"]
    assert len(r["component_max"]) == len(printed)
    assert np.all(np.abs(r["component_max"] - printed) <= 5e-6 * np.abs(printed))


ORIENT_CASES = ["torus", "mobius", "klein", "mixed", "curves"]


@pytest.mark.parametrize("case", ORIENT_CASES)
def test_oracle_orient_matches_reference(case, oracle_lib):
    """MarchingSimplices::Orient on shuffled, partly flipped complexes of known
    topology (tests/golden/orient_cases.npz, written by the reference's own
    Orient): orientable ones come out consistent, and on the Moebius strip and
    the Klein bottle the walk's order decides -- both reproduced."""
    import os
    from conftest import GOLDEN
    g = np.load(os.path.join(GOLDEN, "orient_cases.npz"))
    out = oracle_lib.orient(g[case + "_in"])
    assert np.array_equal(out, g[case + "_out"])
    assert oriented(out) == (case in ("torus", "curves"))


def test_oracle_orient_rejects_three_incidences(oracle_lib):
    S = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], dtype=np.uint32)    # edge 0-1 in three triangles
    with pytest.raises(ValueError, match="Both faces occupied"):
        oracle_lib.orient(S)
