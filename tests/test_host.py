"""CPU-only checks of the boundary: the C-ABI library loads and exports every
entry point of include/epr.h, the PLY helpers round-trip, and the command line
reproduces the reference's argument and error behaviour without touching a GPU."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from exteriorpoissonrecon_amd import samples

HEADER = os.path.join(ROOT, "include", "epr.h")
CLI = os.path.join(ROOT, "exteriorpoissonrecon_amd", "bin", "ExteriorPoissonRecon")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(epr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_surface():
    names = declared()
    for must in ["epr_create", "epr_set_samples", "epr_build", "epr_solve", "epr_gs_sweeps",
                 "epr_energies", "epr_prolong_add", "epr_get_coeffs", "epr_write_tree"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    import ctypes
    from exteriorpoissonrecon_amd import _lib
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value
    assert set(declared()) == set(_lib.EXPORTS)


def test_default_params_match_reference():
    from exteriorpoissonrecon_amd import default_params
    p = default_params()
    # ExteriorPoissonRecon/ExteriorPoissonRecon.cpp:56-72
    assert (p.depth, p.seed, p.tRadius, p.dRadius, p.minSolveDepth, p.gsIters) == (5, 0, 1, 1, 0, 10)
    assert (p.nSigma, p.scale, p.sWeight, p.dWeight, p.iMult) == (0.25, 1.1, 100.0, 0.00625, 2.0)
    assert p.aRadius == 3


def test_ply_roundtrip(tmp_path):
    pos, skew = samples.torus_knot(64)
    path = str(tmp_path / "k.ply")
    samples.write_ply(path, pos, skew)
    p2, s2 = samples.read_ply(path)
    np.testing.assert_array_equal(pos, p2)
    np.testing.assert_array_equal(skew, s2)
    # frames are skew-symmetric with unit-tangent duals
    assert np.abs(skew + np.swapaxes(skew, 1, 2)).max() == 0


def run_cli(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_cli_usage_without_input():
    r = run_cli()
    assert r.returncode == 0 and "--in" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--in", "x.txt"], "Only .ply files supported"),
    (["--in", "missing.ply", "--aRadius", "7"], "Only kernel radius in the range [0,5] supported: 7"),
    (["--in", "missing.ply"], "Invalid sample file"),
    (["--in", "missing.ply", "--project"], "--project is not supported"),
])
def test_cli_errors_exit_zero(args, msg):
    # MK_ERROR_OUT prints to stderr and exits with status 0 (Misha/Exceptions.h:101-108)
    r = run_cli(*args)
    assert r.returncode == 0
    assert "[ERROR]" in r.stderr and msg in r.stderr


# ---- the Extract boundary: argument and file errors, before any device work -----------
EXTRACT = os.path.join(ROOT, "exteriorpoissonrecon_amd", "bin", "Extract")


def test_extract_usage_without_input():
    r = subprocess.run([EXTRACT], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--in" in r.stdout and "--levelSets" in r.stdout


def test_indexer_hash_matches_reference_trees():
    """typeid(...).hash_code() as the reference writes it: the fixtures' trees carry these values."""
    from conftest import EXTRACT_FIXTURES, load_golden
    from exteriorpoissonrecon_amd import load
    for name in EXTRACT_FIXTURES:
        t = load_golden(name)["tree"].tobytes()
        dim = int(np.frombuffer(t, "<u4", 1, 0)[0])
        h = int(np.frombuffer(t, "<u8", 1, 4)[0])
        assert h == load().epr_indexer_hash(dim, int("adapted" in name)), name


def _tree_bytes(name):
    from conftest import load_golden
    return bytearray(load_golden(name)["tree"].tobytes())


@pytest.mark.parametrize("edit,msg", [
    (lambda b: b.__setitem__(slice(4, 12), (12345).to_bytes(8, "little")), "Unrecognized hierarchical indexer hash: 12345"),
    (lambda b: b.__setitem__(slice(0, 4), (5).to_bytes(4, "little")), "Only dimensions 3 and 4 supported: 5"),
    (lambda b: b.__delitem__(slice(len(b) - 100, len(b))), "truncated"),
])
def test_extract_file_errors(edit, msg, tmp_path):
    """Extract.cpp:352-363 rejects unknown indexers and dimensions; a short file is an error too."""
    from exteriorpoissonrecon_amd import EprError
    from exteriorpoissonrecon_amd.extract import IsoTree
    b = _tree_bytes("extract_knot3_d4_adapted")
    edit(b)
    p = str(tmp_path / "bad.tree")
    open(p, "wb").write(bytes(b))
    with pytest.raises(EprError, match=msg.replace("[", r"\[")):
        IsoTree(p)
    r = subprocess.run([EXTRACT, "--in", p, "--out", str(tmp_path / "o.ply")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "[ERROR]" in r.stderr and msg in r.stderr


def test_extract_ply_serialisers_agree(tmp_path):
    """The package's PLY writers produce the bytes the reference's writer does (the tests' restatement)."""
    from conftest import level_sets_ply, mesh_ply
    from exteriorpoissonrecon_amd.extract import level_sets_ply_bytes, mesh_ply_bytes
    rng = np.random.default_rng(0)
    for dim in (3, 4):
        V = rng.normal(size=(7, 3 * dim))
        S = rng.integers(0, 7, size=(5, dim - 1))
        assert mesh_ply_bytes(dim, V, S) == mesh_ply(dim, V, S)
        sets = [(rng.normal(size=(4, dim)), rng.integers(0, 4, size=(3, dim))) for _ in range(2)]
        assert level_sets_ply_bytes(dim, sets) == level_sets_ply(dim, sets)


def test_batch_argument_errors():
    """epr_build_batch / epr_solve_batch reject bad context arrays before any
    device call (context-free errors through epr_last_error(NULL))."""
    import ctypes
    from exteriorpoissonrecon_amd._lib import load
    L = load()
    assert L.epr_build_batch(2, None, 0) != 0
    assert b"null context array" in L.epr_last_error(None)
    arr = (ctypes.c_void_p * 2)(None, None)
    assert L.epr_solve_batch(2, arr, None) != 0
    assert b"context 0 is null" in L.epr_last_error(None)
    assert L.epr_solve_batch(0, None, None) == 0 and L.epr_build_batch(0, None, 0) == 0


def test_batch_workload_shape():
    """The batched serving workload: 256 3D inputs, depths 4-6, 2048-16384
    samples, coprime torus knots or spirals; deterministic for a seed."""
    w = samples.batch_workload(256)
    assert len(w) == 256
    for pos, skew, depth, desc in w:
        assert depth in (4, 5, 6) and 2048 <= len(pos) <= 16384
        assert pos.shape == (len(pos), 3) and skew.shape == (len(pos), 3, 3)
        assert np.allclose(skew, -np.transpose(skew, (0, 2, 1)))
        kind, *args = desc.split(":")
        if kind == "torus_knot":
            p, q = map(int, args)
            assert 2 <= p <= 7 and 2 <= q <= 7 and np.gcd(p, q) == 1
        else:
            assert kind == "spiral" and 2 <= int(args[0]) <= 6
    w2 = samples.batch_workload(4)
    assert all(np.array_equal(a[0], b[0]) for a, b in zip(w[:4], w2))


def test_spiral_frames_are_tangent_duals():
    """spiral(): points on (cos 2 pi r s, 2 s - 1, sin 2 pi r s) and frames
    Dual(unit tangent), the reference's spiral:r curve (Samples.h:321-331)."""
    pos, skew = samples.spiral(512, 4, seed=3)
    s = (pos[:, 1] + 1) / 2
    X = 2 * np.pi * 4
    assert np.allclose(pos[:, 0], np.cos(X * s)) and np.allclose(pos[:, 2], np.sin(X * s))
    tan = np.stack([-X * np.sin(X * s), np.full(len(s), 2.0), X * np.cos(X * s)], -1)
    tan /= np.linalg.norm(tan, axis=1, keepdims=True)
    # Dual(t): n((d+1)%3,(d+2)%3) = t[d]
    for d in range(3):
        assert np.allclose(skew[:, (d + 1) % 3, (d + 2) % 3], tan[:, d])


def test_cli_density_grid_needs_out(tmp_path):
    ply = str(tmp_path / "in.ply")
    pos, skew = samples.torus_knot(64, seed=1)
    samples.write_ply(ply, pos, skew)
    cli = os.path.join(ROOT, "exteriorpoissonrecon_amd", "bin", "ExteriorPoissonRecon")
    res = subprocess.run([cli, "--in", ply, "--densityGrid", str(tmp_path / "d.grid")], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0 and "--densityGrid needs --out" in res.stderr


def test_density_grid_argument_errors():
    from exteriorpoissonrecon_amd._lib import load
    L = load()
    assert L.epr_density_grid(None, 3, 4, b"x.grid") != 0
    assert L.epr_density_grid(b"missing.density.tree", 5, 4, b"x.grid") != 0


def test_ply_writer_matches_reference_sample():
    """samples.read_ply / write_ply against PLY files written by the
    reference's own Sample program (tests/golden/sample_plys.npz, 3D curves
    and 4D surfaces, with and without its noise models): reading and writing
    back gives the reference's bytes."""
    import tempfile
    g = np.load(os.path.join(ROOT, "tests", "golden", "sample_plys.npz"))
    names = [k for k in g.files if not k.endswith("_args")]
    assert len(names) == 5
    for name in names:
        ref = g[name].tobytes()
        with tempfile.TemporaryDirectory() as td:
            p = os.path.join(td, "in.ply")
            open(p, "wb").write(ref)
            pos, skew = samples.read_ply(p)
            dim = pos.shape[1]
            assert dim == (4 if name.startswith(("clifford", "hopf")) else 3), name
            assert np.allclose(skew, -np.transpose(skew, (0, 2, 1)))
            q = os.path.join(td, "out.ply")
            samples.write_ply(q, pos, skew)
            assert open(q, "rb").read() == ref, name


def _sample_args(text):
    """Sample's command line (--type, --res, --aNoise, --pNoise, --sNoise, --regular) as keywords."""
    a = text.split()
    kw = {"res": 1024, "regular": False, "a_noise": 0.0, "p_noise": 0.0, "s_noise": 0.0}
    i = 0
    while i < len(a):
        if a[i] == "--regular":
            kw["regular"] = True
            i += 1
            continue
        key, val = a[i], a[i + 1]
        i += 2
        if key == "--type":
            kw["sample_type"] = val
        elif key == "--res":
            kw["res"] = int(val)
        else:
            kw[{"--aNoise": "a_noise", "--pNoise": "p_noise", "--sNoise": "s_noise"}[key]] = float(val)
    return kw


@pytest.mark.parametrize("fixture", ["sample_plys", "sample_extra"])
def test_sample_generator_writes_the_reference_bytes(fixture, tmp_path):
    """The restated Sample program (epr_sample_generate: Samples.h's curve and
    surface samplers with their AutoDiff derivative arithmetic, the unit-cube
    map and AddNoise of Sample.cpp) regenerates every PLY the reference's own
    Sample wrote (tests/golden/make_sample_golden.py: torus knots incl. a gcd-2
    knot, Borromean rings, spirals, the spherical curve, Clifford and Hopf tori,
    random and regular parameters, positional, angular and spot noise) byte
    for byte, through samples.reference_sample + write_ply and through the
    bin/Sample command line."""
    g = np.load(os.path.join(ROOT, "tests", "golden", fixture + ".npz"))
    names = [k for k in g.files if not k.endswith("_args")]
    assert len(names) >= 5
    exe = os.path.join(ROOT, "exteriorpoissonrecon_amd", "bin", "Sample")
    for name in names:
        ref = g[name].tobytes()
        args = str(g[name + "_args"])
        pos, skew = samples.reference_sample(**_sample_args(args))
        p = str(tmp_path / f"{name}.ply")
        samples.write_ply(p, pos, skew)
        assert open(p, "rb").read() == ref, name
        if os.path.exists(exe):
            q = str(tmp_path / f"{name}_cli.ply")
            subprocess.run([exe, *args.split(), "--out", q], check=True, capture_output=True, timeout=120)
            assert open(q, "rb").read() == ref, name


def test_sample_generator_errors():
    from exteriorpoissonrecon_amd import EprError
    with pytest.raises(EprError, match="Could not parse sample type"):
        samples.reference_sample("mobius_strip", 16)
    with pytest.raises(EprError, match="Could not read"):
        samples.reference_sample("torus_knot:2", 16)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message, before touching a device."""
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 2 and "one GPU" in res.stderr and res.stdout == ""


@pytest.mark.parametrize("kw,msg", [
    (dict(depth=13), "depth must be in"),
    (dict(depth=3, aRadius=6), "Only kernel radius in the range"),
    (dict(depth=3, tRadius=5), "Kernel radius must be in range"),
    (dict(depth=3, dRadius=5), "Only radii"),
])
def test_parameter_range_errors(kw, msg):
    """Parameter ranges are checked before any device call (the reference's own messages)."""
    from exteriorpoissonrecon_amd import EprError, Reconstructor, samples
    pos, skew = samples.torus_knot(256, seed=1)
    with pytest.raises(EprError, match=msg):
        Reconstructor(pos, skew, **kw).build()


def test_private_rand_matches_glibc(tmp_path):
    """The library's private copy of glibc's rand() (csrc/epr_rand.h) gives
    glibc's own sequence for every seed, including 0 (glibc treats it as 1)
    and seeds above INT_MAX, without touching the process-wide state."""
    import subprocess
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include "epr_rand.h"
#include <cstdio>
#include <cstdlib>
int main(){
  unsigned seeds[] = { 0u , 1u , 2u , 7u , 12345u , 2147483647u , 2147483648u , 4294967295u };
  for( unsigned s : seeds ){
    srand( s ); epr::GlibcRand g( s );
    for( int i=0 ; i<20000 ; i++ ){ int a = rand() , b = g.rand(); if( a!=b ){ printf( "seed %u draw %d: %d %d\n" , s , i , a , b ); return 1; } }
  }
  srand( 5 ); int a = rand(); epr::GlibcRand g( 9 ); (void)g.random01(); srand( 5 ); if( rand()!=a ) return 2;
  epr::GlibcRand h( 3 ); srand( 3 );
  for( int i=0 ; i<1000 ; i++ ){
    unsigned long long r1 = rand() , r2 = rand();
    double ref = double( (long long)( r1*( (unsigned long long)RAND_MAX+1 ) + r2 ) ) / double( ( (unsigned long long)RAND_MAX+1 )*( (unsigned long long)RAND_MAX+1 ) - 1 );
    if( h.random01()!=ref ) return 3;
  }
  puts( "ok" ); return 0;
}''')
    exe = tmp_path / "t"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "exteriorpoissonrecon_amd", "csrc"),
                           str(src), "-o", str(exe)])
    assert subprocess.run([str(exe)], capture_output=True, text=True).stdout.strip() == "ok"


def test_hopf_sampler_matches_reference_sample(tmp_path):
    """samples.hopf_map_torus with the reference's curve (hopf_torus:3:0.3, the
    latitude 0.3 cos(6 pi s)) at the parameters the reference Sample drew
    (std::default_random_engine = minstd_rand0 seeded 1, each
    uniform_real_distribution<double> value from two draws,
    generate_canonical<double,53>; s then t) gives the reference file's points
    and frames (tests/golden/sample_plys.npz, written by the reference Sample)."""
    z = np.load(os.path.join(ROOT, "tests", "golden", "sample_plys.npz"), allow_pickle=False)
    assert str(z["hopf4_args"]).split()[1] == "hopf_torus:3:0.3"
    f = tmp_path / "h.ply"
    f.write_bytes(z["hopf4"].tobytes())
    P, K = samples.read_ply(str(f))
    x = 1
    R = 2147483646.0

    def draw():
        nonlocal x
        x = (16807 * x) % 2147483647
        return x - 1

    st = []
    for _ in range(len(P)):
        s = (draw() + draw() * R) / (R * R)
        t = (draw() + draw() * R) / (R * R)
        st.append((s, t))
    s, t = np.array(st).T
    w = 6.0 * np.pi
    pos, skew = samples.hopf_map_torus(s, t, 0.3 * np.cos(w * s), -0.3 * w * np.sin(w * s))
    assert np.abs(pos - P).max() < 1e-12
    unit = lambda k: k / np.linalg.norm(k.reshape(len(k), -1), axis=1)[:, None, None]
    assert np.abs(unit(skew) - unit(K)).max() < 1e-12


def test_fixed_divisor_division_is_exact():
    """The multiply-high division the 32-bit index kernels use (FastDiv in
    csrc/epr_kernels.h: q = hi32(n * ceil(2^32 / d)), one down when q d > n)
    equals n // d for every divisor the kernels take (grid extents 2^r + 1 and
    colour-class extents) and numerators up to 2^31 - 1, boundaries included."""
    rng = np.random.default_rng(7)
    divisors = sorted({1, 2, 3} | {(1 << r) + 1 for r in range(0, 14)} | {(1 << r) // 2 + 1 for r in range(1, 14)}
                      | {(1 << r) // 2 for r in range(1, 14)} | set(rng.integers(1, 1 << 16, 200).tolist()))
    for d in divisors:
        m = np.uint64(((1 << 32) - 1) // d + 1)
        k = np.arange(0, (1 << 31) // d + 2, max(1, ((1 << 31) // d) // 4000), dtype=np.uint64)
        n = np.concatenate([k * np.uint64(d), k * np.uint64(d) - np.uint64(1), k * np.uint64(d) + np.uint64(d - 1),
                            rng.integers(0, 1 << 31, 4000).astype(np.uint64), np.array([(1 << 31) - 1], np.uint64)])
        n = n[(n < np.uint64(1 << 31))]
        if d == 1:
            q = n.copy()
        else:
            q = (n * m) >> np.uint64(32)
            q = q - (q * np.uint64(d) > n).astype(np.uint64)
        assert np.array_equal(q, n // np.uint64(d)), d
