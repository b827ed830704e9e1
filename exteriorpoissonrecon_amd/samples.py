"""Synthetic oriented-sample generators for the benchmark workloads.

Each generator returns ``(positions, skews)`` with ``positions`` of shape
``(n, dim)`` and ``skews`` of shape ``(n, dim, dim)`` holding the full
skew-symmetric frame matrix ``m[c][r]`` that the PLY properties
``skew_c_r`` store (reference reader: ``Misha/PlyVertexData.inl:228-234``,
``ExteriorPoissonRecon/ExteriorPoissonRecon.cpp:282-303``).

For curves in 3D the frame is the Hodge dual of the unit tangent, using the
reference sample convention ``n((d+1)%3,(d+2)%3) = t[d]``
(``Include/Samples.h:74-79``).  For surfaces in 4D the frame is the tangent
2-blade ``(u v^T - v u^T)/2`` of the two tangent vectors.
"""
import numpy as np


def dual3(t):
    """Skew matrix n with n[(d+1)%3,(d+2)%3] = t[d] and n antisymmetric."""
    n = np.zeros(t.shape[:-1] + (3, 3))
    for d in range(3):
        a, b = (d + 1) % 3, (d + 2) % 3
        n[..., a, b] = t[..., d]
        n[..., b, a] = -t[..., d]
    return n


def blade(u, v):
    """Skew matrix (u v^T - v u^T)/2 of two vectors (batched)."""
    return 0.5 * (u[..., :, None] * v[..., None, :] - v[..., :, None] * u[..., None, :])


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def torus_knot(n, p=2, q=3, seed=0):
    """(p,q) torus knot ((2+cos qt)cos pt, (2+cos qt)sin pt, sin qt), t ~ U[0,2pi)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = rng.uniform(0.0, 2.0 * np.pi, n)
    r = 2.0 + np.cos(q * t)
    pos = np.stack([r * np.cos(p * t), r * np.sin(p * t), np.sin(q * t)], axis=-1)
    dr = -q * np.sin(q * t)
    tan = np.stack([dr * np.cos(p * t) - r * p * np.sin(p * t),
                    dr * np.sin(p * t) + r * p * np.cos(p * t),
                    q * np.cos(q * t)], axis=-1)
    return pos, dual3(_unit(tan))


def spiral(n, r=3, seed=0):
    """Spiral with r rotations (cos 2 pi r s, 2 s - 1, sin 2 pi r s), s ~ U[0,1)
    (the curve of the reference's spiral:r sampler, Samples.h:321-331)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = rng.uniform(0.0, 1.0, n)
    X = 2.0 * np.pi * r
    pos = np.stack([np.cos(X * s), -1.0 + 2.0 * s, np.sin(X * s)], axis=-1)
    tan = np.stack([-X * np.sin(X * s), np.full(n, 2.0), X * np.cos(X * s)], axis=-1)
    return pos, dual3(_unit(tan))


def batch_workload(count=256, seed=0, depths=(4, 5, 6)):
    """The batched serving workload: `count` independent 3D inputs, each a
    torus knot with random coprime (p, q) in [2,7]^2 or a spiral with r in
    [2,6] rotations, 2048-16384 samples, depth uniform over `depths`.
    Returns [(pos, skew, depth, description)]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        n = int(rng.integers(2048, 16385))
        depth = int(rng.choice(depths))
        if rng.random() < 0.5:
            while True:
                p, q = (int(v) for v in rng.integers(2, 8, 2))
                if np.gcd(p, q) == 1:
                    break
            pos, skew = torus_knot(n, p, q, seed=seed * 100003 + k)
            desc = f"torus_knot:{p}:{q}"
        else:
            r = int(rng.integers(2, 7))
            pos, skew = spiral(n, r, seed=seed * 100003 + k)
            desc = f"spiral:{r}"
        out.append((pos, skew, depth, desc))
    return out


def _rotation(rng, axis, angle):
    axis = _unit(axis)
    k = np.zeros(axis.shape[:-1] + (3, 3))
    k[..., 0, 1], k[..., 0, 2] = -axis[..., 2], axis[..., 1]
    k[..., 1, 0], k[..., 1, 2] = axis[..., 2], -axis[..., 0]
    k[..., 2, 0], k[..., 2, 1] = -axis[..., 1], axis[..., 0]
    s = np.sin(angle)[..., None, None]
    c = np.cos(angle)[..., None, None]
    eye = np.broadcast_to(np.eye(3), k.shape)
    return eye + s * k + (1 - c) * (k @ k)


def borromean(n=16384, pos_sigma=0.005, ang_deg=5.0, spot=0.1, seed=0):
    """Three mutually orthogonal ellipses (semi-axes 1.0, 0.5) in the xy, yz and
    zx planes with positional/angular noise and a fraction of uniform spot
    samples with random skew frames."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = [n // 3 + (1 if i < n % 3 else 0) for i in range(3)]
    pos_l, tan_l = [], []
    for k, m in enumerate(counts):
        t = rng.uniform(0.0, 2.0 * np.pi, m)
        a, b = np.cos(t), 0.5 * np.sin(t)
        da, db = -np.sin(t), 0.5 * np.cos(t)
        p = np.zeros((m, 3)); tt = np.zeros((m, 3))
        i, j = k, (k + 1) % 3
        p[:, i], p[:, j] = a, b
        tt[:, i], tt[:, j] = da, db
        pos_l.append(p); tan_l.append(tt)
    pos = np.concatenate(pos_l); tan = _unit(np.concatenate(tan_l))
    diam = np.linalg.norm(pos.max(0) - pos.min(0))
    pos = pos + rng.normal(0.0, pos_sigma * diam, pos.shape)
    axis = rng.normal(size=(n, 3))
    ang = np.deg2rad(ang_deg) * rng.uniform(0.0, 1.0, n)
    tan = np.einsum("nij,nj->ni", _rotation(rng, axis, ang), tan)
    skew = dual3(_unit(tan))
    ns = int(np.ceil(n * spot))
    lo, hi = pos.min(0), pos.max(0)
    spos = rng.uniform(lo, hi, (ns, 3))
    sv = rng.uniform(-1.0, 1.0, (ns, 3))
    return np.concatenate([pos, spos]), np.concatenate([skew, dual3(sv)])


def clifford_torus(side=256, ang_deg=2.0, seed=0):
    """(cos u, sin u, cos v, sin v)/sqrt(2), u,v ~ U[0,2pi)^2, side^2 samples,
    frame = tangent 2-blade of (d/du, d/dv), with small angular noise."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = side * side
    u = rng.uniform(0.0, 2.0 * np.pi, n)
    v = rng.uniform(0.0, 2.0 * np.pi, n)
    s = 1.0 / np.sqrt(2.0)
    pos = s * np.stack([np.cos(u), np.sin(u), np.cos(v), np.sin(v)], -1)
    tu = np.stack([-np.sin(u), np.cos(u), 0 * u, 0 * u], -1)
    tv = np.stack([0 * u, 0 * u, -np.sin(v), np.cos(v)], -1)
    if ang_deg > 0:
        eps = np.deg2rad(ang_deg)
        tu = _unit(tu + eps * rng.normal(size=tu.shape) / 2.0)
        tv = _unit(tv + eps * rng.normal(size=tv.shape) / 2.0)
    return pos, blade(tu, tv)


def dual4(t):
    """Hodge dual of a 4x4 skew matrix with the reference's index map
    (Surface::Dual, Include/Samples.h:798-808): n(1,0) = t(2,3), n(2,0) = t(3,1),
    n(3,0) = t(1,2), n(0,1) = t(3,2), n(2,1) = t(0,3), n(3,1) = t(2,0),
    n(0,2) = t(1,3), n(1,2) = t(3,0), n(3,2) = t(0,1), n(0,3) = t(2,1),
    n(1,3) = t(0,2), n(2,3) = t(1,0)."""
    n = np.zeros(t.shape)
    pairs = {(1, 0): (2, 3), (2, 0): (3, 1), (3, 0): (1, 2), (0, 1): (3, 2), (2, 1): (0, 3), (3, 1): (2, 0),
             (0, 2): (1, 3), (1, 2): (3, 0), (3, 2): (0, 1), (0, 3): (2, 1), (1, 3): (0, 2), (2, 3): (1, 0)}
    for (a, b), (c, d) in pairs.items():
        n[..., a, b] = t[..., c, d]
    return n


def hopf_map_torus(s, t, lat, dlat):
    """Points and frames of a Hopf torus at parameters (s, t) in [0,1)^2: the
    spherical curve (cos g cos f, cos g sin f, sin g) with longitude f = 2 pi s
    and latitude g = lat(s) (derivative dlat) is lifted through the Hopf
    fibration as the reference does (Samples.h:1072-1107): with (a, b, c) the
    curve point and T = 2 pi t, the point is ((1+c) cos T, a sin T - b cos T,
    a cos T + b sin T, (1+c) sin T) / sqrt(2 (1+c)); the frame is the Hodge dual
    of the tangent blade (d/ds, d/dt), which is the reference's frame with its
    sign (its derivative along the curve comes from the curve sampler's dual
    frame)."""
    f = 2.0 * np.pi * s
    g, dg = lat, dlat
    a, b, c = np.cos(g) * np.cos(f), np.cos(g) * np.sin(f), np.sin(g)
    da = -np.sin(g) * dg * np.cos(f) - np.cos(g) * np.sin(f) * 2.0 * np.pi
    db = -np.sin(g) * dg * np.sin(f) + np.cos(g) * np.cos(f) * 2.0 * np.pi
    dc = np.cos(g) * dg
    T = 2.0 * np.pi * t
    ct, st = np.cos(T), np.sin(T)
    k = 1.0 / np.sqrt(2.0 * (1.0 + c))
    dk = -dc / (2.0 * (1.0 + c)) ** 1.5
    pos = np.stack([(1 + c) * ct * k, (a * st - b * ct) * k, (a * ct + b * st) * k, (1 + c) * st * k], -1)
    d_t = np.stack([-(1 + c) * st * k, (a * ct + b * st) * k, (-a * st + b * ct) * k, (1 + c) * ct * k], -1)
    d_s = np.stack([dc * ct * k + (1 + c) * ct * dk,
                    (da * st - db * ct) * k + (a * st - b * ct) * dk,
                    (da * ct + db * st) * k + (a * ct + b * st) * dk,
                    dc * st * k + (1 + c) * st * dk], -1)
    return pos, dual4(blade(d_s, d_t))


def hopf_torus(side=256, lobes=3, amp=0.3, latitude="sin", phase=0.0, ang_deg=0.0, seed=0, rotation=None):
    """Hopf torus: the preimage under the Hopf map of a spherical curve of
    `lobes` lobes, side^2 samples with (s, t) ~ U[0,1)^2.  latitude="sin": the
    polar angle pi/2 + amp sin(2 pi lobes s + phase) (the benchmark's curve
    theta = pi/2 + 0.3 sin(3 s), phi = s); latitude="cos": latitude
    amp cos(2 pi lobes s + phase), the reference's hopf_torus:lobes:amp curve
    (Samples.h:596-608).  Optional angular noise of the frame's two tangents
    and a 4x4 rotation of positions and frames.  With latitude="cos" and the
    reference Sample's parameter draws the points and frames are the reference's
    (tests/test_host.py)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = side * side
    s = rng.uniform(0.0, 1.0, n)
    t = rng.uniform(0.0, 1.0, n)
    w = 2.0 * np.pi * lobes
    if latitude == "sin":
        lat, dlat = -amp * np.sin(w * s + phase), -amp * w * np.cos(w * s + phase)
    else:
        lat, dlat = amp * np.cos(w * s + phase), -amp * w * np.sin(w * s + phase)
    pos, skew = hopf_map_torus(s, t, lat, dlat)
    if ang_deg > 0:
        # perturb the normal plane: rotate the frame by a small random rotation
        eps = np.deg2rad(ang_deg)
        A = rng.normal(size=(n, 4, 4)) * eps / 2.0
        A = A - np.swapaxes(A, -1, -2)
        Q = np.eye(4) + A + 0.5 * A @ A
        skew = Q @ skew @ np.swapaxes(Q, -1, -2)
    if rotation is not None:
        pos = pos @ rotation.T
        skew = rotation @ skew @ rotation.T
    return pos, skew


def random_rotation(dim, rng):
    """A uniformly random rotation (Haar measure, QR of a Gaussian matrix)."""
    q, r = np.linalg.qr(rng.normal(size=(dim, dim)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def reference_sample(sample_type, res=1024, regular=False, a_noise=0.0, p_noise=0.0, s_noise=0.0):
    """The reference's Sample program (Sample/Sample.cpp, Include/Samples.h)
    restated in the library (epr_sample_generate): `sample_type` as its
    --type (torus_knot:<p>:<q>, spiral:<r>, borromean_rings, spherical:1:<n>:<l>
    curves in 3D; clifford_torus, hopf_torus:<n>:<l> surfaces in 4D), res
    samples (res^2 for surfaces) at random parameters from
    std::default_random_engine (or the regular (i+0.5)/res grid), mapped to the
    unit cube, with --aNoise (degrees), --pNoise and --sNoise noise, and mapped
    back.  Returns (positions n x dim, frames n x dim x dim) equal bit for bit
    to what the reference writes, so write_ply gives its bytes."""
    import ctypes
    from ._lib import EprError, load
    L = load()
    dim, n = ctypes.c_int(), ctypes.c_size_t()
    args = (sample_type.encode(), int(res), 1 if regular else 0, float(a_noise), float(p_noise), float(s_noise))
    rc = L.epr_sample_generate(*args, ctypes.byref(dim), ctypes.byref(n), None, None, 0)
    if rc != 0:
        msg = L.epr_last_error(None)
        raise EprError(f"epr error {rc}: {msg.decode() if msg else ''}")
    pos = np.zeros((n.value, dim.value))
    skew = np.zeros((n.value, dim.value, dim.value))
    pp = pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    sp = skew.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = L.epr_sample_generate(*args, ctypes.byref(dim), ctypes.byref(n), pp, sp, n.value)
    if rc != 0:
        msg = L.epr_last_error(None)
        raise EprError(f"epr error {rc}: {msg.decode() if msg else ''}")
    return pos, skew


def write_ply(path, pos, skew):
    """Binary little-endian PLY with x,y,z[,w] and skew_c_r (double): the
    bytes the reference's Sample writes for the same values
    (PLY::WritePoints with a position and a "skew" matrix factory,
    Sample/Sample.cpp:225-229)."""
    n, dim = pos.shape
    names = ["x", "y", "z", "w"][:dim]
    props = names + [f"skew_{c}_{r}" for c in range(dim) for r in range(dim)]
    header = "ply\nformat binary_little_endian 1.0\n"
    header += f"element vertex {n}\n"
    header += "".join(f"property double {p}\n" for p in props)
    header += "end_header\n"
    data = np.concatenate([pos, skew.reshape(n, dim * dim)], axis=1).astype("<f8")
    with open(path, "wb") as f:
        f.write(header.encode("ascii"))
        f.write(data.tobytes())


def read_ply(path):
    """Reads a PLY written by write_ply (binary little endian or ASCII)."""
    with open(path, "rb") as f:
        raw = f.read()
    end = raw.index(b"end_header") + len(b"end_header")
    while raw[end:end + 1] in (b"\r", b"\n"):
        end += 1
        if raw[end - 1:end] == b"\n":
            break
    header = raw[:end].decode("ascii").splitlines()
    n = 0; props = []; ascii_fmt = False; in_vertex = False
    for line in header:
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "format":
            ascii_fmt = tok[1] == "ascii"
        elif tok[0] == "element":
            in_vertex = tok[1] == "vertex"
            if in_vertex:
                n = int(tok[2])
        elif tok[0] == "property" and in_vertex:
            props.append((tok[-1], tok[1]))
    types = {"double": "<f8", "float": "<f4", "int": "<i4", "uchar": "u1"}
    dt = np.dtype([(nm, types[t]) for nm, t in props])
    if ascii_fmt:
        body = raw[end:].decode("ascii").split()
        vals = np.array(body[: n * len(props)], dtype=np.float64).reshape(n, len(props))
        rec = {nm: vals[:, i] for i, (nm, _) in enumerate(props)}
    else:
        arr = np.frombuffer(raw, dtype=dt, count=n, offset=end)
        rec = {nm: arr[nm].astype(np.float64) for nm, _ in props}
    dim = 4 if "w" in rec else 3
    pos = np.stack([rec[c] for c in ["x", "y", "z", "w"][:dim]], -1)
    skew = np.zeros((n, dim, dim))
    for c in range(dim):
        for r in range(dim):
            skew[:, c, r] = rec[f"skew_{c}_{r}"]
    return pos, skew
