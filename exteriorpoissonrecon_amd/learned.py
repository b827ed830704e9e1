"""The learned initializer: a BF16 U-Net that predicts a coarse level's
coefficients from the target field (no reference counterpart).

The network, its training and its use all run in libepr_hip.so (C ABI in
include/epr.h, kernels in csrc/epr_unet.hip); this module drives them:

* ``UNet``: the network (create, parameters, forward, loss gradients, device
  training, save/load);
* ``features``: the network input of a built reconstruction;
* ``training_pairs``: (input, target) pairs from the solver itself -- random
  4D Hopf tori and two-circle tori (3D: torus knots and spirals), each
  reconstructed by the reference's cascade up to the network's level, the
  target being that level's solution;
* ``train``: the training recipe that produced the shipped weights;
* ``Reconstructor.solve_learned`` equivalent ``solve_learned`` and the
  time-to-energy measurement used by the tests and bench.

The shipped weights (``weights/unet4d_l4.npz``) are for 4D reconstructions
started at level 4 (17^4 nodes, the coarse level of a depth-6 4D solve).
"""
import ctypes
import os
import time

import numpy as np

from ._lib import EprError, _p, load
from . import samples

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_4D = os.path.join(HERE, "weights", "unet4d_l4.npz")


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class UNet:
    """The network on the node grid of level `level` ((2^level+1)^dim nodes)."""

    def __init__(self, dim=4, level=4, in_channels=None):
        in_channels = in_channels or dim * (dim - 1) // 2
        L = load()
        h = ctypes.c_void_p()
        rc = L.epr_unet_create(dim, level, in_channels, ctypes.byref(h))
        if rc != 0:
            msg = L.epr_unet_last_error(None)
            raise EprError(f"epr error {rc}: {msg.decode() if msg else ''}")
        self._h = h
        self.dim, self.level, self.in_channels = dim, level, in_channels
        self.num_params = L.epr_unet_num_params(h)
        self.num_nodes = L.epr_unet_num_nodes(h)

    def _check(self, rc):
        if rc != 0:
            msg = load().epr_unet_last_error(self._h)
            raise EprError(f"epr error {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None):
            load().epr_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (self.num_params,):
            raise ValueError(f"expected {self.num_params} parameters")
        self._check(load().epr_unet_set_params(self._h, _fp(p)))

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        self._check(load().epr_unet_get_params(self._h, _fp(p)))
        return p

    def _inputs(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim == 2:
            x = x[None]
        if x.shape[1:] != (self.num_nodes, self.in_channels):
            raise ValueError(f"inputs must be batch x {self.num_nodes} x {self.in_channels}")
        return x

    def forward(self, x):
        x = self._inputs(x)
        y = np.zeros((x.shape[0], self.num_nodes, 2), np.float32)
        self._check(load().epr_unet_forward(self._h, x.shape[0], _fp(x), _fp(y)))
        return y

    def loss_gradients(self, x, t):
        x = self._inputs(x)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(x.shape[0], self.num_nodes, 2)
        g = np.zeros(self.num_params, np.float32)
        loss = ctypes.c_double()
        self._check(load().epr_unet_loss_gradients(self._h, x.shape[0], _fp(x), _fp(t), ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def set_dataset(self, x, t):
        x = self._inputs(x)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(x.shape[0], self.num_nodes, 2)
        self._check(load().epr_unet_set_dataset(self._h, x.shape[0], _fp(x), _fp(t)))

    def train(self, steps, batch=8, lr=1e-3, seed=0):
        losses = np.zeros(steps)
        self._check(load().epr_unet_train(self._h, steps, batch, float(lr), seed, _p(losses)))
        return losses

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez_compressed(path, params=self.get_params(), dim=self.dim, level=self.level,
                            in_channels=self.in_channels)

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            net = cls(int(z["dim"]), int(z["level"]), int(z["in_channels"]))
            net.set_params(z["params"])
        return net


def init_params(dim, in_channels, seed=0):
    """He-uniform weights and zero biases in the canonical order (the output
    layer's weights halved)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    T = 3 ** dim
    c = [16, 32, 64, 128]
    layers = [(in_channels, c[0], 1), (c[0], c[0], 1), (c[0], c[1], 1), (c[1], c[1], 1), (c[1], c[2], 1),
              (c[2], c[2], 1), (c[2], c[3], 1), (c[3], c[3], 1), (c[3], c[2], 1), (c[2], c[2], 1),
              (c[2], c[1], 1), (c[1], c[1], 1), (c[1], c[0], 1), (c[0], c[0], 1), (c[0], 2, 0)]
    parts = []
    for ci, co, relu in layers:
        lim = np.sqrt(6.0 / (T * ci)) * (1.0 if relu else 0.5)
        parts.append(rng.uniform(-lim, lim, T * co * ci))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


def forward_flops(dim, level, in_channels):
    """Algorithmic multiply-add flops (2 per MAC) of one forward pass: every
    convolution's nodes x 3^dim taps x input x output channels."""
    c = [16, 32, 64, 128]
    N = [(2 ** (level - k) + 1) ** dim for k in range(4)]
    convs = [(0, in_channels, c[0]), (0, c[0], c[0]), (1, c[0], c[1]), (1, c[1], c[1]), (2, c[1], c[2]),
             (2, c[2], c[2]), (3, c[2], c[3]), (3, c[3], c[3]), (2, c[3], c[2]), (2, c[2], c[2]),
             (1, c[2], c[1]), (1, c[1], c[1]), (0, c[1], c[0]), (0, c[0], c[0]), (0, c[0], 2)]
    return sum(2 * N[l] * 3 ** dim * ci * co for l, ci, co in convs)


def features(rec, level):
    """The network input of a built Reconstructor at `level` (nodes x C(dim,2))."""
    cd = rec.dim * (rec.dim - 1) // 2
    out = np.zeros((rec.num_functions(level), cd), np.float32)
    rec._check(load().epr_learned_features(rec._h, level, _fp(out)))
    return out


def prolong(v, side, dim):
    """Hat prolongation of node values (side^dim, last coordinate fastest) to the
    next level: per axis new[2i] = old[i], new[2i+1] = (old[i] + old[i+1]) / 2."""
    x = v.reshape((side,) * dim)
    for ax in range(dim):
        n = x.shape[ax]
        shp = list(x.shape); shp[ax] = 2 * n - 1
        y = np.zeros(shp)
        ev = [slice(None)] * dim; ev[ax] = slice(0, None, 2)
        od = [slice(None)] * dim; od[ax] = slice(1, None, 2)
        lo = [slice(None)] * dim; lo[ax] = slice(0, n - 1)
        hi = [slice(None)] * dim; hi[ax] = slice(1, n)
        y[tuple(ev)] = x
        y[tuple(od)] = 0.5 * (x[tuple(lo)] + x[tuple(hi)])
        x = y
    return x.ravel()


def regular_coefficients(rec, level):
    """The regular coefficients of a level from the hierarchical ones a solve leaves."""
    x, y = rec.coefficients(0)
    for d in range(level):
        hx, hy = rec.coefficients(d + 1)
        side = 2 ** d + 1
        x = prolong(x, side, rec.dim) + hx
        y = prolong(y, side, rec.dim) + hy
    return x, y


def _unit_rms(v):
    r = np.sqrt(np.mean(v.astype(np.float64) ** 2))
    return (v / r if r > 0 else v).astype(np.float32)


def random_surface(rng, dim, n):
    """A random training shape: 4D: a Hopf torus (2-5 lobes, amplitude
    0.12-0.45, random phase and latitude form) or a two-circle torus
    (r1 cos u, r1 sin u, r2 cos v, r2 sin v), randomly rotated, with small
    angular noise; 3D: a torus knot with coprime (p, q) in [2,7]^2 or a spiral."""
    seed = int(rng.integers(1 << 31))
    if dim == 4:
        R = samples.random_rotation(4, rng)
        side = int(np.sqrt(n))
        if rng.random() < 0.7:
            return samples.hopf_torus(side, lobes=int(rng.integers(2, 6)), amp=float(rng.uniform(0.12, 0.45)),
                                      latitude="sin" if rng.random() < 0.5 else "cos",
                                      phase=float(rng.uniform(0, 2 * np.pi)), ang_deg=float(rng.uniform(0, 3)),
                                      seed=seed, rotation=R)
        r1 = float(rng.uniform(0.45, 0.89))
        r2 = np.sqrt(1 - r1 * r1)
        pos, skew = samples.clifford_torus(side, ang_deg=float(rng.uniform(0, 3)), seed=seed)
        pos = pos * np.sqrt(2.0) * np.array([r1, r1, r2, r2])
        return pos @ R.T, R @ skew @ R.T
    if rng.random() < 0.5:
        while True:
            p, q = (int(v) for v in rng.integers(2, 8, 2))
            if np.gcd(p, q) == 1:
                break
        return samples.torus_knot(n, p, q, seed=seed)
    return samples.spiral(n, int(rng.integers(2, 7)), seed=seed)


def training_pairs(count, dim=4, level=4, depth=5, samples_per_shape=32768, seed=0, group=32, log=None):
    """(inputs, targets): the features of `count` random shapes reconstructed at
    `depth` and the regular coefficients of their cascade's level `level`
    (maxSolveDepth = level), each target scaled to unit RMS.  Builds and solves
    run `group` at a time through the batch API."""
    from .reconstructor import Reconstructor, build_batch, solve_batch
    rng = np.random.Generator(np.random.PCG64(seed))
    N = (2 ** level + 1) ** dim
    cd = dim * (dim - 1) // 2
    X = np.zeros((count, N, cd), np.float32)
    T = np.zeros((count, N, 2), np.float32)
    recs = []
    t0 = time.perf_counter()
    k = 0
    while k < count:
        m = min(group, count - k)
        shapes = [random_surface(rng, dim, samples_per_shape) for _ in range(m)]
        while len(recs) < m:
            recs.append(Reconstructor(*shapes[len(recs)], depth=depth, aRadius=-1, maxSolveDepth=level))
        for r, (p, s) in zip(recs, shapes):
            r.set_samples(p, s)
        build_batch(recs[:m])
        solve_batch(recs[:m])
        for i in range(m):
            x, y = regular_coefficients(recs[i], level)
            X[k + i] = features(recs[i], level)
            T[k + i] = _unit_rms(np.stack([x, y], -1))
        k += m
        if log:
            log(f"[learned] pairs {k}/{count} ({time.perf_counter() - t0:.1f} s)")
    for r in recs:
        r.close()
    return X, T


def train(net, X, T, steps, batch=8, lr=1e-3, seed=0, log=None, chunk=200, warmup=200):
    """Adam on the device over (X, T): the learning rate ramps up linearly over
    the first `warmup` steps (in chunks of 25), then follows a cosine decay to
    a tenth.  Returns every step's loss."""
    net.set_dataset(X, T)
    losses = []
    done = 0
    while done < steps:
        m = min(25 if done < warmup else chunk, steps - done)
        if done < warmup:
            rate = lr * (done + m) / warmup
        else:
            f = (done - warmup) / max(1, steps - warmup)
            rate = lr * (0.1 + 0.9 * 0.5 * (1 + np.cos(np.pi * f)))
        losses.extend(net.train(m, batch, rate, seed + done).tolist())
        done += m
        if log and (done >= warmup or done == m):
            log(f"[learned] step {done}/{steps} lr {rate:.2e} loss {np.mean(losses[-m:]):.4f}")
    return np.array(losses)


def init_learned(rec, net, level):
    """Sets the level's coefficients from the network (epr_init_learned);
    returns (scale, reflected, energy)."""
    info = np.zeros(3)
    rec._check(load().epr_init_learned(rec._h, net._h, level, _p(info)))
    return info


def solve_learned(rec, net):
    """The reference's cascade started at the network's level from its prediction."""
    E = np.zeros(2)
    rec._check(load().epr_solve_learned(rec._h, net._h, _p(E)))
    return E


def reference_iters(rec):
    """The reference's iteration count of every level, ceil(iMult^(D-d) gsIters)."""
    p = rec.params
    return [int(np.ceil(p.iMult ** (rec.depth - d) * p.gsIters)) for d in range(rec.depth + 1)]


def solve_schedule(rec, net, iters):
    """The cascade (epr_solve_schedule) from the random start (net None) or the
    network's start, with iters[d] Gauss-Seidel iterations at level d (all
    depth+1 counts; the levels below the start are ignored).  With
    fp32Sweeps a colour-launch level runs all its iterations but the last in
    FP32.  Returns (E_fit, E_reg) at the finest level."""
    import ctypes
    E = np.zeros(2)
    it = (ctypes.c_int * (rec.depth + 1))(*[int(v) for v in iters])
    rec._check(load().epr_solve_schedule(rec._h, net._h if net is not None else None, it, _p(E)))
    return E


def _first_crossing(rec, net, coarse, max_finest, target, start_level=0):
    """The first finest-iteration count whose energy is <= target: the start
    (the network's at its level, or the random one at start_level), the
    coarse levels with the given counts, then single FP64 finest iterations
    with an energy check after each (None when max_finest do not reach it)."""
    D = rec.depth
    lo = net.level if net is not None else start_level
    if net is None:
        rec.init(lo)
    else:
        init_learned(rec, net, lo)
    for d in range(lo, D):
        if coarse[d]:
            rec.gauss_seidel(d, coarse[d])
        rec.prolong_add(d)
    for k in range(1, max_finest + 1):
        rec.gauss_seidel(D, 1)
        if rec.energies(D).sum() <= target:
            return k
    return None


def time_to_energy(rec, target, net=None, coarse=None, max_finest=40, reps=3):
    """Time to reach the FP64 energy target from the random start at level 0
    (net None) or the network's start at its level.  coarse: the iteration
    counts of the levels below the finest (None: the reference's).  A first
    run checks the energy after every finest iteration (FP64) to find the
    count k that reaches the target; then the whole schedule with k finest
    iterations runs as one epr_solve_schedule call (the precision of the
    reconstruction: with fp32Sweeps all but the last finest iteration are
    FP32), best of reps; if its final energy misses the target, k grows by one.
    Returns (seconds, k, final energy, iteration counts); seconds is None when
    max_finest iterations do not reach the target."""
    coarse = list(reference_iters(rec) if coarse is None else coarse)
    k = _first_crossing(rec, net, coarse, max_finest, target)
    if k is None:
        return None, max_finest, None, coarse
    while k <= max_finest:
        iters = coarse[:rec.depth] + [k]
        best, E = None, None
        for _ in range(reps):
            t0 = time.perf_counter()
            E = float(solve_schedule(rec, net, iters).sum())
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        if E <= target:
            return best, k, E, iters
        k += 1
    return None, max_finest, E, coarse


def search_schedule(rec, net, target, grid=(5, 10, 20, 40), max_finest=40, reps=2):
    """The fastest schedule of the levels between the network's level and the
    finest (every level's count from grid) with which the learned start
    reaches the target, by time_to_energy on each; (net None: the reference's
    schedule with every coarse count scaled by 1, 1/2, 1/4 or 1/8, for the
    plain form).  Returns the best time_to_energy result and all trials."""
    D = rec.depth
    trials = []
    if net is None:
        ref = reference_iters(rec)
        cands = [[max(1, int(np.ceil(v * f))) for v in ref[:D]] + [ref[D]] for f in (1, 0.5, 0.25, 0.125)]
    else:
        import itertools
        L = net.level
        cands = []
        for combo in itertools.product(grid, repeat=D - L):
            c = [0] * L + list(combo) + [0]
            cands.append(c)
    for c in cands:
        r = time_to_energy(rec, target, net, c, max_finest, reps)
        trials.append(r)
    ok = [r for r in trials if r[0] is not None]
    best = min(ok, key=lambda r: r[0]) if ok else None
    return best, trials


def main(argv=None):
    """python -m exteriorpoissonrecon_amd.learned --pairs P --steps S --out weights.npz
    Generates P solver pairs (4D depth-5 reconstructions solved to level 4),
    trains the network on the device, reports the validation loss against a
    held-out set and the zero predictor, and saves the weights."""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2048)
    ap.add_argument("--val", type=int, default=64)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--depth", type=int, default=5)
    ap.add_argument("--level", type=int, default=4)
    ap.add_argument("--samples", type=int, default=32768)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=WEIGHTS_4D)
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    X, T = training_pairs(a.pairs + a.val, dim=4, level=a.level, depth=a.depth, samples_per_shape=a.samples,
                          seed=a.seed, log=log)
    log(f"[learned] {a.pairs + a.val} pairs in {time.perf_counter() - t0:.1f} s; feature max {np.abs(X).max():.2f}")
    Xv, Tv = X[a.pairs:], T[a.pairs:]
    X, T = X[:a.pairs], T[:a.pairs]
    net = UNet(4, a.level)
    net.set_params(init_params(4, net.in_channels, seed=a.seed))

    def val_loss():
        return float(np.mean([net.loss_gradients(Xv[i:i + 8], Tv[i:i + 8])[0] for i in range(0, len(Xv), 8)]))
    log(f"[learned] validation loss at init {val_loss():.4f} (zero predictor {2.0:.4f})")
    t1 = time.perf_counter()
    losses = train(net, X, T, a.steps, a.batch, a.lr, a.seed, log=log)
    log(f"[learned] trained {a.steps} steps in {time.perf_counter() - t1:.1f} s; validation loss {val_loss():.4f}")
    net.save(a.out)
    log(f"[learned] saved {a.out}")
    return losses


if __name__ == "__main__":
    main()
