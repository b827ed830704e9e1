"""The learned initializer's network on the device, through the C ABI.

Bars:
* forward output and every parameter gradient against the CPU oracle
  (oracle/pyunet.py, pinned by tests/test_oracle_unet.py) run with the same
  bfloat16 roundings as the device: relative L2 error <= 1e-2 overall, and per
  layer <= 1e-1 with cosine >= 0.995.  The device sums in the matrix cores'
  FP32 accumulators, the oracle in FP64; a sum landing on the other side of a
  bfloat16 rounding boundary changes one stored value by one bf16 step (2^-8)
  and the coarse grids spread such a step over every node.  Measured: output
  1e-4 (3D) to 3e-3 (4D), error smooth in space and largest in the interior
  (not at the grid's faces, as an index or boundary mistake would make it);
  4D per-layer gradients up to 5e-2 on the coarsest grid (16 nodes, 10 368-term
  sums);
* against the exact FP64 network (no rounding): relative L2 error <= 6e-2 and
  cosine >= 0.998 -- what BF16 storage costs;
* training on the device (Adam) lowers the loss on pairs the solver made;
* the loss-gradient entry point agrees with the loss it reports (a step
  along -grad lowers it).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel(); b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _cos(a, b):
    a = np.asarray(a, np.float64).ravel(); b = np.asarray(b, np.float64).ravel()
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


def _case(dim, level, B, seed):
    from oracle import pyunet
    rng = np.random.Generator(np.random.PCG64(seed))
    in_c = dim * (dim - 1) // 2
    N = (2 ** level + 1) ** dim
    p = pyunet.init_params(dim, in_c, seed=seed)
    k = 0
    for _, ci, co, _ in pyunet.arch(in_c):
        k += 3 ** dim * co * ci
        p[k:k + co] = rng.uniform(-0.05, 0.05, co).astype(np.float32)
        k += co
    x = rng.normal(size=(B, N, in_c)).astype(np.float32)
    t = rng.normal(size=(B, N, 2)).astype(np.float32)
    return p, x, t


@pytest.mark.parametrize("dim,level,B", [(3, 3, 2), (4, 3, 2), (3, 5, 1), (4, 4, 1)])
def test_forward_and_gradients_match_oracle(dim, level, B):
    from exteriorpoissonrecon_amd.learned import UNet
    from oracle import pyunet
    print(f"[gpu] unet dim {dim} level {level} batch {B}", flush=True)
    p, x, t = _case(dim, level, B, seed=dim * 10 + level)
    net = UNet(dim, level)
    assert net.num_params == pyunet.num_params(dim, net.in_channels)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    y = net.forward(x)
    yb, _ = pyunet.forward(p, x, dim, level, bf=True)
    y64, _ = pyunet.forward(p, x, dim, level, bf=False)
    print(f"[gpu]   forward rel (bf16 oracle) {_rel(y, yb):.2e} (fp64) {_rel(y, y64):.2e}", flush=True)
    assert _rel(y, yb) <= 1e-2
    assert _rel(y, y64) <= 6e-2 and _cos(y, y64) >= 0.998
    L, g = net.loss_gradients(x, t)
    Lb, gb, _ = pyunet.loss_and_grad(p, x, t, dim, level, bf=True)
    L64, g64, _ = pyunet.loss_and_grad(p, x, t, dim, level, bf=False)
    print(f"[gpu]   loss {L:.6f} oracle {Lb:.6f} / {L64:.6f}; grad rel (bf16 oracle) {_rel(g, gb):.2e} "
          f"(fp64) {_rel(g, g64):.2e} cos {_cos(g, g64):.5f}", flush=True)
    assert abs(L - Lb) <= 1e-3 * abs(Lb)
    assert _rel(g, gb) <= 1e-2
    assert _rel(g, g64) <= 6e-2 and _cos(g, g64) >= 0.998
    # per layer too: no block may hide behind the others' norm
    k = 0
    for li, (_, ci, co, _) in enumerate(pyunet.arch(net.in_channels)):
        n = 3 ** dim * co * ci + co
        r, c = _rel(g[k:k + n], gb[k:k + n]), _cos(g[k:k + n], gb[k:k + n])
        print(f"[gpu]   layer {li}: rel {r:.2e} cos {c:.5f}", flush=True)
        assert r <= 1e-1 and c >= 0.995, (li, r, c)
        k += n
    # a small step against the gradient lowers the loss
    net.set_params(p - 1e-3 * g / np.abs(g).max())
    L2, _ = net.loss_gradients(x, t)
    assert L2 < L
    net.close()


def test_device_training_lowers_the_loss_on_solver_pairs():
    from exteriorpoissonrecon_amd import learned
    print("[gpu] 24 training pairs (4D depth 5 -> level 4)", flush=True)
    X, T = learned.training_pairs(24, dim=4, level=4, depth=5, samples_per_shape=16384, seed=11, group=12)
    assert np.isfinite(X).all() and np.isfinite(T).all()
    assert np.allclose(np.sqrt((T.astype(np.float64) ** 2).mean(axis=(1, 2))), 1.0, atol=1e-4)
    net = learned.UNet(4, 4)
    net.set_params(learned.init_params(4, 6, seed=3))
    losses = learned.train(net, X, T, steps=60, batch=4, lr=2e-3, seed=5)
    print(f"[gpu]   loss {losses[:5].mean():.4f} -> {losses[-5:].mean():.4f}", flush=True)
    assert np.isfinite(losses).all()
    assert losses[-10:].mean() < 0.8 * losses[:5].mean()
    net.close()
