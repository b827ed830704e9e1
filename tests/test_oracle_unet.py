"""Pins of the learned initializer's CPU oracle (oracle/pyunet.py), on the CPU.

The oracle is the checker of the device network (tests/test_gpu_unet.py), so
it is pinned here by what the mathematics fixes, not by retyping its formulas:
  * the backward pass against central finite differences of the FP64 loss,
    along random directions and per parameter block (3D and 4D);
  * the convolution and its input transpose against brute-force loops over
    nodes, taps and channels (zero outside the grid);
  * up/down against explicit hat-function prolongation matrices (Kronecker
    products of the 1D matrix), and up_t as up's transpose;
  * the loss: invariant under rotations of the prediction's (x, y) plane,
    equal to a brute-force minimum over rotations, zero for a rotated target.
"""
import itertools
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
from oracle import pyunet  # noqa: E402


def _problem(dim, level, B=2, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    in_c = dim * (dim - 1) // 2
    N = (2 ** level + 1) ** dim
    p = pyunet.init_params(dim, in_c, seed=seed).astype(np.float64)
    # non-zero biases so their gradients are exercised
    k = 0
    for _, ci, co, _ in pyunet.arch(in_c):
        k += 3 ** dim * co * ci
        p[k:k + co] = rng.uniform(-0.1, 0.1, co)
        k += co
    x = rng.normal(size=(B, N, in_c))
    t = rng.normal(size=(B, N, 2))
    return p, x, t


@pytest.mark.parametrize("dim,level", [(3, 3), (4, 3)])
def test_backward_matches_finite_differences(dim, level):
    p, x, t = _problem(dim, level)
    L0, g, _ = pyunet.loss_and_grad(p, x, t, dim, level)
    rng = np.random.Generator(np.random.PCG64(7))
    # a small step: ReLU kinks within h |v| of a pre-activation bias a coarser difference
    h = 1e-8
    f = lambda q: pyunet.loss_and_grad(q, x, t, dim, level)[0]
    for _ in range(3):
        v = rng.normal(size=p.shape)
        fd = (f(p + h * v) - f(p - h * v)) / (2 * h)
        assert abs(fd - g @ v) <= 1e-5 * max(1.0, abs(fd)), (fd, g @ v)
    if dim == 4:
        return   # (the per-block check below runs in 3D: same code paths, 1/9 of the nodes)
    # every layer's weights and biases on their own
    k = 0
    for li, (_, ci, co, _) in enumerate(pyunet.arch(x.shape[-1])):
        for n in (3 ** dim * co * ci, co):
            v = np.zeros_like(p)
            v[k:k + n] = rng.normal(size=n)
            fd = (f(p + h * v) - f(p - h * v)) / (2 * h)
            assert abs(fd - g @ v) <= 1e-5 * max(1.0, abs(fd)) + 1e-7, (li, n, fd, g @ v)
            k += n


@pytest.mark.parametrize("dim", [2, 3])
def test_conv_matches_brute_force(dim):
    rng = np.random.Generator(np.random.PCG64(3))
    side, ci, co = 4, 3, 2
    x = rng.normal(size=(1,) + (side,) * dim + (ci,))
    W = rng.normal(size=(3 ** dim, co, ci))
    b = rng.normal(size=co)
    out = pyunet.conv(x, W, b, dim)
    offs = pyunet.taps(dim)
    ref = np.zeros_like(out)
    for n in itertools.product(range(side), repeat=dim):
        for t, off in enumerate(offs):
            m = tuple(np.array(n) + off)
            if all(0 <= q < side for q in m):
                ref[(0,) + n] += W[t] @ x[(0,) + m]
        ref[(0,) + n] += b
    assert np.abs(out - ref).max() < 1e-12
    # the input transpose: <conv(x), d> = <x, conv_input_grad(d)> (bias-free)
    d = rng.normal(size=out.shape)
    lhs = (pyunet.conv(x, W, 0 * b, dim) * d).sum()
    rhs = (x * pyunet.conv_input_grad(d, W, dim)).sum()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # tap order: tap t's offset digits are base 3 with coordinate 0 most significant
    assert list(offs[1]) == [-1] * (dim - 1) + [0] and list(offs[3 ** (dim - 1)]) == [0] + [-1] * (dim - 1)


def _p1(sc):
    sf = 2 * sc - 1
    P = np.zeros((sf, sc))
    for i in range(sc):
        P[2 * i, i] = 1
        if i + 1 < sc:
            P[2 * i + 1, i] = P[2 * i + 1, i + 1] = 0.5
    return P


@pytest.mark.parametrize("dim", [3, 4])
def test_up_down_match_hat_matrices(dim):
    rng = np.random.Generator(np.random.PCG64(5))
    sc, C = 3, 2
    P = _p1(sc)
    for _ in range(dim - 1):
        P = np.kron(P, _p1(sc))
    x = rng.normal(size=(1,) + (sc,) * dim + (C,))
    u = pyunet.up(x, dim)
    assert np.abs(u.reshape(-1, C) - P @ x.reshape(-1, C)).max() < 1e-13
    g = rng.normal(size=u.shape)
    assert np.abs(pyunet.up_t(g, dim).reshape(-1, C) - P.T @ g.reshape(-1, C)).max() < 1e-13
    # down: P^T normalised by its row sums (an average); down_t its transpose
    w = P.sum(0)
    dn = pyunet.down(g, dim).reshape(-1, C)
    assert np.abs(dn - (P.T @ g.reshape(-1, C)) / w[:, None]).max() < 1e-13
    ones = np.ones((1,) + (2 * sc - 1,) * dim + (1,))
    assert np.abs(pyunet.down(ones, dim) - 1).max() < 1e-14
    y = rng.normal(size=dn.shape)
    lhs = (dn * y).sum()
    rhs = (g.reshape(-1, C) * pyunet.down_t(y.reshape((1,) + (sc,) * dim + (C,)), dim).reshape(-1, C)).sum()
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_loss_is_the_distance_after_the_best_rotation():
    rng = np.random.Generator(np.random.PCG64(9))
    y = rng.normal(size=(2, 50, 2))
    t = rng.normal(size=(2, 50, 2))
    L, _ = pyunet.loss(y, t)
    th = np.linspace(0, 2 * np.pi, 20001)
    brute = 0
    for b in range(2):
        c, s = np.cos(th)[:, None], np.sin(th)[:, None]
        rx = c * y[b, :, 0] - s * y[b, :, 1]
        ry = s * y[b, :, 0] + c * y[b, :, 1]
        brute += (((rx - t[b, :, 0]) ** 2 + (ry - t[b, :, 1]) ** 2).sum(1) / 50).min()
    assert abs(L - brute / 2) < 1e-6
    a = 1.234
    R = np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
    assert abs(pyunet.loss(y @ R.T, t)[0] - L) < 1e-12
    assert pyunet.loss(t @ R.T, t)[0] < 1e-12


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -3.14159, 65504.0, 1e-30])
    r = pyunet.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0                  # a tie rounds to even
    assert r[2] == 1.0 + 2 ** -7 and r[3] == 1.0 + 2 ** -6
    assert abs(r[4] + 3.140625) < 1e-12
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)
