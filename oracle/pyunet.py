"""CPU FP64 oracle of the learned initializer's network (test infrastructure).

A numpy restatement of the U-Net of exteriorpoissonrecon_amd/csrc/epr_unet.hip:
the same architecture, tap order, down/up operators, loss and canonical
parameter order, with hand-written backward passes.  `bf16=True` rounds to
bfloat16 (round to nearest even) exactly where the device stores bf16 --
inputs, weights, every stored activation and activation gradient -- and keeps
everything else in FP64, so the device's results differ from it only by its
FP32 accumulation order; `bf16=False` is the exact FP64 network.

Pinned by tests/test_oracle_unet.py: the backward pass against central finite
differences of the FP64 loss, the convolution against a brute-force loop, the
down/up operators against explicit hat-function matrices, and the loss
against its rotation invariance and a brute-force minimum over rotations.
"""
import numpy as np

CT = 16


def pad(c):
    return (c + CT - 1) // CT * CT


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(x))


def arch(in_c):
    """(level, cin, cout, relu) of the 15 convolutions, in the canonical order."""
    c = [16, 32, 64, 128]
    return [(0, in_c, c[0], 1), (0, c[0], c[0], 1), (1, c[0], c[1], 1), (1, c[1], c[1], 1),
            (2, c[1], c[2], 1), (2, c[2], c[2], 1), (3, c[2], c[3], 1), (3, c[3], c[3], 1),
            (2, c[3], c[2], 1), (2, c[2], c[2], 1), (1, c[2], c[1], 1), (1, c[1], c[1], 1),
            (0, c[1], c[0], 1), (0, c[0], c[0], 1), (0, c[0], 2, 0)]


def taps(dim):
    """Offsets of the 3^dim taps: base-3 digits, first coordinate most significant, minus 1."""
    t = np.arange(3 ** dim)
    return np.stack([(t // 3 ** (dim - 1 - d)) % 3 - 1 for d in range(dim)], -1)


def num_params(dim, in_c):
    return sum(3 ** dim * co * ci + co for _, ci, co, _ in arch(in_c))


def unpack(p, dim, in_c):
    out, k = [], 0
    T = 3 ** dim
    for _, ci, co, _ in arch(in_c):
        W = p[k:k + T * co * ci].reshape(T, co, ci); k += T * co * ci
        b = p[k:k + co]; k += co
        out.append((W, b))
    assert k == len(p)
    return out


def pack(grads):
    return np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])


def init_params(dim, in_c, seed=0):
    """He-uniform weights, zero biases (the initialization the training script uses)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    T = 3 ** dim
    parts = []
    for _, ci, co, relu in arch(in_c):
        lim = np.sqrt(6.0 / (T * ci)) * (1.0 if relu else 0.5)
        parts.append(rng.uniform(-lim, lim, T * co * ci))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


# ---- grid operators on arrays of shape (B, side, ..., side, C) ----------------
def shift(x, off):
    """y[n] = x[n + off], zero outside the grid (spatial axes 1..dim)."""
    y = np.zeros_like(x)
    src, dst = [slice(None)], [slice(None)]
    for o in off:
        o = int(o)
        if o >= 0:
            src.append(slice(o, None)); dst.append(slice(0, x.shape[len(dst)] - o))
        else:
            src.append(slice(0, o)); dst.append(slice(-o, None))
    src.append(slice(None)); dst.append(slice(None))
    y[tuple(dst)] = x[tuple(src)]
    return y


def conv(x, W, b, dim):
    """out[n][co] = b[co] + sum_t W[t][co][ci] x[n + off(t)][ci]."""
    out = np.zeros(x.shape[:-1] + (W.shape[1],))
    for t, off in enumerate(taps(dim)):
        out += shift(x, off) @ W[t].T
    return out + b


def conv_input_grad(d, W, dim):
    """Transpose of conv in its input: dx[m][ci] = sum_t W[t][co][ci] d[m - off(t)][co]."""
    out = np.zeros(d.shape[:-1] + (W.shape[2],))
    for t, off in enumerate(taps(dim)):
        out += shift(d, -off) @ W[t]
    return out


def conv_weight_grad(x, d, dim):
    """dW[t][co][ci] = sum_n d[n][co] x[n + off(t)][ci]; db[co] = sum_n d[n][co]."""
    C = d.shape[-1]
    dW = np.stack([d.reshape(-1, C).T @ shift(x, off).reshape(-1, x.shape[-1]) for off in taps(dim)])
    return dW, d.reshape(-1, C).sum(0)


def up(x, dim):
    """Hat prolongation: per axis new[2i] = old[i], new[2i+1] = (old[i] + old[i+1]) / 2."""
    for ax in range(1, dim + 1):
        n = x.shape[ax]
        shp = list(x.shape); shp[ax] = 2 * n - 1
        y = np.zeros(shp)
        ev = [slice(None)] * x.ndim; ev[ax] = slice(0, None, 2)
        od = [slice(None)] * x.ndim; od[ax] = slice(1, None, 2)
        lo = [slice(None)] * x.ndim; lo[ax] = slice(0, n - 1)
        hi = [slice(None)] * x.ndim; hi[ax] = slice(1, n)
        y[tuple(ev)] = x
        y[tuple(od)] = 0.5 * (x[tuple(lo)] + x[tuple(hi)])
        x = y
    return x


def up_t(g, dim):
    """Transpose of up: per axis old[i] = new[2i] + (new[2i-1] + new[2i+1]) / 2."""
    for ax in range(1, dim + 1):
        m = g.shape[ax]
        n = (m + 1) // 2
        shp = list(g.shape); shp[ax] = n
        y = np.zeros(shp)
        ev = [slice(None)] * g.ndim; ev[ax] = slice(0, None, 2)
        y += g[tuple(ev)]
        od = [slice(None)] * g.ndim; od[ax] = slice(1, None, 2)
        o = g[tuple(od)]
        a = [slice(None)] * g.ndim; a[ax] = slice(0, n - 1)
        c = [slice(None)] * g.ndim; c[ax] = slice(1, n)
        y[tuple(a)] += 0.5 * o
        y[tuple(c)] += 0.5 * o
        g = y
    return g


def down_weights(sc, dim):
    """Row sums of up_t (the down operator's normaliser) on the coarse grid of side sc."""
    w1 = np.ones(sc)
    w1[1:] += 0.5
    w1[:-1] += 0.5
    w = w1
    for _ in range(dim - 1):
        w = np.multiply.outer(w, w1)
    return w[None, ..., None]


def down(x, dim):
    sc = (x.shape[1] + 1) // 2
    return up_t(x, dim) / down_weights(sc, dim)


def down_t(g, dim):
    return up(g / down_weights(g.shape[1], dim), dim)


# ---- the network -----------------------------------------------------------------
def forward(params, x, dim, level, bf=False):
    """x: (B, N, in_c) node-major (last coordinate fastest).  Returns (y (B, N, 2), cache)."""
    in_c = x.shape[-1]
    R = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    L = unpack(np.asarray(params, dtype=np.float64), dim, in_c)
    if bf:
        L = [(bf16(W), b.astype(np.float32).astype(np.float64)) for W, b in L]
    B = x.shape[0]
    sides = [2 ** (level - k) + 1 for k in range(4)]
    grid = lambda a, k: a.reshape((B,) + (sides[k],) * dim + (a.shape[-1],))
    c = {"x0": R(grid(np.asarray(x, dtype=np.float64), 0))}

    def cv(k, inp, relu=True):
        v = conv(inp, L[k][0], L[k][1], dim)
        return np.maximum(v, 0) if relu else v

    c["h0"] = R(cv(0, c["x0"])); c["s0"] = R(cv(1, c["h0"]))
    c["in1"] = R(down(c["s0"], dim)); c["h1"] = R(cv(2, c["in1"])); c["s1"] = R(cv(3, c["h1"]))
    c["in2"] = R(down(c["s1"], dim)); c["h2"] = R(cv(4, c["in2"])); c["s2"] = R(cv(5, c["h2"]))
    c["in3"] = R(down(c["s2"], dim)); c["h3"] = R(cv(6, c["in3"])); c["s3"] = R(cv(7, c["h3"]))
    c["u2"] = R(up(c["s3"], dim)); c["g2"] = R(cv(8, c["u2"])); c["a2"] = R(c["g2"] + c["s2"]); c["d2"] = R(cv(9, c["a2"]))
    c["u1"] = R(up(c["d2"], dim)); c["g1"] = R(cv(10, c["u1"])); c["a1"] = R(c["g1"] + c["s1"]); c["d1"] = R(cv(11, c["a1"]))
    c["u0"] = R(up(c["d1"], dim)); c["g0"] = R(cv(12, c["u0"])); c["a0"] = R(c["g0"] + c["s0"]); c["d0"] = R(cv(13, c["a0"]))
    y = cv(14, c["d0"], relu=False)
    if bf:
        y = y.astype(np.float32).astype(np.float64)
    c["L"], c["dim"], c["bf"], c["B"] = L, dim, bf, B
    return y.reshape(B, -1, 2), c


def loss(y, t):
    """Batch mean of (|p|^2 + |t|^2 - 2 sqrt(A^2 + B^2)) / N and its gradient in y."""
    Bn, N = y.shape[0], y.shape[1]
    px, py, tx, ty = y[..., 0], y[..., 1], t[..., 0], t[..., 1]
    pp = (px * px + py * py).sum(1); tt = (tx * tx + ty * ty).sum(1)
    A = (px * tx + py * ty).sum(1); Bv = (px * ty - py * tx).sum(1)
    rho = np.sqrt(A * A + Bv * Bv)
    Lv = ((pp + tt - 2 * rho) / N).mean()
    cs = np.where(rho > 0, A / np.where(rho > 0, rho, 1), 1.0)[:, None]
    sn = np.where(rho > 0, Bv / np.where(rho > 0, rho, 1), 0.0)[:, None]
    f = 2.0 / (N * Bn)
    dy = np.stack([f * (px - cs * tx - sn * ty), f * (py - cs * ty + sn * tx)], -1)
    return Lv, dy


def backward(c, dy):
    """Gradients of the loss in every parameter (canonical flat order)."""
    dim, bf, L = c["dim"], c["bf"], c["L"]
    R = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    grads = [None] * 15
    B = c["B"]
    d = R(dy.reshape(c["d0"].shape[:-1] + (2,)))

    def lb(k, inp, out, dout, relu=True, need_in=True):
        da = dout * (out > 0) if relu else dout
        grads[k] = conv_weight_grad(inp, da, dim)
        return R(conv_input_grad(da, L[k][0], dim)) if need_in else None

    g = lb(14, c["d0"], None, d, relu=False)
    g = lb(13, c["a0"], c["d0"], g)
    gs0 = g.copy()
    g = lb(12, c["u0"], c["g0"], g)
    g = R(up_t(g, dim))
    g = lb(11, c["a1"], c["d1"], g)
    gs1 = g.copy()
    g = lb(10, c["u1"], c["g1"], g)
    g = R(up_t(g, dim))
    g = lb(9, c["a2"], c["d2"], g)
    gs2 = g.copy()
    g = lb(8, c["u2"], c["g2"], g)
    g = R(up_t(g, dim))
    g = lb(7, c["h3"], c["s3"], g)
    g = lb(6, c["in3"], c["h3"], g)
    gs2 = R(gs2 + down_t(g, dim))
    g = lb(5, c["h2"], c["s2"], gs2)
    g = lb(4, c["in2"], c["h2"], g)
    gs1 = R(gs1 + down_t(g, dim))
    g = lb(3, c["h1"], c["s1"], gs1)
    g = lb(2, c["in1"], c["h1"], g)
    gs0 = R(gs0 + down_t(g, dim))
    g = lb(1, c["h0"], c["s0"], gs0)
    lb(0, c["x0"], c["h0"], g, need_in=False)
    del B
    return pack(grads)


def loss_and_grad(params, x, t, dim, level, bf=False):
    y, c = forward(params, x, dim, level, bf)
    Lv, dy = loss(y, t)
    if bf:
        dy = bf16(dy)
    return Lv, backward(c, dy), y
