"""CPU restatement of the learned far field network (the library's
cosmic-collisions_amd/csrc/cc_ff.hip) -- TEST INFRASTRUCTURE, the checker
only: FP64 arithmetic with the device's BF16 roundings at the same places
(weights, stored activations, the gradient fed to the matrix cores), so that
forward outputs and weight gradients can be compared with the GPU's.

Layout and architecture as the library defines them: weights per layer
[27 taps][cin][cout] then [cout] biases, layers in order; activations
channels-last [instance][z][y][x][channel]; taps t = (dz+1)*9 + (dy+1)*3 + (dx+1).
The model has no reference counterpart: it stands in for the far field of
src/glsl-sim.js:263-295.
"""
import numpy as np

LAYERS = [(16, 32, 1)] + [(32, 32, 1)] * 10 + [(32, 16, 0)]
LAYER_S = [16, 16, 8, 8, 4, 4, 2, 2, 4, 8, 16, 16]
SLOPE = 0.25   # the library's default negative slope


def weight_count():
    return sum(27 * ci * co + co for ci, co, _ in LAYERS)


def split(w):
    """The flat master array -> [(W[27][cin][cout], b[cout])] per layer."""
    out, off = [], 0
    for ci, co, _ in LAYERS:
        W = w[off:off + 27 * ci * co].reshape(27, ci, co)
        off += 27 * ci * co
        b = w[off:off + co]
        off += co
        out.append((W, b))
    assert off == w.size
    return out


def bf16(x):
    """Round to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def _cols(x):
    """im2col of [B,S,S,S,C] with zero padding: [B*S^3, 27*C] (tap-major)."""
    B, S, _, _, C = x.shape
    xp = np.zeros((B, S + 2, S + 2, S + 2, C))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    cols = np.empty((B, S, S, S, 27, C))
    for t in range(27):
        dz, dy, dx = t // 9, (t // 3) % 3, t % 3
        cols[..., t, :] = xp[:, dz:dz + S, dy:dy + S, dx:dx + S]
    return cols.reshape(B * S ** 3, 27 * C)


def _col2im(g, B, S, C):
    """Adjoint of _cols: [B*S^3, 27*C] -> [B,S,S,S,C]."""
    g = g.reshape(B, S, S, S, 27, C)
    xp = np.zeros((B, S + 2, S + 2, S + 2, C))
    for t in range(27):
        dz, dy, dx = t // 9, (t // 3) % 3, t % 3
        xp[:, dz:dz + S, dy:dy + S, dx:dx + S] += g[..., t, :]
    return xp[:, 1:-1, 1:-1, 1:-1]


def conv(x, W, b):
    """x [B,S,S,S,cin] -> pre-activation [B,S,S,S,cout] (FP64)."""
    B, S = x.shape[0], x.shape[1]
    y = _cols(x) @ W.reshape(-1, W.shape[2]) + b
    return y.reshape(B, S, S, S, -1)


def leaky(p):
    return np.where(p > 0, p, SLOPE * p)


def pool(x):
    B, S, _, _, C = x.shape
    h = S // 2
    return x.reshape(B, h, 2, h, 2, h, 2, C).mean(axis=(2, 4, 6))


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def _ident(x):
    return np.asarray(x, np.float64)


def forward(w, x0, rounding=True):
    """x0 [B,16,16,16,16] -> (y [B,16,16,16,16], cache).  With rounding, BF16
    where the device stores BF16 (weights, every stored activation) and FP32
    pre-activations; without, plain FP64 (for finite-difference checks)."""
    bf16_, f32_ = (bf16, f32) if rounding else (_ident, _ident)
    P = [(bf16_(W), f32_(b)) for W, b in split(w)]
    c = {"x0": bf16_(x0), "pre": [None] * 12, "inp": [None] * 12, "rounding": rounding}

    def C(l, x):
        c["inp"][l] = x
        p = f32_(conv(x, *P[l]))
        c["pre"][l] = p
        return bf16_(leaky(p)) if LAYERS[l][2] else p
    a0 = C(0, c["x0"])
    e0 = C(1, a0)
    p1 = bf16_(pool(e0))
    e1 = C(3, C(2, p1))
    p2 = bf16_(pool(e1))
    e2 = C(5, C(4, p2))
    p3 = bf16_(pool(e2))
    bot = C(7, C(6, p3))
    d2 = C(8, bf16_(up(bot) + e2))
    d1 = C(9, bf16_(up(d2) + e1))
    d0 = C(10, bf16_(up(d1) + e0))
    y = C(11, d0)
    c["P"] = P
    return y, c


def loss_grad(y, tgt, mask, rounding=True):
    """Masked mean squared error over the ten used channels, per voxel mean."""
    B = y.shape[0]
    m = (mask.reshape(B, 16, 16, 16, 1) > 0) & (np.arange(16) < 10)
    d = np.where(m, y - tgt, 0.0)
    n = B * 4096.0
    return float((d * d).sum() / n), (f32(2.0 * d / n) if rounding else 2.0 * d / n)


def backward(c, gy):
    """Weight gradients (flat, FP64) from the gradient at the output."""
    P = c["P"]
    grads = [None] * 12
    bf16_, f32_ = (bf16, f32) if c["rounding"] else (_ident, _ident)

    def G(l, gout):
        """gradient at layer l's output -> its input; records dW, db."""
        g = gout * np.where(c["pre"][l] > 0, 1.0, SLOPE) if LAYERS[l][2] else gout
        g = bf16_(g)
        B, S = g.shape[0], g.shape[1]
        gf = g.reshape(B * S ** 3, -1)
        grads[l] = ((_cols(c["inp"][l]).T @ gf).reshape(P[l][0].shape), gf.sum(0))
        return f32_(_col2im(gf @ P[l][0].reshape(-1, P[l][0].shape[2]).T, B, S, P[l][0].shape[1]))

    def upb(g):
        B, S, _, _, C = g.shape
        h = S // 2
        return g.reshape(B, h, 2, h, 2, h, 2, C).sum(axis=(2, 4, 6))

    def poolb(g):
        return up(g) / 8.0
    gd0 = G(11, gy)
    gu0 = G(10, gd0)
    ge0 = gu0.copy()
    gu1 = G(9, upb(gu0))
    ge1 = gu1.copy()
    gu2 = G(8, upb(gu1))
    ge2 = gu2.copy()
    gp3 = G(6, G(7, upb(gu2)))
    ge2 = ge2 + poolb(gp3)
    gp2 = G(4, G(5, ge2))
    ge1 = ge1 + poolb(gp2)
    gp1 = G(2, G(3, ge1))
    ge0 = ge0 + poolb(gp1)
    G(0, G(1, ge0))
    return np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
