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


CHAN_W = np.array([1.0] * 3 + [0.25] * 7 + [0.0] * 6)   # the loss's channel weights (cc_ff.hip: ff_chan_weight)


def loss_grad(y, tgt, mask, rounding=True):
    """Weighted mean squared error (each voxel by its mask, each channel by
    CHAN_W), per voxel mean, and its gradient."""
    B = y.shape[0]
    w = np.asarray(mask, np.float64).reshape(B, 16, 16, 16, 1) * CHAN_W
    d = np.where(w > 0, y - tgt, 0.0)
    n = B * 4096.0
    g = 2.0 * w * d / n
    return float((w * d * d).sum() / n), (f32(g) if rounding else g)


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


# ---------------------------------------------------------------- the mesh interface
# Restatements of the library's interface between the mesh and the network
# (cc_ff.hip: k_ff_feat, k_ff_base, k_ff_pair_in, k_ff_pair_tgt, k_ff_out,
# k_ff_judge_sum), in FP32 where the device computes in FP32 and in the same
# operation order, so that features, records and the carried far field can be
# compared bit for bit.  Blocks by voxel v = (z*16 + y)*16 + x.
FH = 2.0 * 6.0 / 64.0
BLOCK_LEN = np.float32(4.0 * FH)
FF_R = 12
_V = np.arange(4096)
VX, VY, VZ = _V & 15, (_V >> 4) & 15, _V >> 8


def coarse_texel(x, y, z):
    return ((z >> 2) * 16 + y) * 64 + (z & 3) * 16 + x


def _by_voxel(mesh64):
    """A 64x64 coarse atlas (float4 texels) -> [4096, 4] by voxel."""
    return np.asarray(mesh64, np.float32).reshape(-1, 4)[coarse_texel(VX, VY, VZ)]


def _centres(L):
    """Each block's centre, FP32 as the device forms it: (4k + 2) FH - L."""
    fh = np.float32(FH)
    return np.stack([np.float32(4 * c + 2) * fh - np.float32(L[k]) for k, c in enumerate((VX, VY, VZ))], 1)


def _offsets(c, L):
    """Centre of mass less centre (occupied blocks), FP32."""
    with np.errstate(invalid="ignore", divide="ignore"):
        return c[:, :3] / c[:, 3:4] - _centres(L)


def features(coarse, mono_w, L):
    """k_ff_feat: ([4096, 4] features, meta[4])."""
    c = _by_voxel(coarse)
    mt = np.float32(max(np.float32(mono_w), np.float32(1e-20)))
    f = np.zeros((4096, 4), np.float32)
    occ = c[:, 3] > 0
    d = _offsets(c, L)
    f[occ, 0] = c[occ, 3] / mt * np.float32(64.0)
    for k in range(3):
        f[occ, 1 + k] = f[occ, 0] * d[occ, k] / BLOCK_LEN
    meta = np.array([L[0], L[1], L[2], mt], np.float32)
    return f, meta


def _centre_occupied(c, L, A, T0, T1):
    """An occupied block's record: the far field carried to its centre to first order."""
    d = _offsets(c, L)
    dx, dy, dz = d[:, 0], d[:, 1], d[:, 2]
    ax = A[:, 0] - (T0[:, 0] * dx + T0[:, 3] * dy + T1[:, 0] * dz)
    ay = A[:, 1] - (T0[:, 3] * dx + T0[:, 1] * dy + T1[:, 1] * dz)
    az = A[:, 2] - (T1[:, 0] * dx + T1[:, 1] * dy + T0[:, 2] * dz)
    p = A[:, 3] + (ax * dx + ay * dy + az * dz)
    z = np.zeros_like(ax)
    return np.stack([ax, ay, az, p, T0[:, 0], T0[:, 1], T0[:, 2], T0[:, 3], T1[:, 0], T1[:, 1], z, z], 1)


def texel_counts(fine):
    """Occupied fine texels of each block, by voxel (a 513x512 fine atlas)."""
    F = np.asarray(fine, np.float32).reshape(-1, 4)[:512 * 512]
    e = np.arange(512 * 512)
    mx, my = e & 511, e >> 9
    c0, c1, c2 = mx & 63, my & 63, (my >> 6) * 8 + (mx >> 6)
    vox = ((c2 >> 2) * 16 + (c1 >> 2)) * 16 + (c0 >> 2)
    return np.bincount(vox[F[:, 3] > 0], minlength=4096)


def centre_record(fine, coarse, L, farA, farT0, farT1):
    """k_ff_base: [4096, 12] record, FP32 as the device computes it: occupied
    blocks' far field carried to their centres, their occupied fine texels in
    channel 10; empty blocks zero."""
    c = _by_voxel(coarse)
    A, T0, T1 = _by_voxel(farA), _by_voxel(farT0), _by_voxel(farT1)
    occ = c[:, 3] > 0
    R = np.zeros((4096, FF_R), np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        R[occ] = _centre_occupied(c, L, A, T0, T1)[occ]
    R[occ, 10] = texel_counts(fine)[occ]
    return R


def pair_input(fT, fB):
    """k_ff_pair_in: (x0 [4096, 16] before the BF16 rounding, s); s in FP64."""
    d = fT.astype(np.float64) - fB
    use = (fT[:, 0] > 0) | (fB[:, 0] > 0)
    ss = (d[use] ** 2).sum()
    s = np.sqrt(ss / (4.0 * use.sum())) if ss > 0 else 1.0
    x = np.zeros((4096, 16))
    x[:, 0:4] = fT
    x[:, 4:8] = d / s
    x[:, 8] = fT[:, 0] > 0
    x[:, 9] = fB[:, 0] > 0
    return x, s


def units(s, mt):
    """Per-channel units of the output: s * mt / 64 over BLOCK_LEN^2, ^1, ^3."""
    u = np.float64(s) * np.float64(mt) / 64.0
    bl = np.float64(BLOCK_LEN)
    return np.array([u / bl ** 2] * 3 + [u / bl] + [u / bl ** 3] * 6)


def pair_target(RT, RB, fT, fB, mt, s):
    """k_ff_pair_tgt: (target [4096, 16], mask [4096]): the change on blocks
    occupied at T and B, weighted by their occupied fine texels at T / 64."""
    occ = (fT[:, 0] > 0) & (fB[:, 0] > 0)
    t = np.zeros((4096, 16))
    t[occ, :10] = (RT[occ, :10].astype(np.float64) - RB[occ, :10]) / units(s, mt)
    m = np.zeros(4096, np.float32)
    m[occ] = RT[occ, 10] / np.float32(64.0)
    return t, m


def carry(y, RB, s, mt, coarse, L):
    """k_ff_out: the network's far field (farA, farT0, farT1, each [4096, 4] by
    voxel) from outputs y [4096, 16], FP32 as the device computes it (blocks
    empty at the base, which k_ff_out leaves to k_far, included)."""
    c = _by_voxel(coarse)
    occ = c[:, 3] > 0
    y = np.asarray(y, np.float32)
    r = np.asarray(RB, np.float32)
    u = np.float32(s) * np.float32(mt) / np.float32(64.0)
    bl = BLOCK_LEN
    uA, uP, uT = u * (np.float32(1.0) / (bl * bl)), u * (np.float32(1.0) / bl), u * (np.float32(1.0) / (bl * bl * bl))
    a = [r[:, k] + y[:, k] * uA for k in range(3)]
    p = r[:, 3] + y[:, 3] * uP
    t = [r[:, 4 + k] + y[:, 4 + k] * uT for k in range(6)]
    txx, tyy, tzz, txy, txz, tyz = t
    d = _offsets(c, L)
    dx, dy, dz = d[:, 0], d[:, 1], d[:, 2]
    sx = a[0] + txx * dx + txy * dy + txz * dz
    sy = a[1] + txy * dx + tyy * dy + tyz * dz
    sz = a[2] + txz * dx + tyz * dy + tzz * dz
    A = np.stack([sx, sy, sz, p - (a[0] * dx + a[1] * dy + a[2] * dz)], 1)
    T0 = np.stack([txx, tyy, tzz, txy], 1)
    T1 = np.stack([txz, tyz, np.zeros_like(txz), np.zeros_like(txz)], 1)
    for X in (A, T0, T1):
        X[~occ] = 0
    return A.astype(np.float32), T0.astype(np.float32), T1.astype(np.float32)


def judge_error(fine, coarse, learned, exact, force, gm, fB):
    """k_ff_judge_sum + k_judge_decide: the relative error, over the force mesh
    (its 512 x 512 texels), of the far-field part the learned far field gives
    each occupied texel of a block occupied at the base (fB: the base's
    features).  learned, exact: (farA, farT0, farT1) by voxel."""
    F = np.asarray(fine, np.float64).reshape(-1, 4)[:512 * 512]
    e = np.arange(512 * 512)
    mx, my = e & 511, e >> 9
    c0, c1, c2 = mx & 63, my & 63, (my >> 6) * 8 + (mx >> 6)
    vox = ((c2 >> 2) * 16 + (c1 >> 2)) * 16 + (c0 >> 2)
    c = _by_voxel(coarse).astype(np.float64)
    occ = (F[:, 3] > 0) & (np.asarray(fB)[vox, 0] > 0)
    com = np.zeros((len(F), 3))
    com[occ] = F[occ, :3] / F[occ, 3:4]
    with np.errstate(invalid="ignore", divide="ignore"):
        bcom = c[:, :3] / c[:, 3:4]
    q = com - bcom[vox]
    dA = (np.asarray(learned[0], np.float64) - exact[0])[vox]
    d0 = (np.asarray(learned[1], np.float64) - exact[1])[vox]
    d1 = (np.asarray(learned[2], np.float64) - exact[2])[vox]
    dx = dA[:, 0] + d0[:, 0] * q[:, 0] + d0[:, 3] * q[:, 1] + d1[:, 0] * q[:, 2]
    dy = dA[:, 1] + d0[:, 3] * q[:, 0] + d0[:, 1] * q[:, 1] + d1[:, 1] * q[:, 2]
    dz = dA[:, 2] + d1[:, 0] * q[:, 0] + d1[:, 1] * q[:, 1] + d0[:, 2] * q[:, 2]
    num = ((dx * dx + dy * dy + dz * dz)[occ]).sum() * gm * gm
    Fm = np.asarray(force, np.float64).reshape(-1, 4)[:512 * 512, :3]
    return float(np.sqrt(num / (Fm * Fm).sum()))
