"""The CPU restatement of the learned far field network (oracle/ff_oracle.py)
pinned by what the mathematics fixes: the convolution against a brute-force
loop, pooling and upsampling against their adjoints, and the whole
backward pass against central finite differences of the loss in FP64; its
interface with the mesh by the round trip of a pair's target through the
output transform back to the exact far field (FAR_FS's restatement), the
centre records against FP64 direct sums at the block centres, and the texel
counts against a direct count (a flipped transport sign, a swapped tensor
component or a wrong unit each fails one of them).  CPU only."""
import numpy as np
import pytest

import golden_util as gu  # noqa: F401  (puts the oracle on the path)
import ff_oracle as fo


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    B, S, ci, co = 2, 3, 2, 3
    x = rng.normal(size=(B, S, S, S, ci))
    W = rng.normal(size=(27, ci, co))
    b = rng.normal(size=co)
    y = fo.conv(x, W, b)
    ref = np.zeros_like(y)
    for n in range(B):
        for z in range(S):
            for yy in range(S):
                for xx in range(S):
                    acc = b.copy()
                    for t in range(27):
                        zz, y2, x2 = z + t // 9 - 1, yy + (t // 3) % 3 - 1, xx + t % 3 - 1
                        if 0 <= zz < S and 0 <= y2 < S and 0 <= x2 < S:
                            acc += x[n, zz, y2, x2] @ W[t]
                    ref[n, z, yy, xx] = acc
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_pool_up_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 4, 4, 4, 3))
    y = rng.normal(size=(2, 2, 2, 2, 3))
    assert np.isclose((fo.pool(x) * y).sum(), (x * fo.up(y) / 8.0).sum())
    assert np.allclose(fo.pool(fo.up(y)), y)


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 0.0], np.float32)
    r = fo.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0            # tie rounds to even
    assert r[2] == 1.0078125                      # above the tie rounds up
    assert abs(r[3] - (-3.140625)) < 1e-12 and r[4] == 0.0


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = rng.normal(scale=0.08, size=fo.weight_count())
    x = np.zeros((1, 16, 16, 16, 16))
    x[..., :4] = rng.normal(size=(1, 16, 16, 16, 4))
    tgt = rng.normal(size=(1, 16, 16, 16, 16))
    mask = (rng.random(4096) < 0.3).astype(np.float64)
    y, c = fo.forward(w, x, rounding=False)
    L, gy = fo.loss_grad(y, tgt, mask, rounding=False)
    g = fo.backward(c, gy)
    # probe weights spread over the layers, and biases
    offs, off = [], 0
    for ci, co, _ in fo.LAYERS:
        offs.append(off + rng.integers(27 * ci * co))
        offs.append(off + 27 * ci * co + rng.integers(co))
        off += 27 * ci * co + co
    for k in offs[::3] + [offs[-1], offs[-2]]:
        e = 1e-5
        wp, wm = w.copy(), w.copy()
        wp[k] += e
        wm[k] -= e
        Lp = fo.loss_grad(fo.forward(wp, x, rounding=False)[0], tgt, mask, rounding=False)[0]
        Lm = fo.loss_grad(fo.forward(wm, x, rounding=False)[0], tgt, mask, rounding=False)[0]
        fd = (Lp - Lm) / (2 * e)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-10), k


# ---- the mesh interface (features, centre records, pairs, the carried far field, the judge)

def _coarse_mesh(rng, occ_frac=0.06, jitter=0.3, L=(6.0, 6.0, 6.0)):
    """A coarse atlas: occupied blocks with masses and centres of mass jittered
    by up to `jitter` blocks from the block centre (xyz mass-weighted)."""
    coarse = np.zeros((64, 64, 4), np.float32)
    flat = coarse.reshape(-1, 4)
    occ = rng.random(4096) < occ_frac
    t = fo.coarse_texel(fo.VX, fo.VY, fo.VZ)
    m = (rng.random(4096) + 0.05).astype(np.float32)
    com = fo._centres(np.array(L, np.float32)).astype(np.float64) + \
        rng.uniform(-jitter, jitter, (4096, 3)) * float(fo.BLOCK_LEN)
    flat[t[occ], :3] = (com[occ] * m[occ, None]).astype(np.float32)
    flat[t[occ], 3] = m[occ]
    return coarse, occ


def test_carried_far_field_round_trip():
    """A pair's target, fed back through the output transform as if the
    network had predicted it exactly, returns the present exact far field (the
    oracle's FAR_FS restatement, pinned against the reference) on every block
    occupied in both passes: the centre record, the units and the carry are
    each other's inverses.  A zero prediction keeps the base's field carried
    to the present centres of mass; the judge of an exact prediction is zero."""
    import oracle
    rng = np.random.default_rng(3)
    L = np.array([6.0, 6.0, 6.0], np.float32)
    cB, occB = _coarse_mesh(rng)
    # the later pass: every mass moved a little, some blocks emptied and some filled
    cT = cB.copy()
    flat = cT.reshape(-1, 4)
    t = fo.coarse_texel(fo.VX, fo.VY, fo.VZ)
    occ = flat[t, 3] > 0
    flat[t[occ], :3] += (rng.normal(scale=0.01, size=(occ.sum(), 3)) * flat[t[occ], 3:4]).astype(np.float32)
    flat[t[occ], 3] *= (1 + rng.normal(scale=0.01, size=occ.sum())).astype(np.float32)
    fine = np.zeros((513, 512, 4), np.float32)
    farB, farT = oracle.far(cB), oracle.far(cT)
    mB, mT = oracle.mono(cB)[3], oracle.mono(cT)[3]
    fB, _ = fo.features(cB, mB, L)
    fT, _ = fo.features(cT, mT, L)
    RB = fo.centre_record(fine, cB, L, *farB)
    RT = fo.centre_record(fine, cT, L, *farT)
    _, s = fo.pair_input(fT, fB)
    tgt, mask = fo.pair_target(RT, RB, fT, fB, mT, s)
    both = (fT[:, 0] > 0) & (fB[:, 0] > 0)
    assert both.sum() > 100 and not mask.any()   # no occupied texels in these meshes: no weight
    got = fo.carry(tgt.astype(np.float32), RB, s, mT, cT, L)
    want = [fo._by_voxel(x) for x in farT]
    for g, w_, what in zip(got, want, ("farA", "farT0", "farT1")):
        scale = np.abs(w_[both]).max()
        err = np.abs(g[both] - w_[both]).max() / scale
        assert err < 1e-5, (what, err)
    # the base alone (a zero prediction) misses by the change, which the round trip recovered
    stale = fo.carry(np.zeros((4096, 16), np.float32), RB, s, mT, cT, L)
    miss = np.abs(stale[0][both, :3] - want[0][both, :3]).max() / np.abs(want[0][both, :3]).max()
    assert miss > 1e-4
    # the judge: an exact prediction scores zero, the base alone its miss
    fineT = np.zeros((513, 512, 4), np.float32)
    e = np.arange(512 * 512)
    mx, my = e & 511, e >> 9
    c0, c1, c2 = mx & 63, my & 63, (my >> 6) * 8 + (mx >> 6)
    vox = ((c2 >> 2) * 16 + (c1 >> 2)) * 16 + (c0 >> 2)
    cfl = fo._by_voxel(cT)
    sel = (cfl[vox, 3] > 0) & (rng.random(len(e)) < 0.2)
    fineT.reshape(-1, 4)[e[sel], :3] = (cfl[vox[sel], :3] / cfl[vox[sel], 3:4]) * 0.5
    fineT.reshape(-1, 4)[e[sel], 3] = 0.5
    force = np.zeros_like(fineT)
    force.reshape(-1, 4)[e[sel], :3] = rng.normal(size=(sel.sum(), 3))
    assert fo.judge_error(fineT, cT, want, want, force, 1.0, fB) == 0.0
    assert fo.judge_error(fineT, cT, stale, want, force, 1.0, fB) > 0.0


def test_centre_record_against_direct_sums():
    """An occupied block's record is its far field carried to the block's
    centre to first order: with centres of mass within 2 % of a block of the
    centres, it equals the FP64 direct sum at the centre over the occupied
    blocks beyond the 5^3 around to the second-order remainder; the texel
    counts are those of a direct count."""
    import oracle
    rng = np.random.default_rng(5)
    L = np.array([6.0, 6.0, 6.0], np.float32)
    c, occ = _coarse_mesh(rng, jitter=0.02)
    fine = np.zeros((513, 512, 4), np.float32)
    cnt = np.zeros(4096, int)
    for v in np.nonzero(occ)[0][:40]:       # a few blocks' texels, counted by hand
        x, y, z = fo.VX[v], fo.VY[v], fo.VZ[v]
        k = rng.integers(1, 64)
        for q in rng.choice(64, k, replace=False):
            fx, fy, fz = 4 * x + (q & 3), 4 * y + ((q >> 2) & 3), 4 * z + (q >> 4)
            fine[(fz >> 3) * 64 + fy, (fz & 7) * 64 + fx, 3] = 1.0
        cnt[v] = k
    R = fo.centre_record(fine, c, L, *oracle.far(c))
    assert np.array_equal(R[:, 10], cnt)
    cv = fo._by_voxel(c).astype(np.float64)
    src = cv[:, :3] / np.where(cv[:, 3:4] > 0, cv[:, 3:4], 1)
    m = cv[:, 3]
    cen = fo._centres(L).astype(np.float64)
    for v in np.nonzero(occ)[0][:60]:
        cheb = np.maximum(np.maximum(abs(fo.VX - fo.VX[v]), abs(fo.VY - fo.VY[v])), abs(fo.VZ - fo.VZ[v]))
        use = (m > 0) & (cheb > 2)
        r = src[use] - cen[v]
        inv = 1.0 / np.sqrt((r * r).sum(1))
        a = ((m[use] * inv ** 3)[:, None] * r).sum(0)
        err = np.linalg.norm(R[v, :3] - a) / np.linalg.norm(a)
        assert err < 2e-3, (v, err)
