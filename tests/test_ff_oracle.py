"""The CPU restatement of the learned far field network (oracle/ff_oracle.py)
pinned by what the mathematics fixes: the convolution against a brute-force
loop, pooling and upsampling against their adjoints, and the whole
backward pass against central finite differences of the loss in FP64.
CPU only."""
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
