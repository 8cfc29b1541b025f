"""The learned far field on the GPU (cc_ff_*): its BF16 matrix-core forward
pass and weight gradients against the CPU restatement (oracle/ff_oracle.py,
FP64 with the device's BF16 roundings), training on samples of the exact far
field, and the judge in the step: with an unmeetable bound every mesh step
falls back to the exact far field and the run is bit-identical to one
without the network.  Needs the GPU.

Bars: forward outputs 1e-2 relative L2; loss 1e-3 relative; each layer's
weight gradient 3e-2 relative L2.  Every stored activation is BF16: a sum
that lands next to a BF16 rounding boundary rounds one way in the device's
FP32 and the other in the oracle's FP64, a 2^-8 change the following
layers carry (measured: 2.4e-3 on the forward output).
"""
import json
import os
import sys

import numpy as np
import pytest

import golden_util as gu
import ff_oracle as fo

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

pytestmark = pytest.mark.gpu


def synthetic(rng, n):
    """Samples shaped like the coarse mesh: a few hundred occupied blocks."""
    xs, ts, ms = [], [], []
    for _ in range(n):
        occ = rng.random(4096) < 0.08
        x = np.zeros((4096, 16), np.float32)
        x[occ, 0] = rng.gamma(2.0, 0.5, occ.sum())
        x[occ, 1:4] = rng.uniform(-0.5, 0.5, (occ.sum(), 3))
        t = np.zeros((4096, 16), np.float32)
        t[occ, :10] = rng.normal(size=(occ.sum(), 10))
        xs.append(x); ts.append(t); ms.append(occ.astype(np.float32))
    return xs, ts, ms


def init_weights(rng):
    w = np.zeros(fo.weight_count(), np.float32)
    off = 0
    for ci, co, act in fo.LAYERS:
        a = np.sqrt(3.0 * (2.0 if act else 1.0) / (27 * ci))
        w[off:off + 27 * ci * co] = rng.uniform(-a, a, 27 * ci * co)
        off += 27 * ci * co
        w[off:off + co] = rng.normal(scale=0.05, size=co)
        off += co
    return w


@pytest.fixture(scope="module")
def net():
    rng = np.random.default_rng(7)
    ff = ccamd.FarField(capacity=8)
    assert ff.nw == fo.weight_count()
    w = init_weights(rng)
    ff.set_weights(w)
    xs, ts, ms = synthetic(rng, 2)
    for x, t, m in zip(xs, ts, ms):
        ff.push(x, t, m)
    yield ff, w, np.stack(xs), np.stack(ts), np.stack(ms)
    ff.close()


def test_ff_forward_vs_oracle(net):
    ff, w, xs, ts, ms = net
    print("ff forward: 2 samples", flush=True)
    ra, rl, out = ff.eval(0, 2, outputs=True)
    y, _ = fo.forward(w.astype(np.float64), xs.reshape(2, 16, 16, 16, 16))
    y = y.reshape(2, 4096, 16)
    err = np.linalg.norm(out - y) / np.linalg.norm(y)
    assert err < 1e-2, err


def test_ff_gradient_vs_oracle(net):
    ff, w, xs, ts, ms = net
    print("ff gradient: 2 samples", flush=True)
    g, loss = ff.gradient(0, 2)
    y, c = fo.forward(w.astype(np.float64), xs.reshape(2, 16, 16, 16, 16))
    L, gy = fo.loss_grad(y, ts.reshape(2, 16, 16, 16, 16), ms)
    ref = fo.backward(c, gy)
    assert loss == pytest.approx(L, rel=1e-3)
    off = 0
    for l, (ci, co, _) in enumerate(fo.LAYERS):
        n = 27 * ci * co + co
        a, b = g[off:off + n], ref[off:off + n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err < 3e-2, (l, err)
        off += n


def make_ctx(fx):
    s = json.loads(str(fx["sim"]))
    ctx = ccamd.Context(ccamd.params(N=s["N"], W=s["W"], G=s["G"], m=s["m"], a=s["a"], cell=s["cell"], dt=s["dt"],
                                     k=s["k"], c=s["c"], creep=s["creep"], damp=s["damp"], tK=s["tK"], kappa=s["kappa"],
                                     spin0=s["spin0"], spin1=s["spin1"]))
    ctx.upload(fx["pos"], fx["vel"], fx["mat"], fx["q"])
    u = gu.draws(fx, "dep")[0]
    ctx.set_boxes([float(x) for x in u["uLA"]], [float(x) for x in u["uLB"]], float(u["uBoxB"][0]) > 0.5)
    ctx.set_pp_table(fx["pptab"])
    ctx.set_pp(fx["pp_index"])
    return ctx


def test_ff_passes_and_pairs_vs_oracle():
    """Recorded passes of a stepped 16k run: features and centre records bit
    for bit the oracle's from the meshes read back; a pair's input (after the
    BF16 rounding) and target as the oracle forms them; training on the pairs
    brings the loss and the error of the predicted change down."""
    fx = gu.fixture("c1_impact")
    ctx = make_ctx(fx)
    ff = ccamd.FarField(capacity=64)
    try:
        args = ccamd.sim_args()
        seq = []
        for k in range(24):
            ctx.step_n(8, args)
            ctx.mesh()
            seq.append(ff.record(ctx, 0))
            if k in (0, 23):
                ms = {w: ctx.read_mesh(0, w) for w in ("fine", "coarse", "mono", "farA", "farT0", "farT1")}
                feat, meta, rec = ff.read_pass(seq[-1])
                of, om = fo.features(ms["coarse"], ms["mono"][3], meta[:3])
                assert np.array_equal(feat, of) and np.array_equal(meta, om)
                assert np.array_equal(rec, fo.centre_record(ms["fine"], ms["coarse"], meta[:3], ms["farA"], ms["farT0"],
                                                            ms["farT1"]))
                assert (of[:, 0] > 0).sum() > 20
        for i in range(len(seq) - 2):
            ff.pair(seq[i], seq[i + 1])
            ff.pair(seq[i], seq[i + 2])
        (fT, mT, RT), (fB, _, RB) = ff.read_pass(seq[5]), ff.read_pass(seq[3])
        x, t, m = ff.sample(2 * 3 + 1)                 # the pair (seq[3], seq[5])
        ox, s_ = fo.pair_input(fT, fB)
        assert np.array_equal(x, fo.bf16(ox).astype(np.float32))
        ot, om = fo.pair_target(RT, RB, fT, fB, mT[3], s_)
        assert np.array_equal(m, om) and (m > 0).sum() > 20
        assert gu.rel_err(t, ot).max() < 1e-5
        print("ff training: %d samples" % ff.count, flush=True)
        before = ff.eval()[0]
        losses = ff.train(400, batch=8, lr=2e-3, seed=5)
        after = ff.eval()[0]
        print("ff: loss %.4g -> %.4g, relative error of the predicted change %.4g -> %.4g"
              % (losses[:5].mean(), losses[-20:].mean(), before, after))
        assert np.isfinite(losses).all()
        assert losses[-20:].mean() < 0.5 * losses[:5].mean()
        assert after < before
    finally:
        ff.close()
        ctx.close()


def test_ff_judge_fallback_is_exact():
    """An unmeetable bound: every judge fails, every mesh step runs the exact
    far field, and the state is bit-identical to a run without the network;
    with a loose bound the network stands in between judges from the second
    judge on (the first has no base to predict from)."""
    fx = gu.fixture("c1_impact")
    plain, strict, loose = make_ctx(fx), make_ctx(fx), make_ctx(fx)
    ff = ccamd.FarField(capacity=0)
    try:
        strict.set_far_field(ff, 16, 1e-12)
        loose.set_far_field(ff, 16, 1e9)
        args = ccamd.sim_args()
        for ctx in (plain, strict, loose):
            ctx.step_no = 0
            ctx.step_n(64, args)
        for a, b in zip(plain.download(), strict.download()):
            assert np.array_equal(a, b)
        st_s, st_l = strict.ff_stats(), loose.ff_stats()
        print("judge stats strict", st_s[0], "loose", st_l[0])
        assert st_s[0, 1] == 4 and st_s[0, 2] == 4 and st_s[0, 3] == 0
        assert st_l[0, 1] == 4 and st_l[0, 2] == 1 and st_l[0, 3] == 3 and st_l[0, 6] == 3
        assert st_l[0, 0] > 0
        assert all(np.isfinite(a).all() for a in loose.download())
    finally:
        for c in (plain, strict, loose):
            c.close()
        ff.close()


def test_ff_carried_far_field_and_judge_vs_oracle():
    """A zero network (its change is exactly zero) stands in: on the mesh step
    after a judge, every block occupied at the base gets the base's far field
    carried to its present centre of mass (the oracle's carry, bit for bit),
    every block filled since gets k_far's exact far field (bit for bit a run
    without the network); at the next judge the error measured is the
    oracle's from the meshes read back."""
    fx = gu.fixture("c1_impact")
    gm = json.loads(str(fx["sim"]))["m"]
    exact, zero = make_ctx(fx), make_ctx(fx)
    ff = ccamd.FarField(capacity=8)
    net = ccamd.FarField(capacity=0)
    try:
        net.set_weights(np.zeros(net.nw, np.float32))
        zero.set_far_field(net, 16, 1e9)
        args = ccamd.sim_args()
        for ctx in (exact, zero):
            ctx.step_no = 0
            ctx.step_n(17, args)        # judges at 0 (no base yet) and 16 (from the base at 0, passes); base 16
        base = ff.record(exact, 0)      # the same pass as the zero run's base at 16
        for ctx in (exact, zero):
            ctx.step_n(8, args)         # the mesh step at 24: the zero network stands in
        assert zero.ff_stats()[0, 3] == 1
        fB, mB, RB = ff.read_pass(base)
        cz = zero.read_mesh(0, "coarse")
        assert np.array_equal(cz, exact.read_mesh(0, "coarse"))   # one state up to the mesh step at 24
        fT, mT = fo.features(cz, zero.read_mesh(0, "mono")[3], mB[:3])
        _, s_ = fo.pair_input(fT, fB)
        occ, old = fT[:, 0] > 0, (fT[:, 0] > 0) & (fB[:, 0] > 0)
        assert old.sum() > 20
        got = [fo._by_voxel(zero.read_mesh(0, w)) for w in ("farA", "farT0", "farT1")]
        want = fo.carry(np.zeros((4096, 16), np.float32), RB, s_, mT[3], cz, mB[:3])
        ex = [fo._by_voxel(exact.read_mesh(0, w)) for w in ("farA", "farT0", "farT1")]
        for g, w_, e in zip(got, want, ex):
            assert np.array_equal(g[old], w_[old])
            assert np.array_equal(g[occ & ~old], e[occ & ~old])
            assert np.all(g[~occ] == 0)
        print("carried: %d blocks from the base, %d filled since" % (old.sum(), (occ & ~old).sum()))
        zero.step_n(9, args)            # to the judge at 32: the base at 16 against the zero run's own exact pass
        st = zero.ff_stats()[0]
        assert st[1] == 3 and st[6] == 2
        ms = {w: zero.read_mesh(0, w) for w in ("fine", "coarse", "mono", "farA", "farT0", "farT1", "force")}
        f32_, m32 = fo.features(ms["coarse"], ms["mono"][3], mB[:3])
        _, s32 = fo.pair_input(f32_, fB)
        learned = fo.carry(np.zeros((4096, 16), np.float32), RB, s32, m32[3], ms["coarse"], mB[:3])
        exact_far = [fo._by_voxel(ms[w]) for w in ("farA", "farT0", "farT1")]
        want_err = fo.judge_error(ms["fine"], ms["coarse"], learned, exact_far, ms["force"], gm, fB)
        print("judge at 32: %.6e, oracle %.6e" % (st[0], want_err))
        assert want_err > 0 and abs(st[0] - want_err) <= 1e-4 * want_err
    finally:
        for c in (exact, zero):
            c.close()
        ff.close()
        net.close()
