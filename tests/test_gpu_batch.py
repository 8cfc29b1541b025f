"""Batched sweeps (cc_batch_*): instances of uneven size stepped together,
one graph launch per mesh cycle for the whole batch, against each instance
stepped alone (cc_step_n): bit-identical states after every call.  One
member has a particle count that is not a power of two (512 x 24).  Needs
the GPU."""
import json
import os
import sys

import numpy as np
import pytest

import golden_util as gu

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

pytestmark = pytest.mark.gpu


def members():
    out = []
    for name, N in (("impact", None), ("twobox", None), ("impact", 512 * 24)):
        fx = gu.fixture("c1_" + name)
        s = json.loads(str(fx["sim"]))
        n = N or s["N"]
        u = gu.draws(fx, "dep")[0]
        out.append({"fx": fx, "s": s, "N": n, "W": 512 if N else s["W"], "pp": fx["pp_index"][fx["pp_index"] < n],
                    "LA": np.array([float(x) for x in u["uLA"]]), "LB": np.array([float(x) for x in u["uLB"]]),
                    "boxB": float(u["uBoxB"][0]), "mono1": gu.mesh(fx, 1)["mono"] if "mono1_idx" in fx else None})
    return out


def make(m):
    s, fx, N = m["s"], m["fx"], m["N"]
    ctx = ccamd.Context(ccamd.params(N=N, W=m["W"], G=s["G"], m=s["m"], a=s["a"], cell=s["cell"], dt=s["dt"], k=s["k"],
                                     c=s["c"], creep=s["creep"], damp=s["damp"], tK=s["tK"], kappa=s["kappa"],
                                     spin0=s["spin0"], spin1=s["spin1"]))
    ctx.upload(fx["pos"][:N], fx["vel"][:N], fx["mat"][:N], fx["q"][:N])
    ctx.set_boxes(m["LA"], m["LB"], m["boxB"] > 0.5)
    if m["mono1"] is not None:
        ctx.set_mono(1, m["mono1"])
    ctx.set_pp_table(fx["pptab"])
    ctx.set_pp(m["pp"])
    return ctx


def boxes(m, start, n):
    k = sum(1 for i in range(n) if (start + i) % 8 == 0)
    return np.array([np.concatenate([m["LA"] + 1e-3 * j, m["LB"] - 2e-3 * j, [m["boxB"]]]) for j in range(k)])


def test_batch_equals_instances_alone():
    ms = members()
    alone = [make(m) for m in ms]
    together = [make(m) for m in ms]
    batch = ccamd.Batch(together)
    try:
        args = ccamd.sim_args()
        for call, n in enumerate((24, 21, 35)):
            start = alone[0].step_no
            print("batch call", call, "steps", n, "from", start, flush=True)
            bl = [boxes(m, start, n) for m in ms]
            for ctx, b in zip(alone, bl):
                ctx.step_n(n, args, b)
            batch.step_n(n, args, bl)
            for i, (a, t) in enumerate(zip(alone, together)):
                assert a.step_no == t.step_no == start + n
                for x, y, what in zip(a.download(), t.download(), ("pos", "vel", "aux", "q", "mat")):
                    assert np.array_equal(x, y), (call, i, what)
            for ctx in alone + together:   # eager work between calls
                ctx.step(0, 0, args)
                ctx.sync()
    finally:
        batch.close()
        for c in alone + together:
            c.close()


def test_batch_with_learned_far_field_equals_instances_alone():
    """One learned far field on every member: the batch predicts every member's
    boxes in one forward pass per mesh step and judges them together after the
    cycle; each member alone predicts and judges its own.  The states are bit
    for bit the same after every call (a loose bound, so the network stands in
    between judges from the second judge on), and so are the judges' verdicts."""
    ms = members()
    alone = [make(m) for m in ms]
    together = [make(m) for m in ms]
    ff = ccamd.FarField(capacity=0)   # its seeded initial weights: any prediction will do
    batch = None
    try:
        for ctx in alone + together:
            ctx.set_far_field(ff, 16, 1e9)
        batch = ccamd.Batch(together)
        args = ccamd.sim_args()
        for call, n in enumerate((24, 21, 35)):
            start = alone[0].step_no
            print("batch with the network, call", call, "steps", n, "from", start, flush=True)
            bl = [boxes(m, start, n) for m in ms]
            for ctx, b in zip(alone, bl):
                ctx.step_n(n, args, b)
            batch.step_n(n, args, bl)
            for i, (a, t) in enumerate(zip(alone, together)):
                for x, y, what in zip(a.download(), t.download(), ("pos", "vel", "aux", "q", "mat")):
                    assert np.array_equal(x, y), (call, i, what)
                sa, st = a.ff_stats(), t.ff_stats()
                assert np.array_equal(sa[:, [1, 2, 3, 6]], st[:, [1, 2, 3, 6]]), (call, i, sa, st)
                assert np.allclose(sa[:, 0], st[:, 0], rtol=1e-9, atol=0), (call, i)
            for ctx in alone + together:   # eager work between calls
                ctx.step(0, 0, args)
                ctx.sync()
        assert all(t.ff_stats()[0, 3] > 0 for t in together)
    finally:
        if batch is not None:
            batch.close()
        for c in alone + together:
            c.close()
        ff.close()


def make_crater(N):
    """A crater instance (constant gravity: no mesh, no pairwise pass) of the first N grains of the reference's
    crater dump after the coupling."""
    fx = gu.fixture("c3_crater")
    h = gu.crater_host(fx, "crater_build")[0]
    s, k = h["sim"], h["consts"]
    ctx = ccamd.Context(ccamd.params(N=N, W=s["W"], G=s["G"], m=1.0, a=s["a"], cell=s["cell"], dt=s["dt"], k=s["k"],
                                     c=s["c"], creep=s["creep"], damp=s["damp"], tK=s["tK"], kappa=s["kappa"],
                                     field=True, gField=(0.0, -k["G_EARTH"] * k["T_UNIT"] ** 2 / k["L_UNIT"], 0.0),
                                     frozen=k["MAT_FROZEN"], cpIron=1.0))
    pos, vel, q = gu.crater_state(fx, "post")
    ctx.upload(pos[:N], vel[:N], fx["mat"][:N], q[:N])
    return ctx


def test_batch_with_constant_gravity_members_equals_instances_alone():
    """Two crater instances (constant gravity: every step a seat rebuild or not and the step pass, no mesh) in
    one batch with an impact instance (its mesh and pairwise passes every eighth step): the fused step and seat
    passes serve both kinds, and every member is bit-identical to itself stepped alone."""
    ms = members()[:1]
    alone = [make_crater(16384), make_crater(128 * 40)] + [make(m) for m in ms]
    together = [make_crater(16384), make_crater(128 * 40)] + [make(m) for m in ms]
    batch = ccamd.Batch(together)
    try:
        args = ccamd.sim_args()
        for call, n in enumerate((24, 21, 35)):
            start = alone[0].step_no
            print("mixed batch call", call, "steps", n, "from", start, flush=True)
            bl = [np.zeros((0, 7)), np.zeros((0, 7))] + [boxes(m, start, n) for m in ms]
            for ctx, b in zip(alone, bl):
                ctx.step_n(n, args, b if len(b) else None)
            batch.step_n(n, args, bl)
            for i, (a, t) in enumerate(zip(alone, together)):
                assert a.step_no == t.step_no == start + n
                for x, y, what in zip(a.download(), t.download(), ("pos", "vel", "aux", "q", "mat")):
                    assert np.array_equal(x, y), (call, i, what)
    finally:
        batch.close()
        for c in alone + together:
            c.close()


def test_batch_rejects_misaligned_members():
    ms = members()[:2]
    ctxs = [make(m) for m in ms]
    batch = ccamd.Batch(ctxs)
    try:
        ctxs[1].step_no = 3
        with pytest.raises(ccamd.CCError, match="mesh cycle"):
            batch.step_n(8)
    finally:
        batch.close()
        for c in ctxs:
            c.close()
