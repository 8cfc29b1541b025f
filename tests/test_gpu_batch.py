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
