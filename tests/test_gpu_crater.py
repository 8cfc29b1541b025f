"""The crater path (chicxulub.html) on the HIP library, through its C ABI and
through the JavaScript crater engine over the N-API addon, against the
reference's crater page on SwiftShader (tests/golden/fixtures/c3_crater.npz)
and the oracle.  Needs the GPU.

Bars as for the giant impact: seats bit-exact, neighbour counts exact, FP32
values 1e-5 relative (|x - ref| <= 1e-5 max(|ref|, 1e-3 rms(ref))); the
coupling's flow direction (acos/tan, implementation-defined in GLSL) 1e-4 rad
against the reference and 1e-5 against the oracle.
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import golden_util as gu
import oracle
from test_oracle_crater import couple_args

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402

pytestmark = pytest.mark.gpu
JS = os.path.join(gu.REPO, "cosmic-collisions_amd", "js")


@pytest.fixture(scope="module")
def fx():
    return gu.fixture("c3_crater")


def crater_params(fx):
    h = gu.crater_host(fx, "crater_build")[0]
    s, k = h["sim"], h["consts"]
    return ccamd.params(N=s["N"], W=s["W"], G=s["G"], m=1.0, a=s["a"], cell=s["cell"], dt=s["dt"], k=s["k"], c=s["c"],
                        creep=s["creep"], damp=s["damp"], tK=s["tK"], kappa=s["kappa"], field=True,
                        gField=(0.0, -k["G_EARTH"] * k["T_UNIT"] ** 2 / k["L_UNIT"], 0.0), frozen=k["MAT_FROZEN"],
                        cpIron=1.0)


@pytest.fixture(scope="module")
def ctx(fx):
    c = ccamd.Context(crater_params(fx))
    yield c
    c.close()


def test_crater_steps(fx, ctx):
    """Seats and the field-gravity step pass through cc_step, in the page's
    schedule (seats on even steps): two settle steps from the build, two full
    steps after the coupling, each from the reference's own input state."""
    N = ctx.N
    for (src, dst, settle), step_no in zip(gu.CRATER_STEPS, (0, 1, 0, 1)):
        print("crater step", src, "->", dst, flush=True)
        pos, vel, q = gu.crater_state(fx, src)
        ctx.upload(pos, vel, fx["mat"], q)
        ctx.step_no = step_no
        ctx.step(settle, 0, ccamd.SimArgs())
        if step_no == 0:
            a, b = ctx.read_seats()
            ra = gu.dense(fx, dst + "_seatA", (ctx.AH, ctx.AW), gu.EMPTY)
            rb = gu.dense(fx, dst + "_seatB", (ctx.AH, ctx.AW), gu.EMPTY)
            assert np.array_equal(a, ra) and np.array_equal(b, rb), "seats " + dst
        P, V, A, Q, M = ctx.download()
        rp, rv, ra_ = (fx["%s_%s" % (dst, n)].reshape(N, 4) for n in ("pos", "vel", "aux"))
        rq = fx[dst + "_q"].reshape(N, 4)[:, 0]
        assert np.array_equal(A[:, 3], ra_[:, 3]), "neighbour counts " + dst
        assert np.array_equal(P[:, 3], rp[:, 3])
        for mine, ref, what in ((A[:, :3], ra_[:, :3], "gravity"), (V[:, :3], rv[:, :3], "velocity"),
                                (P[:, :3], rp[:, :3], "position"), (Q, rq, "heat"), (V[:, 3], rv[:, 3], "temperature")):
            err = gu.rel_err(mine, ref)
            assert err.max() < 1e-5, (dst, what, err.max(), int(err.argmax()))
        assert ctx.step_no == step_no + 1


def test_crater_couple(fx, ctx):
    """COUPLE_FS through cc_couple with the page's host values (doubles)."""
    h = gu.crater_host(fx, "crater_build")[0]
    s, k = h["sim"], h["consts"]
    centre = gu.crater_host(fx, "crater_touchdown")[0]["centre"]
    N = ctx.N
    pre = [fx["pre_" + n].reshape(N, 4) for n in ("pos", "vel", "aux")]
    ctx.upload(pre[0], pre[1], fx["mat"], fx["pre_q"].reshape(N, 4)[:, 0])
    ctx.couple(centre, vimp=s["vImp"], aimp=s["aImp"], c1=k["HH_C1"], invMu=1 / k["HH_MU"], pexp=k["HH_P"],
               stall=0.5 * s["Dtc"], z=k["Z_MODEL"], heat=k["E_KG"] / k["CP_ROCK"], tK=s["tK"], frozen=k["MAT_FROZEN"],
               ambient=300.0, melt=k["MAT_MELT"])
    P, V, A, Q, M = ctx.download()
    rp, rv = (fx["post_" + n].reshape(N, 4) for n in ("pos", "vel"))
    rq = fx["post_q"].reshape(N, 4)[:, 0]
    assert np.array_equal(P, rp), "positions and melt labels"
    sp = np.linalg.norm(V[:, :3].astype(np.float64), axis=1)
    rsp = np.linalg.norm(rv[:, :3].astype(np.float64), axis=1)
    moving = rsp > 1e-3 * np.sqrt(np.mean(rsp ** 2))
    turn = np.linalg.norm(V[moving, :3] / sp[moving, None] - rv[moving, :3] / rsp[moving, None], axis=1)
    assert turn.max() < 1e-4, turn.max()
    for mine, ref, what in ((sp, rsp, "speed"), (V[:, 3], rv[:, 3], "temperature"), (Q, rq, "heat")):
        err = gu.rel_err(mine, ref)
        assert err.max() < 1e-5, (what, err.max())
    # the oracle with the same float uniforms: every component
    u = gu.draws(fx, "couple")[0]
    oP, oV, oA, oQ = oracle.couple(couple_args(u, k["MAT_MELT"]), pre[0], pre[1], pre[2], fx["mat"])
    assert np.array_equal(P, oP)
    assert gu.rel_err_vec(V[:, :3], oV[:, :3]).max() < 1e-5
    assert gu.rel_err(Q, oQ).max() < 1e-5


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_crater_engine_js(fx, tmp_path):
    """The JavaScript crater engine over the addon: its build uploads the
    reference's initial state bit for bit, its settle step matches the
    reference's, the approach takes the page's 75 frames to touchdown, and the
    coupling leaves the frozen shell still and the impactor molten."""
    script = tmp_path / "crater.js"
    script.write_text("""
const cr = require('%s/crater.js'); const fs = require('fs');
const eng = cr.createCraterEngine();
const sim = eng.build(%s);
let st = eng.readState();
fs.writeFileSync('%s/b0_pos', Buffer.from(st.pos.buffer));
eng.settleSome(1);
st = eng.readState();
for (const k of ['pos', 'vel', 'aux', 'q']) fs.writeFileSync('%s/s1_' + k, Buffer.from(st[k].buffer));
let frames = 0;
while (sim.phase !== 'full' && frames < 1000) { eng.approachStep(); frames++; }
st = eng.readState();
for (const k of ['pos', 'vel']) fs.writeFileSync('%s/post_' + k, Buffer.from(st[k].buffer));
for (let i = 0; i < 4; i++) eng.step(false);
st = eng.readState();
fs.writeFileSync('%s/f4_pos', Buffer.from(st.pos.buffer));
console.log(JSON.stringify({ frames: frames, phase: sim.phase, N: sim.N }));
""" % (JS, json.dumps(json.loads(str(fx["opt"]))), *([tmp_path] * 4)))
    print("crater engine (node)", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "120", "node", str(script)], capture_output=True, text=True, cwd=JS)
    assert r.returncode == 0, r.stderr[-3000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    N = out["N"]
    assert out["phase"] == "full" and out["frames"] == 75

    def rd(name):
        return np.fromfile(tmp_path / name, np.float32)
    assert np.array_equal(rd("b0_pos"), fx["b0_pos"].ravel())
    rp, rv, ra = (fx["s1_" + n].reshape(N, 4) for n in ("pos", "vel", "aux"))
    assert np.array_equal(rd("s1_aux").reshape(N, 4)[:, 3], ra[:, 3])
    assert gu.rel_err(rd("s1_vel").reshape(N, 4)[:, :3], rv[:, :3]).max() < 1e-5
    assert gu.rel_err(rd("s1_pos").reshape(N, 4)[:, :3], rp[:, :3]).max() < 1e-5
    post = rd("post_pos").reshape(N, 4)
    f4 = rd("f4_pos").reshape(N, 4)
    frozen = post[:, 3] == 0
    assert frozen.any() and np.array_equal(f4[frozen, :3], post[frozen, :3]), "the frozen shell moved"
    assert (post[:, 3] == 3).sum() >= (fx["b0_pos"].reshape(N, 4)[:, 3] == 4).sum(), "the impactor is melt"
    assert np.isfinite(f4).all()
