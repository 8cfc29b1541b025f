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
fs.writeFileSync('%s/f4_vel', Buffer.from(st.vel.buffer));
fs.writeFileSync('%s/matR', Buffer.from(sim.matR.buffer));
const rep = eng.analyze();
console.log(JSON.stringify({ frames: frames, phase: sim.phase, N: sim.N, rep, m: sim.m, Rbox: sim.Rbox, rhoT: sim.rhoT,
                             t: sim.t, same: rep === eng.report }));
""" % (JS, json.dumps(json.loads(str(fx["opt"]))), *([tmp_path] * 6)))
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
    # analyze(): the page's readout on the state it reads back, with the page's units
    f4v = rd("f4_vel").reshape(N, 4)
    want = ccamd.crater_stats(f4, f4v, rd("matR"), out["m"], Rbox=out["Rbox"], lSI=1e4, vSI=1000.0, gSI=9.81,
                              rhoSI=out["rhoT"], vOrb=7900 / 1000, vEsc=11200 / 1000)
    rep = out["rep"]
    assert out["same"] and rep["t"] == out["t"]
    flt = lambda v: [float(x) for x in v] if isinstance(v, list) else float(v)  # noqa: E731
    for k in want:   # JSON writes integral doubles past 2^53 as integers: compare as doubles
        assert flt(rep[k]) == want[k], k
    assert rep["mass"] > 0 and rep["Tmax"] > 0


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_crater_step_many_equals_single_steps(fx, tmp_path):
    """The crater engine's stepMany (cc_step_n, one graph launch per 8-step
    cycle) against the same full steps taken one call at a time: two engines
    built alike, through touchdown and the coupling, then 21 steps each way
    (starting and ending off a cycle boundary); the states bit-identical."""
    script = tmp_path / "many.js"
    script.write_text("""
const cr = require('%s/crater.js'); const fs = require('fs');
const run = (many) => {
  const eng = cr.createCraterEngine();
  const sim = eng.build(%s);
  eng.settleSome(2);
  while (sim.phase === 'settle') eng.settleSome(64);
  while (sim.phase !== 'full') eng.approachStep();
  eng.step(false);
  if (many) { eng.stepMany(13); eng.stepMany(8); } else for (let i = 0; i < 21; i++) eng.step(false);
  const st = eng.readState();
  for (const k of ['pos', 'vel', 'aux', 'q']) fs.writeFileSync('%s/' + (many ? 'm_' : 's_') + k, Buffer.from(st[k].buffer));
  const n = sim.stepNo;
  eng.destroy();
  return n;
};
console.log(JSON.stringify({ single: run(false), many: run(true) }));
""" % (JS, json.dumps(json.loads(str(fx["opt"]))), tmp_path))
    print("crater engine stepMany vs step (node)", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "180", "node", str(script)], capture_output=True, text=True, cwd=JS)
    assert r.returncode == 0, r.stderr[-3000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["single"] == out["many"] == 22
    for k in ("pos", "vel", "aux", "q"):
        a, b = (np.fromfile(tmp_path / (w + k), np.float32) for w in ("s_", "m_"))
        assert np.array_equal(a, b), k


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_crater_524k(tmp_path):
    """The crater at its configuration's stated size, 524,288 grains: past the
    reference's 262,144, so the seat words carry a 19-bit index (the place
    fields one bit up).  Seats bit-exact and one settle step at 1e-5 against
    the oracle with the same seat format."""
    script = tmp_path / "b.js"
    script.write_text("""
const cr = require('%s/crater.js'); const fs = require('fs');
const r = cr.buildCraterRun({ imp: 14, vimp: 20, angle: 60, sed: 3, dens: 2500, nIdx: 4, rest: 0.25, N: 524288 });
for (const [k, a] of [['pos', r.pos], ['vel', r.vel], ['mat', r.sim.matR]]) fs.writeFileSync('%s/' + k, Buffer.from(a.buffer));
const s = {}; for (const k of Object.keys(r.sim)) { const v = r.sim[k]; if (typeof v === 'number') s[k] = v; }
console.log(JSON.stringify(s));
""" % (JS, tmp_path))
    print("crater 524k: build (node)", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "240", "node", str(script)], capture_output=True, text=True, cwd=JS)
    assert r.returncode == 0, r.stderr[-3000:]
    s = json.loads(r.stdout.strip().splitlines()[-1])
    N = s["N"]
    assert N == 524288 and s["W"] == 512 and s["G"] == 128
    pos = np.fromfile(tmp_path / "pos", np.float32).reshape(N, 4)
    vel = np.fromfile(tmp_path / "vel", np.float32).reshape(N, 4)
    mat = np.fromfile(tmp_path / "mat", np.float32)
    q = np.zeros(N, np.float32)
    gfield = (0.0, -9.81 * 10 ** 2 / 1e4, 0.0)
    ctx = ccamd.Context(ccamd.params(N=N, W=512, G=128, m=1.0, a=s["a"], cell=s["cell"], dt=s["dt"], k=s["k"], c=s["c"],
                                     creep=s["creep"], damp=s["damp"], tK=s["tK"], kappa=s["kappa"], field=True,
                                     gField=gfield, frozen=0.0, cpIron=1.0))
    try:
        ctx.upload(pos, vel, mat, q)
        ctx.step_no = 0
        print("crater 524k: one settle step", flush=True)
        ctx.step(1, 0, ccamd.SimArgs())
        a, b = ctx.read_seats()
        P, V, A, Q, M = ctx.download()
    finally:
        ctx.close()
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    u = {"uSize": [512, N // 512], "uInvCell": [f32(1 / s["cell"])], "uCell": [f32(s["cell"])], "uDt": [f32(s["dt"])],
         "uGDt": [f32(s["dt"])], "uTouch": [f32(2 * s["a"])], "uLink2": [f32(s["cell"] ** 2)],
         "uReach2": [f32((s["cell"] + 0.6 * s["a"]) ** 2)], "uK": [f32(s["k"])], "uC": [f32(s["c"])], "uInvM": [1.0],
         "uDamp": [f32(s["damp"])], "uCreep": [f32(s["creep"])], "uTk": [f32(s["tK"])], "uKappa": [f32(s["kappa"])],
         "uCpIron": [1.0], "uSettle": [1.0], "uKick": [0.0], "uGMask": [127], "uGField": list(gfield), "uFrozen": [0.0]}
    su = gu.sim_u(u, mode="field")
    su.grid = oracle.grid(128, ib=19)
    ra, rb = oracle.slots(pos, su.invCell, su.grid)
    assert np.array_equal(a, ra) and np.array_equal(b, rb), "seats"
    seated = np.concatenate([ra.ravel(), rb.ravel()])
    seated = seated[seated < 0x80000000] & ((1 << 19) - 1)
    assert seated.max() >= (1 << 18), "indices past 2^18 seated"
    z4, z1 = np.zeros((513, 512, 4), np.float32), np.zeros(4, np.float32)
    oP, oV, oA, oQ = oracle.sim(su, pos, vel, mat, q, ra, rb, z4, z1, z4, z1, np.full(N, -1.0, np.float32),
                                np.zeros((256, 128, 4), np.float32))
    assert np.array_equal(A[:, 3], oA[:, 3]), "neighbour counts"
    for mine, ref, what in ((V[:, :3], oV[:, :3], "velocity"), (P[:, :3], oP[:, :3], "position"), (Q, oQ, "heat")):
        err = gu.rel_err(mine, ref)
        assert err.max() < 1e-5, (what, err.max())
