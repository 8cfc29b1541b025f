"""The library's body relaxation (cc_relax_piles, through the N-API addon's
relaxPiles) against the JavaScript relaxation (js/bodies.js relaxPile, the
restatement of src/bodies.js:97-146 that builds the reference's initial
states bit for bit): the same positions and reports, bit for bit, on piles
that reach every branch -- coincident particles (the 1e-9 floor of the
distance), particles outside the spheroid, a ground patch with particles
beyond its wall, above its sky and below its floor, a pile of one, an empty
pile -- and on whole batches of runs (buildRuns, all bodies in one call)
against the runs built one at a time.  CPU only: no device work."""
import json
import os
import shutil
import subprocess

import pytest

import golden_util as gu

JS = os.path.join(gu.REPO, "cosmic-collisions_amd", "js")
ADDON = os.path.join(JS, "cc_addon.node")
NODE = shutil.which("node")
pytestmark = [pytest.mark.skipif(NODE is None, reason="node is not installed"),
              pytest.mark.skipif(not os.path.exists(ADDON), reason="the N-API addon is not built")]


def node(script, tmp_path):
    path = tmp_path / "t.js"
    path.write_text(script)
    r = subprocess.run([NODE, str(path)], capture_output=True, text=True, timeout=600, cwd=JS)
    assert r.returncode == 0, r.stderr
    return r.stdout


PILES = """
const b = require('%(js)s/bodies.js');
const addon = require('%(addon)s');
const rand = b.seeded(17);
function pile(n, spread, A, C, confine, sweeps) {
  const p = new Float32Array(n * 3);
  for (let i = 0; i < n * 3; i++) p[i] = (rand() - 0.5) * spread;
  return { p, n, a: 0.02, A, C, confine, sweeps };
}
const cases = [];
cases.push(pile(4000, 0.6, 0.25, 0.2, b.SPHEROID, 0));        // crowded, many outside the spheroid
const dup = pile(500, 0.2, 0.3, 0.3, b.SPHEROID, 3);
for (let i = 0; i < 40; i++) for (let k = 0; k < 3; k++) dup.p[(2 * i + 1) * 3 + k] = dup.p[2 * i * 3 + k];   // coincident pairs
cases.push(dup);
const g = pile(3000, 0.8, 0.3, 0.25, b.GROUND, 16);            // beyond the wall, above the sky, below the floor
for (let i = 0; i < 3000; i++) g.p[i * 3 + 1] -= 0.1;
g.p[1] = -0.0; g.p[4] = 0.0;                                     // the sky's signed zeros
cases.push(g);
cases.push(pile(1, 0.1, 0.2, 0.2, b.SPHEROID, 0));
const js = cases.map((q) => Object.assign({}, q, { p: Float32Array.from(q.p) }));
const nat = cases.map((q) => Object.assign({}, q, { p: Float32Array.from(q.p) }));
const rj = b.relaxPiles(js), rn = b.relaxPiles(nat, addon);
const out = cases.map((q, k) => ({
  same: Buffer.compare(Buffer.from(js[k].p.buffer), Buffer.from(nat[k].p.buffer)) === 0,
  moved: Buffer.compare(Buffer.from(q.p.buffer), Buffer.from(nat[k].p.buffer)) !== 0,
  rj: rj[k], rn: rn[k] }));
const empty = b.relaxPiles([pile(0, 0.1, 0.2, 0.2, b.GROUND, 0)], addon)[0];   // nothing to relax
console.log(JSON.stringify({ out, empty }));
"""


def test_native_relaxation_matches_javascript(tmp_path):
    res = json.loads(node(PILES % {"js": JS, "addon": ADDON}, tmp_path))
    out = res["out"]
    assert res["empty"] == {"sweeps": 0, "worst": 0}
    for k, c in enumerate(out):
        assert c["same"], k
        assert c["rj"] == c["rn"], k
    assert all(c["moved"] for c in out[:3])
    assert out[0]["rn"]["sweeps"] == 10 and out[1]["rn"]["sweeps"] == 3 and out[2]["rn"]["sweeps"] == 16


def test_native_relaxation_rejects_bad_piles(tmp_path):
    out = node("""
const addon = require('%s');
const msgs = [];
for (const q of [{ p: new Float32Array(3), n: 1, a: 0, A: 1, C: 1, confine: 0 },
                 { p: new Float32Array(3), n: 1, a: 0.1, A: 1, C: 1, confine: 7 },
                 { p: new Float32Array(3), n: 2, a: 0.1, A: 1, C: 1, confine: 0 },
                 { p: new Float32Array([0, 0, 0, 1e6, 1e6, 1e6]), n: 2, a: 0.001, A: 1, C: 1, confine: 0 }]) {
  try { addon.relaxPiles([q]); msgs.push('accepted'); } catch (e) { msgs.push(e.message); }
}
console.log(JSON.stringify(msgs));
""" % ADDON, tmp_path)
    msgs = json.loads(out)
    assert "cc_relax_piles failed (-1)" in msgs[0]
    assert "cc_relax_piles failed (-1)" in msgs[1]
    assert msgs[2] != "accepted"   # a pile longer than its array
    assert "cc_relax_piles failed (-4)" in msgs[3] and "2^31 cells" in msgs[3]


def test_batch_build_matches_one_at_a_time(tmp_path):
    out = node("""
const nb = require('%s/nbody.js');
const addon = require('%s');
const opts = [{ N: 16384, q: 0.1, dens: 0.75, core: 0.3, spin0: 0.18, spin1: 0.2, angle: 30, incl: 0, vfac: 1.2, rest: 0.3, seed: 5 },
              { N: 24576, q: 0.8, dens: 0.75, core: 0.25, spin0: 0.18, spin1: 0.2, angle: 60, incl: 0, vfac: 2.0, rest: 0.3, seed: 9 },
              { N: 8192, q: 0.03, dens: 0.75, core: 0.35, spin0: 0.18, spin1: 0.2, angle: 5, incl: 0, vfac: 1.0, rest: 0.3, seed: 11 }];
const all = nb.buildRuns(opts, addon);
const res = opts.map((o, k) => {
  const one = nb.buildRun(o);
  return ['pos', 'vel'].every((f) => Buffer.compare(Buffer.from(all[k][f].buffer), Buffer.from(one[f].buffer)) === 0) &&
         JSON.stringify(all[k].sim.cen0) === JSON.stringify(one.sim.cen0);
});
console.log(JSON.stringify(res));
""" % (JS, ADDON), tmp_path)
    assert json.loads(out) == [True, True, True]


def test_python_binding_matches_javascript(tmp_path):
    """ccamd.relax_piles (ctypes) on a spheroid pile and a ground pile, against
    the JavaScript relaxation of the same positions."""
    import numpy as np
    import sys
    sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
    import ccamd
    rng = np.random.default_rng(4)
    body = ((rng.random((2000, 3)) - 0.5) * 0.5).astype(np.float32)
    ground = ((rng.random((1500, 3)) - 0.5) * 0.6).astype(np.float32)
    ground[:, 1] -= 0.12
    body.tofile(tmp_path / "body.f32")
    ground.tofile(tmp_path / "ground.f32")
    node("""
const b = require('%s/bodies.js'); const fs = require('fs');
const rd = (f) => { const b = fs.readFileSync('%s/' + f); return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.length)); };
const p1 = rd('body.f32'), p2 = rd('ground.f32');
const r = b.relaxPiles([{ p: p1, n: 2000, a: 0.02, A: 0.2, C: 0.18, confine: b.SPHEROID, sweeps: 0 },
                        { p: p2, n: 1500, a: 0.02, A: 0.25, C: 0.3, confine: b.GROUND, sweeps: 16 }]);
fs.writeFileSync('%s/body.js.f32', Buffer.from(p1.buffer)); fs.writeFileSync('%s/ground.js.f32', Buffer.from(p2.buffer));
fs.writeFileSync('%s/rep.json', JSON.stringify(r));
""" % (JS, tmp_path, tmp_path, tmp_path, tmp_path), tmp_path)
    reps = ccamd.relax_piles([(body, 0.02, 0.2, 0.18, ccamd.PILE_SPHEROID, 0),
                              (ground, 0.02, 0.25, 0.3, ccamd.PILE_GROUND, 16)])
    assert np.array_equal(body.ravel(), np.fromfile(tmp_path / "body.js.f32", np.float32))
    assert np.array_equal(ground.ravel(), np.fromfile(tmp_path / "ground.js.f32", np.float32))
    js = json.loads((tmp_path / "rep.json").read_text())
    assert reps == [(js[0]["sweeps"], js[0]["worst"]), (js[1]["sweeps"], js[1]["worst"])]
