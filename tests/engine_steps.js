// Test driver (GPU box): the JavaScript engine (js/nbody.js over the N-API
// addon) taken through steps of its own step() from a given state, as a page
// drives it (sim.t advanced by the caller after each step).  The job (JSON):
//   opt     build() options (the run's constants come from the engine's build)
//   dir     pos.f32, vel.f32, mat.f32, q.f32, pp.i32: the state and loose list
//   fields  sim fields set before the first step (t, stepNo, soft, impactT, boxes, bulk0/1, cen0/1)
//   steps   how many; dump: after which steps to write the state
// Before the steps the engine's DIAG pass runs (boxes, seats, mesh, pairwise
// pass), as the reference's analysis leaves the mesh and its monopole.
// Writes <out>/stepN.{pos,vel,aux,q}.f32, the mesh step's force0.f32,
// mono0.f32 and ppforce.f32 after step 0, and after.json (sim scalars).
'use strict';
const fs = require('fs');
const path = require('path');
const { createEngine } = require('../cosmic-collisions_amd/js/nbody');

const job = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const out = process.argv[3];
const f32 = (n) => { const b = fs.readFileSync(path.join(job.dir, n)); return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4).slice(); };
const NB = createEngine();
const sim = NB.build(job.opt);
const addon = NB.addon, ctx = NB.ctx, N = sim.N;
const mat = f32('mat.f32');
addon.uploadState(ctx, f32('pos.f32'), f32('vel.f32'), mat, f32('q.f32'), N);
sim.matR = mat;
Object.assign(sim, job.fields);
const b = fs.readFileSync(path.join(job.dir, 'pp.i32'));
const list = new Int32Array(b.buffer, b.byteOffset, b.byteLength / 4).slice();
addon.setPP(ctx, list, list.length);
NB.passes.diag(true);
const put = (name, arr) => fs.writeFileSync(path.join(out, name), Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength));
for (let n = 0; n < job.steps; n++) {
  NB.step(false, false);
  sim.t += sim.dt;
  if (n === 0) {
    const force = new Float32Array(513 * 512 * 4), mono = new Float32Array(4);
    addon.readMesh(ctx, 0, 6, force); addon.readMesh(ctx, 0, 2, mono);
    const l = new Float32Array(8192 * 4), bx = new Float32Array(8192), f = new Float32Array(4 * 8192 * 4);
    addon.readPP(ctx, l, bx, f);
    put('force0.f32', force); put('mono0.f32', mono); put('ppforce.f32', f);
  }
  if (job.dump.includes(n)) {
    const st = NB.readState();
    for (const k of ['pos', 'vel', 'aux', 'q']) put('step' + n + '.' + k + '.f32', st[k]);
  }
}
fs.writeFileSync(path.join(out, 'after.json'), JSON.stringify({ t: sim.t, stepNo: sim.stepNo, dt: sim.dt, k: sim.k, c: sim.c, damp: sim.damp, soft: sim.soft }));
NB.destroy();
console.log('engine steps done:', job.steps);
