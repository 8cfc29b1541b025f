// The batched what-if sweep's instances: `count` giant impacts of uneven size
// (log-uniform 16k-262k particles, multiples of 512), each with its own seeded
// parameters -- mass ratio 0.02-1 (log-uniform), angle 0-75 deg, speed
// 1.0-2.6 v_esc, core fraction 0.2-0.4, its own body seed -- each run through
// the engine's build, settle, approach on rails and kick, then a look (the
// analysis with the Morton sort) and written for the C ABI.  The bodies of all
// instances are built first, their relaxations in one native call.
//   node prep_batch.js <outdir> [count] [seed]
// Files per instance k: <outdir>/k/state.{pos,vel,q,mat}.f32, pptab.f32,
// pplist.i32, run.json; <outdir>/batch.json lists the instances.
'use strict';
const fs = require('fs');
const path = require('path');
const { createEngine, buildRuns } = require('./nbody');
const bodies = require('./bodies');
const addon = require('./cc_addon.node');

const out = process.argv[2];
const count = parseInt(process.argv[3] || '64', 10);
const rand = bodies.seeded(parseInt(process.argv[4] || '20261203', 10));
fs.mkdirSync(out, { recursive: true });
const t0 = Date.now();
const log = (m) => console.log(`[batch ${((Date.now() - t0) / 1000).toFixed(1)}s] ${m}`);
const list = [], opts = [];
for (let k = 0; k < count; k++) {
  const lu = (lo, hi) => Math.exp(Math.log(lo) + rand() * (Math.log(hi) - Math.log(lo)));
  const N = Math.min(262144, Math.max(16384, Math.round(lu(16384, 262144) / 512) * 512));
  opts.push({ N, q: lu(0.02, 1.0), dens: 0.75, core: 0.2 + 0.2 * rand(), spin0: 0.18, spin1: 0.20,
              angle: 75 * rand(), incl: 0, vfac: 1.0 + 1.6 * rand(), rest: 0.3, seed: 1 + Math.floor(rand() * 2147483646) });
}
// every instance's bodies laid out, then all relaxed in one native call on the host's threads
const runs = buildRuns(opts, addon);
log(`built ${count} instances' bodies`);
for (let k = 0; k < count; k++) {
  const opt = opts[k], N = opt.N;
  const NB = createEngine();
  const sim = NB.build(opt, runs[k]);
  runs[k] = null;
  while (sim.phase === 'settle') NB.settleSome(64);
  while (sim.phase === 'approach') NB.rigidStep(0.01);
  const r = NB.sortNow();
  const st = NB.readState();
  const dir = path.join(out, String(k));
  fs.mkdirSync(dir, { recursive: true });
  for (const f of ['pos', 'vel', 'q', 'mat']) fs.writeFileSync(path.join(dir, `state.${f}.f32`), Buffer.from(st[f].buffer));
  fs.writeFileSync(path.join(dir, 'pptab.f32'), Buffer.from(NB.pairTable.buffer));
  fs.writeFileSync(path.join(dir, 'pplist.i32'), Buffer.from(NB.pp.list.buffer.slice(0)));
  const keep = ['N', 'W', 'H', 'G', 'n1', 'n2', 'm', 'a', 'cell', 'dt', 'k', 'c', 'creep', 'damp', 'tK', 'kappa', 'spin0', 'spin1',
                'cen0', 'cen1', 'bulk0', 'bulk1', 't', 'stepNo', 'impactT', 'soft', 'boxes', 'M', 'vc', 'ref'];
  const run = { opt, sim: {}, report: { core: r.core, L: r.L, KE: r.KE, PE: r.PE, EL: r.EL, Q: r.Q, M: r.M, drift: r.drift } };
  for (const f of keep) run.sim[f] = sim[f];
  fs.writeFileSync(path.join(dir, 'run.json'), JSON.stringify(run));
  list.push({ k, N, dir: String(k) });
  log(`instance ${k}: N=${N} q=${opt.q.toFixed(3)} angle=${opt.angle.toFixed(1)} vfac=${opt.vfac.toFixed(2)} core=${opt.core.toFixed(3)}`);
  NB.destroy();
}
fs.writeFileSync(path.join(out, 'batch.json'), JSON.stringify({ count, instances: list }));
log(`written ${count} instances, ${list.reduce((s, x) => s + x.N, 0)} particles`);
