// Runs a configuration through the engine's own phases -- build, settle,
// approach on rails, the kick at first touch, then `steps` full steps and a
// look (the analysis with a Morton sort, as sortNow does) -- and writes the
// state with everything a caller needs to continue it through the C ABI.
//   node prep_state.js '<build opt JSON>' <outdir> [steps]
// Files: state.{pos,vel,q,mat}.f32, pptab.f32, pplist.i32, run.json.
'use strict';
const fs = require('fs');
const path = require('path');
const { createEngine } = require('./nbody');

const opt = JSON.parse(process.argv[2]);
const out = process.argv[3];
const extra = parseInt(process.argv[4] || '0', 10);
fs.mkdirSync(out, { recursive: true });
const NB = createEngine();
const t0 = Date.now();
const log = (m) => console.log(`[prep ${((Date.now() - t0) / 1000).toFixed(1)}s] ${m}`);
const sim = NB.build(opt);
log(`built N=${sim.N} n1=${sim.n1} dt=${sim.dt.toExponential(3)} settle=${sim.settleTotal}`);
while (sim.phase === 'settle') NB.settleSome(64);
log('settled');
let rigid = 0;
while (sim.phase === 'approach') { NB.rigidStep(0.01); rigid++; }
log(`approach: ${rigid} rigid steps, impact at t=${sim.t.toFixed(4)}`);
if (extra > 0) { NB.stepMany(extra); log(`${extra} full steps`); }
const r = NB.sortNow();
log(`look: largest ${r.largest.toFixed(3)} M, loose ${r.pp.loose}, L ${r.L.toFixed(5)}`);
const st = NB.readState();
for (const k of ['pos', 'vel', 'q', 'mat']) fs.writeFileSync(path.join(out, `state.${k}.f32`), Buffer.from(st[k].buffer));
fs.writeFileSync(path.join(out, 'pptab.f32'), Buffer.from(NB.pairTable.buffer));
fs.writeFileSync(path.join(out, 'pplist.i32'), Buffer.from(NB.pp.list.buffer.slice(0)));
const keep = ['N', 'W', 'H', 'G', 'n1', 'n2', 'm', 'a', 'cell', 'dt', 'k', 'c', 'creep', 'damp', 'tK', 'kappa', 'spin0', 'spin1',
              'cen0', 'cen1', 'bulk0', 'bulk1', 't', 'stepNo', 'impactT', 'soft', 'boxes', 'M', 'vc', 'ref'];
const run = { opt, sim: {}, report: { core: r.core, L: r.L, KE: r.KE, PE: r.PE, EL: r.EL, Q: r.Q, M: r.M, drift: r.drift } };
for (const k of keep) run.sim[k] = sim[k];
fs.writeFileSync(path.join(out, 'run.json'), JSON.stringify(run, null, 1));
NB.destroy();
log('written ' + out);
