// The engine driven as the reference's page drives a run (index.html:592-600,
// __impact.runTo): build, settle, approach on rails, the kick at first touch,
// then full steps with a synchronous analysis every `every` steps (the
// groups, books, boxes, loose list and, every eighth look, the Morton sort).
// Prints one JSON line: particle-steps/s of the full steps with and without
// the analyses' share of the time.
//   node run_timed.js '<build opt JSON>' <steps> [every]
'use strict';
const { createEngine } = require('./nbody');

const opt = JSON.parse(process.argv[2]);
const steps = parseInt(process.argv[3] || '3600', 10);
const every = parseInt(process.argv[4] || '360', 10);
const NB = createEngine();
const sim = NB.build(opt);
while (sim.phase === 'settle') NB.settleSome(64);
while (sim.phase === 'approach') NB.rigidStep(0.01);
NB.analyze();
let stepMs = 0, lookMs = 0, done = 0, looks = 0;
const phases = { diag: 0, analyse: 0, cells: 0, copy: 0, host: 0, rest: 0 };
const t0 = process.hrtime.bigint();
while (done < steps) {
  const n = Math.min(every, steps - done);
  const a = process.hrtime.bigint();
  NB.stepMany(n);
  const b = process.hrtime.bigint();
  NB.analyze();
  const c = process.hrtime.bigint();
  stepMs += Number(b - a) / 1e6;
  lookMs += Number(c - b) / 1e6;
  for (const k of Object.keys(phases)) phases[k] += NB.lookMs[k];
  done += n;
  looks++;
}
const wall = Number(process.hrtime.bigint() - t0) / 1e6;
const r = NB.report, ref = sim.ref;
console.log(JSON.stringify({
  particles: sim.N, steps, every, looks,
  with_analysis: sim.N * steps / (wall / 1e3), steps_only: sim.N * steps / (stepMs / 1e3),
  ms_per_step: stepMs / steps, ms_per_look: lookMs / looks,
  look_phases_ms: Object.fromEntries(Object.entries(phases).map(([k, v]) => [k, v / looks])),
  books: { dL: (r.L - ref.L) / ref.L, balance: (r.KE + r.PE + r.EL + r.Q - ref.EQ) / ref.scale, drift: r.drift }
}));
NB.destroy();
