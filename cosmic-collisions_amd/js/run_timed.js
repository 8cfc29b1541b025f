// The engine driven as the reference's page drives a run (index.html:592-600,
// __impact.runTo): build, settle, approach on rails, the kick at first touch,
// then full steps with an analysis every `every` steps (the groups, books,
// boxes, loose list and, every eighth look, the Morton sort).
//   mode sync     each analysis synchronous, as runTo calls analyze(true)
//   mode overlap  each analysis begun (analyze(false)), its host sums running
//                 while the next `lag` steps are taken, then collected -- the
//                 page's asynchronous look with the worker's one-report lag
//                 (src/nbody.js:895-1025)
// Prints one JSON line: particle-steps/s with and without the analyses' share
// of the time, and the step time in each window between looks (the run's
// growth from first touch).
//   node run_timed.js '<build opt JSON>' <steps> [every] [mode] [lag]
'use strict';
const { createEngine } = require('./nbody');

const opt = JSON.parse(process.argv[2]);
const steps = parseInt(process.argv[3] || '3600', 10);
const every = parseInt(process.argv[4] || '360', 10);
const mode = process.argv[5] || 'sync';
const lag = parseInt(process.argv[6] || '64', 10);
const NB = createEngine();
const sim = NB.build(opt);
while (sim.phase === 'settle') NB.settleSome(64);
while (sim.phase === 'approach') NB.rigidStep(0.01);
NB.analyze(true);
const now = () => Number(process.hrtime.bigint()) / 1e6;
let stepMs = 0, lookMs = 0, done = 0, looks = 0;
const windows = [];
const phases = { diag: 0, analyse: 0, cells: 0, copy: 0, host: 0, rest: 0 };
const addPhases = () => { for (const k of Object.keys(phases)) phases[k] += NB.lookMs[k]; };
const t0 = now();
while (done < steps) {
  const n = Math.min(every, steps - done);
  let ms = 0;
  if (mode === 'overlap' && NB.pending) {
    const k = Math.min(lag, n);
    let a = now();
    NB.stepMany(k);
    ms += now() - a;
    a = now();
    if (NB.collect(true)) addPhases();
    lookMs += now() - a;
    a = now();
    if (n > k) NB.stepMany(n - k);
    ms += now() - a;
  } else {
    const a = now();
    NB.stepMany(n);
    ms += now() - a;
  }
  stepMs += ms;
  windows.push(+(ms / n).toFixed(5));
  done += n;
  const a = now();
  if (mode === 'overlap') NB.analyze(false);
  else { NB.analyze(true); addPhases(); }
  lookMs += now() - a;
  looks++;
}
if (mode === 'overlap') { const a = now(); if (NB.collect(true)) addPhases(); lookMs += now() - a; }
const wall = now() - t0;
const r = NB.report, ref = sim.ref;
console.log(JSON.stringify({
  particles: sim.N, steps, every, looks, mode, lag: mode === 'overlap' ? lag : 0,
  with_analysis: sim.N * steps / (wall / 1e3), steps_only: sim.N * steps / (stepMs / 1e3),
  ms_per_step: stepMs / steps, ms_per_look: lookMs / looks,
  look_phases_ms: Object.fromEntries(Object.entries(phases).map(([k, v]) => [k, v / looks])),
  window_ms_per_step: windows,
  books: { dL: (r.L - ref.L) / ref.L, balance: (r.KE + r.PE + r.EL + r.Q - ref.EQ) / ref.scale, drift: r.drift }
}));
NB.destroy();
