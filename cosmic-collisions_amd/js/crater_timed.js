// The crater page's run (chicxulub.html:512-842) driven through the crater
// engine over the N-API addon: build, the settle steps, the impactor on rails
// to touchdown, the coupling pass, then `steps` full steps timed (each 8-step
// cycle one graph launch; the steps after the coupling pass are timed from
// its end), with the page's readout at the end.
// Prints one JSON line: particle-steps/s of the full steps, the settle steps'
// rate, and the readout's crater diameter and airborne mass.
//   node crater_timed.js '<build opt JSON>' <steps>
'use strict';
const path = require('path');
const { createCraterEngine } = require('./crater');
const addon = require(path.join(__dirname, 'cc_addon.node'));

const opt = JSON.parse(process.argv[2]);
const steps = parseInt(process.argv[3] || '2000', 10);
const eng = createCraterEngine();
const now = () => Number(process.hrtime.bigint()) / 1e6;
let t = now();
const sim = eng.build(opt);
const buildMs = now() - t;
const settle = sim.settleLeft;
t = now();
while (sim.phase === 'settle') eng.settleSome(64);
addon.sync(eng.ctx);
const settleMs = now() - t;
let frames = 0;
while (sim.phase !== 'full' && frames < 1000) { eng.approachStep(); frames++; }
eng.stepMany(8);   // the cycle graphs captured before the clock
addon.sync(eng.ctx);
t = now();
let done = 0;
while (done < steps) { const n = Math.min(1000, steps - done); eng.stepMany(n); done += n; }
addon.sync(eng.ctx);
const wall = now() - t;
const rep = eng.analyze();
console.log(JSON.stringify({
  particles: sim.N, steps, ms_per_step: wall / steps, value: sim.N * steps / (wall / 1e3),
  settle_steps: settle, settle_value: settle > 0 ? sim.N * settle / (settleMs / 1e3) : null, build_ms: buildMs,
  approach_frames: frames, crater_D: rep.craterD, airborne: rep.airborne, melt: rep.melt
}));
eng.destroy();
