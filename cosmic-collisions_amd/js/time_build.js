// Body setup timed on the host: the plain twins build (262,144 particles, two
// 0.5 M_earth bodies) with the bodies relaxed in JavaScript and by the
// library (both bodies side by side), and every body of the batched sweep's
// 64 instances relaxed in one library call.  Prints one JSON line.
//   node time_build.js '<twins opt json>' [count] [seed]
'use strict';
const path = require('path');
const nb = require('./nbody');
const bodies = require('./bodies');
const addon = require(path.join(__dirname, 'cc_addon.node'));

const opt = JSON.parse(process.argv[2]);
const count = parseInt(process.argv[3] || '64', 10);
const rand = bodies.seeded(parseInt(process.argv[4] || '20261203', 10));
const ms = (f) => { const t0 = process.hrtime.bigint(); const r = f(); return [Number(process.hrtime.bigint() - t0) / 1e6, r]; };
const [jsMs, a] = ms(() => nb.buildRun(opt));
const [natMs, b] = ms(() => nb.buildRun(opt, addon));
const same = Buffer.compare(Buffer.from(a.pos.buffer), Buffer.from(b.pos.buffer)) === 0;
const opts = [];
for (let k = 0; k < count; k++) {   // the sweep's instances as js/prep_batch.js draws them
  const lu = (lo, hi) => Math.exp(Math.log(lo) + rand() * (Math.log(hi) - Math.log(lo)));
  const N = Math.min(262144, Math.max(16384, Math.round(lu(16384, 262144) / 512) * 512));
  opts.push({ N, q: lu(0.02, 1.0), dens: 0.75, core: 0.2 + 0.2 * rand(), spin0: 0.18, spin1: 0.20,
              angle: 75 * rand(), incl: 0, vfac: 1.0 + 1.6 * rand(), rest: 0.3, seed: 1 + Math.floor(rand() * 2147483646) });
}
const [sweepMs] = ms(() => nb.buildRuns(opts, addon));
console.log(JSON.stringify({
  twins: { particles: a.sim.N, js_ms: jsMs, native_ms: natMs, identical: same },
  sweep: { instances: count, particles: opts.reduce((s, o) => s + o.N, 0), native_ms: sweepMs,
           threads: require('os').cpus().length }
}));
