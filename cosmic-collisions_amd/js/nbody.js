// The engine object of cosmic-collisions on MI355X: a drop-in for the
// reference's CC.nbody (src/nbody.js:1095-1118) whose passes run as gfx950
// HIP kernels behind the C ABI of libcc_hip.so, reached through the N-API
// addon cc_addon.node.  Body setup, the step schedule, the approach on
// rails, the softening and the analysis with its books are host code here,
// as in the reference; what the reference drew with WebGL2 is one addon call.
//
// Units: G = M_earth = R_earth = 1; time unit 805 s, velocity unit 7.91 km/s.
'use strict';
const path = require('path');
const bodies = require('./bodies');

const T_UNIT = 805, V_UNIT = 7.91, E_KG = 6.256e7, CP_ROCK = 1200, CP_IRON = 450, L_EM = 0.1163;
const JAM = 0.96;
const FL = 6.0, FH = 2 * FL / 64;
const PP_CAP = 8192, PP_CUT = 5 * FH, SEAT_SLACK = 0.6;
const N_CHOICES = [16384, 32768, 65536, 131072, 262144];
const RHO_R = 2.5;                                     // core over mantle density

function meanK(r) { return r.Q / r.M * E_KG / (CP_ROCK + r.iron * (CP_IRON - CP_ROCK)); }

// The mesh's own pair force x r^2 and potential x r by separation (src/nbody.js:56-109):
// two unit masses deposited on the 64^3 grid, the near field between node
// centres of mass (softening 0.5 FH^2, own node skipped), gathered at the
// first, less the first alone; 80 seeded placings per bin.
function pairTable(bins, cut) {
  const rand = bodies.seeded(777), eps2 = 0.5 * FH * FH, per = 80;
  const f = new Float32Array(bins), pt = new Float32Array(bins);
  const corners = (p) => {
    const out = [];
    const g = [(p[0] + FL) / FH - 0.5, (p[1] + FL) / FH - 0.5, (p[2] + FL) / FH - 0.5];
    const fl = g.map(Math.floor), t = [g[0] - fl[0], g[1] - fl[1], g[2] - fl[2]];
    for (let k = 0; k < 8; k++) {
      const ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
      out.push([(fl[0] + ox) + ',' + (fl[1] + oy) + ',' + (fl[2] + oz),
                (ox ? t[0] : 1 - t[0]) * (oy ? t[1] : 1 - t[1]) * (oz ? t[2] : 1 - t[2])]);
    }
    return out;
  };
  const field = (parts, at) => {
    const nodes = new Map();
    for (const p of parts) for (const [key, w] of corners(p)) {
      let nd = nodes.get(key);
      if (!nd) { nd = { m: 0, x: 0, y: 0, z: 0, fx: 0, fy: 0, fz: 0, pot: 0 }; nodes.set(key, nd); }
      nd.m += w; nd.x += w * p[0]; nd.y += w * p[1]; nd.z += w * p[2];
    }
    const list = Array.from(nodes.values());
    for (const nd of list) { nd.x /= nd.m; nd.y /= nd.m; nd.z /= nd.m; }
    for (const c of list) for (const o of list) {
      if (o === c) continue;
      const rx = o.x - c.x, ry = o.y - c.y, rz = o.z - c.z;
      const inv = 1 / Math.sqrt(rx * rx + ry * ry + rz * rz + eps2), inv3 = inv * inv * inv;
      c.fx += rx * o.m * inv3; c.fy += ry * o.m * inv3; c.fz += rz * o.m * inv3; c.pot -= o.m * inv;
    }
    const out = [0, 0, 0, 0];
    for (const [key, w] of corners(at)) {
      const nd = nodes.get(key);
      if (!nd) continue;
      out[0] += w * nd.fx; out[1] += w * nd.fy; out[2] += w * nd.fz; out[3] += w * nd.pot;
    }
    return out;
  };
  for (let b = 0; b < bins; b++) {
    const r = (b + 0.5) / bins * cut;
    let sf = 0, sp = 0;
    for (let s = 0; s < per; s++) {
      const p = [(rand() - 0.5) * FH, (rand() - 0.5) * FH, (rand() - 0.5) * FH];
      const z = 2 * rand() - 1, ph = rand() * 6.2831853, rr = Math.sqrt(1 - z * z), u = [rr * Math.cos(ph), rr * Math.sin(ph), z];
      const q = [p[0] + u[0] * r, p[1] + u[1] * r, p[2] + u[2] * r];
      const both = field([p, q], p), alone = field([p], p);
      sf += (both[0] - alone[0]) * u[0] + (both[1] - alone[1]) * u[1] + (both[2] - alone[2]) * u[2];
      sp += both[3] - alone[3];
    }
    f[b] = sf / per * r * r; pt[b] = sp / per * r;
  }
  const tab = new Float32Array(bins * 2);
  for (let i = 0; i < bins; i++) { tab[i * 2] = f[i]; tab[i * 2 + 1] = pt[i]; }
  return tab;
}

// The run from build(opt): two onion bodies, their layers and masses, the
// encounter walked back to the start line, the contact spring, step and
// dashpot (src/nbody.js:114-250).  Host float64 arithmetic only; the result
// is the initial Float32 state and the sim object.
// Batched sweeps (no reference counterpart) add two options: opt.N, a
// particle count outside N_CHOICES (a multiple of 512, at most 262144: 512
// wide), and opt.seed, the bodies' own seed (the reference fixes 1234567).
// native: the addon, whose relaxPiles relaxes the bodies side by side (the
// same bits as relaxing them here in turn).  buildRuns lays out every run's
// bodies first and relaxes them all in one call (the relaxation draws no
// random numbers, so the runs are those buildRun makes one at a time).
function buildRun(opt, native) {
  return buildRuns([opt], native)[0];
}
function buildRuns(opts, native) {
  const laid = opts.map(layRun);
  const reports = bodies.relaxPiles([].concat(...laid.map((l) => l.piles)), native);
  return laid.map((l, k) => l.finish(reports.slice(2 * k, 2 * k + 2)));
}
function layRun(opt) {
  const N = opt.N > 0 ? opt.N : N_CHOICES[opt.nIdx];
  if (opt.N > 0 && (N % 512 !== 0 || N > 262144 || N < 1024)) throw new Error('N must be a multiple of 512 in [1024, 262144]');
  const W = opt.N > 0 ? 512 : N >= 131072 ? 512 : N >= 32768 ? 256 : 128, H = N / W;
  const q = opt.q, dens = opt.dens, coreF = opt.core;
  const coreV = coreF / (RHO_R - coreF * (RHO_R - 1));
  const coreF2 = opt.coreImp === undefined ? coreF : opt.coreImp;
  const coreV2 = coreF2 / (RHO_R - coreF2 * (RHO_R - 1));
  const V1 = 1, V2 = q / dens;
  const vp = (V1 + V2) / N;
  const s = Math.cbrt(16 * Math.PI * vp / 3);
  const a = JAM * s / (2 * Math.SQRT2);
  let n1 = Math.round(N * V1 / (V1 + V2)), n2 = N - n1;
  if (opt.n2 > 0) { n2 = opt.n2; n1 = N - n2; } else if (n2 < 16) { n2 = 16; n1 = N - n2; }
  const R1 = Math.cbrt(V1), R2 = Math.cbrt(V2), Rsum = R1 + R2;
  const w0 = [0, opt.spin0, 0], w1 = [0, opt.spin1 * Math.sqrt(dens), 0];
  const e1 = bodies.spheroidEcc(Math.abs(opt.spin0)), e2 = bodies.spheroidEcc(Math.abs(opt.spin1));
  const Re1 = R1 / Math.pow(1 - e1 * e1, 1 / 6), Re2 = R2 / Math.pow(1 - e2 * e2, 1 / 6);
  const rand = bodies.seeded(opt.seed > 0 ? opt.seed : 1234567);
  const b1 = bodies.onionLay(n1, s, a, rand, e1), b2 = bodies.onionLay(n2, s, a, rand, e2);
  return { piles: [b1.pile, b2.pile], finish: (reports) => {
    [b1.relax, b2.relax] = reports;
    const pos = new Float32Array(N * 4), vel = new Float32Array(N * 4), matR = new Float32Array(N), lay = new Uint8Array(N);
    // layers by whole shells: the core the shells nearest its volume share, the crust the outer shell
    const layer = (off, n, firsts, scale, cv) => {
      const nCore = firsts.reduce((best, c) => (Math.abs(c - cv * n) < Math.abs(best - cv * n) ? c : best), 0);
      const nCrust = n - firsts[firsts.length - 2];
      let sum = 0;
      for (let j = 0; j < n; j++) {
        const L = j < nCore ? 0 : j >= n - nCrust ? 2 : 1;
        lay[off + j] = L;
        matR[off + j] = scale * (L === 0 ? RHO_R : 1);
        sum += matR[off + j];
      }
      return sum;
    };
    const sum1 = layer(0, n1, b1.bounds, 1, coreV);
    let sum2 = layer(n1, n2, b2.bounds, dens, coreV2);
    const mref = 1 / sum1;
    if (opt.mImp > 0) {
      const k2 = opt.mImp / (mref * sum2);
      for (let j = n1; j < N; j++) matR[j] *= k2;
      sum2 *= k2;
    }
    const M1 = 1, M2 = mref * sum2, M = M1 + M2;
    // the encounter, stated at contact and walked back to the start line 3.7 radii off
    const vesc = Math.sqrt(2 * M / Rsum), vc = opt.vfac * vesc;
    const bc = Rsum * Math.sin(opt.angle * Math.PI / 180);
    const D = Rsum + 3.7;
    const vD = Math.sqrt(Math.max(vc * vc - 2 * M * (1 / Rsum - 1 / D), 0.04));
    const bD = Math.min(bc * vc / vD, D * 0.95);
    const xD = -Math.sqrt(D * D - bD * bD);
    const inc = (opt.incl || 0) * Math.PI / 180, bY = bD * Math.sin(inc), bZ = bD * Math.cos(inc);
    const c1 = [-xD * M2 / M, -bY * M2 / M, -bZ * M2 / M];
    const c2 = [xD + c1[0], bY + c1[1], bZ + c1[2]];
    const rel = { r: [xD, bY, bZ], v: [vD, 0, 0] };
    for (let i = 0; i < n1; i++) { const o = i * 4; pos[o] = b1[i * 3] + c1[0]; pos[o + 1] = b1[i * 3 + 1] + c1[1]; pos[o + 2] = b1[i * 3 + 2] + c1[2]; pos[o + 3] = lay[i]; }
    for (let i = 0; i < n2; i++) { const o = (n1 + i) * 4; pos[o] = b2[i * 3] + c2[0]; pos[o + 1] = b2[i * 3 + 1] + c2[1]; pos[o + 2] = b2[i * 3 + 2] + c2[2]; pos[o + 3] = lay[n1 + i] + 4; }
    const home = new Float32Array(N * 4), hrand = bodies.seeded(99);
    for (let i = 0; i < n1; i++) { const o = i * 4; home[o] = b1[i * 3]; home[o + 1] = b1[i * 3 + 1]; home[o + 2] = b1[i * 3 + 2]; home[o + 3] = hrand(); }
    for (let i = 0; i < n2; i++) { const o = (n1 + i) * 4; home[o] = b2[i * 3] + 1.5; home[o + 1] = b2[i * 3 + 1] + 0.7; home[o + 2] = b2[i * 3 + 2] - 1.1; home[o + 3] = hrand(); }
    // contact spring: half a radius of overlap at the reference speed; the step from its period
    const vref = Math.max(1.5, vc);
    const k = 4 * mref * vref * vref / (a * a);
    const dt = 0.35 / Math.sqrt(2 * k / (mref * Math.min(1, dens)));
    const lnE = Math.log(opt.rest), zeta = -lnE / Math.sqrt(Math.PI * Math.PI + lnE * lnE);
    const cd = 2 * zeta * Math.sqrt(k * mref / 2);
    const settle = Math.ceil(0.6 / dt);
    const sim = {
      N, W, H, G: N >= 65536 ? 128 : 64, n1, n2, m: mref, a, cell: 2.6 * a, dt, k, c: cd, creep: 0.12,
      R1, R2, Re1, Re2, M1, M2, M, rel, bulk0: [0, 0, 0], bulk1: [0, 0, 0], vc, vesc,
      spin0: w0, spin1: w1, cen0: c1, cen1: c2, matR, homeArr: home, ref: null,
      t: -approachTime(rel, M, Re1 + Re2 + 2 * a), phase: 'settle', settleLeft: settle, settleTotal: settle,
      impactT: -1, com1: [0, 0, 0], stepNo: 0, damp: Math.exp(-dt * 12), soft: false, gen: 0,
      boxes: [{ on: true, cen: [0, 0, 0], vel: [0, 0, 0], t: 0 }, { on: false, cen: [0, 0, 0], vel: [0, 0, 0], t: 0 }],
      tK: 0.5 * E_KG / CP_ROCK / mref, kappa: 0.3
    };
    return { sim, pos, vel };
  } };
}

function createEngine(opts) {
  opts = opts || {};
  const addon = require(opts.addon || path.join(__dirname, 'cc_addon.node'));
  let ctx = null, sim = null, report = null, MESH_EVERY = 8, analyzedAt = 0, analyses = 0, gen = 0, lookMs = null;
  const PP_TAB = pairTable(64, PP_CUT);
  const pp = { count: 0, loose: 0, list: new Int32Array(0) };
  const taps = { sim: [], report: [] };
  const emit = (kind, v) => { for (const f of taps[kind]) f(v); };

  // the C ABI's constants of a run: the sim fields and what setStaticUniforms derives
  function params() {
    return { N: sim.N, W: sim.W, H: sim.H, G: sim.G, m: sim.m, a: sim.a, cell: sim.cell, dt: sim.dt, k: sim.k, c: sim.c,
             creep: sim.creep, damp: sim.damp, tK: sim.tK, kappa: sim.kappa, cpIron: CP_IRON / CP_ROCK,
             seatSlack: SEAT_SLACK, spin0: sim.spin0, spin1: sim.spin1, meshEvery: MESH_EVERY,
             field: sim.field ? 1 : 0, gField: sim.gField || [0, 0, 0], frozenOn: sim.frozen === undefined ? 0 : 1,
             frozen: sim.frozen || 0 };
  }

  // build(): the run's bodies and constants (buildRun), uploaded to a fresh context
  // run: one buildRuns made already for this opt
  function build(opt, run) {
    run = run || buildRun(opt, addon);
    sim = run.sim;
    sim.gen = ++gen;
    pending = null;                // a look at the last run is no longer wanted
    if (ctx) addon.destroy(ctx);   // a rebuild frees the last run's device objects (src/nbody.js:208-245)
    ctx = addon.create(params());
    addon.uploadState(ctx, run.pos, run.vel, sim.matR, new Float32Array(sim.N), sim.N);
    addon.setPPTable(ctx, PP_TAB);
    sim.initial = { pos: run.pos, vel: run.vel };
    pp.count = 0; pp.loose = 0; pp.list = new Int32Array(0);
    report = null;
    analyzedAt = 0;
    emit('sim', sim); emit('report', null);
    return sim;
  }

  // where the boxes stand this mesh step (src/nbody.js:342-350)
  function boxes() {
    const b = sim.boxes, L = [];
    for (let i = 0; i < 2; i++) { const dt = sim.t - b[i].t; for (let k = 0; k < 3; k++) L.push(FL - b[i].cen[k] - b[i].vel[k] * dt); }
    addon.setBoxes(ctx, L.slice(0, 3), L.slice(3, 6), b[1].on);
    return L;
  }
  // the boxes from a look (src/nbody.js:361-374)
  function placeBoxes(r) {
    const b = sim.boxes, sec = r.second, B = b[1];
    b[0] = { on: true, cen: r.core, vel: r.bodyVel, t: r.t };
    if (sec && sec.mass > 0.05) {
      const d = Math.max(Math.abs(sec.core[0] - r.core[0]), Math.abs(sec.core[1] - r.core[1]), Math.abs(sec.core[2] - r.core[2]));
      const far = FL - 1.3 * Math.cbrt(sec.mass);
      if (!B.on && d > far) {
        B.on = true;
        const w = sec.mass / sim.m, c = sec.core;
        addon.setMono(ctx, 1, new Float32Array([c[0] * w, c[1] * w, c[2] * w, w]));
      } else if (B.on && d < far - 0.5) B.on = false;
      B.cen = sec.core; B.vel = sec.vel; B.t = r.t;
    } else B.on = false;
  }

  function simArgs(settle, kick, gdt) {
    return { settle: !!settle, kick: !!kick, gdt: gdt || 0, bulk0: sim.bulk0, bulk1: sim.bulk1, cen0: sim.cen0, cen1: sim.cen1 };
  }
  function soften() {
    sim.soft = true;
    sim.k *= 0.5; sim.c *= Math.SQRT1_2; sim.dt *= Math.SQRT2;
    sim.damp = Math.exp(-sim.dt * 12);
    addon.setParams(ctx, params());
  }
  // one step (src/nbody.js:551-558); the caller advances sim.t
  function step(settle, kick) {
    if (!settle && !kick && !sim.soft && sim.t - sim.impactT > 1.5) soften();
    addon.setStepNo(ctx, sim.stepNo);
    if (sim.stepNo % MESH_EVERY === 0) boxes();
    addon.step(ctx, simArgs(settle, kick));
    sim.stepNo++;
  }
  // n full steps at once, sim.t advanced here; boxes() for each mesh step in order
  function stepMany(n) {
    if (!sim.soft && sim.t - sim.impactT > 1.5) soften();
    const L = [];
    let t = sim.t;
    for (let i = 0; i < n; i++) {
      if ((sim.stepNo + i) % MESH_EVERY === 0) {
        const b = sim.boxes;
        for (let j = 0; j < 2; j++) { const dt = t - b[j].t; for (let k = 0; k < 3; k++) L.push(FL - b[j].cen[k] - b[j].vel[k] * dt); }
        L.push(b[1].on ? 1 : 0);
      }
      t += sim.dt;
    }
    addon.setStepNo(ctx, sim.stepNo);
    const ms = addon.stepN(ctx, n, simArgs(false, false), L.length ? new Float64Array(L) : null);
    for (let i = 0; i < n; i++) sim.t += sim.dt;
    sim.stepNo += n;
    return ms;
  }
  function settleSome(n) {
    n = Math.min(n, sim.settleLeft);
    for (let i = 0; i < n; i++) step(true, false);
    sim.settleLeft -= n;
    if (sim.settleLeft === 0) { sim.phase = 'approach'; analyzedAt = sim.stepNo; }
  }
  function centres(r) {
    const M = sim.M, M1 = sim.M1, M2 = sim.M2;
    return [[-r[0] * M2 / M, -r[1] * M2 / M, -r[2] * M2 / M], [r[0] * M1 / M, r[1] * M1 / M, r[2] * M1 / M]];
  }
  // the approach: leapfrog of the two-body problem in 8 substeps, the bodies carried rigid (src/nbody.js:591-636)
  function rigidStep(dtF) {
    const r = sim.rel.r, v = sim.rel.v, M = sim.M, M1 = sim.M1, M2 = sim.M2;
    const r0 = [r[0], r[1], r[2]], n = 8, h = dtF / n;
    let cen = centres(r0);
    const acc = () => { const d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], f = -M * Math.pow(d2, -1.5); return [r[0] * f, r[1] * f, r[2] * f]; };
    let a = acc();
    for (let i = 0; i < n; i++) {
      v[0] += a[0] * h * 0.5; v[1] += a[1] * h * 0.5; v[2] += a[2] * h * 0.5;
      r[0] += v[0] * h; r[1] += v[1] * h; r[2] += v[2] * h;
      a = acc();
      v[0] += a[0] * h * 0.5; v[1] += a[1] * h * 0.5; v[2] += a[2] * h * 0.5;
    }
    const mv = [(r[0] - r0[0]) / dtF, (r[1] - r0[1]) / dtF, (r[2] - r0[2]) / dtF];
    sim.bulk0 = [-mv[0] * M2 / M, -mv[1] * M2 / M, -mv[2] * M2 / M];
    sim.bulk1 = [mv[0] * M1 / M, mv[1] * M1 / M, mv[2] * M1 / M];
    addon.rigid(ctx, sim.bulk0, sim.bulk1, cen[0], cen[1], bodies.rotation(sim.spin0, dtF), bodies.rotation(sim.spin1, dtF), dtF);
    sim.t += dtF;
    if (Math.hypot(r[0], r[1], r[2]) <= sim.Re1 + sim.Re2 + 2 * sim.a) {
      sim.bulk0 = [-v[0] * M2 / M, -v[1] * M2 / M, -v[2] * M2 / M];
      sim.bulk1 = [v[0] * M1 / M, v[1] * M1 / M, v[2] * M1 / M];
      cen = centres(r); sim.cen0 = cen[0]; sim.cen1 = cen[1];
      sim.boxes[0] = { on: true, cen: cen[0], vel: sim.bulk0, t: sim.t };
      sim.boxes[1].on = false;
      sim.stepNo = 0;
      step(false, true);
      sim.t += sim.dt;
      sim.phase = 'full';
      sim.impactT = sim.t;
    }
  }
  function readState() {
    const N = sim.N, st = { pos: new Float32Array(N * 4), vel: new Float32Array(N * 4), aux: new Float32Array(N * 4),
                            q: new Float32Array(N), mat: new Float32Array(N) };
    addon.downloadState(ctx, st.pos, st.vel, st.aux, st.q, st.mat, N);
    return st;
  }
  // the read-outs for the books from seats, mesh and pairwise pass rebuilt now (src/nbody.js:488-507);
  // read back to the host unless `keep` (the analysis reads them on the device)
  function passDiag(keep) {
    boxes();
    addon.slots(ctx);
    addon.mesh(ctx);
    addon.pp(ctx);
    if (keep) { addon.diag(ctx, null); return null; }
    const D = new Float32Array(sim.N * 4);
    addon.diag(ctx, D);
    return D;
  }
  function setPP(list, loose) {
    pp.list = list; pp.count = list.length; pp.loose = loose;
    addon.setPP(ctx, list, list.length);
  }
  // a report into the run (applyReport, src/nbody.js:912-928): boxes, the books' reference, every eighth
  // look in the full phase the Morton order -- the GPU's, from the report's core, on the state as it is now --
  // and the loose list
  function applyReport(r, sort) {
    const t2 = Number(process.hrtime.bigint()) / 1e6;
    report = r; sim.com1 = r.com;
    placeBoxes(r);
    const going = sim.phase === 'full' || sim.bulk0[0] !== 0 || sim.bulk0[1] !== 0 || sim.bulk0[2] !== 0;
    if (!sim.ref && going) sim.ref = { E: r.KE + r.PE + r.EL, EQ: r.KE + r.PE + r.EL + r.Q, L: r.L, scale: r.KE - r.PE };
    let list = r.pp.list;
    if (sort) {
      const perm = new Int32Array(sim.N), into = new Int32Array(sim.N);
      addon.morton(ctx, r.core, perm, into);
      const R = new Float32Array(sim.N), H = new Float32Array(sim.N * 4);
      for (let i = 0; i < sim.N; i++) {
        R[i] = sim.matR[perm[i]];
        for (let c = 0; c < 4; c++) H[i * 4 + c] = sim.homeArr[perm[i] * 4 + c];
      }
      sim.matR = R; sim.homeArr = H;
      list = list.map((i) => into[i]);
      r.perm = perm;
    }
    setPP(list, r.pp.loose);
    emit('report', r);
    return Number(process.hrtime.bigint()) / 1e6 - t2;
  }
  // the analysis (src/nbody.js:967-1008): groups, books, boxes, the loose list.  sync: at once, the
  // report returned.  Otherwise (and asyncRead, the default) the device's half and the read-back now,
  // the host's sums on a thread of the library while the caller steps on; collect() applies the
  // report when they are done, one report behind, as the reference's worker does.  A call while one
  // is pending does nothing, as the reference's does.
  let pending = null, asyncRead = true;
  function analyze(sync) {
    if (pending && !sync) return null;
    if (sync) { sim.gen = ++gen; pending = null; }   // a synchronous look drops the pending one (the library does too)
    analyzedAt = sim.stepNo;
    const sort = sim.phase === 'full' && analyses++ % 8 === 0;
    const ms = () => Number(process.hrtime.bigint()) / 1e6;
    const t0 = ms();
    passDiag(true);
    const t1 = ms();
    const args = { m: sim.m, a: sim.a, spring: sim.k, touch: 2 * sim.a, cap: PP_CAP };
    if (!sync && asyncRead) {
      addon.analyseBegin(ctx, args);
      pending = { gen: sim.gen, sort, t: sim.t, diag: t1 - t0, begin: ms() - t1 };
      return null;
    }
    const r = addon.analyse(ctx, args);
    const t2 = ms();
    r.t = sim.t;
    const rest = applyReport(r, sort);
    const am = r.analysis.ms;
    lookMs = { diag: t1 - t0, analyse: t2 - t1, cells: am[0], copy: am[1], host: am[2], rest, sort };
    return r;
  }
  // the pending analysis' report applied if its sums are done (wait: wait for them); null otherwise
  function collect(wait) {
    if (!pending) return null;
    const t0 = Number(process.hrtime.bigint()) / 1e6;
    const r = addon.analyseEnd(ctx, !!wait);
    if (!r) return null;
    const p = pending;
    pending = null;
    if (p.gen !== sim.gen) return null;   // the run was rebuilt, or looked at synchronously, since
    r.t = p.t;
    const t1 = Number(process.hrtime.bigint()) / 1e6;
    const rest = applyReport(r, p.sort);
    const am = r.analysis.ms;
    lookMs = { diag: p.diag, analyse: p.begin + (t1 - t0), cells: am[0], copy: am[1], host: am[2], rest, sort: p.sort,
               overlapped: true };
    return r;
  }

  return {
    init() { return this; },
    // the device objects freed now rather than when the handle is collected
    destroy() { if (ctx) { addon.destroy(ctx); ctx = null; } },
    build, step, stepMany, settleSome, rigidStep, soften,
    analyze, collect, readState,
    sortNow() { analyses = 0; return analyze(true); },
    passes: {
      slots: () => addon.slots(ctx), deposit: (b) => addon.deposit(ctx, b || 0), coarse: (b) => addon.coarse(ctx, b || 0),
      cell: (b) => addon.cell(ctx, b || 0), mesh: () => addon.mesh(ctx), pp: () => addon.pp(ctx),
      sim: (settle, kick, gdt) => addon.sim(ctx, simArgs(settle, kick, gdt)), diag: passDiag
    },
    onSim(f) { taps.sim.push(f); }, onReport(f) { taps.report.push(f); },
    get sim() { return sim; }, get report() { return report; }, get pending() { return pending; },
    get analyzedAt() { return analyzedAt; }, get pp() { return pp; }, get ctx() { return ctx; }, get addon() { return addon; },
    // the last analysis's phases, ms: seats+mesh+pairwise+DIAG; cc_analyse, of it the device's cells and
    // groups, the read-back, the host's sums; the rest (boxes, Morton sort, loose list)
    get lookMs() { return lookMs; },
    get state() { return { analyzing: !!pending, pending: !!pending, worker: true, analyses, analyzedAt, stepNo: sim && sim.stepNo }; },
    set meshEvery(n) { MESH_EVERY = n; if (ctx) addon.setParams(ctx, params()); },
    set asyncRead(v) { asyncRead = !!v; },
    pairTable: PP_TAB,
    T_UNIT, V_UNIT, E_KG, CP_ROCK, CP_IRON, L_EM, JAM, PP_CAP, meanK
  };
}

// how long the rigid flight takes from the start line to contact (src/nbody.js:578-590)
function approachTime(rel, M, touch) {
  const r = rel.r.slice(), v = rel.v.slice(), h = 0.01 / 8;
  let t = 0;
  const acc = () => { const d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], f = -M * Math.pow(d2, -1.5); return [r[0] * f, r[1] * f, r[2] * f]; };
  let a = acc();
  while (Math.hypot(r[0], r[1], r[2]) > touch && t < 200) {
    v[0] += a[0] * h * 0.5; v[1] += a[1] * h * 0.5; v[2] += a[2] * h * 0.5;
    r[0] += v[0] * h; r[1] += v[1] * h; r[2] += v[2] * h;
    a = acc();
    v[0] += a[0] * h * 0.5; v[1] += a[1] * h * 0.5; v[2] += a[2] * h * 0.5;
    t += h;
  }
  return t;
}

module.exports = { createEngine, buildRun, buildRuns, pairTable, approachTime, N_CHOICES, FL, FH, PP_CAP, PP_CUT };
