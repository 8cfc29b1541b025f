// The crater engine: a patch of ground under constant gravity, hit by an
// impactor carried in on rails and coupled to the ground in one pass -- the
// reference's Chicxulub crater view (chicxulub.html:156-842) on the same C ABI
// as the giant impact.  No mesh: gravity is the field g, and the contact
// grid, contacts, heat and integration are the same kernels.
//
// Units: length 10 km, time 10 s, velocity 1 km/s.
'use strict';
const path = require('path');
const bodies = require('./bodies');

const L_UNIT = 1e4, T_UNIT = 10, V_UNIT = 1000, G_EARTH = 9.81, E_KG = V_UNIT * V_UNIT, CP_ROCK = 1200, JAM = 0.96;
const BOX_R = 2.85, BOX_D = 1.55, BOX_FROZEN = 0.26;
const MAT_FROZEN = 0, MAT_BASEMENT = 1, MAT_PLATFORM = 2, MAT_MELT = 3, MAT_IMPACTOR = 4;
const HH_C1 = 0.55, HH_MU = 0.55, HH_P = 0.3, Z_MODEL = 2.7;
const N_CHOICES = [16384, 32768, 65536, 131072, 262144];
const APPROACH_FRAMES = 75;
const PRESETS = {
  chicxulub: { imp: 14, vimp: 20, angle: 60 }, vertical: { imp: 14, vimp: 20, angle: 90 },
  grazing: { imp: 14, vimp: 20, angle: 30 }, small: { imp: 1, vimp: 20, angle: 60 }, big: { imp: 30, vimp: 20, angle: 60 }
};
const DEFAULTS = { imp: 14, vimp: 20, angle: 60, sed: 3, dens: 2500, nIdx: 3, rest: 0.25 };

// Collins, Melosh & Marcus (2005): the gravity-regime transient crater diameter, m
function transientCrater(L, v, rhoI, rhoT, g, thetaDeg) {
  return 1.161 * Math.pow(rhoI / rhoT, 1 / 3) * Math.pow(L, 0.78) * Math.pow(v, 0.44) *
         Math.pow(g, -0.22) * Math.pow(Math.sin(thetaDeg * Math.PI / 180), 1 / 3);
}
const D_SIMPLE = 3200;
function finalCrater(Dtc) { return Dtc < D_SIMPLE ? 1.25 * Dtc : 1.17 * Math.pow(Dtc, 1.13) / Math.pow(D_SIMPLE, 0.13); }
function seismicMagnitude(E) { return 0.67 * Math.log10(E) - 5.87; }

// A jittered face-centred lattice cut to a cylinder of radius R and depth D,
// relaxed inside its walls, floor and flat top (chicxulub.html:512-550); the
// spacing is searched for, sites are dropped at random to leave nWant.
// latticeLay stops short of the relaxation (which draws no random numbers) and
// returns it as .pile, so the ground and the impactor can be relaxed together.
function latticeFill(nWant, R, D, rand) {
  const lat = latticeLay(nWant, R, D, rand);
  lat.relax = bodies.relaxPiles([lat.pile])[0];
  return lat;
}
function latticeLay(nWant, R, D, rand) {
  let s = Math.cbrt(4 * Math.PI * R * R * D / nWant), pts = null;
  for (let attempt = 0; attempt < 8; attempt++) {
    pts = [];
    const h = s * 0.5, R2 = R * R, nx = Math.ceil(R / h);
    for (let i = -nx; i <= nx; i++) for (let k = -nx; k <= nx; k++) {
      const x = i * h, z = k * h;
      if (x * x + z * z > R2) continue;
      for (let j = 0; j * h <= D; j++) {
        if (((i + j + k) & 1) !== 0) continue;
        pts.push(x, -j * h, z);
      }
    }
    if (pts.length / 3 >= nWant) break;
    s *= Math.pow(pts.length / 3 / nWant, 1 / 3) * 0.99;
  }
  const have = pts.length / 3, keep = new Int32Array(have);
  for (let q = 0; q < have; q++) keep[q] = q;
  for (let q = have - 1; q > 0; q--) { const t = (rand() * (q + 1)) | 0, tmp = keep[q]; keep[q] = keep[t]; keep[t] = tmp; }
  const out = new Float32Array(nWant * 3), jit = 0.34 * s;
  for (let q = 0; q < nWant; q++) {
    const o = keep[q] * 3;
    out[q * 3] = pts[o] + (rand() - 0.5) * jit;
    out[q * 3 + 1] = pts[o + 1] + (rand() - 0.5) * jit;
    out[q * 3 + 2] = pts[o + 2] + (rand() - 0.5) * jit;
  }
  const a = JAM * s / (2 * Math.SQRT2);
  return { p: out, s, a, pile: { p: out, n: nWant, a, A: R, C: D, confine: bodies.GROUND, sweeps: 16 } };
}

// The crater run (chicxulub.html:567-674): ground, impactor, and the run's constants
// S.N (no reference counterpart) asks for a grain count outside N_CHOICES: a
// multiple of 512 up to 524,288, 512 wide (the seat words then carry a
// 19-bit index above 262,144).
function buildCraterRun(S, native) {
  const N = S.N > 0 ? S.N : N_CHOICES[S.nIdx];
  if (S.N > 0 && (N % 512 !== 0 || N > 524288 || N < 16384)) throw new Error('N must be a multiple of 512 in [16384, 524288]');
  const W = S.N > 0 || N >= 131072 ? 512 : N >= 32768 ? 256 : 128, H = N / W;
  const rand = bodies.seeded(20260827);
  const rhoT = S.dens, rhoI = 2630;
  const dImp = S.imp * 1000 / L_UNIT, aImp = dImp * 0.5, vImp = S.vimp * 1000 / V_UNIT, theta = S.angle;
  const Ltc = transientCrater(S.imp * 1000, S.vimp * 1000, rhoI, rhoT, G_EARTH, theta);
  const Rtc = 0.5 * Ltc / L_UNIT;
  const Rbox = BOX_R * Rtc, Dbox = BOX_D * Rtc, froz = BOX_FROZEN * Rtc;
  const vBox = Math.PI * Rbox * Rbox * Dbox;
  const vImpactor = 4 / 3 * Math.PI * aImp * aImp * aImp * (rhoI / rhoT);
  const n2 = Math.max(1, Math.min(Math.round(N * vImpactor / (vBox + vImpactor)), Math.floor(N / 4)));
  const n1 = N - n2;
  const lat = latticeLay(n1, Rbox, Dbox, rand);
  const s = lat.s, a = JAM * s / (2 * Math.SQRT2);
  const b2 = bodies.onionLay(n2, s, a, rand, 0);
  [lat.relax, b2.relax] = bodies.relaxPiles([lat.pile, b2.pile], native);
  const pos = new Float32Array(N * 4), vel = new Float32Array(N * 4), matR = new Float32Array(N);
  const sedTop = -S.sed * 1000 / L_UNIT, wallR = Rbox - froz, floorY = -(Dbox - froz);
  for (let i = 0; i < n1; i++) {
    const x = lat.p[i * 3], y = lat.p[i * 3 + 1], z = lat.p[i * 3 + 2];
    const mat = (x * x + z * z > wallR * wallR || y < floorY) ? MAT_FROZEN : (y > sedTop ? MAT_PLATFORM : MAT_BASEMENT);
    const o = i * 4;
    pos[o] = x; pos[o + 1] = y; pos[o + 2] = z; pos[o + 3] = mat;
    matR[i] = 1;
  }
  const th = theta * Math.PI / 180;
  let dir = [-Math.cos(th) * Math.SQRT1_2, -Math.sin(th), -Math.cos(th) * Math.SQRT1_2];
  const dl = Math.hypot(dir[0], dir[1], dir[2]);
  dir = [dir[0] / dl, dir[1] / dl, dir[2] / dl];
  const standoff = Math.max(6, 2.2 * dImp + 4);
  const c2 = [-dir[0] * standoff, -dir[1] * standoff + aImp, -dir[2] * standoff];
  for (let i = 0; i < n2; i++) {
    const o = (n1 + i) * 4;
    pos[o] = b2[i * 3] + c2[0]; pos[o + 1] = b2[i * 3 + 1] + c2[1]; pos[o + 2] = b2[i * 3 + 2] + c2[2];
    pos[o + 3] = MAT_IMPACTOR;
    matR[n1 + i] = rhoI / rhoT;
    vel[o] = dir[0] * vImp; vel[o + 1] = dir[1] * vImp; vel[o + 2] = dir[2] * vImp;
  }
  const mref = vBox / n1 * L_UNIT * L_UNIT * L_UNIT * rhoT;
  const vref = Math.max(2.5, 0.25 * vImp);
  const k = 4 * vref * vref / (a * a);
  const dt = 0.35 / Math.sqrt(2 * k / Math.min(1, rhoI / rhoT));
  const lnE = Math.log(S.rest), zeta = -lnE / Math.sqrt(Math.PI * Math.PI + lnE * lnE);
  const cd = 2 * zeta * Math.sqrt(k / 2);
  const mImp = Math.PI / 6 * Math.pow(S.imp * 1000, 3) * rhoI;
  const energy = 0.5 * mImp * Math.pow(S.vimp * 1000, 2);
  const settle = Math.ceil(0.5 / dt);
  const sim = {
    N, W, H, G: N >= 65536 ? 128 : 64, n1, n2, m: mref, a, cell: 2.6 * a, dt, k, c: cd, creep: 0.10,
    damp: Math.exp(-dt * 12), matR, t: -standoff / vImp, phase: 'settle', settleLeft: settle, settleTotal: settle,
    stepNo: 0, tK: 0.5 * E_KG / CP_ROCK, kappa: 0.3, dir, vImp, aImp, dImp, cen2: c2, standoff,
    Dtc: Ltc / L_UNIT, DtcM: Ltc, Dfr: finalCrater(Ltc), energy, mImp, Mw: seismicMagnitude(energy), rhoT, rhoI, theta,
    sedTop, impactT: 0, Rbox, Dbox, Rtc, view: 'crater', L: L_UNIT, T: T_UNIT, V: V_UNIT
  };
  return { sim, pos, vel, lat };
}

function createCraterEngine(opts) {
  opts = opts || {};
  const addon = require(opts.addon || path.join(__dirname, 'cc_addon.node'));
  let ctx = null, sim = null;
  const zero3 = [0, 0, 0];
  // the crater page's uniforms (chicxulub.html:676-728): no iron, unit mass, g down
  function params() {
    return { N: sim.N, W: sim.W, H: sim.H, G: sim.G, m: 1, a: sim.a, cell: sim.cell, dt: sim.dt, k: sim.k, c: sim.c,
             creep: sim.creep, damp: sim.damp, tK: sim.tK, kappa: sim.kappa, cpIron: 1, seatSlack: 0.6,
             spin0: zero3, spin1: zero3, meshEvery: 8, field: 1, gField: [0, -G_EARTH * T_UNIT * T_UNIT / L_UNIT, 0],
             frozenOn: 1, frozen: MAT_FROZEN };
  }
  function build(S) {
    const run = buildCraterRun(Object.assign({}, DEFAULTS, S || {}), addon);
    sim = run.sim;
    if (ctx) addon.destroy(ctx);   // a rebuild frees the last run's device objects (src/nbody.js:208-245)
    ctx = addon.create(params());
    addon.uploadState(ctx, run.pos, run.vel, sim.matR, new Float32Array(sim.N), sim.N);
    sim.initial = { pos: run.pos, vel: run.vel };
    return sim;
  }
  const args = (settle) => ({ settle: !!settle, kick: false, bulk0: zero3, bulk1: zero3, cen0: zero3, cen1: zero3 });
  // seats every second step, the field every step (chicxulub.html:779-782)
  function step(settle) {
    addon.setStepNo(ctx, sim.stepNo);
    addon.step(ctx, args(settle));
    sim.stepNo++;
  }
  // n full steps at once (cc_step_n: the schedule of step(false) n times, each 8-step cycle one graph launch)
  function stepMany(n) {
    addon.setStepNo(ctx, sim.stepNo);
    const ms = addon.stepN(ctx, n, args(false), null);
    sim.stepNo += n;
    return ms;
  }
  function settleSome(n) {
    n = Math.min(n, sim.settleLeft);
    for (let i = 0; i < n; i++) step(true);
    sim.settleLeft -= n;
    if (sim.settleLeft === 0) sim.phase = 'approach';
  }
  // the impactor on rails, a fixed number of frames (chicxulub.html:795-813)
  function approachStep() {
    const v = sim.vImp, dtF = sim.standoff / v / APPROACH_FRAMES;
    const I = [1, 0, 0, 0, 1, 0, 0, 0, 1];
    addon.rigid(ctx, zero3, [sim.dir[0] * v, sim.dir[1] * v, sim.dir[2] * v], zero3, zero3, I, I, dtF);
    sim.t += dtF;
    for (let i = 0; i < 3; i++) sim.cen2[i] += sim.dir[i] * v * dtF;
    if (sim.cen2[1] <= sim.aImp) couple();
  }
  // the coupling pass from the end of the penetration path (chicxulub.html:820-842)
  function couple() {
    const pen = 1.3 * sim.dImp;
    const centre = [sim.cen2[0] + sim.dir[0] * pen, sim.cen2[1] + sim.dir[1] * pen, sim.cen2[2] + sim.dir[2] * pen];
    addon.couple(ctx, { centre, vimp: sim.vImp, aimp: sim.aImp, c1: HH_C1, invMu: 1 / HH_MU, pexp: HH_P,
                        stall: 0.5 * sim.Dtc, z: Z_MODEL, heat: E_KG / CP_ROCK, tK: sim.tK, frozen: MAT_FROZEN,
                        ambient: 300, melt: MAT_MELT });
    sim.phase = 'full';
    sim.t = 0;
    sim.stepNo = 0;
    sim.centre = centre;
  }
  function readState() {
    const N = sim.N, st = { pos: new Float32Array(N * 4), vel: new Float32Array(N * 4), aux: new Float32Array(N * 4),
                            q: new Float32Array(N), mat: new Float32Array(N) };
    addon.downloadState(ctx, st.pos, st.vel, st.aux, st.q, st.mat, N);
    return st;
  }
  // analyze(): the crater view's readout (analyze()/craterStats of chicxulub.html:848-919, :926-945) on the
  // state read back, computed by the library on the host; report.t the run's time
  let report = null;
  function analyze() {
    const st = readState();
    report = addon.craterStats(st.pos, st.vel, sim.matR, sim.N, sim.m, {
      Rbox: sim.Rbox, lSI: L_UNIT, vSI: V_UNIT, gSI: G_EARTH, rhoSI: sim.rhoT, vOrb: 7900 / V_UNIT, vEsc: 11200 / V_UNIT
    });
    report.t = sim.t;
    return report;
  }
  function destroy() { if (ctx) { addon.destroy(ctx); ctx = null; } }
  return { build, step, stepMany, settleSome, approachStep, couple, readState, analyze, destroy, get sim() { return sim; },
           get ctx() { return ctx; }, get report() { return report; }, PRESETS, DEFAULTS };
}

module.exports = { createCraterEngine, buildCraterRun, latticeFill, latticeLay, transientCrater, finalCrater, seismicMagnitude,
                   MAT_FROZEN, MAT_MELT, MAT_IMPACTOR, HH_C1, HH_MU, HH_P, Z_MODEL, E_KG, CP_ROCK, G_EARTH, T_UNIT, L_UNIT };
