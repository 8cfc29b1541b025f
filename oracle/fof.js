// CPU restatement of the reference's host analysis (fof, src/nbody.js:646-893)
// -- TEST INFRASTRUCTURE, the checker only: the tests run it beside the
// library's analysis (cc_analyse) on the same state.  Friends-of-friends
// groups, the largest and second body's statistics, spin, disk and moon, the
// conservation books and the loose list the pairwise pass corrects (the
// Morton order the reference also computes there is cc_morton's).
//
// Every float64 sum runs in particle-index order, and the union-find walks
// the same hashed cells in the same order, so group choice, the core centre
// and the loose list agree with the reference's analysis bit for bit; the
// reference's own reports on the golden states pin it (tests/test_js_host.py).
'use strict';

const HASH_BITS = 20;

// connected components at linking length L: counting sort into hashed cells,
// union-find over the 27 cells around each particle; the spring energy of
// the touching pairs on the way (each pair once)
function linkGroups(N, P, L, touch, spring) {
  const inv = 1 / L, L2 = L * L, touch2 = touch * touch, T = 1 << HASH_BITS, TM = T - 1;
  const cx = new Int32Array(N), cy = new Int32Array(N), cz = new Int32Array(N), key = new Int32Array(N);
  const start = new Int32Array(T + 1);
  const hash = (x, y, z) => ((x * 73856093) ^ (y * 19349663) ^ (z * 83492791)) & TM;
  for (let i = 0; i < N; i++) {
    const qx = Math.floor(P[i * 4] * inv), qy = Math.floor(P[i * 4 + 1] * inv), qz = Math.floor(P[i * 4 + 2] * inv);
    cx[i] = qx; cy[i] = qy; cz[i] = qz;
    key[i] = hash(qx, qy, qz);
    start[key[i] + 1]++;
  }
  for (let k = 0; k < T; k++) start[k + 1] += start[k];
  const fill = new Int32Array(start.subarray(0, T)), order = new Int32Array(N);
  for (let i = 0; i < N; i++) order[fill[key[i]]++] = i;
  const parent = new Int32Array(N);
  for (let i = 0; i < N; i++) parent[i] = i;
  const find = (x) => { while (parent[x] !== x) { parent[x] = parent[parent[x]]; x = parent[x]; } return x; };
  let elastic = 0;
  for (let i = 0; i < N; i++) {
    const px = P[i * 4], py = P[i * 4 + 1], pz = P[i * 4 + 2];
    for (let dz = -1; dz <= 1; dz++) for (let dy = -1; dy <= 1; dy++) for (let dx = -1; dx <= 1; dx++) {
      const k = hash(cx[i] + dx, cy[i] + dy, cz[i] + dz);
      for (let s = start[k]; s < start[k + 1]; s++) {
        const j = order[s];
        if (j <= i) continue;
        const ex = P[j * 4] - px, ey = P[j * 4 + 1] - py, ez = P[j * 4 + 2] - pz;
        const e2 = ex * ex + ey * ey + ez * ez;
        if (e2 >= L2) continue;
        const ri = find(i), rj = find(j);
        if (ri !== rj) parent[ri] = rj;
        if (e2 < touch2) { const ov = touch - Math.sqrt(e2); elastic += 0.5 * spring * ov * ov; }
      }
    }
  }
  const root = new Int32Array(N), size = new Int32Array(N);
  for (let i = 0; i < N; i++) { root[i] = find(i); size[root[i]]++; }
  let g1 = -1, g2 = -1;
  for (let i = 0; i < N; i++) if (size[i] > 0) {
    if (g1 < 0 || size[i] > size[g1]) { g2 = g1; g1 = i; }
    else if (g2 < 0 || size[i] > size[g2]) g2 = i;
  }
  return { root, size, g1, g2, elastic };
}

const isIron = (w) => w - 4 * Math.floor(w / 4) < 0.5;

// One group's statistics: mass, centre of mass and velocity, the impactor's
// and the iron's shares, the iterated core centre (the centroid of what lies
// within 0.9 of the mass-radius, walked three times from the majority
// material's centre), surface temperature, hot-spot centroid and the edge.
function groupStats(N, mref, P, V, R, root, g) {
  const members = [];
  let mass = 0, x = 0, y = 0, z = 0, vx = 0, vy = 0, vz = 0, imp = 0, iron = 0, ix = 0, iy = 0, iz = 0;
  for (let t = 0; t < N; t++) {
    if (root[t] !== g) continue;
    const m = mref * R[t], w = P[t * 4 + 3];
    members.push(t);
    mass += m;
    x += m * P[t * 4]; y += m * P[t * 4 + 1]; z += m * P[t * 4 + 2];
    vx += m * V[t * 4]; vy += m * V[t * 4 + 1]; vz += m * V[t * 4 + 2];
    if (w >= 3.5) { imp += m; ix += m * P[t * 4]; iy += m * P[t * 4 + 1]; iz += m * P[t * 4 + 2]; }
    if (isIron(w)) iron += m;
  }
  const Rg = Math.cbrt(mass);
  const icx = imp > 0 ? ix / imp : 0, icy = imp > 0 ? iy / imp : 0, icz = imp > 0 ? iz / imp : 0;
  const tm = mass - imp;
  let core = imp > tm ? [icx, icy, icz] : [(x - ix) / tm, (y - iy) / tm, (z - iz) / tm];
  const rc2 = 0.9 * 0.9 * Rg * Rg;
  for (let it = 0; it < 3; it++) {
    let sm = 0, sx = 0, sy = 0, sz = 0;
    for (const t of members) {
      const qx = P[t * 4] - core[0], qy = P[t * 4 + 1] - core[1], qz = P[t * 4 + 2] - core[2];
      if (qx * qx + qy * qy + qz * qz < rc2) {
        const m = mref * R[t];
        sm += m; sx += m * P[t * 4]; sy += m * P[t * 4 + 1]; sz += m * P[t * 4 + 2];
      }
    }
    if (sm > 0) core = [sx / sm, sy / sm, sz / sm];
  }
  const shell = 0.64 * Rg * Rg, NBIN = 96, dr = 3 * Rg / NBIN, hist = new Float64Array(NBIN);
  let Ts = 0, ws = 0, irms = 0, hs = 0, hx = 0, hy = 0, hz = 0, hr2 = 0;
  for (const t of members) {
    const m = mref * R[t];
    const sx = P[t * 4] - core[0], sy = P[t * 4 + 1] - core[1], sz = P[t * 4 + 2] - core[2], r2 = sx * sx + sy * sy + sz * sz;
    const Th = V[t * 4 + 3];
    if (r2 > shell) { const Tt = Math.min(Th, 12000), wt = Tt * Tt; Ts += wt * Tt; ws += wt; }
    if (Th > 1200) { let wh = Math.min(Th, 12000); wh *= wh; hs += wh; hx += wh * P[t * 4]; hy += wh * P[t * 4 + 1]; hz += wh * P[t * 4 + 2]; hr2 += wh * r2; }
    if (P[t * 4 + 3] >= 3.5) { const ex = P[t * 4] - icx, ey = P[t * 4 + 1] - icy, ez = P[t * 4 + 2] - icz; irms += m * (ex * ex + ey * ey + ez * ez); }
    const bi = Math.sqrt(r2) / dr;
    hist[bi < NBIN - 1 ? bi | 0 : NBIN - 1] += m;
  }
  const hotCom = hs > 0 ? [hx / hs, hy / hs, hz / hs] : core.slice();
  let hotR = Rg;
  if (hs > 0) { const ox = hotCom[0] - core[0], oy = hotCom[1] - core[1], oz = hotCom[2] - core[2]; hotR = Math.sqrt(Math.max(hr2 / hs - (ox * ox + oy * oy + oz * oz), 0)); }
  let Redge = 0;
  if (members.length >= 2000) {
    const Rb = Math.cbrt(Math.max(imp, tm)), rho = new Float64Array(NBIN);
    let refS = 0, refN = 0;
    for (let b = 0; b < NBIN; b++) {
      const rm = (b + 0.5) * dr;
      rho[b] = hist[b] / (rm * rm * dr);
      if (rm > 0.6 * Rb && rm < 0.85 * Rb) { refS += rho[b]; refN++; }
    }
    Redge = Rg;
    for (let b = 0; b < NBIN; b++) if ((b + 0.5) * dr > 0.85 * Rb && refN > 0 && rho[b] < 0.35 * refS / refN) { Redge = b * dr; break; }
  }
  return { n: members.length, members, mass, com: [x / mass, y / mass, z / mass], core,
           vel: [vx / mass, vy / mass, vz / mass], imp: imp / mass, iron: iron / mass, R: Rg, Redge,
           Tsurf: ws > 0 ? Ts / ws : 0, hotCom, hotR, impMass: imp, impCom: [icx, icy, icz],
           impRms: imp > 0 ? Math.sqrt(irms / imp) : 0 };
}

// the largest body's spin: angular momentum about its core over the moment
// of inertia about that axis, over what lies within 1.15 of its own radius
function bodySpin(N, mref, P, V, R, root, A, g1) {
  const rb2 = Math.pow(1.15 * Math.cbrt(Math.max(A.impMass, A.mass - A.impMass)), 2);
  const c0 = A.core[0], c1 = A.core[1], c2 = A.core[2];
  const inBody = (i) => {
    if (root[i] !== g1) return false;
    const dx = P[i * 4] - c0, dy = P[i * 4 + 1] - c1, dz = P[i * 4 + 2] - c2;
    return dx * dx + dy * dy + dz * dz <= rb2;
  };
  let bm = 0, bvx = 0, bvy = 0, bvz = 0;
  for (let i = 0; i < N; i++) if (inBody(i)) { const m = mref * R[i]; bm += m; bvx += m * V[i * 4]; bvy += m * V[i * 4 + 1]; bvz += m * V[i * 4 + 2]; }
  if (bm > 0) { bvx /= bm; bvy /= bm; bvz /= bm; }
  let Lx = 0, Ly = 0, Lz = 0;
  for (let i = 0; i < N; i++) if (inBody(i)) {
    const m = mref * R[i], dx = P[i * 4] - c0, dy = P[i * 4 + 1] - c1, dz = P[i * 4 + 2] - c2;
    const ux = V[i * 4] - bvx, uy = V[i * 4 + 1] - bvy, uz = V[i * 4 + 2] - bvz;
    Lx += m * (dy * uz - dz * uy); Ly += m * (dz * ux - dx * uz); Lz += m * (dx * uy - dy * ux);
  }
  const Lm = Math.sqrt(Lx * Lx + Ly * Ly + Lz * Lz), lx = Lx / (Lm + 1e-30), ly = Ly / (Lm + 1e-30), lz = Lz / (Lm + 1e-30);
  const tilt = Math.acos(Math.max(-1, Math.min(1, ly))) * 180 / Math.PI;
  let I = 0;
  for (let i = 0; i < N; i++) if (inBody(i)) {
    const m = mref * R[i], dx = P[i * 4] - c0, dy = P[i * 4 + 1] - c1, dz = P[i * 4 + 2] - c2;
    const along = dx * lx + dy * ly + dz * lz;
    I += m * (dx * dx + dy * dy + dz * dz - along * along);
  }
  return { day: Lm > 1e-9 ? 2 * Math.PI * I / Lm * 805 / 3600 : Infinity, tilt, axis: [lx, ly, lz], bodyVel: [bvx, bvy, bvz] };
}

// the analysis of one state: P, V (Float32 N*4), R (masses), D (the DIAG read-out)
function analyse(N, mref, a, P, V, R, D, spring, touch, cap) {
  const link = linkGroups(N, P, 2.4 * a, touch, spring);
  const { root, g1, g2 } = link;
  const A = groupStats(N, mref, P, V, R, root, g1), B = g2 >= 0 ? groupStats(N, mref, P, V, R, root, g2) : null;
  const spin = bodySpin(N, mref, P, V, R, root, A, g1);
  const energy = (x, y, z, vx, vy, vz) => {
    const r = Math.sqrt((x - A.com[0]) * (x - A.com[0]) + (y - A.com[1]) * (y - A.com[1]) + (z - A.com[2]) * (z - A.com[2]));
    const dv2 = (vx - A.vel[0]) * (vx - A.vel[0]) + (vy - A.vel[1]) * (vy - A.vel[1]) + (vz - A.vel[2]) * (vz - A.vel[2]);
    return 0.5 * dv2 - A.mass / Math.max(r, A.R);
  };
  // bound material off the largest body is the disk; the Moon by Ida, Canup & Stewart's fit
  let orbit = 0, esc = 0, Dx = 0, Dy = 0, Dz = 0;
  for (let i = 0; i < N; i++) {
    if (root[i] === g1) continue;
    const m = mref * R[i];
    if (energy(P[i * 4], P[i * 4 + 1], P[i * 4 + 2], V[i * 4], V[i * 4 + 1], V[i * 4 + 2]) < 0) {
      orbit += m;
      const dx = P[i * 4] - A.com[0], dy = P[i * 4 + 1] - A.com[1], dz = P[i * 4 + 2] - A.com[2];
      const wx = V[i * 4] - A.vel[0], wy = V[i * 4 + 1] - A.vel[1], wz = V[i * 4 + 2] - A.vel[2];
      Dx += m * (dy * wz - dz * wy); Dy += m * (dz * wx - dx * wz); Dz += m * (dx * wy - dy * wx);
    } else esc += m;
  }
  const jD = orbit > 0 ? Math.sqrt(Dx * Dx + Dy * Dy + Dz * Dz) / (orbit * Math.sqrt(A.mass * 2.9)) : 0;
  const moon = Math.min(Math.max(0, orbit * (1.9 * jD - 1.1 - 1.9 * 0.05)), 0.95 * orbit);
  let second = null;
  if (B && B.n >= Math.max(8, N / 1000)) {
    const dAB = Math.sqrt((B.com[0] - A.com[0]) * (B.com[0] - A.com[0]) + (B.com[1] - A.com[1]) * (B.com[1] - A.com[1]) + (B.com[2] - A.com[2]) * (B.com[2] - A.com[2]));
    if (dAB > A.R + B.R + 2 * a) {
      second = { mass: B.mass, dist: dAB, bound: energy(B.com[0], B.com[1], B.com[2], B.vel[0], B.vel[1], B.vel[2]) < 0,
                 imp: B.imp, iron: B.iron, com: B.com, core: B.core, vel: B.vel, R: B.R, Redge: B.Redge,
                 Tsurf: B.Tsurf, hotCom: B.hotCom, hotR: B.hotR };
    }
  }
  // the books (src/nbody.js:831-850)
  let Mt = 0, px = 0, py = 0, pz = 0, Jx = 0, Jy = 0, Jz = 0, ke = 0, pe = 0, Q = 0, unseated = 0, full = 0, ironAll = 0;
  for (let i = 0; i < N; i++) {
    const m = mref * R[i];
    const bx = P[i * 4], by = P[i * 4 + 1], bz = P[i * 4 + 2], bu = V[i * 4], bv = V[i * 4 + 1], bw = V[i * 4 + 2];
    Mt += m; px += m * bu; py += m * bv; pz += m * bw;
    if (isIron(P[i * 4 + 3])) ironAll += m;
    Jx += m * (by * bw - bz * bv); Jy += m * (bz * bu - bx * bw); Jz += m * (bx * bv - by * bu);
    ke += 0.5 * m * (bu * bu + bv * bv + bw * bw);
    const flag = D[i * 4 + 1], off = flag >= 2;
    pe += (off ? 1 : 0.5) * m * D[i * 4];
    Q += 0.5 * D[i * 4 + 3];
    if (flag - (off ? 2 : 0) < 0.5) unseated++;
    if (D[i * 4 + 2] >= 7.5) full++;
  }
  // the loose list: beyond 1.3 radii of the largest body (and of a second of 0.05 M or more), farthest first past cap
  const bigB = !!(B && B.mass > 0.05);
  const far2 = (body, i) => {
    const ex = P[i * 4] - body.com[0], ey = P[i * 4 + 1] - body.com[1], ez = P[i * 4 + 2] - body.com[2];
    return ex * ex + ey * ey + ez * ez;
  };
  const loose = (i) => far2(A, i) > 1.69 * A.R * A.R && !(bigB && far2(B, i) <= 1.69 * B.R * B.R);
  let nLoose = 0, least = 0;
  for (let i = 0; i < N; i++) if (loose(i)) nLoose++;
  if (nLoose > cap) {
    const ds = new Float64Array(nLoose);
    let q = 0;
    for (let i = 0; i < N; i++) if (loose(i)) ds[q++] = far2(A, i);
    ds.sort();
    least = ds[nLoose - cap];
  }
  const list = new Int32Array(Math.min(nLoose, cap));
  for (let i = 0, w = 0; i < N && w < list.length; i++) if (loose(i) && far2(A, i) >= least) list[w++] = i;
  return {
    largest: A.mass, largestImp: A.imp, largestIron: A.iron, day: spin.day, tilt: spin.tilt, axis: spin.axis,
    bodyVel: spin.bodyVel, orbit, escape: esc, second, com: A.com, core: A.core, vel: A.vel, R: A.R, Redge: A.Redge,
    Tsurf: A.Tsurf, hotCom: A.hotCom, hotR: A.hotR, impMass: A.impMass, impCom: A.impCom, impRms: A.impRms,
    M: Mt, iron: ironAll / Mt, drift: Math.sqrt(px * px + py * py + pz * pz) / Mt, L: Math.sqrt(Jx * Jx + Jy * Jy + Jz * Jz),
    KE: ke, PE: pe, EL: link.elastic, Q, unseated: unseated / N, full: full / N,
    pp: { list, loose: nLoose }, jD, moon
  };
}

module.exports = { analyse, linkGroups, groupStats };
