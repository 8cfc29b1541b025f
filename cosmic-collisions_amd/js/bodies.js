// Body setup for the engine object (nbody.js): the seeded random, the onion
// of spiral shells a body is built from, the overlap relaxation, the
// Maclaurin spheroid of a spinning body and a rotation matrix.
//
// The arithmetic is float64 in the order the reference does it
// (src/math.js:8-15, src/bodies.js:23-169), so that a run starts from the same
// Float32 state bit for bit: the initial state is an input to every parity
// check downstream.
'use strict';

// Mulberry32 (src/math.js:8-15): a 32-bit state stepped by a Weyl constant,
// then two multiply-xorshift rounds; returns [0, 1).
function seeded(seed) {
  let st = seed | 0;
  return function next() {
    st = (st + 0x6D2B79F5) | 0;
    let x = Math.imul(st ^ (st >>> 15), st | 1);
    x = (x + Math.imul(x ^ (x >>> 7), x | 61)) ^ x;
    return ((x ^ (x >>> 14)) >>> 0) / 4294967296;
  };
}

// Area under a shell of semi-axes (A, C) from the south pole, in NB bins of
// reduced latitude: the table a shell's points are spread evenly by.
function shellArea(A, C, NB) {
  const cum = new Float64Array(NB + 1);
  for (let b = 0; b < NB; b++) {
    const beta = (b + 0.5) * Math.PI / NB - Math.PI / 2;
    const cb = Math.cos(beta), sb = Math.sin(beta);
    cum[b + 1] = cum[b] + A * cb * Math.sqrt(C * C * cb * cb + A * A * sb * sb);
  }
  return cum;
}

// n particles of lattice spacing s and radius a, as concentric spheroidal
// shells (eccentricity e) outermost first, each holding its volume's share,
// laid by the golden-angle spiral; the innermost remainder fills a seed.
// Returns xyz (Float32Array, inside-out order) with .bounds (each shell's
// first index, ascending) and .pile, what relaxing it takes ({p, n, a, A, C}:
// the spheroid it is held in).  The relaxation draws no random numbers, so a
// caller may lay several bodies first and relax them together.
function onionLay(n, s, a, rand, e) {
  const R = Math.cbrt(3 * n * s * s * s / (16 * Math.PI));
  const ae = R / Math.pow(1 - e * e, 1 / 6), ce = ae * Math.sqrt(1 - e * e);
  const vp = s * s * s / 4, unit = 3 * vp / (4 * Math.PI), jit = 0.06 * s;
  let h = s * Math.sqrt(1 / 3);
  const K = Math.max(1, Math.round(ce / h)), golden = Math.PI * (3 - Math.sqrt(5)), NB = 256;
  h = ce / K;
  const shells = [];
  let total = 0, kept = 0;
  for (let j = 0; j < K; j++) {
    const A = ae - j * h, C = ce - j * h;
    if (C - h * 0.5 < h) break;
    const Am = A - h * 0.5, Cm = C - h * 0.5;
    const want = (A * A * C - (A - h) * (A - h) * (C - h)) / unit;
    shells.push({ A: Am, C: Cm, area: shellArea(Am, Cm, NB), want: want });
    total += want;
    kept = j + 1;
  }
  const As = ae - kept * h, Cs = ce - kept * h;
  total += Math.max(As * As * Cs, 0) / unit;
  const xyz = new Float32Array(n * 3), firsts = [n];
  let free = n, share = 0;
  for (const sh of shells) {
    share += sh.want * n / total;
    const m = Math.min(Math.min(Math.round(share), n) - (n - free), free), lo = free - m;
    const phi0 = rand() * 2 * Math.PI, flip = rand() < 0.5 ? -1 : 1;
    const cum = sh.area, top = cum[NB];
    let bin = 0;
    for (let k = 0; k < m; k++) {
      const target = (k + 0.5) / m * top;
      while (bin < NB - 1 && cum[bin + 1] < target) bin++;
      const lat = (bin + (target - cum[bin]) / (cum[bin + 1] - cum[bin])) * Math.PI / NB - Math.PI / 2;
      const phi = phi0 + k * golden, rc = sh.A * Math.cos(lat), o = (lo + k) * 3;
      xyz[o] = rc * Math.cos(phi) + (rand() - 0.5) * jit;
      xyz[o + 1] = flip * sh.C * Math.sin(lat) + (rand() - 0.5) * jit;
      xyz[o + 2] = rc * Math.sin(phi) + (rand() - 0.5) * jit;
    }
    free = lo;
    firsts.push(lo);
  }
  // the seed: the remainder at the right density anywhere inside the spheroid the shells leave
  const lam = Math.cbrt(free * vp / Math.max(4 / 3 * Math.PI * As * As * Cs, 1e-12));
  for (let k = 0; k < free; k++) {
    let ux, uy, uz, u2;
    do { ux = rand() * 2 - 1; uy = rand() * 2 - 1; uz = rand() * 2 - 1; u2 = ux * ux + uy * uy + uz * uz; } while (u2 > 1);
    xyz[k * 3] = ux * As * lam; xyz[k * 3 + 1] = uy * Cs * lam; xyz[k * 3 + 2] = uz * As * lam;
  }
  firsts.push(0);
  firsts.reverse();
  xyz.bounds = firsts;
  xyz.pile = { p: xyz, n, a, A: ae - h * 0.5, C: ce - h * 0.5, confine: SPHEROID, sweeps: 0 };
  return xyz;
}

// onion() of the reference (src/bodies.js:23-76): laid and relaxed, with
// .relax the relaxation's report.
function onion(n, s, a, rand, e) {
  const xyz = onionLay(n, s, a, rand, e);
  xyz.relax = relaxPiles([xyz.pile])[0];
  return xyz;
}

// Piles relaxed (relaxPile below): with the native addon all in one call,
// side by side on the host's threads (cc_relax_piles, the same float64
// arithmetic in the same order, so the same bits); without it, here, in turn.  A pile is
// {p, n, a, A, C, confine, sweeps}: confine SPHEROID holds it in the spheroid
// of semi-axes A and C, GROUND in the crater's ground patch (radius A, depth C).
const SPHEROID = 0, GROUND = 1;
function relaxPiles(piles, native) {
  if (native) return native.relaxPiles(piles);
  return piles.map((q) => relaxPile(q.p, q.n, q.a, q.A, q.C, q.confine === GROUND ? groundConfine(q.A, q.C) : undefined,
                                    q.sweeps || undefined));
}

// the crater's ground patch (chicxulub.html:543-548): inside the wall, below the flat sky, above the floor
function groundConfine(R, D) {
  const RR = R * R;
  return (held, x, y, z) => {
    const q2 = x * x + z * z;
    if (q2 > RR) { const f = R / Math.sqrt(q2); x *= f; z *= f; }
    held[0] = x; held[1] = Math.max(Math.min(y, 0), -D); held[2] = z;
  };
}

// Inertia-free relaxation (src/bodies.js:97-146): pairs closer than the target
// distance pushed apart by half their overlap, Gauss-Seidel over a cell grid
// rebuilt every fourth sweep; a particle is held inside the spheroid (A, C)
// or by confine(out, x, y, z).
const SWEEPS = 10, TARGET = 0.96, REGRID = 4;
function relaxPile(p, n, a, A, C, confine, sweeps) {
  const D = 2 * a * TARGET, D2 = D * D, cs = 2.2 * a, iA2 = 1 / (A * A), iC2 = 1 / (C * C);
  const nSweeps = sweeps || SWEEPS, held = confine ? [0, 0, 0] : null;
  const lo = [1e9, 1e9, 1e9], hi = [-1e9, -1e9, -1e9];
  for (let i = 0; i < n; i++) for (let k = 0; k < 3; k++) {
    const v = p[i * 3 + k];
    if (v < lo[k]) lo[k] = v;
    if (v > hi[k]) hi[k] = v;
  }
  const gx = Math.floor((hi[0] - lo[0]) / cs) + 3, gy = Math.floor((hi[1] - lo[1]) / cs) + 3,
        gz = Math.floor((hi[2] - lo[2]) / cs) + 3, nc = gx * gy * gz;
  const first = new Int32Array(nc + 1), sorted = new Int32Array(n), home = new Int32Array(n);
  const clampi = (v, top) => Math.min(Math.max(v, 0), top);
  let worst = 0, sweep;
  for (sweep = 0; sweep < nSweeps; sweep++) {
    if (sweep % REGRID === 0) {
      first.fill(0);
      for (let i = 0; i < n; i++) {
        const ix = clampi(Math.floor((p[i * 3] - lo[0]) / cs) + 1, gx - 1);
        const iy = clampi(Math.floor((p[i * 3 + 1] - lo[1]) / cs) + 1, gy - 1);
        const iz = clampi(Math.floor((p[i * 3 + 2] - lo[2]) / cs) + 1, gz - 1);
        home[i] = (ix * gy + iy) * gz + iz;
        first[home[i] + 1]++;
      }
      for (let k = 0; k < nc; k++) first[k + 1] += first[k];
      for (let i = 0; i < n; i++) sorted[first[home[i]]++] = i;
      for (let k = nc; k > 0; k--) first[k] = first[k - 1];
      first[0] = 0;
    }
    worst = 0;
    for (let i = 0; i < n; i++) {
      const c = home[i];
      let x = p[i * 3], y = p[i * 3 + 1], z = p[i * 3 + 2];
      for (let o = 0; o < 27; o++) {
        const nbr = c + ((o % 3) - 1) * gy * gz + ((Math.floor(o / 3) % 3) - 1) * gz + (Math.floor(o / 9) - 1);
        if (nbr < 0 || nbr >= nc) continue;
        for (let t = first[nbr]; t < first[nbr + 1]; t++) {
          const j = sorted[t];
          if (j <= i) continue;
          let dx = p[j * 3] - x, dy = p[j * 3 + 1] - y, dz = p[j * 3 + 2] - z;
          const d2 = dx * dx + dy * dy + dz * dz;
          if (d2 >= D2) continue;
          const d = Math.sqrt(d2) || 1e-9, ov = D - d, f = 0.5 * ov / d;
          if (ov > worst) worst = ov;
          dx *= f; dy *= f; dz *= f;
          x -= dx; y -= dy; z -= dz;
          p[j * 3] += dx; p[j * 3 + 1] += dy; p[j * 3 + 2] += dz;
        }
      }
      if (held) { confine(held, x, y, z); x = held[0]; y = held[1]; z = held[2]; }
      else {
        const qq = (x * x + z * z) * iA2 + y * y * iC2;
        if (qq > 1.0) { const f = 1 / Math.sqrt(qq); x *= f; y *= f; z *= f; }
      }
      p[i * 3] = x; p[i * 3 + 1] = y; p[i * 3 + 2] = z;
    }
  }
  return { sweeps: sweep, worst: worst / a };
}

// Eccentricity of the Maclaurin spheroid spinning at the fraction f of
// breakup (src/bodies.js:152-160): bisection on the equilibrium relation.
function spheroidEcc(f) {
  const want = 4 / 3 * f * f;
  if (want < 1e-6) return 0;
  const g = (e) => { const s2 = 1 - e * e; return 2 * Math.sqrt(s2) / (e * e * e) * (3 - 2 * e * e) * Math.asin(e) - 6 * s2 / (e * e); };
  let lo = 0.01, hi = 0.93;
  if (want >= g(hi)) return hi;
  for (let i = 0; i < 60; i++) { const mid = 0.5 * (lo + hi); if (g(mid) < want) lo = mid; else hi = mid; }
  return 0.5 * (lo + hi);
}

// Column-major rotation matrix of turning at angular velocity w for dt (src/bodies.js:162-169).
function rotation(w, dt) {
  const m = Math.hypot(w[0], w[1], w[2]);
  if (m < 1e-12) return [1, 0, 0, 0, 1, 0, 0, 0, 1];
  const x = w[0] / m, y = w[1] / m, z = w[2] / m, th = m * dt, c = Math.cos(th), s = Math.sin(th), t = 1 - c;
  return [t * x * x + c, t * x * y + s * z, t * x * z - s * y,
          t * x * y - s * z, t * y * y + c, t * y * z + s * x,
          t * x * z + s * y, t * y * z - s * x, t * z * z + c];
}

module.exports = { seeded, onion, onionLay, relaxPile, relaxPiles, groundConfine, SPHEROID, GROUND, spheroidEcc, rotation };
