// The host half of the analysis (fof, src/nbody.js:646-870): which groups
// are the largest and the second, and every FP64 sum of the report in the
// reference's particle-index order.  Plain C++; the device half
// (cc_fof.hip) hands over the state, the DIAG read-out, every particle's
// group label (its component's smallest index) and the touching pairs'
// squared distances in the reference's pair order.
//
// Rounding follows the reference's JavaScript numbers: FP64 throughout, the
// operations in its order, no contraction (-ffp-contract=off), and V8's
// Math.cbrt and Math.acos restated below.
#include "cc_fof_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <thread>
#include <vector>

namespace cc {

// js_cbrt and js_acos below restate two routines of fdlibm, the libm that V8
// carries (FreeBSD's s_cbrt.c, with Bruce D. Evans's optimisation, and
// fdlibm's e_acos.c), whose notice is kept here as their licence asks:
//
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this
//   software is freely granted, provided that this notice
//   is preserved.
//
// (s_cbrt.c additionally: "Optimized by Bruce D. Evans.")

// V8's Math.cbrt: FreeBSD's s_cbrt.c (a 5-bit estimate from the exponent, a
// polynomial to 23 bits, rounding to 21 bits and one Newton step to 53 bits)
double js_cbrt(double x) {
  const uint32_t B1 = 715094163u, B2 = 696219795u;
  const double P0 = 1.87595182427177009643, P1 = -1.88497979543377169875, P2 = 1.621429720105354466140,
               P3 = -0.758397934778766047437, P4 = 0.145996192886612446982;
  uint64_t ux;
  std::memcpy(&ux, &x, 8);
  uint32_t hx = (uint32_t)(ux >> 32), low = (uint32_t)ux, sign = hx & 0x80000000u;
  hx ^= sign;
  if (hx >= 0x7ff00000u) return x + x;
  auto from_hi = [](uint32_t hi) { uint64_t u = (uint64_t)hi << 32; double d; std::memcpy(&d, &u, 8); return d; };
  double t;
  if (hx < 0x00100000u) {
    if ((hx | low) == 0) return x;
    t = from_hi(0x43500000u) * x;   // 2^54 x
    uint64_t ut;
    std::memcpy(&ut, &t, 8);
    t = from_hi(sign | ((((uint32_t)(ut >> 32)) & 0x7fffffffu) / 3 + B2));
  } else {
    t = from_hi(sign | (hx / 3 + B1));
  }
  double r = (t * t) * (t / x);
  t = t * ((P0 + r * (P1 + r * P2)) + ((r * r) * r) * (P3 + r * P4));
  uint64_t ut;
  std::memcpy(&ut, &t, 8);
  ut = (ut + 0x80000000ull) & 0xffffffffc0000000ull;
  std::memcpy(&t, &ut, 8);
  double s = t * t;
  r = x / s;
  double w = t + t;
  r = (r - t) / (w + r);
  return t + t * r;
}

// V8's Math.acos: fdlibm's e_acos.c (a rational approximation of asin about 0,
// and about 1 through the half-angle identity)
static double js_acos(double x) {
  const double pi = 3.14159265358979311600e+00, pio2_hi = 1.57079632679489655800e+00,
               pio2_lo = 6.12323399573676603587e-17, pS0 = 1.66666666666666657415e-01,
               pS1 = -3.25565818622400915405e-01, pS2 = 2.01212532134862925881e-01,
               pS3 = -4.00555345006794114027e-02, pS4 = 7.91534994289814532176e-04,
               pS5 = 3.47933107596021167570e-05, qS1 = -2.40339491173441421878e+00,
               qS2 = 2.02094576023350569471e+00, qS3 = -6.88283971605453293030e-01,
               qS4 = 7.70381505559019352791e-02;
  uint64_t ux;
  std::memcpy(&ux, &x, 8);
  int32_t hx = (int32_t)(ux >> 32), ix = hx & 0x7fffffff;
  uint32_t lo = (uint32_t)ux;
  auto R = [&](double z) {
    double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    return p / q;
  };
  if (ix >= 0x3ff00000) {
    if (((uint32_t)(ix - 0x3ff00000) | lo) == 0) return hx > 0 ? 0.0 : pi + 2.0 * pio2_lo;
    return std::nan("");
  }
  if (ix < 0x3fe00000) {
    if (ix <= 0x3c600000) return pio2_hi + pio2_lo;
    double r = R(x * x);
    return pio2_hi - (x - (pio2_lo - x * r));
  }
  if (hx < 0) {
    double z = (1.0 + x) * 0.5, s = std::sqrt(z), r = R(z);
    double w = r * s - pio2_lo;
    return pi - 2.0 * (s + w);
  }
  double z = (1.0 - x) * 0.5, s = std::sqrt(z);
  uint64_t us;
  std::memcpy(&us, &s, 8);
  us &= 0xffffffff00000000ull;
  double df;
  std::memcpy(&df, &us, 8);
  double c = (z - df * df) / (s + df), r = R(z);
  double w = r * s + c;
  return 2.0 * (df + w);
}

// Math.max / Math.min: NaN if either is, +0 above -0
static double js_max(double a, double b) {
  if (a != a || b != b) return std::nan("");
  if (a == b) return std::signbit(a) ? b : a;
  return a > b ? a : b;
}
static double js_min(double a, double b) {
  if (a != a || b != b) return std::nan("");
  if (a == b) return std::signbit(a) ? a : b;
  return a < b ? a : b;
}

static inline bool is_iron(double w) { return w - 4 * std::floor(w / 4) < 0.5; }

struct Group {   // stats() of src/nbody.js:682-760
  int n = 0;
  double mass = 0, com[3] = {0, 0, 0}, core[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, imp = 0, iron = 0, R = 0, Redge = 0,
         Tsurf = 0, hotCom[3] = {0, 0, 0}, hotR = 0, impMass = 0, impCom[3] = {0, 0, 0}, impRms = 0;
};

// the index-order sums of a group's first pass (src/nbody.js:688-700)
struct GroupSums {
  std::vector<int> members;
  double mass = 0, x = 0, y = 0, z = 0, vx = 0, vy = 0, vz = 0, imp = 0, iron = 0, ix = 0, iy = 0, iz = 0;
  void add(int t, double mref, const F4* P, const F4* V, const float* Rm) {
    double m = mref * (double)Rm[t], w = P[t].w;
    members.push_back(t);
    mass += m;
    x += m * (double)P[t].x; y += m * (double)P[t].y; z += m * (double)P[t].z;
    vx += m * (double)V[t].x; vy += m * (double)V[t].y; vz += m * (double)V[t].z;
    if (w >= 3.5) { imp += m; ix += m * (double)P[t].x; iy += m * (double)P[t].y; iz += m * (double)P[t].z; }
    if (is_iron(w)) iron += m;
  }
};

// what stats() takes from a group's first pass alone: mass, centre, velocity,
// shares, radius, the impactor's centre (src/nbody.js:701-702)
static Group group_basics(const GroupSums& g) {
  Group o;
  o.n = (int)g.members.size();
  o.mass = g.mass;
  o.com[0] = g.x / g.mass; o.com[1] = g.y / g.mass; o.com[2] = g.z / g.mass;
  o.vel[0] = g.vx / g.mass; o.vel[1] = g.vy / g.mass; o.vel[2] = g.vz / g.mass;
  o.imp = g.imp / g.mass;
  o.iron = g.iron / g.mass;
  o.R = js_cbrt(g.mass);
  o.impMass = g.imp;
  o.impCom[0] = g.imp > 0 ? g.ix / g.imp : 0;
  o.impCom[1] = g.imp > 0 ? g.iy / g.imp : 0;
  o.impCom[2] = g.imp > 0 ? g.iz / g.imp : 0;
  return o;
}

// the rest of stats(): the core walk, the surface temperature, the hot spot,
// the impactor's spread and the edge (src/nbody.js:712-760), over the members
static void group_walks(const GroupSums& g, Group& o, double mref, const F4* P, const F4* V, const float* Rm) {
  const std::vector<int>& mem = g.members;
  const int n = o.n;
  const double Rg = o.R, icx = o.impCom[0], icy = o.impCom[1], icz = o.impCom[2];
  double tm = g.mass - g.imp;
  double core[3];
  if (g.imp > tm) { core[0] = icx; core[1] = icy; core[2] = icz; }
  else { core[0] = (g.x - g.ix) / tm; core[1] = (g.y - g.iy) / tm; core[2] = (g.z - g.iz) / tm; }
  double rc2 = 0.9 * 0.9 * Rg * Rg;
  for (int it = 0; it < 3; it++) {
    double sm = 0, sx = 0, sy = 0, sz = 0;
    for (int t : mem) {
      double qx = (double)P[t].x - core[0], qy = (double)P[t].y - core[1], qz = (double)P[t].z - core[2];
      if (qx * qx + qy * qy + qz * qz < rc2) {
        double m = mref * (double)Rm[t];
        sm += m; sx += m * (double)P[t].x; sy += m * (double)P[t].y; sz += m * (double)P[t].z;
      }
    }
    if (sm > 0) { core[0] = sx / sm; core[1] = sy / sm; core[2] = sz / sm; }
  }
  const int NBIN = 96;
  double shell = 0.64 * Rg * Rg, Ts = 0, ws = 0, irms = 0;
  double dr = 3 * Rg / NBIN, hist[NBIN], hs = 0, hx = 0, hy = 0, hz = 0, hr2 = 0;
  for (int b = 0; b < NBIN; b++) hist[b] = 0;
  for (int t : mem) {
    double m = mref * (double)Rm[t];
    double sx = (double)P[t].x - core[0], sy = (double)P[t].y - core[1], sz = (double)P[t].z - core[2];
    double r2s = sx * sx + sy * sy + sz * sz;
    double Th = V[t].w;
    if (r2s > shell) { double Tt = js_min(Th, 12000), wt = Tt * Tt; Ts += wt * Tt; ws += wt; }
    if (Th > 1200) {
      double wh = js_min(Th, 12000);
      wh *= wh;
      hs += wh; hx += wh * (double)P[t].x; hy += wh * (double)P[t].y; hz += wh * (double)P[t].z; hr2 += wh * r2s;
    }
    if ((double)P[t].w >= 3.5) {
      double ex = (double)P[t].x - icx, ey = (double)P[t].y - icy, ez = (double)P[t].z - icz;
      irms += m * (ex * ex + ey * ey + ez * ez);
    }
    double bi = std::sqrt(r2s) / dr;
    hist[bi < NBIN - 1 ? (int)bi : NBIN - 1] += m;
  }
  double hotCom[3];
  if (hs > 0) { hotCom[0] = hx / hs; hotCom[1] = hy / hs; hotCom[2] = hz / hs; }
  else { hotCom[0] = core[0]; hotCom[1] = core[1]; hotCom[2] = core[2]; }
  double hotR = Rg;
  if (hs > 0) {
    double ox = hotCom[0] - core[0], oy = hotCom[1] - core[1], oz = hotCom[2] - core[2];
    hotR = std::sqrt(js_max(hr2 / hs - (ox * ox + oy * oy + oz * oz), 0));
  }
  double Redge = 0;
  if (n >= 2000) {
    double Rb = js_cbrt(js_max(g.imp, tm)), rho[NBIN], refS = 0;
    int refN = 0;
    for (int b = 0; b < NBIN; b++) {
      double rm = (b + 0.5) * dr;
      rho[b] = hist[b] / (rm * rm * dr);
      if (rm > 0.6 * Rb && rm < 0.85 * Rb) { refS += rho[b]; refN++; }
    }
    Redge = Rg;
    for (int b = 0; b < NBIN; b++)
      if ((b + 0.5) * dr > 0.85 * Rb && refN > 0 && rho[b] < 0.35 * refS / refN) { Redge = b * dr; break; }
  }
  for (int k = 0; k < 3; k++) { o.core[k] = core[k]; o.hotCom[k] = hotCom[k]; }
  o.Redge = Redge;
  o.Tsurf = ws > 0 ? Ts / ws : 0;
  o.hotR = hotR;
  o.impRms = g.imp > 0 ? std::sqrt(irms / g.imp) : 0;
}



FofPick fof_pick(const FofHostIn& h) {
  const int N = h.N;
  const int* lab = h.label;
  std::vector<int> size(N, 0);
  for (int i = 0; i < N; i++) size[lab[i]]++;
  FofPick k;
  int top[3] = { -1, -1, -1 };
  for (int i = 0; i < N; i++) {
    if (!size[i]) continue;
    k.groups++;
    int v = size[i];
    if (top[0] < 0 || v > size[top[0]]) { top[2] = top[1]; top[1] = top[0]; top[0] = i; }
    else if (top[1] < 0 || v > size[top[1]]) { top[2] = top[1]; top[1] = i; }
    else if (top[2] < 0 || v > size[top[2]]) top[2] = i;
  }
  // The reference's choice (src/nbody.js:680-684) is the first root in index
  // order with the largest size, then likewise for the second; its roots are
  // its union-find's, not these labels.  Sizes that differ make the choice the
  // same; a tie for the largest, or for the second when the second can enter
  // the report (as `second`, at max(8, N/1000) members, or through bigB, above
  // 0.05 in mass), is settled by replaying its unions.
  float maxR = 0;
  for (int i = 0; i < N; i++) maxR = std::max(maxR, h.R[i]);
  auto matters2 = [&](int n) {
    return n >= js_max(8, N / 1000.0) || (double)n * h.m * (double)maxR * (1 + 1e-9) > 0.05;
  };
  bool tie1 = top[1] >= 0 && size[top[1]] == size[top[0]];
  bool tie2 = top[2] >= 0 && size[top[2]] == size[top[1]] && matters2(size[top[1]]);
  k.replay = (h.flags & CC_ANALYSE_REPLAY_UNIONS) || tie1 || tie2;
  k.g1 = top[0];
  k.g2 = top[1];
  return k;
}

void fof_pick_replay(const FofHostIn& h, const int* linkOff, const int* partner, FofPick& k) {
  // the reference's union-find over every linked pair in its order (src/nbody.js:660-684)
  const int N = h.N;
  std::vector<int> uf(N);
  for (int i = 0; i < N; i++) uf[i] = i;
  auto find = [&](int x) {
    while (uf[x] != x) { uf[x] = uf[uf[x]]; x = uf[x]; }
    return x;
  };
  for (int i = 0; i < N; i++)
    for (int q = linkOff[i]; q < linkOff[i + 1]; q++) {
      int ri = find(i), rj = find(partner[q]);
      if (ri != rj) uf[ri] = rj;
    }
  std::vector<int> rsize(N, 0);
  for (int i = 0; i < N; i++) rsize[find(i)]++;
  int j1 = -1, j2 = -1;
  for (int i = 0; i < N; i++)
    if (rsize[i] > 0) {
      if (j1 < 0 || rsize[i] > rsize[j1]) { j2 = j1; j1 = i; }
      else if (j2 < 0 || rsize[i] > rsize[j2]) j2 = i;
    }
  k.g1 = h.label[j1];
  k.g2 = j2 >= 0 ? h.label[j2] : -1;
  k.replay = true;
}

// the tasks at once, each but the last on a thread of its own; a task whose
// thread cannot be had runs on the caller's, in turn.  A task's exception
// (an allocation that failed) is passed on to the caller once all have ended.
static void run_together(std::vector<std::function<void()>>& tasks) {
  std::vector<std::exception_ptr> err(tasks.size());
  auto guarded = [&](size_t k) {
    try {
      tasks[k]();
    } catch (...) {
      err[k] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  size_t k = 0;
  for (; k + 1 < tasks.size(); k++) {
    try {
      th.emplace_back(guarded, k);
    } catch (...) {
      break;
    }
  }
  for (size_t r = k; r < tasks.size(); r++) guarded(r);
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

void fof_report(const FofHostIn& h, const FofPick& k, cc_report* out, int32_t* list) {
  const int N = h.N;
  const F4 *P = h.P, *V = h.V, *D = h.D;
  const float* Rm = h.R;
  const int* lab = h.label;
  const double mref = h.m;
  const int g1 = k.g1, g2 = k.g2;

  // stats() of the largest and the second group: the first pass of both in one
  // walk, each group's sums in index order (src/nbody.js:688-700)
  GroupSums s1, s2;
  for (int t = 0; t < N; t++) {
    if (lab[t] == g1) s1.add(t, mref, P, V, Rm);
    else if (lab[t] == g2) s2.add(t, mref, P, V, Rm);
  }
  const bool haveB = g2 >= 0;
  Group A = group_basics(s1), B;
  if (haveB) B = group_basics(s2);
  const bool bigB = haveB && B.mass > 0.05;

  // Every sum below is its own chain in the reference's order; chains that do
  // not read each other's results run side by side: the elastic energy, the
  // second group's walks, the disk with the books, the loose list, and on this
  // thread the largest group's walks and then its spin.
  double elastic = 0;
  double orbit = 0, esc = 0, Dx = 0, Dy = 0, Dz = 0;
  double Mt = 0, px = 0, py = 0, pz = 0, Jx = 0, Jy = 0, Jz = 0, ke = 0, pe = 0, Q = 0, ironAll = 0;
  int unseated = 0, full = 0, nLoose = 0, wr = 0;
  double bvx = 0, bvy = 0, bvz = 0, lx = 0, ly = 0, lz = 0, tilt = 0, day = 0;

  // energy in the largest body's frame (src/nbody.js:797-801)
  auto energy = [&](double x, double y, double z, double vx, double vy, double vz) {
    double r = std::sqrt((x - A.com[0]) * (x - A.com[0]) + (y - A.com[1]) * (y - A.com[1]) +
                         (z - A.com[2]) * (z - A.com[2]));
    double dv2 = (vx - A.vel[0]) * (vx - A.vel[0]) + (vy - A.vel[1]) * (vy - A.vel[1]) +
                 (vz - A.vel[2]) * (vz - A.vel[2]);
    return 0.5 * dv2 - A.mass / js_max(r, A.R);
  };
  auto far2 = [&](const Group& b, int i) {
    double ex = (double)P[i].x - b.com[0], ey = (double)P[i].y - b.com[1], ez = (double)P[i].z - b.com[2];
    return ex * ex + ey * ey + ez * ez;
  };

  std::vector<std::function<void()>> tasks;
  // the contacts' spring energy, summed in the reference's pair order (src/nbody.js:674)
  tasks.push_back([&] {
    for (size_t q = 0; q < h.nTouch; q++) {
      double ov = h.touch - std::sqrt(h.e2Touch[q]);
      elastic += 0.5 * h.spring * ov * ov;
    }
  });
  if (haveB) tasks.push_back([&] { group_walks(s2, B, mref, P, V, Rm); });
  // the disk (src/nbody.js:807-815) and the books (src/nbody.js:831-850), one walk
  tasks.push_back([&] {
    for (int i = 0; i < N; i++) {
      double m = mref * (double)Rm[i];
      double bx = P[i].x, by = P[i].y, bz = P[i].z, bu = V[i].x, bv = V[i].y, bw = V[i].z;
      if (lab[i] != g1) {
        if (energy(bx, by, bz, bu, bv, bw) < 0) {
          orbit += m;
          double dx = bx - A.com[0], dy = by - A.com[1], dz = bz - A.com[2];
          double wx = bu - A.vel[0], wy = bv - A.vel[1], wz = bw - A.vel[2];
          Dx += m * (dy * wz - dz * wy); Dy += m * (dz * wx - dx * wz); Dz += m * (dx * wy - dy * wx);
        } else {
          esc += m;
        }
      }
      Mt += m; px += m * bu; py += m * bv; pz += m * bw;
      if (is_iron(P[i].w)) ironAll += m;
      Jx += m * (by * bw - bz * bv); Jy += m * (bz * bu - bx * bw); Jz += m * (bx * bv - by * bu);
      ke += 0.5 * m * (bu * bu + bv * bv + bw * bw);
      double flag = D[i].y;
      bool off = flag >= 2;
      pe += (off ? 1 : 0.5) * m * (double)D[i].x;
      Q += 0.5 * (double)D[i].w;
      if (flag - (off ? 2 : 0) < 0.5) unseated++;
      if ((double)D[i].z >= 7.5) full++;
    }
  });
  // the loose list (src/nbody.js:857-870): beyond 1.3 radii of the largest
  // body, and of the second when it is a body of 0.05 or more; farthest first
  // past the cap
  tasks.push_back([&] {
    auto loose = [&](int i) { return far2(A, i) > 1.69 * A.R * A.R && !(bigB && far2(B, i) <= 1.69 * B.R * B.R); };
    double least = 0;
    std::vector<double> ds;
    for (int i = 0; i < N; i++)
      if (loose(i)) { nLoose++; ds.push_back(far2(A, i)); }
    if (nLoose > h.cap) {
      std::nth_element(ds.begin(), ds.begin() + (nLoose - h.cap), ds.end());
      least = ds[nLoose - h.cap];
    }
    int count = std::min(nLoose, h.cap);
    for (int i = 0; i < N && wr < count; i++)
      if (loose(i) && far2(A, i) >= least) list[wr++] = i;
  });
  // the largest group's walks, then its body's spin (src/nbody.js:770-795), its members in index order
  tasks.push_back([&] {
    group_walks(s1, A, mref, P, V, Rm);
    const double rbq = 1.15 * js_cbrt(js_max(A.impMass, A.mass - A.impMass)), rb2 = rbq * rbq;
    const double c0 = A.core[0], c1 = A.core[1], c2 = A.core[2];
    double bm = 0;
    for (int t : s1.members) {
      double dx = (double)P[t].x - c0, dy = (double)P[t].y - c1, dz = (double)P[t].z - c2;
      if (dx * dx + dy * dy + dz * dz > rb2) continue;
      double m = mref * (double)Rm[t];
      bm += m; bvx += m * (double)V[t].x; bvy += m * (double)V[t].y; bvz += m * (double)V[t].z;
    }
    if (bm > 0) { bvx /= bm; bvy /= bm; bvz /= bm; }
    double Lx = 0, Ly = 0, Lz = 0, I = 0;
    for (int t : s1.members) {
      double dx = (double)P[t].x - c0, dy = (double)P[t].y - c1, dz = (double)P[t].z - c2;
      if (dx * dx + dy * dy + dz * dz > rb2) continue;
      double m = mref * (double)Rm[t];
      double ux = (double)V[t].x - bvx, uy = (double)V[t].y - bvy, uz = (double)V[t].z - bvz;
      Lx += m * (dy * uz - dz * uy); Ly += m * (dz * ux - dx * uz); Lz += m * (dx * uy - dy * ux);
    }
    double Lm = std::sqrt(Lx * Lx + Ly * Ly + Lz * Lz);
    lx = Lx / (Lm + 1e-30); ly = Ly / (Lm + 1e-30); lz = Lz / (Lm + 1e-30);
    tilt = js_acos(js_max(-1, js_min(1, ly))) * 180 / M_PI;
    for (int t : s1.members) {
      double dx = (double)P[t].x - c0, dy = (double)P[t].y - c1, dz = (double)P[t].z - c2;
      if (dx * dx + dy * dy + dz * dz > rb2) continue;
      double m = mref * (double)Rm[t];
      double along = dx * lx + dy * ly + dz * lz;
      I += m * (dx * dx + dy * dy + dz * dz - along * along);
    }
    day = Lm > 1e-9 ? 2 * M_PI * I / Lm * 805 / 3600 : INFINITY;
  });
  run_together(tasks);

  double jD = orbit > 0 ? std::sqrt(Dx * Dx + Dy * Dy + Dz * Dz) / (orbit * std::sqrt(A.mass * 2.9)) : 0;
  double moon = js_min(js_max(0, orbit * (1.9 * jD - 1.1 - 1.9 * 0.05)), 0.95 * orbit);

  std::memset(out, 0, sizeof *out);
  out->hasSecond = 0;
  if (haveB && B.n >= js_max(8, N / 1000.0)) {
    double dAB = std::sqrt((B.com[0] - A.com[0]) * (B.com[0] - A.com[0]) + (B.com[1] - A.com[1]) * (B.com[1] - A.com[1]) +
                           (B.com[2] - A.com[2]) * (B.com[2] - A.com[2]));
    if (dAB > A.R + B.R + 2 * h.a) {
      cc_second& q = out->second;
      out->hasSecond = 1;
      q.mass = B.mass; q.dist = dAB;
      q.bound = energy(B.com[0], B.com[1], B.com[2], B.vel[0], B.vel[1], B.vel[2]) < 0;
      q.imp = B.imp; q.iron = B.iron; q.R = B.R; q.Redge = B.Redge; q.Tsurf = B.Tsurf; q.hotR = B.hotR;
      for (int k = 0; k < 3; k++) { q.com[k] = B.com[k]; q.core[k] = B.core[k]; q.vel[k] = B.vel[k]; q.hotCom[k] = B.hotCom[k]; }
    }
  }

  out->largest = A.mass; out->largestImp = A.imp; out->largestIron = A.iron; out->day = day; out->tilt = tilt;
  out->axis[0] = lx; out->axis[1] = ly; out->axis[2] = lz;
  out->bodyVel[0] = bvx; out->bodyVel[1] = bvy; out->bodyVel[2] = bvz;
  out->orbit = orbit; out->escape = esc;
  for (int k = 0; k < 3; k++) {
    out->com[k] = A.com[k]; out->core[k] = A.core[k]; out->vel[k] = A.vel[k]; out->hotCom[k] = A.hotCom[k];
    out->impCom[k] = A.impCom[k];
  }
  out->R = A.R; out->Redge = A.Redge; out->Tsurf = A.Tsurf; out->hotR = A.hotR; out->impMass = A.impMass;
  out->impRms = A.impRms;
  out->M = Mt; out->iron = ironAll / Mt; out->drift = std::sqrt(px * px + py * py + pz * pz) / Mt;
  out->L = std::sqrt(Jx * Jx + Jy * Jy + Jz * Jz);
  out->KE = ke; out->PE = pe; out->EL = elastic; out->Q = Q;
  out->unseated = (double)unseated / N; out->full = (double)full / N;
  out->jD = jD; out->moon = moon;
  out->ppCount = wr; out->ppLoose = nLoose;
  out->largestN = A.n; out->secondN = haveB ? B.n : 0;
  out->groups = k.groups;
  out->replayed = k.replay ? 1 : 0;
}

}  // namespace cc
