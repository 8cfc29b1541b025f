// The crater view's readout (craterStats, chicxulub.html:848-919): from the
// state on the host, float64 in the reference's order, what the page reports
// of a crater -- its diameter, rim, depth, the excavated, thrown, airborne and
// escaping mass, melt and vapour, the energies and the melt pool as one light.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/cc_hip.h"
#include "cc_fof_host.h"

// Math.min(NB - 1, Math.floor(r / binW)) as an index, or -1 where r is NaN
// (the page's typed array then ignores the add); r is never negative.
static int ring(double r, double binW, int NB) {
  const double f = std::floor(r / binW);
  if (!(f >= 0)) return -1;
  return f >= NB - 1 ? NB - 1 : (int)f;
}

extern "C" int cc_crater_stats(int N, double mref, const float* P, const float* V, const float* R,
                               const cc_crater_args* o, cc_crater_report* out) {
  if (N < 0 || !out || !o || (N > 0 && (!P || !V || !R)) || !(o->Rbox > 0) || !(o->rhoSI > 0)) return CC_EINVAL;
  const int NB = 96;
  const double binW = o->Rbox / NB;
  std::vector<double> top(NB, -1e9), above(NB, 0.0);
  std::vector<unsigned char> any(NB, 0);
  double exc = 0, excPlat = 0, air = 0, fastOrb = 0, fastEsc = 0, melt = 0, vap = 0, Tmax = 0;
  double ke = 0, pe = 0, px = 0, py = 0, pz = 0, mass = 0;
  double hotM = 0, hx = 0, hy = 0, hz = 0, Tw = 0, Tws = 0;
  for (size_t i = 0; i < (size_t)N; i++) {
    const float* p = P + 4 * i;
    const float* v = V + 4 * i;
    const double x = p[0], y = p[1], z = p[2], w = p[3];
    const double vx = v[0], vy = v[1], vz = v[2], T = v[3];
    if (w < 0.5) continue;   // the frozen shell is not part of the answer
    const double m = mref * R[i];
    const double sp2 = vx * vx + vy * vy + vz * vz;
    mass += m;
    ke += 0.5 * m * sp2 * o->vSI * o->vSI;   // joules
    pe += m * o->gSI * y * o->lSI;
    px += m * vx; py += m * vy; pz += m * vz;
    if (T > Tmax) Tmax = T;
    if (w > 2.5 && w < 3.5) melt += m;       // what the coupling's shock made melt
    if (T > 2500) vap += m;
    if (T > 1200) {
      hotM += m; hx += m * x; hy += m * y; hz += m * z;
      const double Tc = std::min(T, 12000.0), wt = Tc * Tc;
      Tw += wt * Tc; Tws += wt;
    }
    const double r = std::sqrt(x * x + z * z);
    if (y > 0) {
      exc += m;
      const int b = ring(r, binW, NB);
      if (b >= 0) above[b] += m;
      if (w > 1.5 && w < 2.5) excPlat += m;
      if (sp2 > 0.0225) {   // still moving: in the curtain, not on the blanket
        air += m;
        const double sp = std::sqrt(sp2);
        if (sp > o->vOrb) fastOrb += m;
        if (sp > o->vEsc) fastEsc += m;
      }
    }
    if (sp2 < 0.0225 && r < o->Rbox) {   // settled: it can speak for the surface
      const int b = ring(r, binW, NB);
      if (y > top[b]) { top[b] = y; any[b] = 1; }
    }
  }
  // the rim, and where the surface crosses back through where the ground was
  double rimH = -1e9, rimR = 0, floorY = 0, craterR = 0;
  bool seen = false;
  for (int b = 0; b < NB; b++) {
    if (!any[b]) continue;
    if (top[b] > rimH) { rimH = top[b]; rimR = (b + 0.5) * binW; }
    if (top[b] < floorY) floorY = top[b];
    if (top[b] < -0.05) { craterR = (b + 1) * binW; seen = true; }
    else if (seen && top[b] >= -0.05) break;
  }
  const double poolR = cc::js_cbrt(3 * (hotM / o->rhoSI) / (4 * M_PI)) / o->lSI;   // Math.cbrt's roundings
  double thrown = 0;
  for (int b = (int)std::ceil(rimR / binW); b < NB; b++) thrown += above[b];
  out->craterD = 2 * craterR; out->rimR = rimR; out->rimH = rimH; out->depth = -floorY;
  out->excavated = exc; out->thrown = thrown; out->platform = excPlat; out->airborne = air;
  out->fastOrb = fastOrb; out->fastEsc = fastEsc; out->melt = melt; out->vapour = vap; out->Tmax = Tmax;
  out->mass = mass; out->KE = ke; out->PE = pe; out->drift = std::sqrt(px * px + py * py + pz * pz) / mass;
  out->Tsurf = Tws > 0 ? Tw / Tws : 0;
  out->R = std::max(poolR, 1e-3);
  out->com[0] = hotM > 0 ? hx / hotM : 0; out->com[1] = hotM > 0 ? hy / hotM : 0; out->com[2] = hotM > 0 ? hz / hotM : 0;
  return 0;
}
