// The host half of the analysis (cc_fof_host.cpp): plain C++, shared by the
// library and by anything that hands it the device half's results.
#pragma once
#include <cstddef>
#include <cstdint>

#include "../../include/cc_hip.h"

namespace cc {

// V8's Math.cbrt, for the host readouts whose radii must round as the JavaScript page's do
double js_cbrt(double x);

struct F4 { float x, y, z, w; };

struct FofHostIn {
  int N;
  const F4* P;              // positions, w = layer + 4 body
  const F4* V;              // velocities, w = temperature
  const F4* D;              // DIAG read-out
  const float* R;           // masses in units of m
  const int* label;         // each particle's group: its component's smallest index
  const double* e2Touch;    // squared distances of the touching pairs, in the reference's pair order
  size_t nTouch;
  double m, a, spring, touch;
  int cap, flags;
};

struct FofPick {
  int g1 = -1, g2 = -1;     // labels of the reference's largest and second group
  int groups = 0;
  bool replay = false;      // the labels alone cannot settle the choice: replay the unions
};

// the largest and second group from the labels' sizes
FofPick fof_pick(const FofHostIn& h);
// the choice the reference's own union-find makes, from every linked pair in its order
// (partner[linkOff[i] .. linkOff[i+1]) are particle i's)
void fof_pick_replay(const FofHostIn& h, const int* linkOff, const int* partner, FofPick& pick);
// the report and the loose list (out->ms is left to the caller)
void fof_report(const FofHostIn& h, const FofPick& pick, cc_report* out, int32_t* list);

}  // namespace cc
