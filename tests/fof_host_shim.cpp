// Test harness: the library's host half of the analysis (cc_fof_host.cpp)
// behind one C function, fed by the oracle's restatement of the reference's
// linking in place of the device half.  Built by tests/test_fof_host.py.
#include "../cosmic-collisions_amd/csrc/cc_fof_host.h"

extern "C" int fofh_run(int N, const float* P, const float* V, const float* D, const float* R, const int* label,
                        const double* e2, long nT, double m, double a, double spring, double touch, int cap, int flags,
                        const int* linkOff, const int* partner, cc_report* out, int* list) {
  cc::FofHostIn h{ N, reinterpret_cast<const cc::F4*>(P), reinterpret_cast<const cc::F4*>(V),
                   reinterpret_cast<const cc::F4*>(D), R, label, e2, (size_t)nT, m, a, spring, touch, cap, flags };
  cc::FofPick k = cc::fof_pick(h);
  if (k.replay) {
    if (!linkOff) return -1;
    cc::fof_pick_replay(h, linkOff, partner, k);
  }
  cc::fof_report(h, k, out, list);
  return 0;
}
