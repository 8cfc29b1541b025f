// Body setup on the host: the overlap relaxation of a freshly laid pile of
// particles (src/bodies.js:97-146), the piles of a call spread over threads.
//
// The arithmetic is float64 in the reference's order and the positions stay
// FP32 between updates, as the reference's Float32Array keeps them: only
// + - * /, sqrt and floor, so the result is the reference's bit for bit
// (compiled without contraction).  Gauss-Seidel: a particle's pushes land at
// once and the particles after it see them, which makes one pile sequential;
// independent piles (the two bodies of an impact, a ground patch and its
// impactor) run side by side.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cc_hip.h"

namespace {

constexpr int SWEEPS = 10;          // RELAX_SWEEPS
constexpr double TARGET = 0.96;     // RELAX_TARGET
constexpr int REGRID = 4;           // RELAX_REGRID

std::string g_bodies_err;

void relax_one(const cc_pile& pl, cc_pile_report& rep) {
  float* p = pl.p;
  const int n = pl.n;
  const double a = pl.a, A = pl.A, C = pl.C;
  const double D = 2 * a * TARGET, D2 = D * D, cs = 2.2 * a, iA2 = 1 / (A * A), iC2 = 1 / (C * C);
  const double RR = A * A;   // the ground patch's wall radius, squared
  const int nSweeps = pl.sweeps > 0 ? pl.sweeps : SWEEPS;
  double lo[3] = { 1e9, 1e9, 1e9 }, hi[3] = { -1e9, -1e9, -1e9 };
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      double v = p[i * 3 + k];
      if (v < lo[k]) lo[k] = v;
      if (v > hi[k]) hi[k] = v;
    }
  if (n == 0) { rep.sweeps = 0; rep.worst = 0; return; }
  double ext[3];
  for (int k = 0; k < 3; k++) ext[k] = std::floor((hi[k] - lo[k]) / cs) + 3;
  if (!(ext[0] * ext[1] * ext[2] < 2147483647.0)) throw std::length_error("pile spread over more than 2^31 cells");
  const int64_t gx = (int64_t)ext[0], gy = (int64_t)ext[1], gz = (int64_t)ext[2], nc = gx * gy * gz;
  // The positions live in cell order between regrids (pc[t] is particle sorted[t]), so a
  // neighbourhood is nine runs of three cells each in memory; slot[i] finds particle i.
  std::vector<int32_t> first((size_t)nc + 1), sorted((size_t)n), slot((size_t)n);
  std::vector<int64_t> home((size_t)n);
  std::vector<float> pc((size_t)n * 3);
  auto cell_of = [&](double v, int k, int64_t top) {
    int64_t c = (int64_t)std::floor((v - lo[k]) / cs) + 1;
    return c < 0 ? 0 : c > top ? top : c;
  };
  auto scatter = [&]() {
    for (int t = 0; t < n; t++) {
      const int i = sorted[t];
      p[i * 3] = pc[t * 3]; p[i * 3 + 1] = pc[t * 3 + 1]; p[i * 3 + 2] = pc[t * 3 + 2];
    }
  };
  double worst = 0;
  int sweep;
  for (sweep = 0; sweep < nSweeps; sweep++) {
    if (sweep % REGRID == 0) {   // the cells, counted and filled in index order
      if (sweep > 0) scatter();
      std::fill(first.begin(), first.end(), 0);
      for (int i = 0; i < n; i++) {
        int64_t ix = cell_of(p[i * 3], 0, gx - 1), iy = cell_of(p[i * 3 + 1], 1, gy - 1), iz = cell_of(p[i * 3 + 2], 2, gz - 1);
        home[i] = (ix * gy + iy) * gz + iz;
        first[home[i] + 1]++;
      }
      for (int64_t k = 0; k < nc; k++) first[k + 1] += first[k];
      for (int i = 0; i < n; i++) sorted[first[home[i]]++] = i;
      for (int64_t k = nc; k > 0; k--) first[k] = first[k - 1];
      first[0] = 0;
      for (int t = 0; t < n; t++) {
        const int i = sorted[t];
        slot[i] = t;
        pc[t * 3] = p[i * 3]; pc[t * 3 + 1] = p[i * 3 + 1]; pc[t * 3 + 2] = p[i * 3 + 2];
      }
    }
    worst = 0;
    for (int i = 0; i < n; i++) {
      const int64_t c = home[i];
      float* own = &pc[(size_t)slot[i] * 3];
      double x = own[0], y = own[1], z = own[2];
      for (int o = 0; o < 27; o++) {
        const int64_t nbr = c + (int64_t)((o % 3) - 1) * gy * gz + (int64_t)(((o / 3) % 3) - 1) * gz + (o / 9 - 1);
        if (nbr < 0 || nbr >= nc) continue;
        for (int32_t t = first[nbr]; t < first[nbr + 1]; t++) {
          const int j = sorted[t];
          if (j <= i) continue;
          float* q = &pc[(size_t)t * 3];
          double dx = (double)q[0] - x, dy = (double)q[1] - y, dz = (double)q[2] - z;
          const double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 >= D2) continue;
          double d = std::sqrt(d2);
          if (!(d > 0)) d = 1e-9;   // sqrt(d2) || 1e-9: zero (or NaN) takes the floor
          const double ov = D - d, f = 0.5 * ov / d;
          if (ov > worst) worst = ov;
          dx *= f; dy *= f; dz *= f;
          x -= dx; y -= dy; z -= dz;
          q[0] = (float)((double)q[0] + dx);
          q[1] = (float)((double)q[1] + dy);
          q[2] = (float)((double)q[2] + dz);
        }
      }
      if (pl.confine == CC_PILE_GROUND) {   // crater.js's confine: inside the wall, below the sky, above the floor
        const double q2 = x * x + z * z;
        if (q2 > RR) { const double f = A / std::sqrt(q2); x *= f; z *= f; }
        const double ymin = 0.0 < y ? 0.0 : y;   // Math.min(y, 0), -0 kept
        y = ymin < -C ? -C : ymin;                // Math.max(., -D)
      } else {
        const double qq = (x * x + z * z) * iA2 + y * y * iC2;
        if (qq > 1.0) { const double f = 1 / std::sqrt(qq); x *= f; y *= f; z *= f; }
      }
      own[0] = (float)x; own[1] = (float)y; own[2] = (float)z;
    }
  }
  if (sweep > 0) scatter();
  rep.sweeps = sweep;
  rep.worst = worst / a;
}

}  // namespace

extern "C" {

int cc_relax_piles(const cc_pile* piles, int count, cc_pile_report* reports) {
  if (count < 0 || (count > 0 && (!piles || !reports))) { g_bodies_err = "piles and reports must be given"; return CC_EINVAL; }
  for (int k = 0; k < count; k++) {
    const cc_pile& pl = piles[k];
    if (!pl.p || pl.n < 0 || !(pl.a > 0) || !(pl.A > 0) || !(pl.C > 0) ||
        (pl.confine != CC_PILE_SPHEROID && pl.confine != CC_PILE_GROUND)) {
      g_bodies_err = "pile " + std::to_string(k) + ": positions, n >= 0, a, A, C > 0 and a known confinement needed";
      return CC_EINVAL;
    }
  }
  std::vector<std::exception_ptr> err((size_t)count);
  std::atomic<int> next(0);
  auto work = [&]() {   // piles taken in turn until none is left
    for (int k; (k = next.fetch_add(1)) < count;) {
      try {
        relax_one(piles[k], reports[k]);
      } catch (...) {
        err[k] = std::current_exception();
      }
    }
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const int workers = std::min(count, (int)std::max(1u, std::min(hw, 64u)));
  std::vector<std::thread> th;
  for (int w = 1; w < workers; w++) {   // the caller's thread is one of the workers; a thread not had is not needed
    try {
      th.emplace_back(work);
    } catch (...) {
      break;
    }
  }
  work();
  for (auto& t : th) t.join();
  for (int r = 0; r < count; r++)
    if (err[r]) {
      try {
        std::rethrow_exception(err[r]);
      } catch (const std::exception& x) {
        g_bodies_err = std::string("pile ") + std::to_string(r) + ": " + x.what();
      }
      return CC_EHOST;
    }
  return 0;
}

const char* cc_bodies_last_error(void) { return g_bodies_err.c_str(); }

}  // extern "C"
