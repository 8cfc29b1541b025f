// The analysis of a state (fof, src/nbody.js:646-870): friends-of-friends
// groups at linking length 2.4 a, the largest and the second group's
// statistics, the largest body's spin, the disk and the moon, the books and
// the loose list for the pairwise pass.
//
// The split between the device and the host follows what each result
// depends on:
//   device  the hashed cells (the reference's hash, its 2^20-cell table and
//           its within-cell order: particle index ascending), every linked
//           pair enumerated in the reference's order (i ascending, the 27
//           cells dz, dy, dx from -1 to 1, a cell's members in order), the
//           components by lock-free union-find that links the larger root
//           under the smaller (so a component's label is its smallest
//           index, whatever order the links land in), and the squared
//           distances of touching pairs in enumeration order;
//   host    every FP64 sum, in particle-index order as the reference runs
//           it, since a sum's rounding depends on its order.
// A group's members do not depend on how union-find labels it, but which of
// two equally large groups the reference calls the largest (or the second)
// depends on its union-find roots, which depend on the order of its unions.
// When such a tie can change the report, the host replays the reference's
// unions over the device's pair list in the reference's order.
//
// FP64 on both sides without contraction (the build passes -ffp-contract=off
// to the host and the device compiler), so products and sums round as the
// reference's JavaScript numbers do; Math.cbrt and Math.acos are V8's, whose
// algorithms are restated below.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cc_hip.h"
#include "cc_fof_host.h"
#include "cc_internal.h"

namespace cc {

constexpr int FOF_BITS = 20;
constexpr int FOF_T = 1 << FOF_BITS;   // hashed cells (src/nbody.js:648)

// ECMAScript ToInt32: the integer part modulo 2^32, as a signed 32-bit value;
// NaN and infinities give 0.  JavaScript's ^ and Int32Array stores apply it.
__host__ __device__ inline int32_t js_int32(double d) {
  if (!(d == d) || d == INFINITY || d == -INFINITY) return 0;
  if (d > -9.2e18 && d < 9.2e18) return (int32_t)(uint32_t)(uint64_t)(int64_t)d;   // truncates toward zero
  double m = fmod(trunc(d), 4294967296.0);
  if (m < 0) m += 4294967296.0;
  return (int32_t)(uint32_t)m;
}

// the cell hash of src/nbody.js:654 and :667: products in FP64, each operand of ^ taken to int32
__host__ __device__ inline int fof_hash(double x, double y, double z) {
  return (js_int32(x * 73856093.0) ^ js_int32(y * 19349663.0) ^ js_int32(z * 83492791.0)) & (FOF_T - 1);
}

// ------------------------------------------------------------- device part --

// cells of every particle (src/nbody.js:651-656): the cell coordinates as the
// reference's Int32Arrays hold them, the hashed cell from the FP64 floors, the
// cell's count
__global__ void __launch_bounds__(256) k_fof_keys(const float4* __restrict__ pos, int N, double inv, int4* cell,
                                                  int* key, int* cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float4 p = pos[i];
  double qx = floor((double)p.x * inv), qy = floor((double)p.y * inv), qz = floor((double)p.z * inv);
  cell[i] = make_int4(js_int32(qx), js_int32(qy), js_int32(qz), 0);
  int k = fof_hash(qx, qy, qz);
  key[i] = k;
  atomicAdd(&cnt[k], 1);
}

// each particle to a place in its cell's run of `order`; start[k] advances to
// the end of cell k's run (so cell k spans [k ? start[k-1] : 0, start[k]))
__global__ void __launch_bounds__(256) k_fof_place(int N, const int* __restrict__ key, int* start, int* order) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  order[atomicAdd(&start[key[i]], 1)] = i;
}

// a cell's members by index, as the reference's counting sort leaves them
// (src/nbody.js:658-659); a cell holds a handful, insertion sort in place
__global__ void __launch_bounds__(256) k_fof_order(const int* __restrict__ end, int* order) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= FOF_T) return;
  int lo = k ? end[k - 1] : 0, hi = end[k];
  for (int s = lo + 1; s < hi; s++) {
    int v = order[s], t = s - 1;
    while (t >= lo && order[t] > v) { order[t + 1] = order[t]; t--; }
    order[t + 1] = v;
  }
}

__device__ __forceinline__ int uf_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the root of x, halving the path on the way; parent[x] <= x always holds
__device__ inline int uf_find(int* parent, int x) {
  int p = uf_load(&parent[x]);
  while (p != x) {
    int g = uf_load(&parent[p]);
    if (g != p) __hip_atomic_store(&parent[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
    p = g;
  }
  return x;
}

// the larger root under the smaller; a failed exchange hands back the node's
// new parent, from which the walk goes on up
__device__ inline void uf_union(int* parent, int a, int b) {
  a = uf_find(parent, a);
  b = uf_find(parent, b);
  while (a != b) {
    if (a < b) { int t = a; a = b; b = t; }
    int seen = atomicCAS(&parent[a], a, b);
    if (seen == a) return;
    a = uf_find(parent, seen);
    b = uf_find(parent, b);
  }
}

// Every linked pair (j > i, |p_j - p_i|^2 < L^2 in FP64) of particle i, in the
// reference's order (src/nbody.js:663-677).  mode 0: join the pair's groups and
// count i's linked and touching pairs; mode 1: write the touching pairs'
// squared distances (and, with pairs, every linked partner) at i's offsets.
template <int MODE>
__global__ void __launch_bounds__(256) k_fof_pairs(const float4* __restrict__ pos, int N,
                                                   const int4* __restrict__ cell, const int* __restrict__ end,
                                                   const int* __restrict__ order, double L2, double touch2,
                                                   int* parent, int* nLink, int* nTouch,
                                                   const int* __restrict__ linkOff, const int* __restrict__ touchOff,
                                                   int* partner, double* e2Touch) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float4 pi = pos[i];
  double px = pi.x, py = pi.y, pz = pi.z;
  int4 c = cell[i];
  int nl = 0, nt = 0;
  int lo0 = 0, to0 = 0;
  if (MODE == 1) {
    to0 = touchOff[i];
    if (partner) lo0 = linkOff[i];
  }
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        int k = fof_hash((double)c.x + dx, (double)c.y + dy, (double)c.z + dz);
        int hi = end[k];
        for (int s = k ? end[k - 1] : 0; s < hi; s++) {
          int j = order[s];
          if (j <= i) continue;
          float4 pj = pos[j];
          double ex = (double)pj.x - px, ey = (double)pj.y - py, ez = (double)pj.z - pz;
          double e2 = ex * ex + ey * ey + ez * ez;
          if (!(e2 < L2)) continue;
          if (MODE == 0) {
            uf_union(parent, i, j);
          } else if (partner) {
            partner[lo0 + nl] = j;
          }
          if (e2 < touch2) {
            if (MODE == 1) e2Touch[to0 + nt] = e2;
            nt++;
          }
          nl++;
        }
      }
  if (MODE == 0) {
    nLink[i] = nl;
    nTouch[i] = nt;
  }
}

// every particle's label: its component's root, the component's smallest index
__global__ void __launch_bounds__(256) k_fof_label(int N, int* parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int r = uf_find(parent, i);
  __hip_atomic_store(&parent[i], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_fof_iota(int N, int* p) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) p[i] = i;
}

// --------------------------------------------------------------- host part --

struct FofWork {
  int N = 0;
  // device
  int4* cell = nullptr;
  int *key = nullptr, *end = nullptr, *order = nullptr, *parent = nullptr;
  int *nLink = nullptr, *nTouch = nullptr, *blockSums = nullptr;
  int* partner = nullptr;
  double* e2 = nullptr;
  size_t partnerCap = 0, e2Cap = 0;
  // pinned host
  float4 *hP = nullptr, *hV = nullptr, *hD = nullptr;
  float* hR = nullptr;
  int *hLab = nullptr, *hCounts = nullptr;
  double* hE2 = nullptr;
  size_t hE2Cap = 0;
  // a begun analysis (fof_begin): its host half on a thread of its own, over
  // the pinned copies above, filling rep and list; fof_end joins it
  std::thread host;
  bool pending = false;
  std::atomic<bool> done{ false };   // the host half has finished
  FofHostIn hin{};
  FofPick pick;
  cc_report rep{};
  std::vector<int32_t> list;
  std::string hostErr;
  double msCells = 0.0, msCopy = 0.0, msHost = 0.0;
};

#define FCHK(call)                                                                      \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) { err = std::string(#call) + ": " + hipGetErrorString(_e); return CC_EDEVICE; } \
  } while (0)

void fof_destroy(FofWork* w) {
  if (!w) return;
  if (w->host.joinable()) w->host.join();   // a begun analysis' host half reads the pinned buffers
  for (void* p : { (void*)w->cell, (void*)w->key, (void*)w->end, (void*)w->order, (void*)w->parent,
                   (void*)w->nLink, (void*)w->nTouch, (void*)w->blockSums, (void*)w->partner, (void*)w->e2 })
    if (p) (void)hipFree(p);
  for (void* p : { (void*)w->hP, (void*)w->hV, (void*)w->hD, (void*)w->hR, (void*)w->hLab, (void*)w->hCounts,
                   (void*)w->hE2 })
    if (p) (void)hipHostFree(p);
  delete w;
}

static int fof_alloc(FofWork* w, int N, std::string& err) {
  if (w->N == N) return 0;
  if (w->N) { err = "the analysis workspace was made for another N"; return CC_ESTATE; }
  size_t n1 = (size_t)N + 1;
  FCHK(hipMalloc((void**)&w->cell, (size_t)N * sizeof(int4)));
  FCHK(hipMalloc((void**)&w->key, (size_t)N * 4));
  FCHK(hipMalloc((void**)&w->end, ((size_t)FOF_T + 1) * 4));
  FCHK(hipMalloc((void**)&w->order, (size_t)N * 4));
  FCHK(hipMalloc((void**)&w->parent, (size_t)N * 4));
  FCHK(hipMalloc((void**)&w->nLink, n1 * 4));
  FCHK(hipMalloc((void**)&w->nTouch, n1 * 4));
  FCHK(hipMalloc((void**)&w->blockSums, 4096 * 4));
  FCHK(hipHostMalloc((void**)&w->hP, (size_t)N * 16, hipHostMallocDefault));
  FCHK(hipHostMalloc((void**)&w->hV, (size_t)N * 16, hipHostMallocDefault));
  FCHK(hipHostMalloc((void**)&w->hD, (size_t)N * 16, hipHostMallocDefault));
  FCHK(hipHostMalloc((void**)&w->hR, (size_t)N * 4, hipHostMallocDefault));
  FCHK(hipHostMalloc((void**)&w->hLab, (size_t)N * 4, hipHostMallocDefault));
  FCHK(hipHostMalloc((void**)&w->hCounts, 4 * 4, hipHostMallocDefault));
  w->N = N;
  return 0;
}

static int grow_dev(void** p, size_t& cap, size_t need, size_t elem, std::string& err) {
  if (need <= cap) return 0;
  size_t c = std::max(need + need / 4, (size_t)1 << 16);
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  cap = 0;
  FCHK(hipMalloc(p, c * elem));
  cap = c;
  return 0;
}
static int grow_host(void** p, size_t& cap, size_t need, size_t elem, std::string& err) {
  if (need <= cap) return 0;
  size_t c = std::max(need + need / 4, (size_t)1 << 16);
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  cap = 0;
  FCHK(hipHostMalloc(p, c * elem, hipHostMallocDefault));
  cap = c;
  return 0;
}

FofWork* fof_create() { return new FofWork(); }

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// The device half: the cells, the components and the touching pairs on the
// GPU, the arrays the host sums over copied to pinned memory, and which groups
// are the largest and the second (with the reference's unions replayed when a
// tie can reach the report).  Synchronous on stream s; afterwards w->hin and
// w->pick hold everything the host half needs and the device may move on.
static int fof_device(FofWork* w, const FofIn& in, hipStream_t s, std::string& err) {
  const int N = in.N;
  auto tA = std::chrono::steady_clock::now();
  int r = fof_alloc(w, N, err);
  if (r) return r;
  const double L = 2.4 * in.a, L2 = L * L, inv = 1 / L, touch2 = in.touch * in.touch;
  const int blocks = (N + 255) / 256;

  // the hashed cells and their members in index order
  launch_fill_u32(reinterpret_cast<uint32_t*>(w->end), (size_t)FOF_T + 1, 0u, s);
  hipLaunchKernelGGL(k_fof_keys, dim3(blocks), dim3(256), 0, s, in.pos, N, inv, w->cell, w->key, w->end);
  launch_scan_excl(w->end, FOF_T + 1, w->blockSums, s);
  hipLaunchKernelGGL(k_fof_place, dim3(blocks), dim3(256), 0, s, N, w->key, w->end, w->order);
  hipLaunchKernelGGL(k_fof_order, dim3(FOF_T / 256), dim3(256), 0, s, w->end, w->order);
  // the components, and the counts of every particle's linked and touching pairs
  hipLaunchKernelGGL(k_fof_iota, dim3(blocks), dim3(256), 0, s, N, w->parent);
  hipLaunchKernelGGL(k_fof_pairs<0>, dim3(blocks), dim3(256), 0, s, in.pos, N, w->cell, w->end, w->order, L2,
                     touch2, w->parent, w->nLink, w->nTouch, nullptr, nullptr, nullptr, nullptr);
  hipLaunchKernelGGL(k_fof_label, dim3(blocks), dim3(256), 0, s, N, w->parent);
  FCHK(hipMemsetAsync(w->nLink + N, 0, 4, s));
  FCHK(hipMemsetAsync(w->nTouch + N, 0, 4, s));
  launch_scan_excl(w->nTouch, N + 1, w->blockSums, s);
  FCHK(hipGetLastError());
  FCHK(hipMemcpyAsync(w->hCounts, w->nTouch + N, 4, hipMemcpyDeviceToHost, s));
  FCHK(hipStreamSynchronize(s));
  w->msCells = ms_since(tA);
  auto tB = std::chrono::steady_clock::now();
  // the touching pairs' squared distances in the reference's order, and the
  // arrays the host sums over
  const size_t nT = (size_t)w->hCounts[0];
  if ((r = grow_dev((void**)&w->e2, w->e2Cap, nT, 8, err))) return r;
  if ((r = grow_host((void**)&w->hE2, w->hE2Cap, nT, 8, err))) return r;
  hipLaunchKernelGGL(k_fof_pairs<1>, dim3(blocks), dim3(256), 0, s, in.pos, N, w->cell, w->end, w->order, L2,
                     touch2, w->parent, nullptr, nullptr, nullptr, w->nTouch, nullptr, w->e2);
  FCHK(hipGetLastError());
  if (nT) FCHK(hipMemcpyAsync(w->hE2, w->e2, nT * 8, hipMemcpyDeviceToHost, s));
  FCHK(hipMemcpyAsync(w->hLab, w->parent, (size_t)N * 4, hipMemcpyDeviceToHost, s));
  FCHK(hipMemcpyAsync(w->hP, in.pos, (size_t)N * 16, hipMemcpyDeviceToHost, s));
  FCHK(hipMemcpyAsync(w->hV, in.vel, (size_t)N * 16, hipMemcpyDeviceToHost, s));
  FCHK(hipMemcpyAsync(w->hR, in.mat, (size_t)N * 4, hipMemcpyDeviceToHost, s));
  if (in.hostDiag) std::memcpy(w->hD, in.hostDiag, (size_t)N * 16);
  else FCHK(hipMemcpyAsync(w->hD, in.diag, (size_t)N * 16, hipMemcpyDeviceToHost, s));
  FCHK(hipStreamSynchronize(s));
  w->msCopy = ms_since(tB);
  auto tC = std::chrono::steady_clock::now();

  w->hin = FofHostIn{ N, reinterpret_cast<const F4*>(w->hP), reinterpret_cast<const F4*>(w->hV),
                      reinterpret_cast<const F4*>(w->hD), w->hR, w->hLab, w->hE2, nT, in.m, in.a, in.spring,
                      in.touch, in.cap, in.flags };
  const FofHostIn& h = w->hin;
  try {
    w->pick = fof_pick(h);
    if (w->pick.replay) {
      // every linked partner in the reference's order, for the host to replay its unions
      launch_scan_excl(w->nLink, N + 1, w->blockSums, s);
      FCHK(hipGetLastError());
      FCHK(hipMemcpyAsync(w->hCounts + 1, w->nLink + N, 4, hipMemcpyDeviceToHost, s));
      FCHK(hipStreamSynchronize(s));
      const int nLk = w->hCounts[1];
      if ((r = grow_dev((void**)&w->partner, w->partnerCap, (size_t)nLk, 4, err))) return r;
      hipLaunchKernelGGL(k_fof_pairs<1>, dim3(blocks), dim3(256), 0, s, in.pos, N, w->cell, w->end, w->order, L2,
                         touch2, w->parent, nullptr, nullptr, w->nLink, w->nTouch, w->partner, w->e2);
      FCHK(hipGetLastError());
      std::vector<int> off(N + 1), partner((size_t)nLk);
      FCHK(hipMemcpyAsync(off.data(), w->nLink, (size_t)(N + 1) * 4, hipMemcpyDeviceToHost, s));
      if (nLk) FCHK(hipMemcpyAsync(partner.data(), w->partner, (size_t)nLk * 4, hipMemcpyDeviceToHost, s));
      FCHK(hipStreamSynchronize(s));
      fof_pick_replay(h, off.data(), partner.data(), w->pick);
    }
  } catch (const std::exception& x) {
    err = std::string("the analysis' host half: ") + x.what();
    return CC_EDEVICE;
  }
  w->msHost = ms_since(tC);
  return 0;
}

// the host half: every FP64 sum of the report, over the pinned copies
static void fof_host(FofWork* w) {
  auto t0 = std::chrono::steady_clock::now();
  w->hostErr.clear();
  try {
    fof_report(w->hin, w->pick, &w->rep, w->list.data());
  } catch (const std::exception& x) {
    w->hostErr = std::string("the analysis' host half: ") + x.what();
  }
  w->msHost += ms_since(t0);
  w->done.store(true, std::memory_order_release);
}

// a host half still running from fof_begin, ended and its result dropped
static void fof_drop(FofWork* w) {
  if (w->host.joinable()) w->host.join();
  w->pending = false;
}

static int fof_deliver(FofWork* w, cc_report* out, int32_t* list, std::string& err) {
  if (!w->hostErr.empty()) { err = w->hostErr; return CC_EHOST; }
  *out = w->rep;
  if (list && w->rep.ppCount > 0) std::memcpy(list, w->list.data(), (size_t)w->rep.ppCount * 4);
  out->ms[0] = w->msCells;
  out->ms[1] = w->msCopy;
  out->ms[2] = w->msHost;
  return 0;
}

int fof_analyse(FofWork* w, const FofIn& in, cc_report* out, int32_t* list, hipStream_t s, std::string& err) {
  fof_drop(w);   // a synchronous analysis supersedes one begun before it (the reference drops the worker's)
  int r = fof_device(w, in, s, err);
  if (r) return r;
  w->list.assign((size_t)std::max(in.cap, 1), 0);
  fof_host(w);
  return fof_deliver(w, out, list, err);
}

int fof_begin(FofWork* w, const FofIn& in, hipStream_t s, std::string& err) {
  if (w->pending) { err = "an analysis is still running: collect it (cc_analyse_end) first"; return CC_ESTATE; }
  int r = fof_device(w, in, s, err);
  if (r) return r;
  w->list.assign((size_t)std::max(in.cap, 1), 0);
  w->done.store(false, std::memory_order_relaxed);
  try {
    w->host = std::thread(fof_host, w);
  } catch (const std::exception& x) {
    err = std::string("the analysis' host thread: ") + x.what();
    return CC_EHOST;
  }
  w->pending = true;
  return 0;
}

int fof_end(FofWork* w, cc_report* out, int32_t* list, bool wait, bool* ready, std::string& err) {
  *ready = false;
  if (!w->pending) { err = "no analysis was begun"; return CC_ESTATE; }
  if (!wait && !w->done.load(std::memory_order_acquire)) return 0;
  w->host.join();
  w->pending = false;
  *ready = true;
  return fof_deliver(w, out, list, err);
}

bool fof_pending(const FofWork* w) { return w && w->pending; }

}  // namespace cc
