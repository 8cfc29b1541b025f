// The C ABI (include/cc_hip.h): device buffers of one simulation instance,
// uniform derivation as the reference's setStaticUniforms does it, and the
// pass sequence of its step() -- all queued on one HIP stream.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cc_hip.h"
#include "cc_internal.h"

using namespace cc;

// One captured step cycle (meshEvery steps from a mesh step) and the host
// bookkeeping it leaves: the key is everything its launches bake in.
struct CycleGraph {
  std::vector<double> key;
  hipGraphExec_t exec = nullptr;
  int curAfter = 0, gridCurAfter = 0, countedAfter = 0;
};

constexpr int SCHED_CAP = 4096;   // step cycles per schedule upload
// cycles kept captured: a run's looks alternate the seat buffers' parity and whether the first cycle after
// a look takes its rebuild's count, and a Morton sort moves the state's buffers, so keys recur in a handful
constexpr int CYCLE_GRAPHS = 8;

struct cc_ctx {
  cc_params p;
  hipStream_t s = nullptr;
  int N = 0;
  Grid grid;
  size_t cells = 0;
  float4* pos[2] = { nullptr, nullptr };
  float4* vel[2] = { nullptr, nullptr };
  float4* aux[2] = { nullptr, nullptr };
  float* q[2] = { nullptr, nullptr };
  float* mat = nullptr;
  float* matTmp = nullptr;
  int cur = 0;
  GridBuf gb;
  MeshBuf mesh[2];
  int* ppIndex = nullptr;
  float4* ppList = nullptr;
  float* ppBox = nullptr;
  float4* ppBounds = nullptr;   // the pairwise pass's tile boxes
  float4* ppForce = nullptr;
  float* ppSlot = nullptr;
  float2* ppTab = nullptr;
  int ppCount = 0;
  int* ppCountDev = nullptr;   // the loose list's length on the device (what a captured cycle reads)
  float4* diag = nullptr;
  int *mkeys = nullptr, *mslot = nullptr, *mhist = nullptr, *mblock = nullptr, *mperm = nullptr, *minto = nullptr;
  double *bpart = nullptr, *bout = nullptr;
  Box box;                 // host copy of the boxes last set (decides whether box B's mesh is built)
  Box* boxDev = nullptr;   // the boxes the kernels read
  Box* boxSched = nullptr; // a step_n call's boxes, one per step cycle, taken in turn by k_box_next
  int* boxIdx = nullptr;
  Box* boxHost = nullptr;  // pinned staging of the schedule
  CycleGraph graphs[CYCLE_GRAPHS];   // captured step cycles, by what their launches bake in
  int nextGraph = 0;
  int useGraphs = 1;
  // the step pass's visit order (a permutation that only the caches see), re-sorted at a cycle's start once
  // visitEvery steps have passed since the last sort (visitEvery 0: index order)
  int* visit = nullptr;
  int visitEvery = 256;
  int visitAt = -(1 << 30);
  // launch shapes that cannot change a result (cc_autotune, cc_set_tuning): the cell pass's
  // work-groups, the step pass's work-group order
  int cellBlocks = 2048;
  int simOrder = 0;
  // the step pass before a rebuild takes the rebuild's count (fuseCount), and whether the
  // counters of parity gb.cur ^ 1 now hold a count of the present positions (counted)
  int fuseCount = 1;
  int counted = 0;
  // the learned far field, when attached: network, workspace, judge cadence and bound, and per box the gate
  // (1: the network stands in), stats {last error, judged, fallbacks, learned mesh steps}, error sums
  FFNet* ff = nullptr;
  FFWork* ffWork = nullptr;
  int ffJudge = 16;
  float ffTol = 1e-3f;
  int* ffGate = nullptr;
  float* ffStats = nullptr;
  double* ffSums = nullptr;
  float4* ffFar = nullptr;     // per box 3 x 4096 (FF_FAR): the network's farA, farT0, farT1 (judge steps)
  bool ffBatched = false;      // while a batch captures a cycle whose inference and judge it runs itself
  int headN = 0;               // the step whose head (one_step part 1) ran last
  // while a batch that fuses its instances' passes captures a cycle: the step's step pass and seat pass
  // recorded here instead of launched (seatWaits: the recorded seat pass seats a count launched on this
  // ctx's stream in the same step)
  std::vector<SimJob>* simJobs = nullptr;
  std::vector<SeatJob>* seatJobs = nullptr;
  bool seatWaits = false;
  // a judge runs on the second stream beside the steps after it: forked after the exact pass, joined by the end
  // of its cycle (or of the call that forked it), ffPending while it is not joined
  hipEvent_t ffFork = nullptr, ffJoin = nullptr;
  bool ffPending = false;
  int stepNo = 0;
  // the pairwise pass of a mesh step runs on its own stream beside the seats and the mesh
  hipStream_t s2 = nullptr;
  hipEvent_t ppFork = nullptr, ppJoin = nullptr;
  FofWork* fof = nullptr;  // the analysis' workspace, made on its first call
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double lastMs = 0.0;
  std::vector<void*> allocs;
  std::string err;
};

static std::string g_err;
static const int SIM_PLANES_MAX_N = 98304;   // the default step-pass shapes' crossovers (cc_create)
static const int SIM_ROWS_MAX_N = 16384;     // 256 work-groups of 64 particles: one a CU

#define FAIL(ctx, code, ...)                                      \
  do {                                                            \
    char _b[512];                                                 \
    snprintf(_b, sizeof _b, __VA_ARGS__);                         \
    if (ctx) (ctx)->err = _b; else g_err = _b;                    \
    return code;                                                  \
  } while (0)
#define HIPCHK(ctx, call)                                                                  \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess) FAIL(ctx, CC_EDEVICE, "%s: %s", #call, hipGetErrorString(_e));  \
  } while (0)
#define LAUNCHCHK(ctx)                                                                              \
  do {                                                                                              \
    hipError_t _e = hipGetLastError();                                                              \
    if (_e != hipSuccess) FAIL(ctx, CC_EDEVICE, "kernel launch: %s", hipGetErrorString(_e));       \
  } while (0)

static int dalloc(cc_ctx* c, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) FAIL(c, CC_EDEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  c->allocs.push_back(*p);
  return 0;
}
#define ALLOC(ptr, bytes) \
  do { int _r = dalloc(c, (void**)&(ptr), (bytes)); if (_r) return _r; } while (0)

static void grid_make(int G, Grid& g) {   // src/gl.js:73-83
  g.G = G;
  g.GShift = G == 128 ? 7 : 6;
  g.GMask = G - 1;
  g.SX = G == 128 ? 16 : 8;
  g.SXShift = g.SX == 16 ? 4 : 3;
  g.SXMask = g.SX - 1;
  g.AW = g.SX * G;
  g.AH = (G / g.SX) * G;
  g.AWShift = G == 128 ? 11 : 9;
  g.cells = g.AW * g.AH;
}

// The HIP runtime this process runs the library on must be the one it was built
// against, or newer.  A process can hold another copy of libamdhip64: a Python
// wheel (PyTorch's) ships its own, and whichever copy a process loads first
// serves every library that asks for "libamdhip64.so.7" after it.  The graph
// capture and replay paths are tested on the runtime of the build; an older
// copy is refused here rather than run.
static int check_runtime() {
  static int checked = 0, have = 0;   // the loaded runtime cannot change within a process
  if (!checked) {
    if (hipRuntimeGetVersion(&have) != hipSuccess) have = 0;
    checked = 1;
  }
  if (have / 100000 < HIP_VERSION / 100000) {
    char b[320];
    snprintf(b, sizeof b,
             "HIP runtime %d.%d loaded in this process, older than the %d.%d this library was built against (another "
             "copy of libamdhip64 was loaded first, e.g. a Python wheel's: load this library before it)",
             have / 10000000, (have / 100000) % 100, HIP_VERSION_MAJOR, HIP_VERSION_MINOR);
    g_err = b;
    return CC_EDEVICE;
  }
  return 0;
}

static int check_params(cc_ctx* c, const cc_params* p) {
  if (!p) FAIL(c, CC_EINVAL, "params is NULL");
  if (p->N <= 0 || p->W <= 0 || p->H <= 0 || (long long)p->W * p->H != p->N)
    FAIL(c, CC_EINVAL, "N must equal W*H (N=%d W=%d H=%d)", p->N, p->W, p->H);
  if (p->N > (1 << 19))
    FAIL(c, CC_EINVAL, "N=%d exceeds the seat word's 19-bit particle index (524288)", p->N);
  if (p->G != 64 && p->G != 128) FAIL(c, CC_EINVAL, "contact grid G must be 64 or 128, got %d", p->G);
  if (p->meshEvery <= 0) FAIL(c, CC_EINVAL, "meshEvery must be positive");
  if (!(p->m > 0) || !(p->a > 0) || !(p->cell > 0) || !(p->dt > 0)) FAIL(c, CC_EINVAL, "m, a, cell, dt must be > 0");
  return 0;
}

// the step pass's uniforms as setStaticUniforms + passSim set them
static SimU sim_uniforms(const cc_ctx* c, const cc_sim_args* a) {
  const cc_params& p = c->p;
  SimU u;
  memset(&u, 0, sizeof u);
  u.N = c->N;
  u.invCell = (float)(1.0 / p.cell);
  u.cell = (float)p.cell;
  u.dt = (float)p.dt;
  u.touch = (float)(2.0 * p.a);
  u.link2 = (float)(p.cell * p.cell);
  double reach = p.cell + p.seatSlack * p.a;
  u.reach2 = (float)(reach * reach);
  {
    // the prefilter on the 1/480-cell lattice: the reach, plus the half-unit per
    // axis (sqrt(3)/2 in all) that rounding the particle onto the lattice moves it
    double rq = reach * SEAT_Q / p.cell + 1.0;
    u.reachQ2 = (int)(rq * rq);
  }
  u.k = (float)p.k;
  u.c = (float)p.c;
  u.invM = (float)(1.0 / p.m);
  u.damp = (float)p.damp;
  u.creep = (float)p.creep;
  u.tK = (float)p.tK;
  u.kappa = (float)p.kappa;
  u.cpIron = (float)p.cpIron;
  for (int k = 0; k < 3; k++) { u.spin0[k] = (float)p.spin0[k]; u.spin1[k] = (float)p.spin1[k]; }
  u.grid = c->grid;
  u.order = c->simOrder;
  u.visit = c->visitEvery > 0 ? c->visit : nullptr;
  u.gm = (float)p.m;
  u.eps2 = (float)(FH * FH);
  u.field = p.field;
  for (int k = 0; k < 3; k++) u.gField[k] = (float)p.gField[k];
  u.frozenOn = p.frozenOn;
  u.frozen = (float)p.frozen;
  if (a) {
    u.settle = a->settle ? 1.0f : 0.0f;
    u.kick = a->kick ? 1.0f : 0.0f;
    u.gdt = (float)a->gdt;
    for (int k = 0; k < 3; k++) {
      u.bulk0[k] = (float)a->bulk0[k]; u.bulk1[k] = (float)a->bulk1[k];
      u.cen0[k] = (float)a->cen0[k]; u.cen1[k] = (float)a->cen1[k];
    }
  }
  return u;
}

// A count taken for the next rebuild goes stale when the particles move (or
// the cell size changes) before it is seated: its counters back to zero.
static int discard_count(cc_ctx* c) {
  if (!c->counted) return 0;
  launch_slots_discard(c->N, c->gb, c->s);
  LAUNCHCHK(c);
  c->counted = 0;
  return 0;
}

// The runtime's first instantiation of a graph that forks onto a second stream costs milliseconds once per
// process and device (6-9 ms measured, against 0.1-0.5 ms for every later one): a step cycle with the
// pairwise pass beside the mesh is such a graph, so a run would otherwise pay it at its first mesh step with a
// loose list.  The first context on each device pays it at creation instead, with a two-node graph of its own
// (a fill of the context's box index on each stream, the second after the first).
static int warm_graphs(cc_ctx* c) {
  static std::mutex mu;
  static std::vector<int> warmed;
  int dev = 0;
  HIPCHK(c, hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  for (int d : warmed)
    if (d == dev) return 0;
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
  launch_fill_u32(reinterpret_cast<uint32_t*>(c->boxIdx), 1, 0u, c->s);
  hipError_t e = hipEventRecord(c->ppFork, c->s);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->s2, c->ppFork, 0);
  if (e == hipSuccess) launch_fill_u32(reinterpret_cast<uint32_t*>(c->boxIdx), 1, 0u, c->s2);
  if (e == hipSuccess) e = hipEventRecord(c->ppJoin, c->s2);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->s, c->ppJoin, 0);
  hipError_t ee = hipStreamEndCapture(c->s, &g);
  if (e == hipSuccess) e = ee;
  if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipGraphLaunch(x, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (x) (void)hipGraphExecDestroy(x);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess) FAIL(c, CC_EDEVICE, "graph warm-up: %s", hipGetErrorString(e));
  warmed.push_back(dev);
  return 0;
}

extern "C" {

const char* cc_last_error(cc_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int cc_create(const cc_params* p, cc_ctx** out) {
  if (!out) FAIL((cc_ctx*)nullptr, CC_EINVAL, "out is NULL");
  *out = nullptr;
  int r = check_runtime();
  if (r) return r;
  r = check_params(nullptr, p);
  if (r) return r;
  cc_ctx* c = new cc_ctx();
  c->p = *p;
  c->N = p->N;
  // below about 1.5 wavefronts a SIMD of step pass the GPU is not full: a wavefront per z-plane is faster
  // there (57,856: 15 against 21 us), above it the one-wavefront walk (262,144: 36 against 51 us); while the
  // work-groups of nine wavefronts, one per row of three cells, fit the CUs at once, those (16,384: 11 us)
  c->simOrder = p->N <= SIM_ROWS_MAX_N ? 3 : p->N <= SIM_PLANES_MAX_N ? 2 : 0;
  grid_make(p->G, c->grid);
  c->grid.IB = p->N > (1 << 18) ? 19 : 18;
  c->cells = (size_t)c->grid.AW * c->grid.AH;
  hipError_t e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking);
  if (e != hipSuccess) { g_err = std::string("hipStreamCreate: ") + hipGetErrorString(e); delete c; return CC_EDEVICE; }
  size_t N = (size_t)c->N;
  int rr = 0;
  auto step = [&]() -> int {
    for (int b = 0; b < 2; b++) {
      ALLOC(c->pos[b], N * 16); ALLOC(c->vel[b], N * 16); ALLOC(c->aux[b], N * 16); ALLOC(c->q[b], N * 4);
    }
    ALLOC(c->mat, N * 4); ALLOC(c->matTmp, N * 4);
    ALLOC(c->gb.seats, (c->cells + 1) * SLOTS * 4);
    ALLOC(c->gb.cnt[0], c->cells * 4); ALLOC(c->gb.cnt[1], c->cells * 4);
    ALLOC(c->gb.bucket, c->cells * BUCKET_CAP * 4);
    ALLOC(c->gb.occ[0], N * 4); ALLOC(c->gb.occ[1], N * 4);
    ALLOC(c->gb.segN, 2 * ((N + 63) / 64) * 4);
    ALLOC(c->gb.ovf, N * 8); ALLOC(c->gb.ovfN, 8);
    c->gb.idxMask = (1u << c->grid.IB) - 1u;
    for (int b = 0; b < 2; b++) {
      MeshBuf& m = c->mesh[b];
      ALLOC(m.acc, (size_t)FINE_TEXELS * 4 * 8);
      ALLOC(m.fine, (size_t)FINE_TEXELS * 16); ALLOC(m.force, (size_t)FINE_TEXELS * 16);
      ALLOC(m.com, (size_t)FINE_TEXELS * 16); ALLOC(m.ccom, 4096 * 16);
      ALLOC(m.coarse, 4096 * 16); ALLOC(m.farA, 4096 * 16); ALLOC(m.farT0, 4096 * 16); ALLOC(m.farT1, 4096 * 16);
      ALLOC(m.mono, 16);
      ALLOC(m.occ, (size_t)SCRATCH * 4); ALLOC(m.occN, 4); ALLOC(m.coarseDone, 4);
    }
    ALLOC(c->ppIndex, PP_CAP * 4); ALLOC(c->ppList, PP_CAP * 16); ALLOC(c->ppBox, PP_CAP * 4);
    ALLOC(c->ppCountDev, 4);
    ALLOC(c->ppBounds, PP_TILES_MAX * 32);
    ALLOC(c->ppForce, 4 * PP_CAP * 16); ALLOC(c->ppSlot, N * 4); ALLOC(c->ppTab, 64 * 8);
    ALLOC(c->diag, N * 16);
    ALLOC(c->mkeys, N * 4); ALLOC(c->mslot, N * 4); ALLOC(c->mhist, ((1 << 21) + 1) * 4);
    ALLOC(c->mblock, 4096 * 4); ALLOC(c->mperm, N * 4); ALLOC(c->minto, N * 4); ALLOC(c->visit, N * 4);
    ALLOC(c->bpart, 256 * 13 * 8); ALLOC(c->bout, 13 * 8);
    ALLOC(c->boxDev, sizeof(Box)); ALLOC(c->boxIdx, 4);
    ALLOC(c->boxSched, (size_t)SCHED_CAP * sizeof(Box));
    HIPCHK(c, hipHostMalloc((void**)&c->boxHost, (size_t)SCHED_CAP * sizeof(Box), hipHostMallocDefault));
    // the empty state: seats empty, counters and accumulators zero, meshes zero
    launch_fill_u32(c->gb.seats, (c->cells + 1) * SLOTS, EMPTY_SEAT, c->s);
    for (int b = 0; b < 2; b++) HIPCHK(c, hipMemsetAsync(c->gb.cnt[b], 0, c->cells * 4, c->s));
    HIPCHK(c, hipMemsetAsync(c->gb.segN, 0, 2 * ((N + 63) / 64) * 4, c->s));
    HIPCHK(c, hipMemsetAsync(c->gb.ovfN, 0, 8, c->s));
    for (int b = 0; b < 2; b++) {
      MeshBuf& m = c->mesh[b];
      HIPCHK(c, hipMemsetAsync(m.acc, 0, (size_t)FINE_TEXELS * 32, c->s));
      HIPCHK(c, hipMemsetAsync(m.fine, 0, (size_t)FINE_TEXELS * 16, c->s));
      HIPCHK(c, hipMemsetAsync(m.com, 0, (size_t)FINE_TEXELS * 16, c->s));
      HIPCHK(c, hipMemsetAsync(m.ccom, 0, 4096 * 16, c->s));
      HIPCHK(c, hipMemsetAsync(m.force, 0, (size_t)FINE_TEXELS * 16, c->s));
      for (float4* t : { m.coarse, m.farA, m.farT0, m.farT1 }) HIPCHK(c, hipMemsetAsync(t, 0, 4096 * 16, c->s));
      HIPCHK(c, hipMemsetAsync(m.mono, 0, 16, c->s));
      HIPCHK(c, hipMemsetAsync(m.occN, 0, 4, c->s));
      HIPCHK(c, hipMemsetAsync(m.coarseDone, 0, 4, c->s));
    }
    for (int b = 0; b < 2; b++) {
      HIPCHK(c, hipMemsetAsync(c->aux[b], 0, N * 16, c->s));
      HIPCHK(c, hipMemsetAsync(c->q[b], 0, N * 4, c->s));
    }
    launch_fill_u32(reinterpret_cast<uint32_t*>(c->ppIndex), PP_CAP, 0xFFFFFFFFu, c->s);
    launch_fill_f32(c->ppSlot, N, -1.0f, c->s);
    launch_iota(c->visit, N, c->s);
    HIPCHK(c, hipMemsetAsync(c->ppForce, 0, 4 * PP_CAP * 16, c->s));
    HIPCHK(c, hipMemsetAsync(c->ppList, 0, PP_CAP * 16, c->s));
    HIPCHK(c, hipMemsetAsync(c->ppTab, 0, 64 * 8, c->s));
    HIPCHK(c, hipMemsetAsync(c->ppCountDev, 0, 4, c->s));
    HIPCHK(c, hipMemsetAsync(c->diag, 0, N * 16, c->s));
    LAUNCHCHK(c);
    HIPCHK(c, hipEventCreate(&c->ev0));
    HIPCHK(c, hipEventCreate(&c->ev1));
    HIPCHK(c, hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->ppFork, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ppJoin, hipEventDisableTiming));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return warm_graphs(c);
  };
  rr = step();
  if (rr) { g_err = c->err; cc_destroy(c); return rr; }
  memset(&c->box, 0, sizeof c->box);
  c->box.invH = (float)(1.0 / FH);
  launch_put_box(c->box, c->boxDev, c->s);
  if (hipStreamSynchronize(c->s) != hipSuccess) { g_err = "box upload failed"; cc_destroy(c); return CC_EDEVICE; }
  *out = c;
  return 0;
}

int cc_destroy(cc_ctx* c) {
  if (!c) return 0;
  if (c->s) hipStreamSynchronize(c->s);
  if (c->s2) hipStreamSynchronize(c->s2);   // both streams idle before their buffers go
  for (CycleGraph& g : c->graphs) if (g.exec) hipGraphExecDestroy(g.exec);
  if (c->ffWork) ffwork_destroy(c->ffWork);
  if (c->fof) fof_destroy(c->fof);
  if (c->boxHost) hipHostFree(c->boxHost);
  for (void* p : c->allocs) hipFree(p);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->ppFork) hipEventDestroy(c->ppFork);
  if (c->ppJoin) hipEventDestroy(c->ppJoin);
  if (c->ffFork) hipEventDestroy(c->ffFork);
  if (c->ffJoin) hipEventDestroy(c->ffJoin);
  if (c->s2) hipStreamDestroy(c->s2);
  if (c->s) hipStreamDestroy(c->s);
  delete c;
  return 0;
}

int cc_set_params(cc_ctx* c, const cc_params* p) {
  if (!c) return CC_EINVAL;
  int r = check_params(c, p);
  if (r) return r;
  if (p->N != c->p.N || p->W != c->p.W || p->G != c->p.G) FAIL(c, CC_EINVAL, "N, W and G are fixed at create");
  if (p->cell != c->p.cell && (r = discard_count(c))) return r;
  c->p = *p;
  return 0;
}

int cc_upload_state(cc_ctx* c, const float* pos, const float* vel, const float* mat, const float* q) {
  if (!c) return CC_EINVAL;
  int r;
  if (pos && (r = discard_count(c))) return r;
  size_t N = (size_t)c->N;
  if (pos) HIPCHK(c, hipMemcpyAsync(c->pos[c->cur], pos, N * 16, hipMemcpyHostToDevice, c->s));
  if (vel) HIPCHK(c, hipMemcpyAsync(c->vel[c->cur], vel, N * 16, hipMemcpyHostToDevice, c->s));
  if (mat) HIPCHK(c, hipMemcpyAsync(c->mat, mat, N * 4, hipMemcpyHostToDevice, c->s));
  if (q) HIPCHK(c, hipMemcpyAsync(c->q[c->cur], q, N * 4, hipMemcpyHostToDevice, c->s));
  // new positions: the visit order is re-sorted at the next cycle's start, and a learned far field's bases,
  // recorded from the old ones, serve no prediction before the next judge
  if (pos && c->ffWork) ff_forget_bases(c->ffWork, c->s);
  HIPCHK(c, hipStreamSynchronize(c->s));
  if (pos) c->visitAt = -(1 << 30);
  return 0;
}

int cc_download_state(cc_ctx* c, float* pos, float* vel, float* aux, float* q, float* mat) {
  if (!c) return CC_EINVAL;
  size_t N = (size_t)c->N;
  if (pos) HIPCHK(c, hipMemcpyAsync(pos, c->pos[c->cur], N * 16, hipMemcpyDeviceToHost, c->s));
  if (vel) HIPCHK(c, hipMemcpyAsync(vel, c->vel[c->cur], N * 16, hipMemcpyDeviceToHost, c->s));
  if (aux) HIPCHK(c, hipMemcpyAsync(aux, c->aux[c->cur], N * 16, hipMemcpyDeviceToHost, c->s));
  if (q) HIPCHK(c, hipMemcpyAsync(q, c->q[c->cur], N * 4, hipMemcpyDeviceToHost, c->s));
  if (mat) HIPCHK(c, hipMemcpyAsync(mat, c->mat, N * 4, hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

int cc_set_boxes(cc_ctx* c, const double LA[3], const double LB[3], int boxB) {
  if (!c || !LA || !LB) return CC_EINVAL;
  for (int k = 0; k < 3; k++) { c->box.LA[k] = (float)LA[k]; c->box.LB[k] = (float)LB[k]; }
  c->box.boxB = boxB ? 1.0f : 0.0f;
  c->box.invH = (float)(1.0 / FH);
  launch_put_box(c->box, c->boxDev, c->s);
  LAUNCHCHK(c);
  return 0;
}

static float4* mesh_ptr(cc_ctx* c, int box, int which, size_t* bytes) {
  MeshBuf& m = c->mesh[box];
  switch (which) {
    case 0: *bytes = (size_t)FINE_TEXELS * 16; return m.fine;
    case 1: *bytes = 4096 * 16; return m.coarse;
    case 2: *bytes = 16; return m.mono;
    case 3: *bytes = 4096 * 16; return m.farA;
    case 4: *bytes = 4096 * 16; return m.farT0;
    case 5: *bytes = 4096 * 16; return m.farT1;
    case 6: *bytes = (size_t)FINE_TEXELS * 16; return m.force;
    default: return nullptr;
  }
}

int cc_set_mono(cc_ctx* c, int box, const float mono[4]) { return cc_write_mesh(c, box, 2, mono); }

int cc_read_mesh(cc_ctx* c, int box, int which, float* out) {
  if (!c || !out || box < 0 || box > 1) return CC_EINVAL;
  size_t bytes;
  float4* p = mesh_ptr(c, box, which, &bytes);
  if (!p) FAIL(c, CC_EINVAL, "unknown mesh texture %d", which);
  HIPCHK(c, hipMemcpyAsync(out, p, bytes, hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

int cc_write_mesh(cc_ctx* c, int box, int which, const float* in) {
  if (!c || !in || box < 0 || box > 1) return CC_EINVAL;
  size_t bytes;
  float4* p = mesh_ptr(c, box, which, &bytes);
  if (!p) FAIL(c, CC_EINVAL, "unknown mesh texture %d", which);
  HIPCHK(c, hipMemcpyAsync(p, in, bytes, hipMemcpyHostToDevice, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  if (which == 0 || which == 1) {   // a written fine or coarse mesh: its centres of mass as a deposit would leave them
    MeshBuf& m = c->mesh[box];
    if (which == 0) launch_centres(m.fine, m.com, FINE_TEXELS, c->s);
    else launch_centres(m.coarse, m.ccom, 4096, c->s);
    LAUNCHCHK(c);
  }
  if (which == 0) {   // a written fine mesh: its occupied list as a deposit would leave it
    std::vector<int> occ;
    const float* f = in;
    for (int t = 0; t < SCRATCH; t++) if (f[t * 4 + 3] > 0.0f) occ.push_back(t);
    int n = (int)occ.size();
    if (n) HIPCHK(c, hipMemcpyAsync(c->mesh[box].occ, occ.data(), occ.size() * 4, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(c->mesh[box].occN, &n, 4, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
  }
  return 0;
}

int cc_slots(cc_ctx* c) {
  if (!c) return CC_EINVAL;
  int r;
  if ((r = discard_count(c))) return r;
  launch_slots(c->pos[c->cur], c->N, (float)(1.0 / c->p.cell), c->grid, c->gb, c->s);
  LAUNCHCHK(c);
  return 0;
}

int cc_read_seats(cc_ctx* c, uint32_t* seatA, uint32_t* seatB) {
  if (!c || !seatA || !seatB) return CC_EINVAL;
  std::vector<uint32_t> s(c->cells * SLOTS);
  HIPCHK(c, hipMemcpyAsync(s.data(), c->gb.seats, s.size() * 4, hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  // a cell's eight words are its even seats (atlas A's four channels) then its odd ones (atlas B's)
  for (size_t h = 0; h < c->cells; h++)
    for (int k = 0; k < 4; k++) { seatA[h * 4 + k] = s[h * 8 + k]; seatB[h * 4 + k] = s[h * 8 + 4 + k]; }
  return 0;
}

int cc_deposit(cc_ctx* c, int box) {
  if (!c || box < 0 || box > 1) return CC_EINVAL;
  launch_deposit(c->pos[c->cur], c->mat, c->N, c->visitEvery > 0 ? c->visit : nullptr, c->boxDev, box,
                 (float)(FH * FH), c->mesh[box], false, c->s);
  LAUNCHCHK(c);
  return 0;
}
int cc_coarse(cc_ctx* c, int box) {
  if (!c || box < 0 || box > 1) return CC_EINVAL;
  launch_coarse(c->mesh[box], c->s);
  LAUNCHCHK(c);
  return 0;
}
int cc_cell(cc_ctx* c, int box) {
  if (!c || box < 0 || box > 1) return CC_EINVAL;
  launch_cell(c->mesh[box], (float)(0.5 * FH * FH), (float)c->p.m, c->cellBlocks, false, c->s);
  LAUNCHCHK(c);
  return 0;
}
constexpr int FF_FAR = 3 * 4096;   // float4 per box of ffFar

// A judge still running on the second stream joined into stream s (work on s
// after this waits for it).
static int ff_join(cc_ctx* c, hipStream_t s) {
  if (!c->ffPending) return 0;
  HIPCHK(c, hipStreamWaitEvent(s, c->ffJoin, 0));
  c->ffPending = false;
  return 0;
}

// A box's deposit and coarse mesh (the deposit also zeroes the force mesh's
// empty texels, which the cell pass leaves).
static int box_head(cc_ctx* c, int b) {
  MeshBuf& m = c->mesh[b];
  launch_deposit(c->pos[c->cur], c->mat, c->N, c->visitEvery > 0 ? c->visit : nullptr, c->boxDev, b,
                 (float)(FH * FH), m, true, c->s);
  LAUNCHCHK(c);
  return cc_coarse(c, b);
}
// A box's far field and per-cell pass, with a learned far field attached
// decided by stepIdx (see mesh_all).  ffBatched: the batch predicts and judges
// every instance's boxes in passes of its own, so only the exact far field (as
// the gate leaves it) and the cell pass run here.
static int box_tail(cc_ctx* c, int b, int stepIdx) {
  MeshBuf& m = c->mesh[b];
  float eps2 = (float)(0.5 * FH * FH), gm = (float)c->p.m;
  if (!c->ff || stepIdx < 0) {
    launch_cell(m, eps2, gm, c->cellBlocks, true, c->s);
  } else if (stepIdx % c->ffJudge == 0) {
    launch_far(m, m.farA, m.farT0, m.farT1, nullptr, c->s);
    launch_cell_only(m, m.farA, m.farT0, m.farT1, m.force, eps2, gm, c->cellBlocks, true, c->s);
    if (!c->ffBatched) {
      // The judge: the network's prediction from the box's base, its error against this exact pass and the
      // verdict, then this pass becomes the base.  Nothing of this step waits for it, so it runs on the second
      // stream (after the pairwise pass there) beside the steps that follow, joined by the end of the cycle;
      // it reads this pass's mesh and box, which only the next cycle overwrites.
      float4 *A = c->ffFar + FF_FAR * b, *T0 = A + 4096, *T1 = A + 8192;
      HIPCHK(c, hipEventRecord(c->ffFork, c->s));
      HIPCHK(c, hipStreamWaitEvent(c->s2, c->ffFork, 0));
      ff_predict(c->ff, c->ffWork, b, m.coarse, m.mono, c->boxDev, A, T0, T1, nullptr, c->s2);
      ff_judge_rebase(c->ffWork, b, m.fine, m.com, m.ccom, m.coarse, A, m.farA, m.farT0, m.farT1, m.force,
                      c->boxDev, gm, c->ffSums + 2 * b, c->ffTol, c->ffStats + FF_STATS * b, c->ffGate + b, c->s2);
      HIPCHK(c, hipEventRecord(c->ffJoin, c->s2));
      c->ffPending = true;
    }
  } else {
    // the prediction first: it clears the gate when the base cannot serve; then the exact far field, for
    // every block when the network does not stand in, else for the blocks it has no base for
    if (!c->ffBatched)
      ff_predict(c->ff, c->ffWork, b, m.coarse, m.mono, c->boxDev, m.farA, m.farT0, m.farT1, c->ffGate + b, c->s);
    launch_far(m, m.farA, m.farT0, m.farT1, c->ffGate + b, c->s, ff_base_features(c->ffWork, b));
    launch_count_learned(c->ffGate + b, c->ffStats + FF_STATS * b, c->s);
    launch_cell_only(m, m.farA, m.farT0, m.farT1, m.force, eps2, gm, c->cellBlocks, true, c->s);
  }
  LAUNCHCHK(c);
  return 0;
}
static int boxes_up(const cc_ctx* c) { return c->box.boxB > 0.5f ? 2 : 1; }
// Deposit, coarse and per-cell passes of each box that is up.  With a
// learned far field attached, stepIdx (the step of this mesh pass) decides:
// on a judge step the exact far field makes the force mesh, the network's
// prediction from the box's base is measured against it (which sets the gate),
// and the exact pass becomes the base; otherwise the gate picks the exact far
// field or the network's.  stepIdx < 0: exact only.
static int mesh_all(cc_ctx* c, int stepIdx = -1) {
  int r;
  for (int b = 0; b < boxes_up(c); b++)
    if ((r = box_head(c, b)) || (r = box_tail(c, b, stepIdx))) return r;
  return 0;
}

int cc_mesh(cc_ctx* c) {
  if (!c) return CC_EINVAL;
  return mesh_all(c);
}

int cc_set_pp_table(cc_ctx* c, const float* tab) {
  if (!c || !tab) return CC_EINVAL;
  HIPCHK(c, hipMemcpyAsync(c->ppTab, tab, 64 * 8, hipMemcpyHostToDevice, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

int cc_pp(cc_ctx* c) {
  if (!c) return CC_EINVAL;
  if (!c->ppCount) return 0;
  launch_pp(c->pos[c->cur], c->mat, c->N, c->boxDev, c->ppIndex, c->ppCountDev, c->ppCount, (float)c->p.m, (float)(5.0 * FH),
            (float)(c->p.a * c->p.a), c->ppTab, c->ppList, c->ppBox, c->ppBounds, c->ppForce, c->s);
  LAUNCHCHK(c);
  return 0;
}

int cc_set_pp(cc_ctx* c, const int32_t* list, int count) {
  if (!c || count < 0 || count > PP_CAP || (count > 0 && !list)) FAIL(c, CC_EINVAL, "loose list of %d (cap %d)", count, PP_CAP);
  std::vector<int32_t> idx(PP_CAP, -1);
  for (int k = 0; k < count; k++) {
    if (list[k] < 0 || list[k] >= c->N) FAIL(c, CC_EINVAL, "loose list entry %d = %d out of range", k, list[k]);
    idx[k] = list[k];
  }
  HIPCHK(c, hipMemcpyAsync(c->ppIndex, idx.data(), PP_CAP * 4, hipMemcpyHostToDevice, c->s));
  launch_pp_slot(c->N, c->ppIndex, count, c->ppSlot, c->s);
  LAUNCHCHK(c);
  HIPCHK(c, hipMemcpyAsync(c->ppCountDev, &count, 4, hipMemcpyHostToDevice, c->s));
  c->ppCount = count;
  HIPCHK(c, hipStreamSynchronize(c->s));
  return count ? cc_pp(c) : 0;
}

int cc_read_pp(cc_ctx* c, float* list, float* box, float* force) {
  if (!c) return CC_EINVAL;
  if (list) HIPCHK(c, hipMemcpyAsync(list, c->ppList, PP_CAP * 16, hipMemcpyDeviceToHost, c->s));
  if (box) HIPCHK(c, hipMemcpyAsync(box, c->ppBox, PP_CAP * 4, hipMemcpyDeviceToHost, c->s));
  if (force) HIPCHK(c, hipMemcpyAsync(force, c->ppForce, 4 * PP_CAP * 16, hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

// The step pass; count: it also takes the next rebuild's count on the
// positions it writes (the counters of parity gb.cur ^ 1, zero before it).
static int sim_pass(cc_ctx* c, const cc_sim_args* a, bool count = false) {
  int r;
  if ((r = discard_count(c))) return r;   // the particles move: a count of theirs is stale
  SimU u = sim_uniforms(c, a);
  int s = c->cur, d = c->cur ^ 1;
  SlotsOut o = slots_out(c->gb, c->N, c->gb.cur ^ 1);
  if (c->simJobs) {
    c->simJobs->push_back(sim_job(u, c->boxDev, c->pos[s], c->vel[s], c->mat, c->q[s], c->gb.seats, c->mesh[0],
                                  c->mesh[1], c->ppSlot, c->ppForce, c->pos[d], c->vel[d], c->aux[d], c->q[d],
                                  count ? &o : nullptr));
  } else {
    launch_sim(u, c->boxDev, c->pos[s], c->vel[s], c->mat, c->q[s], c->gb.seats, c->mesh[0], c->mesh[1], c->ppSlot,
               c->ppForce, c->pos[d], c->vel[d], c->aux[d], c->q[d], count ? &o : nullptr, c->s);
    LAUNCHCHK(c);
  }
  c->cur = d;
  c->counted = count ? 1 : 0;
  return 0;
}

int cc_sim(cc_ctx* c, const cc_sim_args* a) {
  if (!c || !a) return CC_EINVAL;
  return sim_pass(c, a);
}

// the visit order re-sorted when due, at a cycle's start (outside any captured graph: the sort runs on the
// stream between cycles, and the graphs read the order's buffer at their replay)
static int visit_due(cc_ctx* c, hipStream_t s) {
  if (c->visitEvery <= 0 || c->stepNo % c->p.meshEvery != 0 || c->stepNo - c->visitAt < c->visitEvery) return 0;
  launch_visit(c->pos[c->cur], c->N, (float)(1.0 / c->p.cell), c->mkeys, c->mslot, c->mhist, c->mblock, c->visit, s);
  LAUNCHCHK(c);
  c->visitAt = c->stepNo;
  return 0;
}

// One step.  part 0: the whole step; 1: a mesh step up to its boxes' coarse
// meshes; 2: the rest of the mesh step whose head ran last (a batch predicts
// every instance's far field in one pass between the two).
static int one_step(cc_ctx* c, int settle, int kick, const cc_sim_args* a, int part = 0) {
  const int n = part == 2 ? c->headN : c->stepNo++;
  int r;
  bool mesh = !c->p.field && n % c->p.meshEvery == 0;
  // The pairwise pass reads the positions, masses and boxes and writes only its
  // own buffers, as the seats and the mesh do theirs: on a mesh step it runs on
  // the second stream beside them (a fork and a join, which a captured cycle
  // keeps as graph edges), and the step pass waits for all of them.
  bool side = mesh && c->ppCount > 0;
  if (part != 2) {
    if (mesh && (r = ff_join(c, c->s))) return r;
    if (side) {
      HIPCHK(c, hipEventRecord(c->ppFork, c->s));
      HIPCHK(c, hipStreamWaitEvent(c->s2, c->ppFork, 0));
      launch_pp(c->pos[c->cur], c->mat, c->N, c->boxDev, c->ppIndex, c->ppCountDev, c->ppCount, (float)c->p.m, (float)(5.0 * FH),
                (float)(c->p.a * c->p.a), c->ppTab, c->ppList, c->ppBox, c->ppBounds, c->ppForce, c->s2);
      LAUNCHCHK(c);
      HIPCHK(c, hipEventRecord(c->ppJoin, c->s2));
    }
    if (n % 2 == 0) {   // a rebuild: seated from the count the last step pass took, or counted and seated here
      if (c->counted) {
        if (c->seatJobs) {
          c->seatJobs->push_back(seat_job(c->N, c->gb));
        } else {
          launch_slots_seat(c->N, c->gb, c->s);
          LAUNCHCHK(c);
        }
        c->counted = 0;
      } else if (c->seatJobs) {
        launch_slots_count(c->pos[c->cur], c->N, (float)(1.0 / c->p.cell), c->grid, c->gb, c->s);
        LAUNCHCHK(c);
        c->seatJobs->push_back(seat_job(c->N, c->gb));
        c->seatWaits = true;
      } else if ((r = cc_slots(c))) {
        return r;
      }
    }
  }
  const bool count = c->fuseCount && n % 2 == 1;   // the next step rebuilds: this step pass takes its count
  if (c->p.field) {   // constant gravity: no mesh, and the field acts on every step
    cc_sim_args b = *a;
    b.settle = settle;
    b.kick = kick;
    b.gdt = c->p.dt;
    return sim_pass(c, &b, count);
  }
  if (mesh) {
    if (part == 1) {
      for (int bx = 0; bx < boxes_up(c); bx++)
        if ((r = box_head(c, bx))) return r;
      c->headN = n;
      return 0;
    }
    if (part == 2) {
      for (int bx = 0; bx < boxes_up(c); bx++)
        if ((r = box_tail(c, bx, n))) return r;
    } else if ((r = mesh_all(c, n))) {
      return r;
    }
  }
  if (side) HIPCHK(c, hipStreamWaitEvent(c->s, c->ppJoin, 0));
  cc_sim_args b = *a;
  b.settle = settle;
  b.kick = kick;
  b.gdt = mesh ? c->p.meshEvery * c->p.dt : 0.0;
  if ((r = sim_pass(c, &b, count))) return r;
  // a judge forked on this cycle's mesh step is joined by the cycle's end (inside a captured cycle too)
  return n % c->p.meshEvery == c->p.meshEvery - 1 ? ff_join(c, c->s) : 0;
}

int cc_step(cc_ctx* c, int settle, int kick, const cc_sim_args* a) {
  if (!c || !a) return CC_EINVAL;
  int r = visit_due(c, c->s);
  if (!r) r = one_step(c, settle, kick, a);
  if (!r) r = ff_join(c, c->s);
  return r;
}

// What a captured cycle's launches bake in: the uniforms (params and the
// call's settle/kick/bulk/centre values), the loose-list count (its grid), box
// B (whether its mesh is built), the parities of the state and seat-list
// double buffers and the mass array (swapped by the Morton permutation), and
// the step number's parity (which steps of the cycle rebuild the seats and
// which take the next rebuild's count: with an odd meshEvery a cycle can
// start on either).
static std::vector<double> cycle_key(const cc_ctx* c, const cc_sim_args* a) {
  const cc_params& p = c->p;
  std::vector<double> k = { (double)p.N, (double)p.G, p.m, p.a, p.cell, p.dt, p.k, p.c, p.creep, p.damp, p.tK,
                            p.kappa, p.cpIron, p.seatSlack, (double)p.field, (double)p.frozenOn, p.frozen,
                            (double)p.meshEvery, (double)a->settle, (double)a->kick };
  for (int i = 0; i < 3; i++) {
    k.push_back(p.spin0[i]); k.push_back(p.spin1[i]); k.push_back(p.gField[i]);
    k.push_back(a->bulk0[i]); k.push_back(a->bulk1[i]); k.push_back(a->cen0[i]); k.push_back(a->cen1[i]);
  }
  k.push_back((double)(c->ppCount > 0));   // the length itself is read on the device
  k.push_back((double)(c->visitEvery > 0));
  k.push_back((double)(c->box.boxB > 0.5f));
  k.push_back((double)c->cur);
  k.push_back((double)c->gb.cur);
  k.push_back((double)(uintptr_t)c->mat);
  k.push_back((double)(uintptr_t)c->ff);
  k.push_back((double)c->cellBlocks);
  k.push_back((double)c->simOrder);
  k.push_back((double)c->fuseCount);
  k.push_back((double)c->counted);
  k.push_back((double)(c->stepNo & 1));
  if (c->ff) {
    k.push_back((double)c->ffJudge);
    k.push_back((double)c->ffTol);
    k.push_back((double)(c->stepNo % c->ffJudge));
  }
  return k;
}

// The cycle graph for the ctx's present state, captured on first use: the
// box of the cycle (k_box_next), then meshEvery steps from a mesh step.  The
// capture runs the host side of the steps without executing them, so the
// host bookkeeping is put back and the graph's net effect recorded.
static int cycle_graph(cc_ctx* c, const cc_sim_args* a, CycleGraph** out) {
  std::vector<double> key = cycle_key(c, a);
  for (CycleGraph& g : c->graphs)
    if (g.exec && g.key == key) { *out = &g; return 0; }
  CycleGraph& g = c->graphs[c->nextGraph];
  c->nextGraph = (c->nextGraph + 1) % CYCLE_GRAPHS;
  if (g.exec) {   // its last launch may still run: let it end first
    HIPCHK(c, hipStreamSynchronize(c->s));
    hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
  }
  int stepNo = c->stepNo, cur = c->cur, gcur = c->gb.cur, counted = c->counted;
  HIPCHK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
  launch_box_next(c->boxSched, c->boxIdx, c->boxDev, c->s);
  int r = 0;
  for (int i = 0; i < c->p.meshEvery && !r; i++) r = one_step(c, a->settle, a->kick, a);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(c->s, &graph);
  g.curAfter = c->cur;
  g.gridCurAfter = c->gb.cur;
  g.countedAfter = c->counted;
  c->stepNo = stepNo; c->cur = cur; c->gb.cur = gcur; c->counted = counted;
  if (r) { if (graph) hipGraphDestroy(graph); return r; }
  if (e != hipSuccess) FAIL(c, CC_EDEVICE, "step cycle capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) { g.exec = nullptr; FAIL(c, CC_EDEVICE, "step cycle instantiate: %s", hipGetErrorString(e)); }
  g.key = key;
  *out = &g;
  return 0;
}

// n steps: eager steps up to the next mesh step, whole cycles as replays of
// the captured cycle graph (one launch per meshEvery steps, each cycle's box
// from the schedule uploaded once for the call), eager steps for the rest.
int cc_step_n(cc_ctx* c, int n, const cc_sim_args* a, const double* boxL, int nBoxL) {
  if (!c || !a || n < 0 || nBoxL < 0) return CC_EINVAL;
  HIPCHK(c, hipEventRecord(c->ev0, c->s));
  const int M = c->p.meshEvery;
  int used = 0, i = 0, r;
  auto next_box = [&](Box& b) {
    b = c->box;
    if (boxL && used < nBoxL) {
      const double* L = boxL + 7 * used++;
      for (int k = 0; k < 3; k++) { b.LA[k] = (float)L[k]; b.LB[k] = (float)L[3 + k]; }
      b.boxB = L[6] > 0.5 ? 1.0f : 0.0f;
      b.invH = (float)(1.0 / FH);
    }
  };
  auto eager = [&]() -> int {
    if (c->stepNo % M == 0 && boxL && used < nBoxL) {
      const double* L = boxL + 7 * used++;
      int e = cc_set_boxes(c, L, L + 3, L[6] > 0.5);
      if (e) return e;
    }
    int e = visit_due(c, c->s);
    if (e) return e;
    return one_step(c, a->settle, a->kick, a);
  };
  while (i < n && c->stepNo % M != 0) { if ((r = eager())) return r; i++; }
  int cycles = c->useGraphs ? (n - i) / M : 0;
  while (cycles > 0) {   // in chunks of the schedule's capacity; the staging is reused once the last upload is done
    int chunk = cycles < SCHED_CAP ? cycles : SCHED_CAP;
    HIPCHK(c, hipStreamSynchronize(c->s));
    for (int k = 0; k < chunk; k++) next_box(c->boxHost[k]);
    HIPCHK(c, hipMemcpyAsync(c->boxSched, c->boxHost, (size_t)chunk * sizeof(Box), hipMemcpyHostToDevice, c->s));
    launch_fill_u32(reinterpret_cast<uint32_t*>(c->boxIdx), 1, 0u, c->s);
    LAUNCHCHK(c);
    for (int k = 0; k < chunk; k++) {
      c->box = c->boxHost[k];
      if ((r = visit_due(c, c->s))) return r;
      if ((r = ff_join(c, c->s))) return r;   // outside any capture
      CycleGraph* g = nullptr;
      if ((r = cycle_graph(c, a, &g))) return r;
      HIPCHK(c, hipGraphLaunch(g->exec, c->s));
      c->cur = g->curAfter;
      c->gb.cur = g->gridCurAfter;
      c->counted = g->countedAfter;
      c->stepNo += M;
    }
    i += chunk * M;
    cycles -= chunk;
  }
  while (i < n) { if ((r = eager())) return r; i++; }
  if ((r = ff_join(c, c->s))) return r;
  HIPCHK(c, hipEventRecord(c->ev1, c->s));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->lastMs = ms;
  return 0;
}

int cc_time_pass(cc_ctx* c, int pass, int reps, double* ms) {
  if (!c || !ms || reps <= 0 || pass < 0 || pass > 6) return CC_EINVAL;
  cc_sim_args a;
  memset(&a, 0, sizeof a);
  HIPCHK(c, hipEventRecord(c->ev0, c->s));
  for (int i = 0; i < reps; i++) {
    int r = 0;
    if (pass == 6) {   // the state's positions into the spare aux buffer: N x 16 B read, N x 16 B written
      launch_copy4(c->pos[c->cur], c->aux[c->cur ^ 1], c->N, c->s);
      LAUNCHCHK(c);
    } else {
      r = pass == 0 ? sim_pass(c, &a) : pass == 1 ? cc_slots(c) : pass == 2 ? mesh_all(c) : pass == 3 ? cc_pp(c)
        : mesh_all(c, pass == 4 ? c->p.meshEvery : 0);
      if (!r) r = ff_join(c, c->s);   // a judge pass is timed whole
    }
    if (r) return r;
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->s));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float t = 0;
  HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = t / reps;
  return 0;
}

static const int CELL_BLOCKS[4] = { 512, 1024, 2048, 4096 };

int cc_set_tuning(cc_ctx* c, int cellBlocks, int simOrder, int fuseCount) {
  if (!c) return CC_EINVAL;
  bool known = false;
  for (int b : CELL_BLOCKS) known = known || b == cellBlocks;
  if (!known || simOrder < 0 || simOrder > 3 || (fuseCount != 0 && fuseCount != 1))
    FAIL(c, CC_EINVAL, "unknown tuning (%d, %d, %d)", cellBlocks, simOrder, fuseCount);
  c->cellBlocks = cellBlocks;
  c->simOrder = simOrder;
  c->fuseCount = fuseCount;
  return 0;
}

// Each knob timed on the present state in isolation (the mesh pass for the
// cell pass's grid, the step pass for its order), the others held; a
// candidate replaces the current choice only when 2 % faster, so timing noise
// does not move it.  The state is left as it was: the step pass is timed
// writing the spare buffers without taking them as the state, and the mesh
// pass only where the next step is a mesh step, which rebuilds the mesh
// buffers before anything reads them -- all but the monopoles, which the next
// deposit reads (the strays' pull), so they are put back -- and with no
// learned far field attached, whose judge keeps statistics across passes.
int cc_autotune(cc_ctx* c, int reps, cc_tune_report* out) {
  if (!c || !out || reps <= 0) return CC_EINVAL;
  if (c->p.field || c->ff || c->stepNo % c->p.meshEvery != 0)
    FAIL(c, CC_ESTATE, "tune at a mesh step, without constant gravity or a learned far field attached");
  memset(out, 0, sizeof *out);
  auto timed = [&](auto&& body, double* ms) -> int {
    HIPCHK(c, hipEventRecord(c->ev0, c->s));
    for (int i = 0; i < reps; i++) {
      int r = body();
      if (r) return r;
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->s));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t / reps;
    return 0;
  };
  int r;
  float4 mono[2];
  for (int b = 0; b < 2; b++) HIPCHK(c, hipMemcpyAsync(&mono[b], c->mesh[b].mono, 16, hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  r = 0;
  for (int k = 0; k < 4 && !r; k++) {
    c->cellBlocks = CELL_BLOCKS[k];
    r = timed([&]() { return mesh_all(c); }, &out->cellMs[k]);
  }
  for (int b = 0; b < 2; b++) HIPCHK(c, hipMemcpyAsync(c->mesh[b].mono, &mono[b], 16, hipMemcpyHostToDevice, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  if (r) return r;
  int best = 2;
  for (int k = 0; k < 4; k++)
    if (out->cellMs[k] < 0.98 * out->cellMs[best]) best = k;
  c->cellBlocks = CELL_BLOCKS[best];
  cc_sim_args a;
  memset(&a, 0, sizeof a);
  for (int k = 0; k < 4; k++) {
    c->simOrder = k;
    auto once = [&]() -> int {
      SimU u = sim_uniforms(c, &a);
      int s0 = c->cur, d = c->cur ^ 1;
      launch_sim(u, c->boxDev, c->pos[s0], c->vel[s0], c->mat, c->q[s0], c->gb.seats, c->mesh[0], c->mesh[1],
                 c->ppSlot, c->ppForce, c->pos[d], c->vel[d], c->aux[d], c->q[d], nullptr, c->s);
      LAUNCHCHK(c);
      return 0;
    };
    if ((r = timed(once, &out->simMs[k]))) return r;
  }
  // a wavefront per z-plane or per row must win by 2 % (the lists' join is work the one-wavefront walk does
  // not do); launch order fetches about four times the HBM bytes (neighbours' data spread over every XCD's
  // L2), which the pass, bound by issue, hardly feels alone but concurrent work does: it must win by 5 %
  c->simOrder = 0;
  for (int k = 2; k < 4; k++)
    if (out->simMs[k] < 0.98 * out->simMs[c->simOrder]) c->simOrder = k;
  if (out->simMs[1] < 0.95 * out->simMs[c->simOrder]) c->simOrder = 1;
  out->cellBlocks = c->cellBlocks;
  out->simOrder = c->simOrder;
  return 0;
}

int cc_set_far_field(cc_ctx* c, cc_ff* f, int judgeEvery, double tol) {
  if (!c || (f && (judgeEvery <= 0 || judgeEvery % c->p.meshEvery != 0 || !(tol > 0))))
    FAIL(c, CC_EINVAL, "judge cadence must be a positive multiple of meshEvery and the bound positive");
  HIPCHK(c, hipStreamSynchronize(c->s));
  if (f && !c->ffWork) {
    std::string err;
    c->ffWork = ffwork_create(err);
    if (!c->ffWork) FAIL(c, CC_EDEVICE, "learned far field workspace: %s", err.c_str());
    ALLOC(c->ffGate, 2 * 4); ALLOC(c->ffStats, 2 * FF_STATS * 4); ALLOC(c->ffSums, 4 * 8);
    ALLOC(c->ffFar, 2 * FF_FAR * 16);
    HIPCHK(c, hipEventCreateWithFlags(&c->ffFork, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ffJoin, hipEventDisableTiming));
  }
  if (f) {
    HIPCHK(c, hipMemsetAsync(c->ffGate, 0, 8, c->s));
    HIPCHK(c, hipMemsetAsync(c->ffStats, 0, 2 * FF_STATS * 4, c->s));
    HIPCHK(c, hipMemsetAsync(c->ffSums, 0, 32, c->s));
    ff_forget_bases(c->ffWork, c->s);
    LAUNCHCHK(c);
    HIPCHK(c, hipStreamSynchronize(c->s));
  }
  c->ff = f ? ff_net(f) : nullptr;
  c->ffJudge = judgeEvery;
  c->ffTol = (float)tol;
  return 0;
}

int cc_ff_stats(cc_ctx* c, double* out) {
  if (!c || !out) return CC_EINVAL;
  float st[2 * FF_STATS] = {};
  if (c->ffStats) {
    HIPCHK(c, hipMemcpyAsync(st, c->ffStats, sizeof(st), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
  }
  for (int k = 0; k < 2 * FF_STATS; k++) out[k] = st[k];
  return 0;
}

int cc_ff_record(cc_ff* f, cc_ctx* c, int box) {
  if (!f || !c || box < 0 || box > 1) return CC_EINVAL;
  MeshBuf& m = c->mesh[box];
  if (ff_record_pass(ff_net(f), m.fine, m.coarse, m.mono, m.farA, m.farT0, m.farT1, c->boxDev, box, c->s) < 0)
    FAIL(c, CC_ESTATE, "the learned far field's pass record is full");
  LAUNCHCHK(c);
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

int cc_set_visit(cc_ctx* c, int every) {
  if (!c || every < 0) return CC_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->s));
  c->visitEvery = every;
  c->visitAt = -(1 << 30);
  if (every == 0) {   // index order again
    launch_iota(c->visit, c->N, c->s);
    LAUNCHCHK(c);
    HIPCHK(c, hipStreamSynchronize(c->s));
  }
  return 0;
}

int cc_set_graphs(cc_ctx* c, int on) {
  if (!c) return CC_EINVAL;
  c->useGraphs = on ? 1 : 0;
  return 0;
}

int cc_get_step_no(cc_ctx* c, int* stepNo) {
  if (!c || !stepNo) return CC_EINVAL;
  *stepNo = c->stepNo;
  return 0;
}
int cc_set_step_no(cc_ctx* c, int stepNo) {
  if (!c) return CC_EINVAL;
  c->stepNo = stepNo;
  return 0;
}

int cc_rigid(cc_ctx* c, const double bulk0[3], const double bulk1[3], const double cen0[3], const double cen1[3],
             const double rot0[9], const double rot1[9], double dtF) {
  if (!c) return CC_EINVAL;
  float b0[3], b1[3], c0[3], c1[3], s0[3], s1[3], r0[9], r1[9];
  for (int k = 0; k < 3; k++) {
    b0[k] = (float)bulk0[k]; b1[k] = (float)bulk1[k]; c0[k] = (float)cen0[k]; c1[k] = (float)cen1[k];
    s0[k] = (float)c->p.spin0[k]; s1[k] = (float)c->p.spin1[k];
  }
  for (int k = 0; k < 9; k++) { r0[k] = (float)rot0[k]; r1[k] = (float)rot1[k]; }
  int r;
  if ((r = discard_count(c))) return r;
  int s = c->cur, d = c->cur ^ 1;
  launch_rigid(c->N, c->pos[s], c->vel[s], c->aux[s], c->q[s], b0, b1, c0, c1, s0, s1, r0, r1, (float)dtF, c->pos[d],
               c->vel[d], c->aux[d], c->q[d], c->s);
  LAUNCHCHK(c);
  c->cur = d;
  return 0;
}

int cc_couple(cc_ctx* c, const cc_couple_args* a) {
  if (!c || !a) return CC_EINVAL;
  CoupleU u;
  for (int k = 0; k < 3; k++) u.centre[k] = (float)a->centre[k];
  u.vimp = (float)a->vimp; u.aimp = (float)a->aimp; u.c1 = (float)a->c1; u.invMu = (float)a->invMu;
  u.pexp = (float)a->pexp; u.stall = (float)a->stall; u.z = (float)a->z; u.heat = (float)a->heat;
  u.tK = (float)a->tK; u.frozen = (float)a->frozen; u.ambient = (float)a->ambient; u.melt = (float)a->melt;
  int r;
  if ((r = discard_count(c))) return r;
  int s = c->cur, d = c->cur ^ 1;
  launch_couple(c->N, u, c->pos[s], c->vel[s], c->aux[s], c->mat, c->pos[d], c->vel[d], c->aux[d], c->q[d], c->s);
  LAUNCHCHK(c);
  c->cur = d;
  return 0;
}

int cc_diag(cc_ctx* c, float* out) {
  if (!c) return CC_EINVAL;
  SimU u = sim_uniforms(c, nullptr);
  launch_diag(u, c->boxDev, c->pos[c->cur], c->q[c->cur], c->gb.seats, c->mesh[0], c->mesh[1], c->ppSlot,
              c->ppForce, c->diag, c->s);
  LAUNCHCHK(c);
  if (out) {
    HIPCHK(c, hipMemcpyAsync(out, c->diag, (size_t)c->N * 16, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
  }
  return 0;
}

int cc_morton(cc_ctx* c, const double core[3], int32_t* perm, int32_t* into) {
  if (!c || !core) return CC_EINVAL;
  int r0;
  if ((r0 = discard_count(c))) return r0;
  int s = c->cur, d = c->cur ^ 1;
  launch_morton(c->pos[s], c->N, core[0], core[1], core[2], c->mkeys, c->mslot, c->mhist, c->mblock, c->mperm,
                c->minto, c->s);
  launch_permute(c->N, c->mperm, c->pos[s], c->vel[s], c->aux[s], c->q[s], c->mat, c->pos[d], c->vel[d], c->aux[d],
                 c->q[d], c->matTmp, c->s);
  LAUNCHCHK(c);
  std::swap(c->mat, c->matTmp);
  c->cur = d;
  c->visitAt = -(1 << 30);   // the visit order names particles by their old places: re-sorted at the next cycle
  int r = cc_slots(c);   // applyPerm rebuilds the seats: the old ones name particles by their old places
  if (r) return r;
  if (perm) HIPCHK(c, hipMemcpyAsync(perm, c->mperm, (size_t)c->N * 4, hipMemcpyDeviceToHost, c->s));
  if (into) HIPCHK(c, hipMemcpyAsync(into, c->minto, (size_t)c->N * 4, hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

int cc_analyse(cc_ctx* c, const cc_analyse_args* a, cc_report* out, int32_t* list) {
  if (!c || !a || !out) return CC_EINVAL;
  if (a->cap < 0 || a->cap > PP_CAP) FAIL(c, CC_EINVAL, "cap %d outside 0..%d", a->cap, PP_CAP);
  if (a->cap > 0 && !list) FAIL(c, CC_EINVAL, "list is NULL");
  if (!(a->a > 0) || !(a->touch >= 0) || !(a->m > 0)) FAIL(c, CC_EINVAL, "m, a and touch must be positive");
  if (!c->fof) c->fof = fof_create();
  FofIn in{ c->N, c->pos[c->cur], c->vel[c->cur], c->mat, c->diag, a->diag, a->m, a->a, a->spring, a->touch,
            a->cap, a->flags };
  int r = fof_analyse(c->fof, in, out, list, c->s, c->err);
  if (r) return r;
  LAUNCHCHK(c);
  return 0;
}

// The analysis in two halves, the reference's analyze(false) and collect()
// (src/nbody.js:895-1025): the device's part and the read-back now, the host's
// FP64 sums on a thread of their own while the caller steps on.
int cc_analyse_begin(cc_ctx* c, const cc_analyse_args* a) {
  if (!c || !a) return CC_EINVAL;
  if (a->cap < 0 || a->cap > PP_CAP) FAIL(c, CC_EINVAL, "cap %d outside 0..%d", a->cap, PP_CAP);
  if (!(a->a > 0) || !(a->touch >= 0) || !(a->m > 0)) FAIL(c, CC_EINVAL, "m, a and touch must be positive");
  if (!c->fof) c->fof = fof_create();
  FofIn in{ c->N, c->pos[c->cur], c->vel[c->cur], c->mat, c->diag, a->diag, a->m, a->a, a->spring, a->touch,
            a->cap, a->flags };
  int r = fof_begin(c->fof, in, c->s, c->err);
  if (r) return r;
  LAUNCHCHK(c);
  return 0;
}

int cc_analyse_end(cc_ctx* c, int wait, cc_report* out, int32_t* list) {
  if (!c || !out) return CC_EINVAL;
  if (!c->fof || !fof_pending(c->fof)) FAIL(c, CC_ESTATE, "no analysis was begun");
  bool ready = false;
  int r = fof_end(c->fof, out, list, wait != 0, &ready, c->err);
  if (r) return r;
  return ready ? 0 : 1;
}

int cc_books(cc_ctx* c, double* out) {
  if (!c || !out) return CC_EINVAL;
  launch_books(c->N, c->p.m, c->pos[c->cur], c->vel[c->cur], c->mat, c->diag, c->bpart, c->bout, c->s);
  LAUNCHCHK(c);
  HIPCHK(c, hipMemcpyAsync(out, c->bout, 13 * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

int cc_sync(cc_ctx* c) {
  if (!c) return CC_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

int cc_last_step_ms(cc_ctx* c, double* ms) {
  if (!c || !ms) return CC_EINVAL;
  *ms = c->lastMs;
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ batch --
// Independent instances stepped together: each instance's step cycle is one
// branch of a single captured graph (forked from and joined to the batch's
// stream), so a cycle of meshEvery steps of every instance is one launch.
// A captured batch cycle: what it bakes in (each instance's cycle key), and each instance's
// buffer parities and count state after it.
struct BatchGraph {
  std::vector<double> key;
  hipGraphExec_t exec = nullptr;
  std::vector<int> curAfter, gridCurAfter, countedAfter;
  // a fused cycle's jobs: per step of the cycle, every instance's step pass and seat pass (instance order)
  SimJob* simTab = nullptr;
  SeatJob* seatTab = nullptr;
};
struct cc_batch {
  std::vector<cc_ctx*> ctx;
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> join, mid;
  hipEvent_t fork2 = nullptr;
  // the last few captured cycles, replaced in turn: cycles alternate between keys (a learned far
  // field's judge and the steps between, the step parity with an odd meshEvery)
  BatchGraph graphs[FF_BATCH_TABLES];
  // with one learned far field attached to every instance: one inference over all instances' boxes per
  // mesh step and one judge of the judged ones (the graph cache's slot tables, one per entry)
  FFBatchWork* ffw = nullptr;
  // the fused passes' work-group maps (work-group -> instance) and each instance's first work-group, for the
  // step pass (SIM_JOB_BLOCK particles a work-group) and the seat pass (SEAT_JOB_BLOCK)
  uint16_t* simMap = nullptr;
  uint16_t* seatMap = nullptr;
  int simBlocks = 0, seatBlocks = 0;
  std::vector<int> simBlock0, seatBlock0;
  int nextGraph = 0;
  double lastMs = 0.0;
  std::string err;
};

#define BFAIL(b, code, ...)                       \
  do {                                            \
    char _b[512];                                 \
    snprintf(_b, sizeof _b, __VA_ARGS__);         \
    (b)->err = _b;                                \
    return code;                                  \
  } while (0)
#define BCHK(b, call)                                                                         \
  do {                                                                                        \
    hipError_t _e = (call);                                                                   \
    if (_e != hipSuccess) BFAIL(b, CC_EDEVICE, "%s: %s", #call, hipGetErrorString(_e));       \
  } while (0)

// One learned far field on every instance (none under constant gravity): its inference then runs once for the
// whole batch at the mesh step (every instance's boxes after their coarse meshes) and its judge once after it.
static FFNet* batch_net(const cc_batch* b) {
  FFNet* net = b->ctx[0]->ff;
  for (const cc_ctx* c : b->ctx)
    if (c->ff != net || c->p.field) return nullptr;
  return net;
}

// Slot table t for the mesh step every instance is at: a slot per box that is up (a judge step's prediction into
// the instance's own buffer, else into its far field behind the gate), a judged slot per box on a judge step.
static int batch_ff_table(cc_batch* b, int t, int* nSlots, int* nJudged) {
  size_t n = b->ctx.size();
  if (!b->ffw) {
    std::string err;
    b->ffw = ffbatch_create((int)(2 * n), err);
    if (!b->ffw) BFAIL(b, CC_EDEVICE, "batched far field workspace: %s", err.c_str());
  }
  std::vector<FFSlot> slots;
  std::vector<FFJudgeSlot> judged;
  for (size_t i = 0; i < n; i++) {
    cc_ctx* c = b->ctx[i];
    const bool judge = c->stepNo % c->ffJudge == 0;
    for (int bx = 0; bx < boxes_up(c); bx++) {
      MeshBuf& m = c->mesh[bx];
      float4* A = c->ffFar + FF_FAR * bx;
      if (judge) {
        slots.push_back(ff_slot(c->ffWork, bx, m.coarse, m.mono, c->boxDev, nullptr, A, A + 4096, A + 8192));
        judged.push_back(ff_judge_slot(c->ffWork, bx, m.fine, m.com, m.ccom, m.coarse, A, m.farA, m.farT0, m.farT1,
                                       m.force, c->boxDev, (float)c->p.m, c->ffSums + 2 * bx, c->ffTol,
                                       c->ffStats + FF_STATS * bx, c->ffGate + bx));
      } else {
        slots.push_back(ff_slot(c->ffWork, bx, m.coarse, m.mono, c->boxDev, c->ffGate + bx, m.farA, m.farT0,
                                m.farT1));
      }
    }
  }
  std::string err;
  if (ffbatch_set(b->ffw, t, slots, judged, err)) BFAIL(b, CC_EDEVICE, "batched far field: %s", err.c_str());
  *nSlots = (int)slots.size();
  *nJudged = (int)judged.size();
  return 0;
}

// The instances' step and seat passes fuse into one launch each per step when every instance takes the next
// rebuild's count the same way (the launch is one kernel, with the count or without) and the instances fit the
// maps' 16-bit entries.
static bool batch_fusable(const cc_batch* b) {
  if (b->ctx.size() > 65535) return false;
  for (const cc_ctx* c : b->ctx)
    if (c->fuseCount != b->ctx[0]->fuseCount) return false;
  return true;
}

// The work-group maps of the fused passes, made once (the instances' sizes are fixed for the batch's life).
static int batch_maps(cc_batch* b) {
  if (b->simMap) return 0;
  const size_t n = b->ctx.size();
  std::vector<uint16_t> sm, tm;
  b->simBlock0.assign(n, 0);
  b->seatBlock0.assign(n, 0);
  for (size_t i = 0; i < n; i++) {
    const int N = b->ctx[i]->N;
    b->simBlock0[i] = (int)sm.size();
    b->seatBlock0[i] = (int)tm.size();
    sm.insert(sm.end(), (size_t)((N + SIM_JOB_BLOCK - 1) / SIM_JOB_BLOCK), (uint16_t)i);
    tm.insert(tm.end(), (size_t)((N + SEAT_JOB_BLOCK - 1) / SEAT_JOB_BLOCK), (uint16_t)i);
  }
  if (sm.size() > (size_t)(1 << 30)) BFAIL(b, CC_EINVAL, "batch too large to fuse (%zu work-groups)", sm.size());
  BCHK(b, hipMalloc(&b->simMap, sm.size() * sizeof(uint16_t)));
  BCHK(b, hipMalloc(&b->seatMap, tm.size() * sizeof(uint16_t)));
  BCHK(b, hipMemcpy(b->simMap, sm.data(), sm.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->seatMap, tm.data(), tm.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  b->simBlocks = (int)sm.size();
  b->seatBlocks = (int)tm.size();
  return 0;
}

// A fused cycle, issued into the batch stream's capture after its first fork.  Per step: each instance's own
// work of the step (a mesh step's passes with the pairwise pass beside them, the cycle's box, a count the last
// step pass did not take, the judge's join at the cycle's end) on its own stream, forked from the batch stream
// after the step before; the step's seat pass of every instance in one launch (beside that work, which does not
// read the seats, unless it counted for them); then, after every instance's work, the step pass of every
// instance in one launch.  With one learned far field on every instance the mesh step is split as in the
// unfused batched cycle (heads, one inference over every box, tails) and the judged boxes are judged after the
// cycle.  The jobs are collected into simAll / seatAll (step-major, instance order) for the graph's tables.
static int batch_capture_fused(cc_batch* b, const cc_sim_args* args, BatchGraph& g, FFNet* net, int t, int nSlots,
                               int nJudged, std::vector<SimJob>& simAll, std::vector<SeatJob>& seatAll,
                               hipError_t& e) {
  const size_t n = b->ctx.size();
  const int M = b->ctx[0]->p.meshEvery;
  simAll.assign((size_t)M * n, SimJob{});
  seatAll.assign((size_t)M * n, SeatJob{});
  std::vector<SimJob> sim;
  std::vector<SeatJob> seat;
  int r = 0;
  for (size_t i = 0; i < n; i++) b->ctx[i]->ffBatched = net != nullptr;
  for (int k = 0; k < M && !r && e == hipSuccess; k++) {
    sim.clear();
    seat.clear();
    bool waits = false;
    const bool heads = net && k == 0;
    if (k > 0) e = hipEventRecord(b->fork, b->s);
    for (size_t i = 0; i < n && !r && e == hipSuccess; i++) {
      cc_ctx* c = b->ctx[i];
      e = hipStreamWaitEvent(c->s, b->fork, 0);
      if (e != hipSuccess) break;
      if (k == 0) launch_box_next(c->boxSched, c->boxIdx, c->boxDev, c->s);
      c->simJobs = &sim; c->seatJobs = &seat; c->seatWaits = false;
      r = one_step(c, args[i].settle, args[i].kick, &args[i], heads ? 1 : 0);
      waits = waits || c->seatWaits;
      c->simJobs = nullptr; c->seatJobs = nullptr;
      if (r) { b->err = c->err; break; }
      e = hipEventRecord(heads ? b->mid[i] : b->join[i], c->s);
    }
    if (r || e != hipSuccess) break;
    if (!seat.empty() && seat.size() != n) { b->err = "fused batch: the instances rebuild on different steps"; r = CC_ESTATE; break; }
    SimJob* simTab = g.simTab + (size_t)k * n;
    SeatJob* seatTab = g.seatTab + (size_t)k * n;
    if (!seat.empty() && !waits) launch_slots_seat_jobs(seatTab, b->seatMap, b->seatBlocks, b->s);
    if (heads) {
      for (size_t i = 0; i < n && e == hipSuccess; i++) e = hipStreamWaitEvent(b->s, b->mid[i], 0);
      if (e != hipSuccess) break;
      ff_predict_batch(net, b->ffw, t, nSlots, b->s);
      e = hipEventRecord(b->fork2, b->s);
      for (size_t i = 0; i < n && !r && e == hipSuccess; i++) {
        cc_ctx* c = b->ctx[i];
        e = hipStreamWaitEvent(c->s, b->fork2, 0);
        if (e != hipSuccess) break;
        c->simJobs = &sim;
        r = one_step(c, args[i].settle, args[i].kick, &args[i], 2);
        c->simJobs = nullptr;
        if (r) { b->err = c->err; break; }
        e = hipEventRecord(b->join[i], c->s);
      }
      if (r || e != hipSuccess) break;
    }
    for (size_t i = 0; i < n && e == hipSuccess; i++) e = hipStreamWaitEvent(b->s, b->join[i], 0);
    if (e != hipSuccess) break;
    if (!seat.empty() && waits) launch_slots_seat_jobs(seatTab, b->seatMap, b->seatBlocks, b->s);
    if (sim.size() != n) { b->err = "fused batch: an instance took no step pass"; r = CC_ESTATE; break; }
    const bool count = sim[0].cnt.cnt != nullptr;
    for (size_t i = 0; i < n; i++) {
      if ((sim[i].cnt.cnt != nullptr) != count) { b->err = "fused batch: the instances count on different steps"; r = CC_ESTATE; break; }
      sim[i].block0 = b->simBlock0[i];
      simAll[(size_t)k * n + i] = sim[i];
      if (!seat.empty()) {
        seat[i].block0 = b->seatBlock0[i];
        seatAll[(size_t)k * n + i] = seat[i];
      }
    }
    if (r) break;
    launch_sim_jobs(simTab, b->simMap, b->simBlocks, count, b->s);
  }
  if (!r && e == hipSuccess && net) ff_judge_batch(b->ffw, t, nJudged, b->s);
  for (size_t i = 0; i < n; i++) b->ctx[i]->ffBatched = false;
  return r;
}

static int batch_graph(cc_batch* b, const cc_sim_args* args, BatchGraph** out) {
  std::vector<double> key;
  for (size_t i = 0; i < b->ctx.size(); i++) {
    std::vector<double> k = cycle_key(b->ctx[i], &args[i]);
    key.insert(key.end(), k.begin(), k.end());
  }
  for (BatchGraph& g : b->graphs)
    if (g.exec && g.key == key) { *out = &g; return 0; }
  const int t = b->nextGraph;
  BatchGraph& g = b->graphs[t];
  b->nextGraph = (b->nextGraph + 1) % FF_BATCH_TABLES;
  if (g.exec) {   // its last launch may still run: let it end first
    BCHK(b, hipStreamSynchronize(b->s));
    hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
  }
  size_t n = b->ctx.size();
  FFNet* net = batch_net(b);
  const bool batched = net != nullptr;
  int nSlots = 0, nJudged = 0, r;
  if (batched && (r = batch_ff_table(b, t, &nSlots, &nJudged))) return r;
  const bool fused = batch_fusable(b);
  if (fused) {
    if ((r = batch_maps(b))) return r;
    const size_t jobs = (size_t)b->ctx[0]->p.meshEvery * n;
    if (!g.simTab) BCHK(b, hipMalloc(&g.simTab, jobs * sizeof(SimJob)));
    if (!g.seatTab) BCHK(b, hipMalloc(&g.seatTab, jobs * sizeof(SeatJob)));
  }
  std::vector<SimJob> simAll;
  std::vector<SeatJob> seatAll;
  std::vector<int> stepNo(n), cur(n), gcur(n), counted(n);
  for (size_t i = 0; i < n; i++) {
    stepNo[i] = b->ctx[i]->stepNo; cur[i] = b->ctx[i]->cur; gcur[i] = b->ctx[i]->gb.cur; counted[i] = b->ctx[i]->counted;
  }
  BCHK(b, hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipEventRecord(b->fork, b->s);
  r = 0;
  if (fused) {
    if (e == hipSuccess) r = batch_capture_fused(b, args, g, net, t, nSlots, nJudged, simAll, seatAll, e);
  } else if (!batched) {
    for (size_t i = 0; i < n && !r && e == hipSuccess; i++) {
      cc_ctx* c = b->ctx[i];
      e = hipStreamWaitEvent(c->s, b->fork, 0);
      if (e != hipSuccess) break;
      launch_box_next(c->boxSched, c->boxIdx, c->boxDev, c->s);
      for (int k = 0; k < c->p.meshEvery && !r; k++) r = one_step(c, args[i].settle, args[i].kick, &args[i]);
      if (r) { b->err = c->err; break; }
      e = hipEventRecord(b->join[i], c->s);
      if (e == hipSuccess) e = hipStreamWaitEvent(b->s, b->join[i], 0);
    }
  } else {
    for (size_t i = 0; i < n; i++) b->ctx[i]->ffBatched = true;
    // every instance's mesh step up to its coarse meshes
    for (size_t i = 0; i < n && !r && e == hipSuccess; i++) {
      cc_ctx* c = b->ctx[i];
      e = hipStreamWaitEvent(c->s, b->fork, 0);
      if (e != hipSuccess) break;
      launch_box_next(c->boxSched, c->boxIdx, c->boxDev, c->s);
      r = one_step(c, args[i].settle, args[i].kick, &args[i], 1);
      if (r) { b->err = c->err; break; }
      e = hipEventRecord(b->mid[i], c->s);
      if (e == hipSuccess) e = hipStreamWaitEvent(b->s, b->mid[i], 0);
    }
    // one inference over every box of every instance
    if (!r && e == hipSuccess) {
      ff_predict_batch(net, b->ffw, t, nSlots, b->s);
      e = hipEventRecord(b->fork2, b->s);
    }
    // the rest of each instance's cycle
    for (size_t i = 0; i < n && !r && e == hipSuccess; i++) {
      cc_ctx* c = b->ctx[i];
      e = hipStreamWaitEvent(c->s, b->fork2, 0);
      if (e != hipSuccess) break;
      r = one_step(c, args[i].settle, args[i].kick, &args[i], 2);
      for (int k = 1; k < c->p.meshEvery && !r; k++) r = one_step(c, args[i].settle, args[i].kick, &args[i]);
      if (r) { b->err = c->err; break; }
      e = hipEventRecord(b->join[i], c->s);
      if (e == hipSuccess) e = hipStreamWaitEvent(b->s, b->join[i], 0);
    }
    // the judged boxes, their exact passes done
    if (!r && e == hipSuccess) ff_judge_batch(b->ffw, t, nJudged, b->s);
    for (size_t i = 0; i < n; i++) b->ctx[i]->ffBatched = false;
  }
  hipGraph_t graph = nullptr;
  hipError_t ee = hipStreamEndCapture(b->s, &graph);
  g.curAfter.assign(n, 0);
  g.gridCurAfter.assign(n, 0);
  g.countedAfter.assign(n, 0);
  for (size_t i = 0; i < n; i++) {
    cc_ctx* c = b->ctx[i];
    g.curAfter[i] = c->cur; g.gridCurAfter[i] = c->gb.cur; g.countedAfter[i] = c->counted;
    c->stepNo = stepNo[i]; c->cur = cur[i]; c->gb.cur = gcur[i]; c->counted = counted[i];
  }
  if (r || e != hipSuccess || ee != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    if (r) return r;
    BFAIL(b, CC_EDEVICE, "batch capture: %s", hipGetErrorString(e != hipSuccess ? e : ee));
  }
  if (fused) {   // the jobs the captured launches read (no replay of this entry's last graph is running)
    e = hipMemcpy(g.simTab, simAll.data(), simAll.size() * sizeof(SimJob), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g.seatTab, seatAll.data(), seatAll.size() * sizeof(SeatJob), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipGraphDestroy(graph);
      BFAIL(b, CC_EDEVICE, "fused batch jobs: %s", hipGetErrorString(e));
    }
  }
  if (fused) {   // the jobs the captured launches read (no replay of this entry's last graph is running)
    e = hipMemcpy(g.simTab, simAll.data(), simAll.size() * sizeof(SimJob), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g.seatTab, seatAll.data(), seatAll.size() * sizeof(SeatJob), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipGraphDestroy(graph);
      BFAIL(b, CC_EDEVICE, "fused batch jobs: %s", hipGetErrorString(e));
    }
  }
  e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) { g.exec = nullptr; BFAIL(b, CC_EDEVICE, "batch instantiate: %s", hipGetErrorString(e)); }
  g.key = key;
  *out = &g;
  return 0;
}

// One step of every instance outside a whole cycle (a call that starts or ends mid-cycle).  Each instance's
// launches go to its own stream, forked from the batch stream and joined back, so the instances' passes overlap
// as in a captured cycle.  A mesh step of a batch with one learned far field runs as a captured cycle's does:
// every instance up to its coarse meshes, one inference over all their boxes (the eager slot table), the rest of
// the step, then the judge of the judged boxes.
static int batch_step_eager(cc_batch* b, const cc_sim_args* args, const std::vector<const double*>& L,
                            const std::vector<int>& nL, std::vector<int>& used) {
  const size_t n = b->ctx.size();
  const int M = b->ctx[0]->p.meshEvery;
  const bool mesh = b->ctx[0]->stepNo % M == 0;
  FFNet* net = mesh ? batch_net(b) : nullptr;
  int nSlots = 0, nJudged = 0, r;
  if (net) {
    BCHK(b, hipStreamSynchronize(b->s));   // the eager table's last inference and judge have ended
    if ((r = batch_ff_table(b, FF_EAGER_TABLE, &nSlots, &nJudged))) return r;
  }
  BCHK(b, hipEventRecord(b->fork, b->s));
  for (size_t i = 0; i < n; i++) {
    cc_ctx* c = b->ctx[i];
    BCHK(b, hipStreamWaitEvent(c->s, b->fork, 0));
    if (mesh && L[i] && used[i] < nL[i]) {
      const double* q = L[i] + 7 * used[i]++;
      if ((r = cc_set_boxes(c, q, q + 3, q[6] > 0.5))) { b->err = c->err; return r; }
    }
    if ((r = visit_due(c, c->s))) { b->err = c->err; return r; }
    c->ffBatched = net != nullptr;
    r = one_step(c, args[i].settle, args[i].kick, &args[i], net ? 1 : 0);
    c->ffBatched = false;
    if (r) { b->err = c->err; return r; }
    hipEvent_t ev = net ? b->mid[i] : b->join[i];
    BCHK(b, hipEventRecord(ev, c->s));
    BCHK(b, hipStreamWaitEvent(b->s, ev, 0));
  }
  if (!net) return 0;
  ff_predict_batch(net, b->ffw, FF_EAGER_TABLE, nSlots, b->s);
  BCHK(b, hipEventRecord(b->fork2, b->s));
  for (size_t i = 0; i < n; i++) {
    cc_ctx* c = b->ctx[i];
    BCHK(b, hipStreamWaitEvent(c->s, b->fork2, 0));
    c->ffBatched = true;
    r = one_step(c, args[i].settle, args[i].kick, &args[i], 2);
    c->ffBatched = false;
    if (r) { b->err = c->err; return r; }
    BCHK(b, hipEventRecord(b->join[i], c->s));
    BCHK(b, hipStreamWaitEvent(b->s, b->join[i], 0));
  }
  ff_judge_batch(b->ffw, FF_EAGER_TABLE, nJudged, b->s);
  return 0;
}

extern "C" {

int cc_batch_create(cc_ctx* const* ctxs, int n, cc_batch** out) {
  if (!out) return CC_EINVAL;
  *out = nullptr;
  if (!ctxs || n <= 0) { g_err = "batch needs at least one instance"; return CC_EINVAL; }
  for (int i = 0; i < n; i++) {
    if (!ctxs[i]) { g_err = "batch instance is NULL"; return CC_EINVAL; }
    if (ctxs[i]->p.meshEvery != ctxs[0]->p.meshEvery) { g_err = "batch instances differ in meshEvery"; return CC_EINVAL; }
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) { g_err = "an instance appears twice in the batch"; return CC_EINVAL; }
  }
  cc_batch* b = new cc_batch();
  b->ctx.assign(ctxs, ctxs + n);
  b->join.assign(n, nullptr);
  b->mid.assign(n, nullptr);
  hipError_t e = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork2, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  for (int i = 0; i < n && e == hipSuccess; i++) e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
  for (int i = 0; i < n && e == hipSuccess; i++) e = hipEventCreateWithFlags(&b->mid[i], hipEventDisableTiming);
  if (e != hipSuccess) { g_err = std::string("batch create: ") + hipGetErrorString(e); cc_batch_destroy(b); return CC_EDEVICE; }
  *out = b;
  return 0;
}

int cc_batch_destroy(cc_batch* b) {
  if (!b) return 0;
  if (b->s) hipStreamSynchronize(b->s);
  for (BatchGraph& g : b->graphs) {
    if (g.exec) hipGraphExecDestroy(g.exec);
    if (g.simTab) hipFree(g.simTab);
    if (g.seatTab) hipFree(g.seatTab);
  }
  if (b->simMap) hipFree(b->simMap);
  if (b->seatMap) hipFree(b->seatMap);
  for (hipEvent_t e : b->join) if (e) hipEventDestroy(e);
  for (hipEvent_t e : b->mid) if (e) hipEventDestroy(e);
  if (b->fork2) hipEventDestroy(b->fork2);
  if (b->ffw) ffbatch_destroy(b->ffw);
  if (b->fork) hipEventDestroy(b->fork);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->s) hipStreamDestroy(b->s);
  delete b;
  return 0;
}

const char* cc_batch_last_error(cc_batch* b) { return b ? b->err.c_str() : g_err.c_str(); }

int cc_batch_step_n(cc_batch* b, int n, const cc_sim_args* args, const double* boxL, const int* nBoxL) {
  if (!b || !args || n < 0) return CC_EINVAL;
  size_t m = b->ctx.size();
  const int M = b->ctx[0]->p.meshEvery;
  int phase = b->ctx[0]->stepNo % M;
  for (size_t i = 1; i < m; i++)
    if (b->ctx[i]->stepNo % M != phase) BFAIL(b, CC_ESTATE, "instance %zu is at another step of the mesh cycle", i);
  std::vector<const double*> L(m, nullptr);
  std::vector<int> nL(m, 0), used(m, 0);
  if (boxL && nBoxL) {
    size_t off = 0;
    for (size_t i = 0; i < m; i++) {
      if (nBoxL[i] < 0) BFAIL(b, CC_EINVAL, "negative box count");
      L[i] = boxL + 7 * off; nL[i] = nBoxL[i]; off += (size_t)nBoxL[i];
    }
  }
  BCHK(b, hipEventRecord(b->ev0, b->s));
  for (size_t i = 0; i < m; i++) {   // earlier work of each instance before anything of this call
    BCHK(b, hipEventRecord(b->join[i], b->ctx[i]->s));
    BCHK(b, hipStreamWaitEvent(b->s, b->join[i], 0));
  }
  // steps outside a whole cycle: one at a time, every instance on its own stream
  int i0 = 0, r;
  while (i0 < n && b->ctx[0]->stepNo % M != 0) {
    if ((r = batch_step_eager(b, args, L, nL, used))) return r;
    i0++;
  }
  int cycles = (n - i0) / M;
  while (cycles > 0) {
    int chunk = cycles < SCHED_CAP ? cycles : SCHED_CAP;
    BCHK(b, hipStreamSynchronize(b->s));
    for (size_t i = 0; i < m; i++) {
      cc_ctx* c = b->ctx[i];
      for (int k = 0; k < chunk; k++) {
        Box& x = c->boxHost[k];
        x = c->box;
        if (L[i] && used[i] < nL[i]) {
          const double* q = L[i] + 7 * used[i]++;
          for (int t = 0; t < 3; t++) { x.LA[t] = (float)q[t]; x.LB[t] = (float)q[3 + t]; }
          x.boxB = q[6] > 0.5 ? 1.0f : 0.0f;
          x.invH = (float)(1.0 / FH);
        }
      }
      BCHK(b, hipMemcpyAsync(c->boxSched, c->boxHost, (size_t)chunk * sizeof(Box), hipMemcpyHostToDevice, b->s));
      launch_fill_u32(reinterpret_cast<uint32_t*>(c->boxIdx), 1, 0u, b->s);
    }
    for (int k = 0; k < chunk; k++) {
      for (size_t i = 0; i < m; i++) {
        b->ctx[i]->box = b->ctx[i]->boxHost[k];
        if ((r = visit_due(b->ctx[i], b->s))) { b->err = b->ctx[i]->err; return r; }
      }
      for (size_t i = 0; i < m; i++)   // outside any capture
        if ((r = ff_join(b->ctx[i], b->s))) { b->err = b->ctx[i]->err; return r; }
      BatchGraph* g = nullptr;
      if ((r = batch_graph(b, args, &g))) return r;
      BCHK(b, hipGraphLaunch(g->exec, b->s));
      for (size_t i = 0; i < m; i++) {
        cc_ctx* c = b->ctx[i];
        c->cur = g->curAfter[i]; c->gb.cur = g->gridCurAfter[i]; c->counted = g->countedAfter[i]; c->stepNo += M;
      }
    }
    i0 += chunk * M;
    cycles -= chunk;
  }
  for (; i0 < n; i0++)
    if ((r = batch_step_eager(b, args, L, nL, used))) return r;
  for (size_t i = 0; i < m; i++)
    if ((r = ff_join(b->ctx[i], b->s))) { b->err = b->ctx[i]->err; return r; }
  BCHK(b, hipEventRecord(b->ev1, b->s));
  BCHK(b, hipEventSynchronize(b->ev1));
  float ms = 0;
  BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->lastMs = ms;
  return 0;
}

int cc_batch_last_ms(cc_batch* b, double* ms) {
  if (!b || !ms) return CC_EINVAL;
  *ms = b->lastMs;
  return 0;
}

}  // extern "C"

// ------------------------------------------- the pairwise pass, no allocation --
// cc_pp_noalloc: the pairwise pass on the caller's device arrays, for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops).  Its workspace -- the mesh boxes as the kernels read them and the
// gathered loose list with each seat's box -- is `scratch`, a file-scope
// device array of a fixed size (PP_CAP seats: 256 + 8192 x 20 bytes), one copy
// per GPU for the life of the process.  Host code never names `scratch` in a
// launch or a copy: its address on each device comes from hipGetSymbolAddress,
// taken once per device with that device current and kept in g_scratch.
// Every call on a GPU shares that GPU's copy: calls that overlap on one GPU
// (two streams, two threads) race.  Each call writes every scratch byte it
// reads (the boxes, all 8192 seats, the tile boxes of its own list's length)
// before reading it, so nothing needs clearing between calls.
namespace {
constexpr size_t NA_BOX = 0;                                  // Box (32 B), 256-byte aligned
constexpr size_t NA_LIST = 256;                               // float4[PP_CAP]
constexpr size_t NA_SEATBOX = NA_LIST + (size_t)PP_CAP * 16;  // float[PP_CAP]
constexpr size_t NA_BOUNDS = NA_SEATBOX + (size_t)PP_CAP * 4;  // float4[2 * PP_TILES_MAX]
constexpr size_t NA_BYTES = NA_BOUNDS + (size_t)PP_TILES_MAX * 32;
constexpr int NA_DEVICES = 64;
std::mutex g_scratchLock;
unsigned char* g_scratch[NA_DEVICES];   // each device's address of `scratch` (null until looked up)
}  // namespace
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// the current device's copy of `scratch`, looked up on the device's first call
static int scratch_here(unsigned char** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) FAIL((cc_ctx*)nullptr, CC_EDEVICE, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= NA_DEVICES) FAIL((cc_ctx*)nullptr, CC_EDEVICE, "device %d past the %d this library serves", dev, NA_DEVICES);
  std::lock_guard<std::mutex> hold(g_scratchLock);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      FAIL((cc_ctx*)nullptr, CC_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

extern "C" int cc_pp_noalloc(const float* pos, const float* mat, int N, const double LA[3], const double LB[3],
                             int boxB, const int32_t* index, int count, double m, double a, const float* tab,
                             float* force, void* stream) {
  if (!pos || !mat || !index || !tab || !force || !LA || !LB)
    FAIL((cc_ctx*)nullptr, CC_EINVAL, "cc_pp_noalloc: a NULL array");
  if (N <= 0) FAIL((cc_ctx*)nullptr, CC_EINVAL, "cc_pp_noalloc: N=%d", N);
  if (count < 0 || count > PP_CAP)
    FAIL((cc_ctx*)nullptr, CC_EINVAL, "cc_pp_noalloc: a loose list of %d seats (its workspace holds at most %d)", count,
         PP_CAP);
  if (!(m > 0) || !(a > 0)) FAIL((cc_ctx*)nullptr, CC_EINVAL, "cc_pp_noalloc: m and a must be > 0");
  int r = check_runtime();
  if (r) return r;
  unsigned char* w = nullptr;
  r = scratch_here(&w);
  if (r) return r;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Box b;
  b.invH = (float)(1.0 / FH);
  for (int k = 0; k < 3; k++) { b.LA[k] = (float)LA[k]; b.LB[k] = (float)LB[k]; }
  b.boxB = boxB ? 1.0f : 0.0f;
  Box* bx = reinterpret_cast<Box*>(w + NA_BOX);
  launch_put_box(b, bx, s);
  launch_pp(reinterpret_cast<const float4*>(pos), mat, N, bx, index, nullptr, count, (float)m, (float)(5.0 * FH), (float)(a * a),
            reinterpret_cast<const float2*>(tab), reinterpret_cast<float4*>(w + NA_LIST),
            reinterpret_cast<float*>(w + NA_SEATBOX), reinterpret_cast<float4*>(w + NA_BOUNDS),
            reinterpret_cast<float4*>(force), s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) FAIL((cc_ctx*)nullptr, CC_EDEVICE, "cc_pp_noalloc launch: %s", hipGetErrorString(e));
  return 0;
}
