// Internal types shared by the kernels (cc_kernels.hip) and the C-ABI
// (cc_api.hip). Device layouts (HBM):
//   particle state : pos/vel/aux float4[N], q float[N] -- two copies (src/dst),
//                    mat float[N]; index i = y*W + x of the reference's W x H
//                    state textures.
//   contact grid   : seats uint32[cells*8] -- the eight seat words of hashed
//                    cell h (h = atlas texel y*AW + x) contiguous, so one cell
//                    is two 16-byte loads (one more cell, always empty, past
//                    the grid); cnt int[cells] x 2 and occupied-cell lists x 2
//                    (one of each per rebuild parity), bucket uint32[cells*CAP]
//                    (the arrivals' seat words of the current rebuild), an
//                    overflow list.
//   mesh (per box) : acc int64[513*512*4] fixed-point deposit accumulator,
//                    fine/force float4[513*512] in the reference atlas layout
//                    (8x8 slices of 64x64, scratch texel (0,512)),
//                    coarse/farA/farT0/farT1 float4[64*64], mono float4.
//   loose list     : index int[8192], list float4[8192], box float[8192],
//                    force float4[4*8192] (slice s of seat j at s*8192+j),
//                    slot float[N] (seat of particle, or -1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

struct cc_ff;
struct cc_report;

namespace cc {

constexpr int SLOTS = 8;             // seats per contact cell
constexpr int SEAT_Q = 480;          // the seat prefilter's lattice: 1/480 of a cell (32 per fifteenth)
constexpr int BUCKET_CAP = 16;       // arrivals kept per cell before the overflow list
constexpr int PP_CAP = 8192;
constexpr int PP_TILE = 48;          // entries per tile of a slice in the pairwise pass
constexpr int PP_TILES_MAX = 4 * ((PP_CAP / 4 + PP_TILE - 1) / PP_TILE);   // tiles of the four slices at the cap
constexpr int FINE_TEXELS = 513 * 512;
constexpr int SCRATCH = 512 * 512;   // texel (0, 512)
constexpr uint32_t EMPTY_SEAT = 0xBF800000u;
constexpr double FL = 6.0;
constexpr double FH = 2.0 * FL / 64.0;
// fixed-point scales of the deposit accumulator: the mesh sums (weights up to
// a few thousand, mass-weighted positions up to a few 1e4) and the strays'
// pull in the scratch texel (bounded by N * 2.5 / (0.385 FH^2) < 2e7)
constexpr double DEP_SCALE = 1099511627776.0;     // 2^40
constexpr double SCR_SCALE = 16777216.0;          // 2^24

struct Grid {
  int G, GShift, GMask, SX, SXShift, SXMask, AW, AH;
  int AWShift;          // log2 AW (the atlas is G x SX wide, both powers of two)
  int cells;            // AW x AH; the seat array holds one more cell, always empty, after them
  // bits of the particle index in a seat word: 18 as the reference packs it
  // (idx | q << 18 | q << 22 | q << 26 | 1 << 30, src/glsl-sim.js:103-110,
  // :143-144), 19 above 262,144 particles, where the place fields move up one
  // bit and take bit 30 (no reference counterpart: its index stops at 2^18)
  int IB;
};

struct Box {            // the two mesh boxes as the shaders see them
  float invH;
  float LA[3], LB[3];
  float boxB;
};

struct SimU {           // the step pass's uniforms, already FP32
  int N;
  float invCell, cell, dt, gdt, touch, link2, reach2, k, c, invM, damp, creep, tK, kappa, cpIron;
  float settle, kick;
  float bulk0[3], bulk1[3], cen0[3], cen1[3], spin0[3], spin1[3];
  Grid grid;
  float gm, eps2;       // G m, monopole softening FH^2
  int field;            // 1: constant gravity gField
  float gField[3];
  int frozenOn;
  float frozen;
  int reachQ2;          // the seat prefilter's bound in 1/480-cell units, squared (see k_sim)
  int order;            // 0: work-group b takes particles (b % 8) * (grid / 8) + b / 8 (consecutive
                        // work-groups' neighbourhoods spread over the XCDs' L2s), 1: particles b
  const int* visit;     // the order the particles are taken in (a permutation of 0..N-1), or null: index order
};

struct CoupleU {        // COUPLE_FS's uniforms (chicxulub.html:260-276), FP32
  float centre[3];
  float vimp, aimp, c1, invMu, pexp, stall, z, heat, tK, frozen, ambient, melt;
};

struct MeshBuf {
  unsigned long long* acc;   // int64 two's complement, [FINE_TEXELS*4]
  float4* fine;
  float4* com;               // each fine texel's centre of mass and mass (fine.xyz / fine.w, fine.w)
  float4* force;
  float4* coarse;
  float4* ccom;              // each coarse block's centre of mass and mass
  float4* farA;
  float4* farT0;
  float4* farT1;
  float4* mono;
  int* occ;                  // occupied fine texels of the last deposit
  int* occN;
  int* coarseDone;           // the coarse pass's count of finished work-groups (the last sums the monopole, resets it)
};

struct GridBuf {
  uint32_t* seats;           // [(cells + 1) * 8]: the last cell stays empty (a safe target for skipped loads)
  int* cnt[2];               // [cells] each: arrivals per cell, one set per rebuild parity
  uint32_t* bucket;          // [cells*BUCKET_CAP]: the arrivals' seat words (the index is their low IB bits)
  int* occ[2];               // occupied cells, one list per parity, in 64-entry segments (one per wavefront)
  int* segN;                 // [2][ceil(N / 64)]: entries of each segment
  int2* ovf;                 // (cell, seat word) past a cell's bucket
  int* ovfN;                 // [2]: the overflow count of each rebuild parity
  uint32_t idxMask;          // (1 << IB) - 1
  int cur;                   // the parity of the last rebuild (whose seats are in place)
};

// where a count writes: the arrays of one parity
struct SlotsOut {
  int* cnt;
  uint32_t* bucket;
  int* occ;
  int* segN;
  int2* ovf;
  int* ovfN;
};

// One instance's step pass or seat pass as a job of a launch that serves several instances (a batch's
// fused passes): the arguments its own launch would take, and block0, its first work-group in the launch.
struct SimJob {
  SimU u;
  const Box* bx;
  const float4 *pos, *vel;
  const float *mat, *q;
  const uint32_t* seats;
  const float4 *force0, *mono0, *force1, *mono1;
  const float* ppSlot;
  const float4* ppForce;
  float4 *oPos, *oVel, *oAux;
  float* oQ;
  SlotsOut cnt;
  float touch2r;
  int block0;
};
struct SeatJob {
  const int *occ, *segN, *occPrev, *segPrev;
  int N;
  const int* cnt;
  int* cntPrev;
  const uint32_t* bucket;
  const int2* ovf;
  const int* ovfN;
  int* ovfNOther;
  uint32_t idxMask;
  uint32_t* seats;
  int block0;
};
constexpr int SIM_JOB_BLOCK = 64;    // a step-pass job's work-group: particles
constexpr int SEAT_JOB_BLOCK = 256;  // a seat-pass job's work-group: threads

// ---- launches (cc_kernels.hip) ----
void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
void launch_fill_f32(float* p, size_t n, float v, hipStream_t s);
// a rebuild: the count (into parity cur ^ 1) and the seat pass (which makes it cur)
void launch_slots(const float4* pos, int N, float invCell, const Grid& g, GridBuf& gb, hipStream_t s);
void launch_slots_count(const float4* pos, int N, float invCell, const Grid& g, GridBuf& gb, hipStream_t s);
void launch_slots_seat(int N, GridBuf& gb, hipStream_t s);
// the seat pass as a job (and the parity flipped as launch_slots_seat flips it); jobs[blockJob[b]] for
// work-groups b < blocks of SEAT_JOB_BLOCK threads
SeatJob seat_job(int N, GridBuf& gb);
void launch_slots_seat_jobs(const SeatJob* jobs, const uint16_t* blockJob, int blocks, hipStream_t s);
// a count of parity cur ^ 1 that will not be seated: its counters back to zero
void launch_slots_discard(int N, GridBuf& gb, hipStream_t s);
// the arrays a count of that parity writes
SlotsOut slots_out(GridBuf& gb, int N, int parity);
void launch_copy4(const float4* src, float4* dst, int n, hipStream_t s);
void launch_put_box(const Box& b, Box* dst, hipStream_t s);
void launch_box_next(const Box* sched, int* idx, Box* dst, hipStream_t s);
// zeroForce: also zero the force mesh's empty texels (a cell pass follows that writes the rest)
void launch_deposit(const float4* pos, const float* mat, int N, const int* visit, const Box* b, int box, float eps2,
                    MeshBuf& m, bool zeroForce, hipStream_t s);
void launch_coarse(MeshBuf& m, hipStream_t s);
// the cell pass over the occupied fine texels, `blocks` work-groups of four wavefronts striding over them;
// zeroed: the force mesh's empty texels already zero (the deposit's finish leaves them so)
void launch_cell(MeshBuf& m, float eps2, float gm, int blocks, bool zeroed, hipStream_t s);
// the exact far field (k_far); with *gate set (a learned far field standing in) only the blocks occupied now
// and empty in baseFeat, the network's base (none without baseFeat)
void launch_far(const MeshBuf& m, float4* farA, float4* farT0, float4* farT1, const int* gate, hipStream_t s,
                const float4* baseFeat = nullptr);
// the centres of mass of a mesh written from outside: com[t] for mesh[t], t < n
void launch_centres(const float4* mesh, float4* com, int n, hipStream_t s);
// the per-cell pass from given far-field buffers
void launch_cell_only(const MeshBuf& m, const float4* farA, const float4* farT0, const float4* farT1, float4* force,
                      float eps2, float gm, int blocks, bool zeroed, hipStream_t s);
// the pairwise pass: the loose list's gather, its tiles' boxes (bounds: 2 x PP_TILES_MAX), then each seat's
// four slices (force[slice * PP_CAP + seat]); the list's length *countDev where given (a captured cycle reads
// the present list's), else count
void launch_pp(const float4* pos, const float* mat, int N, const Box* b, const int* index, const int* countDev,
               int count, float gm, float cut, float a2, const float2* tab, float4* list, float* box, float4* bounds,
               float4* force, hipStream_t s);
// the step pass's visit order: the particles sorted by the Morton code of their contact cell (coordinates
// modulo 128); keys, slot: N ints, hist: KEYS + 1 ints, blockSums: scan scratch
void launch_visit(const float4* pos, int N, float invCell, int* keys, int* slot, int* hist, int* blockSums,
                  int* visit, hipStream_t s);
void launch_iota(int* a, int n, hipStream_t s);
// count: the next rebuild's count taken on the positions this pass writes (null: none)
void launch_sim(const SimU& u, const Box* bx, const float4* pos, const float4* vel, const float* mat, const float* q,
                const uint32_t* seats, const MeshBuf& m0, const MeshBuf& m1, const float* ppSlot,
                const float4* ppForce, float4* oPos, float4* oVel, float4* oAux, float* oQ, const SlotsOut* count,
                hipStream_t s);
// the step pass as a job (the same arguments); jobs[blockJob[L]] for logical work-groups L < blocks of
// SIM_JOB_BLOCK particles, every job of one launch with the count or every one without
SimJob sim_job(const SimU& u, const Box* bx, const float4* pos, const float4* vel, const float* mat, const float* q,
               const uint32_t* seats, const MeshBuf& m0, const MeshBuf& m1, const float* ppSlot,
               const float4* ppForce, float4* oPos, float4* oVel, float4* oAux, float* oQ, const SlotsOut* count);
void launch_sim_jobs(const SimJob* jobs, const uint16_t* blockJob, int blocks, bool count, hipStream_t s);
void launch_rigid(int N, const float4* pos, const float4* vel, const float4* aux, const float* q,
                  const float* bulk0, const float* bulk1, const float* cen0, const float* cen1,
                  const float* spin0, const float* spin1, const float* rot0, const float* rot1, float dt,
                  float4* oPos, float4* oVel, float4* oAux, float* oQ, hipStream_t s);
void launch_couple(int N, const CoupleU& c, const float4* pos, const float4* vel, const float4* aux, const float* mat,
                   float4* oPos, float4* oVel, float4* oAux, float* oQ, hipStream_t s);
void launch_diag(const SimU& u, const Box* bx, const float4* pos, const float* q, const uint32_t* seats, const MeshBuf& m0,
                 const MeshBuf& m1, const float* ppSlot, const float4* ppForce, float4* out, hipStream_t s);
void launch_pp_slot(int N, const int* index, int count, float* slot, hipStream_t s);
// exclusive prefix sum of data[0..n) in place; n <= 4 * 1024 * 1024, blockSums holds ceil(n / 1024) ints
void launch_scan_excl(int* data, int n, int* blockSums, hipStream_t s);
// Morton order: keys from FP64 arithmetic, a counting sort over 2^21 keys made
// stable by ranking each key's members by index; perm[dst] = src, into[src] = dst
void launch_morton(const float4* pos, int N, double cx, double cy, double cz, int* keys, int* slot, int* hist,
                   int* blockSums, int* perm, int* into, hipStream_t s);
void launch_permute(int N, const int* perm, const float4* pos, const float4* vel, const float4* aux, const float* q,
                    const float* mat, float4* oPos, float4* oVel, float4* oAux, float* oQ, float* oMat,
                    hipStream_t s);
void launch_books(int N, double mref, const float4* pos, const float4* vel, const float* mat, const float4* diag,
                  double* partial, double* out, hipStream_t s);

// ---- the learned far field (cc_ff.hip) ----
// per box: last error, judged, fallbacks, learned mesh steps, largest / summed / counted error from a usable base, spare
constexpr int FF_STATS = 8;
struct FFNet;
struct FFWork;
// One (instance, box) of a batched inference: its mesh and box, its present features and base (from its
// FFWork), its gate (a mesh step the network may stand in for) or none (a judge step), and where the
// prediction goes.
struct FFSlot {
  const float4* coarse;
  const float4* mono;
  const Box* box;
  int which;
  float4* featT;
  float* metaT;
  const float4* featB;
  const float* metaB;
  const int* validB;
  int* okBase;
  int* gate;
  const float* RB;
  float4 *A, *T0, *T1;
};
// One (instance, box) judged in a batch: its exact pass (mesh, far field, force mesh), the prediction
// (learned: farA, farT0, farT1 of 4096 each), the judge's sums, bound, statistics and gate, and the base
// the exact pass becomes.
struct FFJudgeSlot {
  const float4 *fine, *com, *ccom, *coarse, *learned, *farA, *farT0, *farT1, *force;
  const Box* box;
  int which;
  float gm, tol;
  double* sums;
  float* stats;
  int* gate;
  int* okBase;
  float* RB;
  const float4* featT;
  const float* metaT;
  float4* featB;
  float* metaB;
  int* validB;
};
size_t ff_weight_count();
FFNet* ff_create(int cap, std::string& err);
void ff_destroy(FFNet* n);
FFWork* ffwork_create(std::string& err);
void ffwork_destroy(FFWork* w);
// the network's far field for box `which` at a mesh step (the box's base plus the predicted change) into
// farA/T0/T1; gate (between judges): cleared when the base cannot serve, the network skipped unless set
void ff_predict(const FFNet* n, FFWork* wk, int which, const float4* coarse, const float4* mono, const Box* bx,
                float4* farA, float4* farT0, float4* farT1, int* gate, hipStream_t s);
// a judge step after ff_predict (into learned: 3 x 4096) and the exact far field and force mesh: the judge's
// verdict (gate, stats) for the coming mesh steps, then the exact pass becomes the box's base
void ff_judge_rebase(FFWork* wk, int which, const float4* fine, const float4* com, const float4* ccom, const float4* coarse,
                     const float4* learned, const float4* farA, const float4* farT0, const float4* farT1,
                     const float4* force, const Box* bx, float gm, double* sums, float tol, float* stats, int* gate,
                     hipStream_t s);
// both boxes' bases forgotten (a new state): no mesh step stands in before the next judge records one
void ff_forget_bases(FFWork* w, hipStream_t s);
// the features of box `which`'s base by block (what k_far reads to find the blocks the network has no base for)
const float4* ff_base_features(const FFWork* w, int which);
// a ctx's slot of a batched inference for box `which` (gate: a mesh step between judges; null: a judge step),
// and its judged slot
FFSlot ff_slot(FFWork* wk, int which, const float4* coarse, const float4* mono, const Box* box, int* gate, float4* A,
               float4* T0, float4* T1);
FFJudgeSlot ff_judge_slot(FFWork* wk, int which, const float4* fine, const float4* com, const float4* ccom,
                          const float4* coarse, const float4* learned, const float4* farA, const float4* farT0,
                          const float4* farT1, const float4* force, const Box* box, float gm, double* sums, float tol,
                          float* stats, int* gate);
// a batch's workspace for batched inference: activations for up to cap slots, a slot table per captured cycle
// (FF_BATCH_TABLES of them) and one more (FF_EAGER_TABLE) for a mesh step outside a whole cycle
constexpr int FF_BATCH_TABLES = 4;
constexpr int FF_EAGER_TABLE = FF_BATCH_TABLES;
struct FFBatchWork;
FFBatchWork* ffbatch_create(int cap, std::string& err);
void ffbatch_destroy(FFBatchWork* w);
int ffbatch_cap(const FFBatchWork* w);
// table t's slots (synchronous upload; no launch of a cycle using table t may be running)
int ffbatch_set(FFBatchWork* w, int t, const std::vector<FFSlot>& slots, const std::vector<FFJudgeSlot>& judge,
                std::string& err);
// one forward pass over table t's nSlots slots; the judge of table t's nJudge judged slots
void ff_predict_batch(const FFNet* n, FFBatchWork* w, int t, int nSlots, hipStream_t s);
void ff_judge_batch(FFBatchWork* w, int t, int nJudge, hipStream_t s);
// a recorded pass of box `which`'s exact far field (features, centre record); -1 when the record is full
int ff_record_pass(FFNet* n, const float4* fine, const float4* coarse, const float4* mono, const float4* farA, const float4* farT0, const float4* farT1, const Box* bx, int which,
                   hipStream_t s);
void launch_count_learned(const int* gate, float* stats, hipStream_t s);

// ---- the analysis (cc_fof.hip) ----
struct FofWork;
struct FofIn {
  int N;
  const float4* pos;       // device state
  const float4* vel;
  const float* mat;
  const float4* diag;      // the DIAG read-out on the device ...
  const float* hostDiag;   // ... or this host copy of it when not null
  double m, a, spring, touch;
  int cap, flags;
};
FofWork* fof_create();
void fof_destroy(FofWork* w);
// fills *out (a cc_report) and list; synchronous on stream s
int fof_analyse(FofWork* w, const FofIn& in, cc_report* out, int32_t* list, hipStream_t s, std::string& err);
// the same in two halves: the device's part and the read-back (synchronous on s), then the host's
// sums on a thread of their own; fof_end waits for them and delivers the report
int fof_begin(FofWork* w, const FofIn& in, hipStream_t s, std::string& err);
int fof_end(FofWork* w, cc_report* out, int32_t* list, bool wait, bool* ready, std::string& err);
bool fof_pending(const FofWork* w);

}  // namespace cc

cc::FFNet* ff_net(cc_ff* f);
