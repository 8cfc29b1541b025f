// gfx950 kernels of the cosmic-collisions simulation step.
//
// Every kernel restates one pass of the reference's WebGL2 pipeline
// (src/glsl-sim.js) as a compute kernel over linear HBM arrays.  The FP32
// arithmetic of each output follows the shader's operation order (the build
// uses -ffp-contract=off and correctly rounded division and square root), so
// the step pass matches the reference per particle.  Where the reference
// relies on fixed-function hardware the kernels use what CDNA4 has instead:
//   * the depth-peeled contact seats (8 point draws under a depth test) are
//     an atomic-counter bucket per hashed cell and a selection of the eight
//     lowest indices -- the same seat words, bit for bit;
//   * the additively blended mass deposit is an int64 fixed-point atomic sum
//     -- deterministic, and exact up to 2^-40 per contribution;
//   * the loops that one fragment ran alone (the far field, the per-cell
//     near field) are spread over the 64 lanes of a wavefront and reduced
//     with cross-lane shuffles.
#include <cmath>
#include <cstring>

#include "cc_fpmath.h"
#include "cc_internal.h"

namespace cc {

// ---------------------------------------------------------------- helpers --
// GLSL inversesqrt as 1/sqrt(x), both correctly rounded, for a normal x > 0 (a squared distance plus a softening)
__device__ __forceinline__ float rsq(float x) { return qrcp(qsqrt(x)); }
__device__ __forceinline__ float gmax(float a, float b) { return a > b ? a : b; }  // GLSL max
__device__ __forceinline__ float gmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
__device__ __forceinline__ float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return ax * bx + ay * by + az * bz;
}
__device__ __forceinline__ int shl(int v, int s) { return (int)((uint32_t)v << s); }
__device__ __forceinline__ int fine_at(int cx, int cy, int cz) { return ((cz >> 3) * 64 + cy) * 512 + ((cz & 7) * 64 + cx); }
__device__ __forceinline__ int coarse_at(int cx, int cy, int cz) { return ((cz >> 2) * 16 + cy) * 64 + ((cz & 3) * 16 + cx); }
__device__ __forceinline__ float mod4(float x) { return x - 4.0f * floorf(x / 4.0f); }

// blocks are dispatched to the 8 XCDs round-robin; hand each XCD a contiguous
// run of blocks so that neighbouring particles (Morton order) share its L2
__device__ __forceinline__ int xcd_block() {
  int nb = gridDim.x, b = blockIdx.x;
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// src/glsl-sim.js:79-91 hashAt
__device__ __forceinline__ int hash_cell(const Grid& g, int cx, int cy, int cz) {
  int px = cx >> g.GShift, py = cy >> g.GShift, pz = cz >> g.GShift;
  cy += shl(px, g.GShift - 1) + shl(pz, g.GShift - 2);
  cz += shl(py, g.GShift - 1);
  cx &= g.GMask; cy &= g.GMask; cz &= g.GMask;
  int ax = shl(cz & g.SXMask, g.GShift) + cx;
  int ay = shl(cz >> g.SXShift, g.GShift) + cy;
  return (ay << g.AWShift) + ax;
}

// src/glsl-sim.js:25-37 boxOf
__device__ __forceinline__ int box_of(const Box& b, float px, float py, float pz, float& gx, float& gy, float& gz) {
  float ax = (px + b.LA[0]) * b.invH - 0.5f, ay = (py + b.LA[1]) * b.invH - 0.5f, az = (pz + b.LA[2]) * b.invH - 0.5f;
  float bx = (px + b.LB[0]) * b.invH - 0.5f, by = (py + b.LB[1]) * b.invH - 0.5f, bz = (pz + b.LB[2]) * b.invH - 0.5f;
  bool inA = ax >= 0.0f && ay >= 0.0f && az >= 0.0f && ax <= 63.0f && ay <= 63.0f && az <= 63.0f;
  bool inB = b.boxB > 0.5f && bx >= 0.0f && by >= 0.0f && bz >= 0.0f && bx <= 63.0f && by <= 63.0f && bz <= 63.0f;
  if (inA && inB) {
    float mA = gmax(fabsf(ax - 31.5f), gmax(fabsf(ay - 31.5f), fabsf(az - 31.5f)));
    float mB = gmax(fabsf(bx - 31.5f), gmax(fabsf(by - 31.5f), fabsf(bz - 31.5f)));
    inA = mA <= mB;
    inB = !inA;
  }
  gx = inA ? ax : bx; gy = inA ? ay : by; gz = inA ? az : bz;
  return inA ? 0 : inB ? 1 : -1;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// one atomic per wavefront for an append to a shared list; every lane of the
// wave must call it (want = whether this lane appends); returns the slot or -1
__device__ __forceinline__ int wave_append(int* counter, bool want) {
  unsigned long long mask = __ballot(want);
  if (mask == 0ull) return -1;
  int lane = (int)__lane_id();
  int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(mask));
  base = __shfl(base, leader, 64);
  return want ? base + __popcll(mask & ((1ull << lane) - 1ull)) : -1;
}

// ------------------------------------------------------------------ fills --
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}
void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  int blocks = (int)((n + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, n, v);
}
void launch_fill_f32(float* p, size_t n, float v, hipStream_t s) {
  uint32_t bits;
  std::memcpy(&bits, &v, 4);
  launch_fill_u32(reinterpret_cast<uint32_t*>(p), n, bits, s);
}

// ----------------------------------------------------- contact grid (A1-A3) --
// A rebuild (src/glsl-sim.js:133-147, the depth-peeled seat passes) in two
// steps: a count and a seat pass.  The count takes each particle's hashed
// cell and seat word (index, place in the cell to a fifteenth, bit 30), an
// arrival in the cell's counter and bucket (arrivals past the bucket to the
// overflow list), and lists the cell on its first arrival in the wavefront's
// 64-entry segment of the occupied list; the bucket holds the seat words
// themselves (the index is their low bits).  The count runs as a pass of its own
// or at the end of the step pass before a rebuild, on the positions that step
// writes (k_sim<true>): the same positions, the same counts.  Counters, lists
// and overflow counters come in two sets used by alternate rebuilds.

// the particle at (px, py, pz) with index i: its hashed cell h and seat word
__device__ __forceinline__ uint32_t seat_word(float px, float py, float pz, float invCell, const Grid& g, int i,
                                              int& h) {
  float gx = px * invCell, gy = py * invCell, gz = pz * invCell;
  float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
  h = hash_cell(g, (int)fx, (int)fy, (int)fz);
  uint32_t qx = (uint32_t)gclamp((gx - fx) * 15.0f, 0.0f, 14.0f);
  uint32_t qy = (uint32_t)gclamp((gy - fy) * 15.0f, 0.0f, 14.0f);
  uint32_t qz = (uint32_t)gclamp((gz - fz) * 15.0f, 0.0f, 14.0f);
  int B = g.IB;
  return (uint32_t)i | (qx << B) | (qy << (B + 4)) | (qz << (B + 8)) | (B == 18 ? 0x40000000u : 0u);
}

// One particle's arrival, by a lane of a wavefront whose lanes hold particles
// i = 64 s .. 64 s + 63 (segment s); lanes past N must have left.
__device__ __forceinline__ void count_arrival(int i, uint32_t w, int h, const SlotsOut& o) {
  int old = atomicAdd(&o.cnt[h], 1);
  const unsigned long long first = __ballot(old == 0);
  const int seg = i >> 6;
  if (old == 0)   // a cell's first arrival lists it, in lane order within the segment
    o.occ[(seg << 6) + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(first >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)first, 0u))] = h;
  if ((i & 63) == 0) o.segN[seg] = __popcll(first);
  if (old < BUCKET_CAP) o.bucket[(size_t)h * BUCKET_CAP + old] = w;
  else o.ovf[atomicAdd(o.ovfN, 1)] = make_int2(h, (int)w);
}

__global__ void __launch_bounds__(256) k_slots_count(const float4* __restrict__ pos, int N, float invCell, Grid g,
                                                     SlotsOut o) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float4 p = pos[i];
  int h;
  uint32_t w = seat_word(p.x, p.y, p.z, invCell, g, i, h);
  count_arrival(i, w, h, o);
}

// The seat pass.  Per cell of this rebuild: the eight lowest arrival indices,
// ascending -- what the reference's depth peel leaves in seats 0..7.  Per cell
// of the previous rebuild: its seats back to empty where nothing arrived this
// time (a cell with arrivals has all eight seats written), and its counter of
// that rebuild back to zero for the rebuild after this one.  Neither job
// writes what the other reads.
// (t: the thread's place among the instance's N)
__device__ __forceinline__ void seat_body(const int* __restrict__ occ, const int* __restrict__ segN,
                                          const int* __restrict__ occPrev, const int* __restrict__ segPrev, int N,
                                          const int* __restrict__ cnt, int* __restrict__ cntPrev,
                                          const uint32_t* __restrict__ bucket, const int2* __restrict__ ovf,
                                          const int* __restrict__ ovfN, int* __restrict__ ovfNOther, uint32_t idxMask,
                                          uint32_t* __restrict__ seats, int t) {
  if (t == 0) *ovfNOther = 0;   // the overflow counter of the next rebuild
  if (t >= N) return;
  // The loads that wait on nothing but each other, issued before any store: the segment lengths, the
  // cells, their counters and the first four arrivals (an occupied cell has at least one).
  const int sp = segPrev[t >> 6], sn = segN[t >> 6];
  const bool prev = (t & 63) < sp, cur = (t & 63) < sn;
  const int hp = prev ? occPrev[t] : 0, h = cur ? occ[t] : 0;
  const int np = prev ? cnt[hp] : 1, n = cur ? cnt[h] : 0;
  const uint4* b4 = reinterpret_cast<const uint4*>(bucket + (size_t)h * BUCKET_CAP);
  const uint4 v0 = cur ? b4[0] : make_uint4(0u, 0u, 0u, 0u);
  if (prev) {
    if (np == 0) {
      uint4 e = make_uint4(EMPTY_SEAT, EMPTY_SEAT, EMPTY_SEAT, EMPTY_SEAT);
      uint4* sv = reinterpret_cast<uint4*>(seats + (size_t)hp * SLOTS);
      sv[0] = e; sv[1] = e;
    }
    cntPrev[hp] = 0;
  }
  if (!cur) return;
  const int m = n < BUCKET_CAP ? n : BUCKET_CAP;
  uint32_t wd[BUCKET_CAP];   // the arrivals' seat words
  int ix[BUCKET_CAP];        // and their particle indices (past the arrivals: above any index)
#pragma unroll
  for (int q = 0; q < BUCKET_CAP / 4; q++) {
    uint4 v = q == 0 ? v0 : (q * 4 < m) ? b4[q] : make_uint4(0u, 0u, 0u, 0u);
    wd[q * 4] = v.x; wd[q * 4 + 1] = v.y; wd[q * 4 + 2] = v.z; wd[q * 4 + 3] = v.w;
#pragma unroll
    for (int u = 0; u < 4; u++) ix[q * 4 + u] = q * 4 + u < m ? (int)(wd[q * 4 + u] & idxMask) : 0x7fffffff;
  }
  int nov = n > BUCKET_CAP ? *ovfN : 0;
  uint32_t out[SLOTS];
  int last = -1;
#pragma unroll
  for (int k = 0; k < SLOTS; k++) {
    int best = 0x7fffffff;
    uint32_t bw = EMPTY_SEAT;
#pragma unroll
    for (int q = 0; q < BUCKET_CAP; q++) {
      bool take = ix[q] > last && ix[q] < best;
      best = take ? ix[q] : best;
      bw = take ? wd[q] : bw;
    }
    for (int o = 0; o < nov; o++) {       // a cell past its bucket: its other members
      int2 e = ovf[o];
      int v = (int)((uint32_t)e.y & idxMask);
      if (e.x == h && v > last && v < best) { best = v; bw = (uint32_t)e.y; }
    }
    out[k] = bw;
    if (best != 0x7fffffff) last = best;
  }
  uint4* sv = reinterpret_cast<uint4*>(seats + (size_t)h * SLOTS);
  sv[0] = make_uint4(out[0], out[2], out[4], out[6]);
  sv[1] = make_uint4(out[1], out[3], out[5], out[7]);
}

__global__ void __launch_bounds__(256) k_slots_seat(const int* __restrict__ occ, const int* __restrict__ segN,
                                                    const int* __restrict__ occPrev, const int* __restrict__ segPrev,
                                                    int N, const int* __restrict__ cnt, int* __restrict__ cntPrev,
                                                    const uint32_t* __restrict__ bucket, const int2* __restrict__ ovf,
                                                    const int* __restrict__ ovfN, int* __restrict__ ovfNOther,
                                                    uint32_t idxMask, uint32_t* __restrict__ seats) {
  seat_body(occ, segN, occPrev, segPrev, N, cnt, cntPrev, bucket, ovf, ovfN, ovfNOther, idxMask, seats,
            blockIdx.x * blockDim.x + threadIdx.x);
}

// The seat passes of several instances in one launch: work-group b serves job blockJob[b], its local block
// b - block0 (a batch whose instances step together launches one per step instead of one per instance).
__global__ void __launch_bounds__(SEAT_JOB_BLOCK) k_slots_seat_jobs(const SeatJob* __restrict__ jobs,
                                                         const uint16_t* __restrict__ blockJob, int blocks) {
  const int b = (int)blockIdx.x;
  if (b >= blocks) return;
  const SeatJob& J = jobs[blockJob[b]];
  seat_body(J.occ, J.segN, J.occPrev, J.segPrev, J.N, J.cnt, J.cntPrev, J.bucket, J.ovf, J.ovfN, J.ovfNOther,
            J.idxMask, J.seats, (b - J.block0) * SEAT_JOB_BLOCK + (int)threadIdx.x);
}

// A count that will not be seated (the particles moved since): its counters
// and overflow counter back to zero.
__global__ void __launch_bounds__(256) k_slots_discard(const int* __restrict__ occ, const int* __restrict__ segN, int N,
                                                       int* cnt, int* ovfN) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) *ovfN = 0;
  if (t < N && (t & 63) < segN[t >> 6]) cnt[occ[t]] = 0;
}

SlotsOut slots_out(GridBuf& gb, int N, int parity) {
  SlotsOut o;
  o.cnt = gb.cnt[parity];
  o.bucket = gb.bucket;
  o.occ = gb.occ[parity];
  o.segN = gb.segN + (size_t)parity * ((N + 63) / 64);
  o.ovf = gb.ovf;
  o.ovfN = gb.ovfN + parity;
  return o;
}

void launch_slots_count(const float4* pos, int N, float invCell, const Grid& g, GridBuf& gb, hipStream_t s) {
  int blocks = (N + 255) / 256;
  blocks = (blocks + 7) & ~7;   // a multiple of the 8 XCDs, so the block remap applies
  hipLaunchKernelGGL(k_slots_count, dim3(blocks), dim3(256), 0, s, pos, N, invCell, g, slots_out(gb, N, gb.cur ^ 1));
}

void launch_slots_seat(int N, GridBuf& gb, hipStream_t s) {
  int prev = gb.cur, next = gb.cur ^ 1, blocks = (N + 255) / 256;
  SlotsOut n = slots_out(gb, N, next), p = slots_out(gb, N, prev);
  hipLaunchKernelGGL(k_slots_seat, dim3(blocks), dim3(256), 0, s, n.occ, n.segN, p.occ, p.segN, N, n.cnt, p.cnt,
                     gb.bucket, gb.ovf, n.ovfN, p.ovfN, gb.idxMask, gb.seats);
  gb.cur = next;
}

SeatJob seat_job(int N, GridBuf& gb) {
  int prev = gb.cur, next = gb.cur ^ 1;
  SlotsOut n = slots_out(gb, N, next), p = slots_out(gb, N, prev);
  SeatJob j;
  j.occ = n.occ; j.segN = n.segN; j.occPrev = p.occ; j.segPrev = p.segN; j.N = N; j.cnt = n.cnt; j.cntPrev = p.cnt;
  j.bucket = gb.bucket; j.ovf = gb.ovf; j.ovfN = n.ovfN; j.ovfNOther = p.ovfN; j.idxMask = gb.idxMask;
  j.seats = gb.seats; j.block0 = 0;
  gb.cur = next;
  return j;
}

void launch_slots_seat_jobs(const SeatJob* jobs, const uint16_t* blockJob, int blocks, hipStream_t s) {
  if (blocks > 0)
    hipLaunchKernelGGL(k_slots_seat_jobs, dim3(blocks), dim3(SEAT_JOB_BLOCK), 0, s, jobs, blockJob, blocks);
}

void launch_slots_discard(int N, GridBuf& gb, hipStream_t s) {
  SlotsOut n = slots_out(gb, N, gb.cur ^ 1);
  hipLaunchKernelGGL(k_slots_discard, dim3((N + 255) / 256), dim3(256), 0, s, n.occ, n.segN, N, n.cnt, n.ovfN);
}

void launch_slots(const float4* pos, int N, float invCell, const Grid& g, GridBuf& gb, hipStream_t s) {
  launch_slots_count(pos, N, invCell, g, gb, s);
  launch_slots_seat(N, gb, s);
}

// ----------------------------------------------------------- deposit (A4) --
__device__ __forceinline__ unsigned long long to_fixed(float v, double scale) {
  return (unsigned long long)(long long)rint((double)v * scale);
}

// One particle per lane: its eight cloud-in-cell corners into its own box's
// mesh, or -- a stray in neither box -- its pull on the mesh into the scratch
// texel (src/glsl-sim.js:168-219).  The work-group's corners are privatised
// in an LDS tile when its particles fall in a small block of cells (Morton
// order makes that the common case); each touched cell is flushed to the
// mesh with one atomic add per channel.  Fixed-point int64 throughout, so
// the sums are exact and independent of the order the adds land in.
constexpr int TILE = 8;
__global__ void __launch_bounds__(256) k_deposit(const float4* __restrict__ pos, const float* __restrict__ mat, int N,
                                                 const int* __restrict__ visit, const Box* __restrict__ bxp, int which,
                                                 const float4* __restrict__ mono, float eps2, unsigned long long* acc,
                                                 int* occN) {
  __shared__ unsigned long long tile[TILE * TILE * TILE * 4];
  __shared__ unsigned long long stray[4];   // the work-group's strays, summed before one atomic per channel
  __shared__ int sum[4];                    // the deposited particles' cells, summed, and their count
  // the particles in the step pass's visit order (or index order): a work-group's particles then lie in a
  // small block of cells, whatever the index order; the sums are integers, so the order changes nothing
  const int at = xcd_block() * blockDim.x + threadIdx.x;
  if (at == 0) *occN = 0;   // the occupied list k_deposit_finish appends to
  bool live = at < N;
  const int i = live && visit ? visit[at] : at;
  float4 p = live ? pos[i] : make_float4(0, 0, 0, 0);
  float mr = live ? mat[i] : 0.0f;
  float gx = 0, gy = 0, gz = 0;
  const Box b = *bxp;
  int mine = live ? box_of(b, p.x, p.y, p.z, gx, gy, gz) : -2;
  bool dep = mine == which;
  int cx = 0, cy = 0, cz = 0;
  float tx = 0, ty = 0, tz = 0;
  if (dep) {
    float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    tx = gx - fx; ty = gy - fy; tz = gz - fz;
    cx = (int)fx; cy = (int)fy; cz = (int)fz;
  }
  if (threadIdx.x < 4) { sum[threadIdx.x] = 0; stray[threadIdx.x] = 0ull; }
  for (int t = threadIdx.x; t < TILE * TILE * TILE * 4; t += blockDim.x) tile[t] = 0ull;
  __syncthreads();
  {   // the cells summed per wavefront first: one LDS atomic per wavefront and component, not one per lane
    int sx = wave_sum(dep ? cx : 0), sy = wave_sum(dep ? cy : 0), sz = wave_sum(dep ? cz : 0), sn = wave_sum(dep ? 1 : 0);
    if ((threadIdx.x & 63) == 0 && sn > 0) {
      atomicAdd(&sum[0], sx); atomicAdd(&sum[1], sy); atomicAdd(&sum[2], sz); atomicAdd(&sum[3], sn);
    }
  }
  __syncthreads();
  // The tile is centred on the work-group's mean cell; a particle whose corners
  // fall outside it (an outlier of a work-group the motion has spread since the
  // last Morton sort) adds to the mesh directly, the rest through the tile.
  int nd = max(sum[3], 1);
  int ox = sum[0] / nd - (TILE / 2 - 1), oy = sum[1] / nd - (TILE / 2 - 1), oz = sum[2] / nd - (TILE / 2 - 1);
  bool tiled = dep && cx >= ox && cy >= oy && cz >= oz && cx + 1 < ox + TILE && cy + 1 < oy + TILE && cz + 1 < oz + TILE;
  if (dep) {
    // Each lane walks its eight corners from its own starting corner: the lanes of a wavefront lie in a few
    // cells, and in step they would all add to the same few texels at once, which the LDS serialises.  The
    // sums are integers, so the order of the adds changes nothing.
    const int k0 = threadIdx.x & 7;
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
      const int k = (kk + k0) & 7;
      int o0 = k & 1, o1 = (k >> 1) & 1, o2 = k >> 2;
      int c0 = cx + o0, c1 = cy + o1, c2 = cz + o2;
      if (c0 > 63 || c1 > 63 || c2 > 63) continue;
      float w = gmix(1.0f - tx, tx, (float)o0) * gmix(1.0f - ty, ty, (float)o1) * gmix(1.0f - tz, tz, (float)o2) * mr;
      unsigned long long v0 = to_fixed(p.x * w, DEP_SCALE), v1 = to_fixed(p.y * w, DEP_SCALE);
      unsigned long long v2 = to_fixed(p.z * w, DEP_SCALE), v3 = to_fixed(1.0f * w, DEP_SCALE);
      if (tiled) {
        int l = ((c2 - oz) * TILE + (c1 - oy)) * TILE + (c0 - ox);
        atomicAdd(&tile[l * 4], v0); atomicAdd(&tile[l * 4 + 1], v1);
        atomicAdd(&tile[l * 4 + 2], v2); atomicAdd(&tile[l * 4 + 3], v3);
      } else {
        unsigned long long* d = acc + (size_t)fine_at(c0, c1, c2) * 4;
        atomicAdd(d, v0); atomicAdd(d + 1, v1); atomicAdd(d + 2, v2); atomicAdd(d + 3, v3);
      }
    }
  } else if (live && mine < 0) {
    float4 mo = mono[0];
    const float w = gmax(mo.w, 1e-6f), yw = qrcp(w);
    float rx = qdiv_by(mo.x, w, yw) - p.x, ry = qdiv_by(mo.y, w, yw) - p.y, rz = qdiv_by(mo.z, w, yw) - p.z;
    float inv = rsq(dot3(rx, ry, rz, rx, ry, rz) + eps2);
    float s = mr * inv * inv * inv;
    atomicAdd(&stray[0], to_fixed(rx * s, SCR_SCALE)); atomicAdd(&stray[1], to_fixed(ry * s, SCR_SCALE));
    atomicAdd(&stray[2], to_fixed(rz * s, SCR_SCALE)); atomicAdd(&stray[3], to_fixed(mr, SCR_SCALE));
  }
  __syncthreads();
  // integer sums: the same whatever order the work-groups' adds land in
  if (threadIdx.x < 4 && stray[threadIdx.x] != 0ull) atomicAdd(acc + (size_t)SCRATCH * 4 + threadIdx.x, stray[threadIdx.x]);
  for (int l = threadIdx.x; l < TILE * TILE * TILE; l += blockDim.x) {
    unsigned long long w = tile[l * 4 + 3], a0 = tile[l * 4], a1 = tile[l * 4 + 1], a2 = tile[l * 4 + 2];
    if ((w | a0 | a1 | a2) == 0ull) continue;
    int c0 = ox + l % TILE, c1 = oy + (l / TILE) % TILE, c2 = oz + l / (TILE * TILE);
    if (c0 > 63 || c1 > 63 || c2 > 63) continue;
    unsigned long long* d = acc + (size_t)fine_at(c0, c1, c2) * 4;
    atomicAdd(d, a0); atomicAdd(d + 1, a1); atomicAdd(d + 2, a2); atomicAdd(d + 3, w);
  }
}

// A mesh texel's centre of mass (its mass-weighted sums over its mass, each
// quotient as the cell and far-field passes of the reference take it) and its
// mass: computed once per deposit, where every cell pass used to recompute it
// for each of the ~100 texels that read it.
__device__ __forceinline__ float4 centre_of(float4 v) {
  if (!(v.w > 0.0f)) return make_float4(0.0f, 0.0f, 0.0f, v.w);
  const float y = qrcp(v.w);   // the three quotients share 1/w
  return make_float4(qdiv_by(v.x, v.w, y), qdiv_by(v.y, v.w, y), qdiv_by(v.z, v.w, y), v.w);
}

// The accumulator to the FP32 mesh (and back to zero), each texel's centre of
// mass; the occupied texels listed.  The list's order is immaterial (each
// entry's cell pass writes its own texel); a work-group of 1024 texels takes
// its places in it with one atomic, where one per wavefront had thousands of
// them queue on the counter once the ejecta spread over the mesh.
constexpr int FINISH_BLOCK = 1024;
__global__ void __launch_bounds__(FINISH_BLOCK) k_deposit_finish(unsigned long long* acc, float4* fine, float4* com,
                                                                 int* occ, int* occN, float4* force) {
  __shared__ int waveBase[FINISH_BLOCK / 64];
  __shared__ int groupBase;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool app = false;
  if (t < FINE_TEXELS) {
    unsigned long long* a = acc + (size_t)t * 4;
    double sc = t == SCRATCH ? 1.0 / SCR_SCALE : 1.0 / DEP_SCALE;
    float4 v = make_float4((float)((double)(long long)a[0] * sc), (float)((double)(long long)a[1] * sc),
                           (float)((double)(long long)a[2] * sc), (float)((double)(long long)a[3] * sc));
    a[0] = 0; a[1] = 0; a[2] = 0; a[3] = 0;
    fine[t] = v;
    com[t] = centre_of(v);
    app = t < SCRATCH && v.w > 0.0f;
    // given a force mesh: its empty texels, which the cell pass (it writes the occupied ones and the scratch row) leaves
    if (!app && force && t < SCRATCH) force[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  const unsigned long long m = __ballot(app);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) waveBase[wave] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int n = 0;
    for (int w = 0; w < FINISH_BLOCK / 64; w++) { const int c = waveBase[w]; waveBase[w] = n; n += c; }
    groupBase = n ? atomicAdd(occN, n) : 0;
  }
  __syncthreads();
  if (app)
    occ[groupBase + waveBase[wave] +
        (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = t;
}
// the centres of a mesh written from outside (fine or coarse)
__global__ void __launch_bounds__(256) k_centres(const float4* __restrict__ mesh, float4* com, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) com[t] = centre_of(mesh[t]);
}
void launch_centres(const float4* mesh, float4* com, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_centres, dim3((n + 255) / 256), dim3(256), 0, s, mesh, com, n);
}


void launch_deposit(const float4* pos, const float* mat, int N, const int* visit, const Box* b, int box, float eps2,
                    MeshBuf& m, bool zeroForce, hipStream_t s) {
  int blocks = (N + 255) / 256;
  hipLaunchKernelGGL(k_deposit, dim3(blocks), dim3(256), 0, s, pos, mat, N, visit, b, box, m.mono, eps2, m.acc,
                     m.occN);
  hipLaunchKernelGGL(k_deposit_finish, dim3((FINE_TEXELS + FINISH_BLOCK - 1) / FINISH_BLOCK), dim3(FINISH_BLOCK), 0, s,
                     m.acc, m.fine, m.com, m.occ,
                     m.occN, zeroForce ? m.force : nullptr);
}

// A plain copy of n float4 (read once, written once): the calibration of the
// HBM traffic counters against a known byte count.
__global__ void __launch_bounds__(256) k_copy4(const float4* __restrict__ src, float4* dst, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
void launch_copy4(const float4* src, float4* dst, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy4, dim3((n + 255) / 256), dim3(256), 0, s, src, dst, n);
}

// The mesh boxes live in device memory so that a captured step cycle reads
// the box of the cycle it is replayed for: k_put_box writes one (eager
// steps), k_box_next takes the next of a call's schedule (graph replays).
__global__ void k_put_box(Box b, Box* dst) {
  if (threadIdx.x == 0) *dst = b;
}
__global__ void k_box_next(const Box* __restrict__ sched, int* idx, Box* dst) {
  if (threadIdx.x == 0) { int k = *idx; *dst = sched[k]; *idx = k + 1; }
}
void launch_put_box(const Box& b, Box* dst, hipStream_t s) { hipLaunchKernelGGL(k_put_box, dim3(1), dim3(64), 0, s, b, dst); }
void launch_box_next(const Box* sched, int* idx, Box* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_box_next, dim3(1), dim3(64), 0, s, sched, idx, dst);
}

// ------------------------------------------------------ coarse + mono (A5) --
// The coarse mesh (each 4^3 block of fine texels summed in order) and, by the
// last of the 16 work-groups to finish, the monopole: the coarse texels summed
// sixteen a thread, then across the wavefronts and the work-group -- the
// monopole pass of its own before, now without a launch of its own (the
// counter that finds the last work-group is put back to zero by it).
__global__ void __launch_bounds__(256) k_coarse(const float4* __restrict__ fine, float4* coarse, float4* ccom,
                                                float4* mono, int* done) {
  __shared__ float4 part[4];
  __shared__ int last;
  int t = blockIdx.x * blockDim.x + threadIdx.x;   // the grid is 16 x 256: one thread per coarse texel
  int mx = t & 63, my = t >> 6;
  int c0 = mx & 15, c1 = my & 15, c2 = (my >> 4) * 4 + (mx >> 4);
  float4 s = make_float4(0, 0, 0, 0);
  for (int z = 0; z < 4; z++) for (int y = 0; y < 4; y++) for (int x = 0; x < 4; x++) {
    float4 f = fine[fine_at(c0 * 4 + x, c1 * 4 + y, c2 * 4 + z)];
    s.x += f.x; s.y += f.y; s.z += f.z; s.w += f.w;
  }
  coarse[t] = s;
  ccom[t] = centre_of(s);
  __threadfence();   // this work-group's texels visible before it is counted
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(done, 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();   // every other work-group's texels seen (the L1 is invalidated)
  float4 m = make_float4(0, 0, 0, 0);
  for (int u = threadIdx.x; u < 4096; u += 256) {
    float4 c = coarse[u];
    m.x += c.x; m.y += c.y; m.z += c.z; m.w += c.w;
  }
  m.x = wave_sum(m.x); m.y = wave_sum(m.y); m.z = wave_sum(m.z); m.w = wave_sum(m.w);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float4 r = part[0];
    for (int w = 1; w < 4; w++) { r.x += part[w].x; r.y += part[w].y; r.z += part[w].z; r.w += part[w].w; }
    mono[0] = r;
    *done = 0;
  }
}

// -------------------------------------------------------- far field (A6) --
// One wavefront per coarse block; its lanes share the 4096 source blocks.
// While a learned far field stands in (*gate set) only the blocks it has no
// base for run: occupied now and empty in baseFeat (by block, (z*16 + y)*16 + x).
__global__ void __launch_bounds__(256) k_far(const float4* __restrict__ ccom, float4* farA, float4* farT0,
                                             float4* farT1, const int* __restrict__ gate,
                                             const float4* __restrict__ baseFeat) {
  const bool learned = gate && *gate != 0;
  if (learned && !baseFeat) return;
  int lane = threadIdx.x & 63;
  int me = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (me >= 4096) return;
  int mx = me & 63, my = me >> 6;
  int c0 = mx & 15, c1 = my & 15, c2 = (my >> 4) * 4 + (mx >> 4);
  if (learned && baseFeat[(c2 * 16 + c1) * 16 + c0].x > 0.0f) return;
  float4 mc = ccom[me];
  if (mc.w <= 0.0f) {
    if (!learned && lane == 0) {
      farA[me] = make_float4(0, 0, 0, 0); farT0[me] = make_float4(0, 0, 0, 0); farT1[me] = make_float4(0, 0, 0, 0);
    }
    return;
  }
  float x0 = mc.x, x1 = mc.y, x2 = mc.z;
  float a0 = 0, a1 = 0, a2 = 0, pot = 0, d0 = 0, d1 = 0, d2 = 0, o0 = 0, o1 = 0, o2 = 0;
  // four source blocks' loads in flight at a time; each lane's sum in the same block order
  for (int k4 = lane; k4 < 4096; k4 += 256) {
    float4 mq[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int k = k4 + 64 * u;
      mq[u] = ccom[coarse_at(k & 15, (k >> 4) & 15, k >> 8)];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int k = k4 + 64 * u;
      int kx = k & 15, ky = (k >> 4) & 15, kz = k >> 8;
      int dx = abs(kx - c0), dy = abs(ky - c1), dz = abs(kz - c2);
      int dm = max(dx, max(dy, dz));
      if (dm <= 2) continue;
      float4 m = mq[u];
      if (m.w <= 0.0f) continue;
      float r0 = m.x - x0, r1 = m.y - x1, r2 = m.z - x2;
      float inv2 = qrcp(dot3(r0, r1, r2, r0, r1, r2)), inv = qsqrt(inv2), inv3 = inv * inv2, inv5 = inv3 * inv2;
      float s = m.w * inv3;
      a0 += r0 * s; a1 += r1 * s; a2 += r2 * s;
      pot -= m.w * inv;
      d0 += m.w * (3.0f * r0 * r0 * inv5 - inv3);
      d1 += m.w * (3.0f * r1 * r1 * inv5 - inv3);
      d2 += m.w * (3.0f * r2 * r2 * inv5 - inv3);
      o0 += m.w * (3.0f * (r0 * r1) * inv5);
      o1 += m.w * (3.0f * (r0 * r2) * inv5);
      o2 += m.w * (3.0f * (r1 * r2) * inv5);
    }
  }
  a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2); pot = wave_sum(pot);
  d0 = wave_sum(d0); d1 = wave_sum(d1); d2 = wave_sum(d2);
  o0 = wave_sum(o0); o1 = wave_sum(o1); o2 = wave_sum(o2);
  if (lane == 0) {
    farA[me] = make_float4(a0, a1, a2, pot);
    farT0[me] = make_float4(d0, d1, d2, o0);
    farT1[me] = make_float4(o1, o2, 0.0f, 0.0f);
  }
}

// ------------------------------------------------- per-cell gravity (A7) --
// 1/sqrt(x) for a normal x > 0: the hardware estimate and one Newton step
// (within an ulp; the cell pass sums in its own order, so the reference's
// 1.0 / sqrt() need not be rounded as it rounds).
__device__ __forceinline__ float rsqrt_nr(float x) {
  const float y = __builtin_amdgcn_rsqf(x), h = 0.5f * x;
  const float e = fmaf(-h * y, y, 0.5f);
  return fmaf(y, e, y);
}

// One wavefront per occupied fine cell; lane l takes fine cell l of each of
// the 5^3 surrounding coarse blocks; the far field comes from the block's
// tidal tensor (src/glsl-sim.js:303-352).
constexpr int CELL_AHEAD = 4;   // blocks whose cells are loaded together
__global__ void __launch_bounds__(256) k_cell(const float4* __restrict__ fine, const float4* __restrict__ com,
                                              const float4* __restrict__ coarse, const float4* __restrict__ ccom,
                                              const float4* __restrict__ farA, const float4* __restrict__ farT0,
                                              const float4* __restrict__ farT1, const int* __restrict__ occ,
                                              const int* occN, float eps2, float gm, float4* force) {
  int lane = threadIdx.x & 63;
  int wave = blockIdx.x * 4 + (threadIdx.x >> 6), nwave = gridDim.x * 4;
  int n = *occN;
  if (blockIdx.x == 0) {           // the scratch row is carried through
    for (int x = threadIdx.x; x < 512; x += blockDim.x) force[SCRATCH + x] = fine[SCRATCH + x];
  }
  int fx = lane & 3, fy = (lane >> 2) & 3, fz = lane >> 4;
  for (int w = wave; w < n; w += nwave) {
    int me = occ[w];
    int mx = me & 511, my = me >> 9;
    int c0 = mx & 63, c1 = my & 63, c2 = (my >> 6) * 8 + (mx >> 6);
    float4 mc = com[me];
    float x0 = mc.x, x1 = mc.y, x2 = mc.z;
    int q0 = c0 >> 2, q1 = c1 >> 2, q2 = c2 >> 2;
    float a0 = 0, a1 = 0, a2 = 0, pot = 0;
    // The 5^3 coarse blocks around, b = (kz+2)*25 + (ky+2)*5 + (kx+2): which hold
    // mass, read all at once (two blocks a lane) into two wavefront masks; the
    // walk then loads the cells of CELL_AHEAD held blocks together before
    // taking their terms, in the blocks' order.
    auto held = [&](int b) {
      if (b >= 125) return false;
      int k0 = q0 + b % 5 - 2, k1 = q1 + (b / 5) % 5 - 2, k2 = q2 + b / 25 - 2;
      if (k0 < 0 || k1 < 0 || k2 < 0 || k0 > 15 || k1 > 15 || k2 > 15) return false;
      return coarse[coarse_at(k0, k1, k2)].w > 0.0f;
    };
    unsigned long long m0 = __ballot(held(lane)), m1 = __ballot(held(lane + 64));
    while (m0 | m1) {
      float4 mf[CELL_AHEAD];
      bool self[CELL_AHEAD];
#pragma unroll
      for (int k = 0; k < CELL_AHEAD; k++) {
        int b = -1;
        if (m0) { b = __ffsll((long long)m0) - 1; m0 &= m0 - 1; }
        else if (m1) { b = __ffsll((long long)m1) - 1 + 64; m1 &= m1 - 1; }
        self[k] = true;
        mf[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (b >= 0) {
          int k0 = q0 + b % 5 - 2, k1 = q1 + (b / 5) % 5 - 2, k2 = q2 + b / 25 - 2;
          int f0 = k0 * 4 + fx, f1 = k1 * 4 + fy, f2 = k2 * 4 + fz;
          self[k] = f0 == c0 && f1 == c1 && f2 == c2;
          mf[k] = com[fine_at(f0, f1, f2)];
        }
      }
#pragma unroll
      for (int k = 0; k < CELL_AHEAD; k++) {
        if (self[k] || mf[k].w <= 0.0f) continue;
        float r0 = mf[k].x - x0, r1 = mf[k].y - x1, r2 = mf[k].z - x2;
        float inv = rsqrt_nr(dot3(r0, r1, r2, r0, r1, r2) + eps2);
        float s = mf[k].w * inv * inv * inv;
        a0 += r0 * s; a1 += r1 * s; a2 += r2 * s;
        pot -= mf[k].w * inv;
      }
    }
    a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2); pot = wave_sum(pot);
    if (lane == 0) {
      int ca = coarse_at(q0, q1, q2);
      float4 mb = ccom[ca], fa = farA[ca], t0 = farT0[ca], t1 = farT1[ca];
      float d0 = x0 - mb.x, d1 = x1 - mb.y, d2 = x2 - mb.z;
      float T0 = t0.x * d0 + t0.w * d1 + t1.x * d2;
      float T1 = t0.w * d0 + t0.y * d1 + t1.y * d2;
      float T2 = t1.x * d0 + t1.y * d1 + t0.z * d2;
      a0 += fa.x + T0; a1 += fa.y + T1; a2 += fa.z + T2;
      pot += fa.w - dot3(fa.x, fa.y, fa.z, d0, d1, d2) - 0.5f * dot3(d0, d1, d2, T0, T1, T2);
      force[me] = make_float4(a0 * gm, a1 * gm, a2 * gm, pot * gm);
    }
  }
}

void launch_coarse(MeshBuf& m, hipStream_t s) {
  hipLaunchKernelGGL(k_coarse, dim3(16), dim3(256), 0, s, m.fine, m.coarse, m.ccom, m.mono, m.coarseDone);
}
void launch_far(const MeshBuf& m, float4* farA, float4* farT0, float4* farT1, const int* gate, hipStream_t s,
                const float4* baseFeat) {
  hipLaunchKernelGGL(k_far, dim3(1024), dim3(256), 0, s, m.ccom, farA, farT0, farT1, gate, baseFeat);
}
void launch_cell_only(const MeshBuf& m, const float4* farA, const float4* farT0, const float4* farT1, float4* force,
                      float eps2, float gm, int blocks, bool zeroed, hipStream_t s) {
  if (!zeroed) launch_fill_f32(reinterpret_cast<float*>(force), (size_t)FINE_TEXELS * 4, 0.0f, s);
  hipLaunchKernelGGL(k_cell, dim3(blocks), dim3(256), 0, s, m.fine, m.com, m.coarse, m.ccom, farA, farT0, farT1, m.occ,
                     m.occN, eps2, gm, force);
}
void launch_cell(MeshBuf& m, float eps2, float gm, int blocks, bool zeroed, hipStream_t s) {
  launch_far(m, m.farA, m.farT0, m.farT1, nullptr, s);
  launch_cell_only(m, m.farA, m.farT0, m.farT1, m.force, eps2, gm, blocks, zeroed, s);
}

// ------------------------------------------------------ pairwise P3M (A11) --
// The loose list's seats (PPGATHER_FS, src/glsl-sim.js:676-698): seat s < count
// holds particle index[s]; the rest, and an entry outside [0, N), are empty.
// (the list's length: *countDev where given -- a captured cycle then reads the length of the present list --
// else countArg; the same in the three passes)
__global__ void __launch_bounds__(256) k_pp_gather(const float4* __restrict__ pos, const float* __restrict__ mat,
                                                   int N, const Box* __restrict__ bxp, const int* __restrict__ index,
                                                   const int* __restrict__ countDev, int countArg, float4* list,
                                                   float* box) {
  const int count = countDev ? *countDev : countArg;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= PP_CAP) return;
  int i = s < count ? index[s] : -1;
  if (i < 0 || i >= N) { list[s] = make_float4(0, 0, 0, 0); box[s] = -1.0f; return; }
  float4 p = pos[i];
  float gx, gy, gz;
  const Box b = *bxp;
  box[s] = (float)box_of(b, p.x, p.y, p.z, gx, gy, gz);
  list[s] = make_float4(p.x, p.y, p.z, mat[i]);
}

// The pairwise pass (src/glsl-sim.js:712-754): each seat's four slices, slice
// s the sum over the s-th quarter of the list in the reference's order -- one
// sequential FP32 chain per seat, slice and component, as the fragment's loop
// adds it, whatever the list's length.  A wavefront (a work-group of its own)
// owns PP_ROWS seats and one slice and walks the slice in tiles of PP_COLS
// entries:
//   1. its lanes clear the tile's terms, take the tile's distances (a seat
//      against every eighth entry) and list the pairs within the cut, by
//      ballot -- the other pairs' terms stay zero, as the reference skips them;
//   2. the listed pairs' terms, 64 at a time with no lane idle on a pair
//      beyond the cut, into the LDS tile, component-major;
//   3. lane (component, seat) adds the tile's terms in entry order to its
//      running sum, four entries per LDS read.
// Every pair is evaluated once for each end, as the reference evaluates it.
constexpr int PP_ROWS = 8;
constexpr int PP_COLS = PP_TILE;
constexpr int PP_CHAIN = PP_COLS + 4;   // floats per (component, seat) row of the term tile: 16-byte reads of
                                        // eight consecutive rows fall in distinct banks
// (8 seats x 48 entries keep a work-group's LDS under 10 KiB: sixteen of them, a CU's limit, fit at once)

// Each tile's bounding box (its entries with mass; an empty tile's box is empty), so a wavefront can pass
// over a tile whose box lies beyond the cut from its seats' box: every distance is at least the boxes'
// gap, and FP32 rounding keeps that order (the differences, squares and sums are monotone in their
// non-negative operands), so none of the tile's pairs is within the cut and each of its terms would be
// zero.  A running sum that starts at +0 never becomes -0, so leaving out adds of +0 or -0 changes nothing.
__global__ void __launch_bounds__(64) k_pp_bounds(const float4* __restrict__ list, const int* __restrict__ countDev,
                                                  int countArg, float4* bounds) {
  const int count = countDev ? *countDev : countArg;
  const int lane = threadIdx.x, share = (count + 3) >> 2, tps = (share + PP_COLS - 1) / PP_COLS;
  if ((int)blockIdx.x >= 4 * tps) return;   // a grid for the cap: the tiles past this list's
  const int slice = blockIdx.x / tps, t = blockIdx.x % tps;
  const int j0 = slice * share + t * PP_COLS, j1 = min(min(j0 + PP_COLS, (slice + 1) * share), count);
  const int j = j0 + lane;
  float4 q = (lane < PP_COLS && j < j1) ? list[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const bool has = q.w != 0.0f;
  float lx = has ? q.x : INFINITY, ly = has ? q.y : INFINITY, lz = has ? q.z : INFINITY;
  float hx = has ? q.x : -INFINITY, hy = has ? q.y : -INFINITY, hz = has ? q.z : -INFINITY;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, o)); ly = fminf(ly, __shfl_xor(ly, o)); lz = fminf(lz, __shfl_xor(lz, o));
    hx = fmaxf(hx, __shfl_xor(hx, o)); hy = fmaxf(hy, __shfl_xor(hy, o)); hz = fmaxf(hz, __shfl_xor(hz, o));
  }
  if (lane == 0) {
    bounds[2 * blockIdx.x] = make_float4(lx, ly, lz, 0.0f);
    bounds[2 * blockIdx.x + 1] = make_float4(hx, hy, hz, 0.0f);
  }
}

// the squared gap between two boxes along the axes, in the order dot3 takes a distance
__device__ __forceinline__ float box_gap2(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  const float gx = fmaxf(fmaxf(blo.x - ahi.x, alo.x - bhi.x), 0.0f);
  const float gy = fmaxf(fmaxf(blo.y - ahi.y, alo.y - bhi.y), 0.0f);
  const float gz = fmaxf(fmaxf(blo.z - ahi.z, alo.z - bhi.z), 0.0f);
  return dot3(gx, gy, gz, gx, gy, gz);
}

__device__ __forceinline__ float4 pp_term(float4 P, float bi, float4 Q, float bj, float cut, float a2,
                                          float a, float ya, float yCut, float yTap, float tapD, const float2* st) {
  float d0 = Q.x - P.x, d1 = Q.y - P.y, d2 = Q.z - P.z;
  float r2 = dot3(d0, d1, d2, d0, d1, d2);   // below cut2: only such pairs are listed
  float r = qsqrt(r2);
  float x = gclamp(qdiv_by(r, cut, yCut) * 64.0f - 0.5f, 0.0f, 63.0f);
  int k = (int)x, k1 = min(k + 1, 63);
  float t = x - (float)k;
  float invN = qrcp(qsqrt(r2 + a2));
  // smoothstep(cut, 0.8 cut, r): (r - cut) / (0.8 cut - cut), both signs turned (exactly) to divide by a positive
  float taper = gclamp(qdiv_by(cut - r, tapD, yTap), 0.0f, 1.0f);
  taper = taper * taper * (3.0f - 2.0f * taper);
  float same = (bi >= 0.0f && bj == bi) ? 1.0f : 0.0f;
  float2 tk = st[k], tk1 = st[k1];
  float Tx = gmix(tk.x, tk1.x, t) * same, Ty = gmix(tk.y, tk1.y, t) * same;
  float fn = invN * invN * invN * r - qdiv(Tx, gmax(r2, a2));
  float rr = gmax(r, 1e-6f), yr = qrcp(rr);   // 1/rr correctly rounded, shared by four quotients
  float n0 = qdiv_by(d0, rr, yr), n1 = qdiv_by(d1, rr, yr), n2 = qdiv_by(d2, rr, yr);
  // max(r, a) is rr where r > a (a > 1e-6), else the uniform a with its reciprocal ya
  float ph = -invN - qdiv_by(Ty, gmax(r, a), r > a ? yr : ya);
  float s = Q.w * fn * taper;
  return make_float4(n0 * s, n1 * s, n2 * s, Q.w * ph * taper);
}

__global__ void __launch_bounds__(64) k_pp_force(const float4* __restrict__ list, const float* __restrict__ box,
                                                 const float4* __restrict__ bounds, const int* __restrict__ countDev,
                                                 int countArg, float gm, float cut, float a2,
                                                 const float2* __restrict__ tab, float4* force) {
  const int count = countDev ? *countDev : countArg;
  __shared__ __align__(16) float term[4 * PP_ROWS * PP_CHAIN];
  __shared__ float4 lP[PP_ROWS], lQ[PP_COLS];
  __shared__ float bP[PP_ROWS], bQ[PP_COLS];
  __shared__ unsigned short todo[PP_ROWS * PP_COLS];
  __shared__ float2 st[64];
  const int lane = threadIdx.x, slice = blockIdx.y, r0 = blockIdx.x * PP_ROWS;
  const int sc = lane / PP_ROWS, sr = lane % PP_ROWS;   // step 3 (lanes below 4 * PP_ROWS): component sc of seat r0 + sr
  float* out = reinterpret_cast<float*>(force + (size_t)slice * PP_CAP + r0 + sr) + sc;
  if (r0 >= count) {   // seats past the list: zero (the whole wavefront)
    if (lane < 4 * PP_ROWS) *out = 0.0f;
    return;
  }
  const int share = (count + 3) >> 2, j0 = slice * share, j1 = min(j0 + share, count);
  st[lane] = tab[lane];
  if (lane < PP_ROWS) {
    const int i = r0 + lane;
    lP[lane] = i < count ? list[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bP[lane] = i < count ? box[i] : -1.0f;
  }
  const int er = lane % PP_ROWS, ec = lane / PP_ROWS;   // step 1: seat r0 + er, entries ec, ec + 8, ...
  const int i = r0 + er;
  const float4 P = i < count ? list[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const bool live = i < count && P.w != 0.0f;
  const float cut2 = cut * cut, a = sqrtf(a2), tapD = cut - 0.8f * cut;
  const float yCut = qrcp(cut), yTap = qrcp(tapD), ya = qrcp(a);
  const float4* chain = reinterpret_cast<const float4*>(term + lane * PP_CHAIN);   // step 3's row (lane = sc * PP_ROWS + sr)
  float run = 0.0f;
  // the seats' box (the live ones; lanes differing in their low three bits hold the eight seats), the same
  // in every lane, and this slice's tile boxes
  float4 rlo = make_float4(live ? P.x : INFINITY, live ? P.y : INFINITY, live ? P.z : INFINITY, 0.0f);
  float4 rhi = make_float4(live ? P.x : -INFINITY, live ? P.y : -INFINITY, live ? P.z : -INFINITY, 0.0f);
#pragma unroll
  for (int o = 1; o < PP_ROWS; o <<= 1) {
    rlo.x = fminf(rlo.x, __shfl_xor(rlo.x, o)); rlo.y = fminf(rlo.y, __shfl_xor(rlo.y, o));
    rlo.z = fminf(rlo.z, __shfl_xor(rlo.z, o));
    rhi.x = fmaxf(rhi.x, __shfl_xor(rhi.x, o)); rhi.y = fmaxf(rhi.y, __shfl_xor(rhi.y, o));
    rhi.z = fmaxf(rhi.z, __shfl_xor(rhi.z, o));
  }
  const float4* tb = bounds + 2 * slice * ((share + PP_COLS - 1) / PP_COLS);
  // the tile's entries, loaded one tile ahead (lanes below PP_COLS)
  float4 nQ = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float nB = -1.0f;
  if (lane < PP_COLS && j0 + lane < j1) { nQ = list[j0 + lane]; nB = box[j0 + lane]; }
  for (int jt = j0, tt = 0; jt < j1; jt += PP_COLS, tt++) {
    const bool beyond = box_gap2(rlo, rhi, tb[2 * tt], tb[2 * tt + 1]) >= cut2;   // every lane alike
    if (lane < PP_COLS) {
      if (!beyond) {
        lQ[lane] = nQ;
        bQ[lane] = nB;
      }
      const int j = jt + PP_COLS + lane;
      nQ = j < j1 ? list[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      nB = j < j1 ? box[j] : -1.0f;
    }
    if (beyond) continue;
    float4* t4 = reinterpret_cast<float4*>(term);
#pragma unroll
    for (int k = lane; k < PP_ROWS * PP_CHAIN; k += 64) t4[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    __syncthreads();
    int n = 0;
#pragma unroll
    for (int k = 0; k < PP_COLS / 8; k++) {
      const int jl = ec + 8 * k, j = jt + jl;   // ec = lane / 8 < 8
      const float4 Q = lQ[jl];
      const float d0 = Q.x - P.x, d1 = Q.y - P.y, d2 = Q.z - P.z;
      const bool in = live && j < j1 && j != i && dot3(d0, d1, d2, d0, d1, d2) < cut2;
      const unsigned long long m = __ballot(in);
      if (in)
        todo[n + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] =
            (unsigned short)(er | (jl << 3));
      n += __popcll(m);
    }
    __syncthreads();
    for (int t = lane; t < n; t += 64) {
      const int e = todo[t], r = e & (PP_ROWS - 1), jl = e >> 3;
      const float4 v = pp_term(lP[r], bP[r], lQ[jl], bQ[jl], cut, a2, a, ya, yCut, yTap, tapD, st);
      float* w = term + r * PP_CHAIN + jl;
      w[0] = v.x;
      w[PP_ROWS * PP_CHAIN] = v.y;
      w[2 * PP_ROWS * PP_CHAIN] = v.z;
      w[3 * PP_ROWS * PP_CHAIN] = v.w;
    }
    __syncthreads();
    if (lane < 4 * PP_ROWS) {
      float4 v[PP_COLS / 4];
#pragma unroll
      for (int c = 0; c < PP_COLS / 4; c++) v[c] = chain[c];
#pragma unroll
      for (int c = 0; c < PP_COLS / 4; c++) { run += v[c].x; run += v[c].y; run += v[c].z; run += v[c].w; }
    }
    __syncthreads();
  }
  if (lane < 4 * PP_ROWS) {
    const int ir = r0 + sr;
    *out = (ir < count && list[ir].w != 0.0f) ? run * gm : 0.0f;
  }
}

void launch_pp(const float4* pos, const float* mat, int N, const Box* b, const int* index, const int* countDev,
               int count, float gm, float cut, float a2, const float2* tab, float4* list, float* box, float4* bounds,
               float4* force, hipStream_t s) {
  hipLaunchKernelGGL(k_pp_gather, dim3(PP_CAP / 256), dim3(256), 0, s, pos, mat, N, b, index, countDev, count, list,
                     box);
  // with the length on the device the tiles' grid is the cap's, the work-groups past the list's tiles leaving
  const int tiles = countDev ? PP_TILES_MAX : 4 * (((count + 3) / 4 + PP_COLS - 1) / PP_COLS);
  if (tiles > 0) hipLaunchKernelGGL(k_pp_bounds, dim3(tiles), dim3(64), 0, s, list, countDev, count, bounds);
  hipLaunchKernelGGL(k_pp_force, dim3(PP_CAP / PP_ROWS, 4), dim3(64), 0, s, list, box, bounds, countDev, count, gm, cut,
                     a2, tab, force);
}

__global__ void k_pp_slot(int N, const int* __restrict__ index, int count, float* slot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) slot[i] = -1.0f;
}
__global__ void k_pp_slot_set(const int* __restrict__ index, int count, float* slot) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) slot[index[k]] = (float)k;
}
void launch_pp_slot(int N, const int* index, int count, float* slot, hipStream_t s) {
  hipLaunchKernelGGL(k_pp_slot, dim3((N + 255) / 256), dim3(256), 0, s, N, index, count, slot);
  if (count > 0) hipLaunchKernelGGL(k_pp_slot_set, dim3((count + 255) / 256), dim3(256), 0, s, index, count, slot);
}

// -------------------------------------------------------- the step (A12) --
__device__ __forceinline__ void mass_at(float px, float py, float pz, float4 mo, float gm, float eps2, float& ox,
                                        float& oy, float& oz) {
  const float w = gmax(mo.w, 1e-6f), yw = qrcp(w);   // the centre's three quotients share 1/w
  float rx = qdiv_by(mo.x, w, yw) - px, ry = qdiv_by(mo.y, w, yw) - py, rz = qdiv_by(mo.z, w, yw) - pz;
  float inv = rsq(dot3(rx, ry, rz, rx, ry, rz) + eps2);
  float s = gm * mo.w * inv * inv * inv;
  ox = rx * s; oy = ry * s; oz = rz * s;
}
__device__ __forceinline__ float pot_at(float px, float py, float pz, float4 mo, float gm, float eps2) {
  const float w = gmax(mo.w, 1e-6f), yw = qrcp(w);   // the centre's three quotients share 1/w
  float rx = qdiv_by(mo.x, w, yw) - px, ry = qdiv_by(mo.y, w, yw) - py, rz = qdiv_by(mo.z, w, yw) - pz;
  return -gm * mo.w * rsq(dot3(rx, ry, rz, rx, ry, rz) + eps2);
}
__device__ __forceinline__ void cross_at(float4 mo, float4 mm, float gm, float eps2, float& ox, float& oy, float& oz) {
  const float wo = gmax(mo.w, 1e-6f), wm = gmax(mm.w, 1e-6f), yo = qrcp(wo), ym = qrcp(wm);
  float dx = qdiv_by(mo.x, wo, yo) - qdiv_by(mm.x, wm, ym);
  float dy = qdiv_by(mo.y, wo, yo) - qdiv_by(mm.y, wm, ym);
  float dz = qdiv_by(mo.z, wo, yo) - qdiv_by(mm.z, wm, ym);
  float inv = rsq(dot3(dx, dy, dz, dx, dy, dz) + eps2);
  float s = gm * mo.w * inv * inv * inv;
  ox = dx * s; oy = dy * s; oz = dz * s;
}

// Contacts over the 27 cells around (spring-dashpot with the no-tension
// clamp, dissipated power, heat exchange), gravity from the mesh of the
// particle's box (or the monopoles, or a constant field), then symplectic
// Euler (src/glsl-sim.js:362-533).
//
// One lane per particle, its contact sums in the reference's order (cells dz,
// dy, dx; seats ascending), which keeps every partial sum the reference's.
// A wavefront runs the union of its lanes' work, so the contacts are found
// in three compacted phases instead of eight predicated seat slots per cell:
//   A  the seats that pass the cell and seat-place prefilters, in order,
//      into the lane's LDS list (the particle's own seat among them);
//   B  their positions: linked pairs other than the particle itself counted
//      (the reference skips its own index first), touching pairs kept in order
//      (r^2 < touch2r: the r^2 below which sqrt(r^2) rounds under the contact
//      distance -- the reference's touch test without its square root);
//   C  the touching pairs' velocities and masses, and the sums.
// A lane whose list would overflow finishes alone on the direct path.
constexpr int SIM_BLOCK = 64;
static_assert(SIM_BLOCK == SIM_JOB_BLOCK, "a step-pass job's work-group is the pass's own");
constexpr int CAND_CAP = 31;                 // a list longer than this takes the direct path
constexpr int CAND_ROWS = CAND_CAP + 9;      // A's writes run up to a cell's eight seats past the row after the
                                             // cap; B reads eight rows at a time from below the cap

struct Contact {
  float pushx, pushy, pushz, nb, lost, flow;
};

// The pair's dashpot and conduction factors depend only on the two masses and
// heat capacities; for a partner of the particle's own kind they are the
// particle's own (the same operations on the same operands), taken once.
struct PairSelf {
  float cii, hii;
};
__device__ __forceinline__ PairSelf pair_self(const SimU& u, float ri, float ci) {
  return { u.c * qsqrt(qdiv(2.0f * ri * ri, ri + ri)), qdiv(2.0f * ci * ci, ci + ci) };
}

// one candidate pair's share, exactly as the step pass computes it; the
// factors are computed when any lane of the wavefront has a partner of
// another kind, and taken from `self` otherwise
__device__ __forceinline__ void contact_pair(const SimU& u, float px, float py, float pz, float vx, float vy, float vz,
                                             float Ti, float ri, float ci, const PairSelf& self, float4 Pj, float4 Vj,
                                             float rj, Contact& a) {
  float ddx = Pj.x - px, ddy = Pj.y - py, ddz = Pj.z - pz;
  float r = qsqrt(dot3(ddx, ddy, ddz, ddx, ddy, ddz));
  float rr = gmax(r, 1e-6f);
  const float yr = qrcp(rr);   // 1/rr correctly rounded, shared by the three quotients
  float nx = qdiv_by(ddx, rr, yr), ny = qdiv_by(ddy, rr, yr), nz = qdiv_by(ddz, rr, yr);
  float ov = u.touch - r;
  float cij = self.cii;
  if (__any(rj != ri)) cij = u.c * qsqrt(qdiv(2.0f * ri * rj, ri + rj));
  float vn = dot3(Vj.x - vx, Vj.y - vy, Vj.z - vz, nx, ny, nz);
  float fs = u.k * ov - cij * vn, fn = gmax(fs, 0.0f), on = fs < 0.0f ? 0.0f : 1.0f;
  a.pushx -= nx * fn; a.pushy -= ny * fn; a.pushz -= nz * fn;
  a.lost += cij * vn * vn * on + u.k * ov * vn * (1.0f - on);
  float cj = rj * (mod4(Pj.w) < 0.5f ? u.cpIron : 1.0f);
  float hij = self.hii;
  if (__any(cj != ci)) hij = qdiv(2.0f * ci * cj, ci + cj);
  a.flow += (Vj.w - Ti) * hij;
}

// The seat prefilter (seatPos and the reach test, src/glsl-sim.js:109,
// :426-427): the seat's recorded place, (q + 0.5)/15 of a cell into it, within
// reach (cell + 0.6a, src/nbody.js:41, :297) of p.  It only prunes: the 0.6a
// of slack over the link distance is the reference's bound on how far a
// particle sits from its seat's place (a fifteenth of a cell and a step of
// motion), so a seat near the edge of reach holds a particle beyond the link
// test, and how the reach test rounds cannot change which pairs are summed.
// Here it runs in integers on a lattice of 1/480 cell: the seat's place is
// exactly 32 q + 16 from its cell's corner, the particle is rounded onto the
// lattice (half a unit per axis at most) and the bound widened by that
// (reachQ2), so the test keeps every seat the reference keeps and the few
// more it admits are dropped by the link test.  Three shifts, three
// multiply-adds and a compare, where the places in FP32 took a dozen.
// Buffer-resource loads (a 32-bit byte offset from a base in scalar
// registers, no 64-bit address arithmetic per load); gfx9's descriptor word 3
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint4 bload_u4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ float4 bload_f4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ float bload_f1(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// The lattice offsets lie within +-2048, so the squares are exact in the
// 24-bit multiply-add, which reads its operands' low 24 bits as signed: the
// sign extension that __mul24 spells out is a no-op here and is left out.
__device__ __forceinline__ int seat_d2(uint32_t sv, int B, int bx, int by, int bz) {
  int ex = (int)(__builtin_amdgcn_ubfe(sv, (uint32_t)B, 4u) << 5) + bx;
  int ey = (int)(__builtin_amdgcn_ubfe(sv, (uint32_t)B + 4u, 4u) << 5) + by;
  int ez = (int)(__builtin_amdgcn_ubfe(sv, (uint32_t)B + 8u, 4u) << 5) + bz;
  int d2;
  asm("v_mul_i32_i24 %0, %1, %1\n\tv_mad_i32_i24 %0, %2, %2, %0\n\tv_mad_i32_i24 %0, %3, %3, %0"
      : "=&v"(d2) : "v"(ez), "v"(ey), "v"(ex));
  return d2;
}

// The pass for the particle at place `at` of the visit order, by lane threadIdx.x % 64 (cand: the
// work-group's candidate lists, entry-major: a lane's list is a column, no bank conflicts).  PLANES 1: a
// one-wavefront work-group, the wavefront walks the whole neighbourhood.  PLANES 3: three wavefronts for the
// same 64 particles, wavefront w walks the z-plane dz = w - 1 into list w (rows w * CAND_ROWS on); PLANES 9:
// nine, wavefront w the row dz = w / 3 - 1, dy = w % 3 - 1.  The lists are joined in that order (the
// reference's cell order) behind list 0 -- the same list, so the same sums -- and wavefront 1 takes the
// gravity while wavefront 0 takes the contacts: more wavefronts in flight where the particles are too few
// to fill the GPU.
template <bool COUNT, int PLANES = 1>
__device__ __forceinline__ void sim_body(int (*cand)[SIM_BLOCK], const SimU& u, const Box* __restrict__ bxp,
                                         float touch2r, const float4* __restrict__ pos,
                                         const float4* __restrict__ vel, const float* __restrict__ mat,
                                         const float* __restrict__ q, const uint32_t* __restrict__ seats,
                                         const float4* __restrict__ force0, const float4* __restrict__ mono0,
                                         const float4* __restrict__ force1, const float4* __restrict__ mono1,
                                         const float* __restrict__ ppSlot, const float4* __restrict__ ppForce,
                                         float4* oPos, float4* oVel, float4* oAux, float* oQ, const SlotsOut& cnt,
                                         int at) {
  const int lane = (int)threadIdx.x & (SIM_BLOCK - 1), w = PLANES == 1 ? 0 : (int)threadIdx.x / SIM_BLOCK;
  if (at >= u.N) return;
  const int me = u.visit ? u.visit[at] : at;
  float4 P = pos[me];
  float px = P.x, py = P.y, pz = P.z;
  Contact acc = { 0, 0, 0, 0, 0, 0 };
  int cx = (int)floorf(px * u.invCell), cy = (int)floorf(py * u.invCell), cz = (int)floorf(pz * u.invCell);
  float c0x = (float)cx * u.cell, c0y = (float)cy * u.cell, c0z = (float)cz * u.cell;
  // A: candidates in order.  A z-plane's nine cells have their seat loads
  // issued together (each under the cell's cull test), then are walked in
  // order.  Every seat's word is written at the list's end and kept by
  // counting it when it passes; a wavefront leaves a cell at the first seat
  // that is empty for all its lanes (seats fill from the first, so an empty
  // seat ends its cell).
  const __amdgpu_buffer_rsrc_t seatR = buf(seats, (uint32_t)(u.grid.cells + 1) * (SLOTS * 4));
  const __amdgpu_buffer_rsrc_t posR = buf(pos, (uint32_t)u.N * 16), velR = buf(vel, (uint32_t)u.N * 16),
                               matR = buf(mat, (uint32_t)u.N * 4);
  const float qScale = (float)SEAT_Q / u.cell;
  // (the particle lies in its own cell, 0..480 on the lattice; the clamp only tells the compiler the range)
  const int pqx = min(max((int)rintf((px - c0x) * qScale), -1024), 1024),
            pqy = min(max((int)rintf((py - c0y) * qScale), -1024), 1024),
            pqz = min(max((int)rintf((pz - c0z) * qScale), -1024), 1024);
  const int B = u.grid.IB;
  const uint32_t idxMask = (1u << B) - 1u;
  // the list's next entry as a byte offset into the LDS array (row stride SIM_BLOCK ints), held at the row
  // past the cap after each cell
  const uint32_t rowB = SIM_BLOCK * 4u, wa0 = (uint32_t)lane * 4u, waEnd = (CAND_CAP + 1) * rowB + wa0;
  char* const candB = reinterpret_cast<char*>(&cand[w * CAND_ROWS][0]);
  uint32_t wa = wa0;
  // the cells a wavefront walks in a plane: all nine, or (PLANES 9) its row of three, dy = w % 3 - 1
  constexpr int NC = PLANES == 9 ? 3 : 9;
  const int dzw = PLANES == 9 ? w / 3 - 1 : w - 1;
  for (int dz = PLANES == 1 ? -1 : dzw; dz <= (PLANES == 1 ? 1 : dzw); dz++) {
    uint4 sa[NC], sb[NC];
    bool ok[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
      int dy = PLANES == 9 ? w % 3 - 1 : i / 3 - 1, dx = PLANES == 9 ? i - 1 : i % 3 - 1;
      float lox = c0x + (float)dx * u.cell, loy = c0y + (float)dy * u.cell, loz = c0z + (float)dz * u.cell;
      float qx = gclamp(px, lox, lox + u.cell) - px, qy = gclamp(py, loy, loy + u.cell) - py,
            qz = gclamp(pz, loz, loz + u.cell) - pz;
      ok[i] = !(dot3(qx, qy, qz, qx, qy, qz) > u.link2);
      // a culled cell reads the always-empty cell past the grid: no branch, and its lanes' addresses coincide
      uint32_t off = (uint32_t)(ok[i] ? hash_cell(u.grid, cx + dx, cy + dy, cz + dz) : u.grid.cells) * (SLOTS * 4);
      sa[i] = bload_u4(seatR, off);
      sb[i] = bload_u4(seatR, off + 16);
    }
    int bz = SEAT_Q * dz + 16 - pqz;
    asm volatile("" : "+v"(bz));   // held in a register: one shift-add per seat, not two
#pragma unroll
    for (int i = 0; i < NC; i++) {
      const unsigned long long okm = __ballot(ok[i]);
      if (okm == 0ull) continue;
      const int bx = SEAT_Q * ((PLANES == 9 ? i : i % 3) - 1) + 16 - pqx,
                by = SEAT_Q * ((PLANES == 9 ? w % 3 : i / 3) - 1) + 16 - pqy;
      uint32_t sv8[8] = { sa[i].x, sb[i].x, sa[i].y, sb[i].y, sa[i].z, sb[i].z, sa[i].w, sb[i].w };
#pragma unroll
      for (int k = 0; k < SLOTS; k++) {
        uint32_t sv = sv8[k];
        if ((__ballot((int)sv >= 0) & okm) == 0ull) break;
        const bool open = ok[i] & ((int)sv >= 0);
        // (the particle's own seat passes here and is dropped by the link test below)
        const bool pass = open & (seat_d2(sv, B, bx, by, bz) <= u.reachQ2);
        *reinterpret_cast<uint32_t*>(candB + wa) = sv;   // the seat word; its index is taken below
        wa += pass ? rowB : 0u;
      }
      wa = min(wa, waEnd);
    }
  }
  // gravity at the particle: the constant field, or the mesh of its box (the monopoles outside both) and
  // the pairwise pass's four slices
  auto gravity = [&](float& gvx, float& gvy, float& gvz) {
    gvx = 0; gvy = 0; gvz = 0;
    if (u.field) {
      gvx = u.gField[0]; gvy = u.gField[1]; gvz = u.gField[2];
    } else {
      float gx, gy, gz;
      const Box B = *bxp;
      int box = box_of(B, px, py, pz, gx, gy, gz);
      float4 m0 = mono0[0], m1 = mono1[0];
      if (box >= 0) {
        const float4* F = box == 1 ? force1 : force0;
        float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
        float tx = gx - fx, ty = gy - fy, tz = gz - fz;
        int ix = (int)fx, iy = (int)fy, iz = (int)fz;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          int o0 = k & 1, o1 = (k >> 1) & 1, o2 = k >> 2;
          float wt = gmix(1.0f - tx, tx, (float)o0) * gmix(1.0f - ty, ty, (float)o1) * gmix(1.0f - tz, tz, (float)o2);
          float4 f = F[fine_at(min(ix + o0, 63), min(iy + o1, 63), min(iz + o2, 63))];
          gvx += f.x * wt; gvy += f.y * wt; gvz += f.z * wt;
        }
        float4 sc = F[SCRATCH];
        gvx -= sc.x * u.gm; gvy -= sc.y * u.gm; gvz -= sc.z * u.gm;
        if (B.boxB > 0.5f) {
          float ox, oy, oz;
          if (box == 0) cross_at(m1, m0, u.gm, u.eps2, ox, oy, oz); else cross_at(m0, m1, u.gm, u.eps2, ox, oy, oz);
          gvx += ox; gvy += oy; gvz += oz;
        }
      } else {
        mass_at(px, py, pz, m0, u.gm, u.eps2, gvx, gvy, gvz);
        if (B.boxB > 0.5f) {
          float ox, oy, oz;
          mass_at(px, py, pz, m1, u.gm, u.eps2, ox, oy, oz);
          gvx += ox; gvy += oy; gvz += oz;
        }
      }
      float seat = ppSlot[me];
      if (seat >= 0.0f) {
        int si = (int)seat;
        for (int s = 0; s < 4; s++) {
          float4 pf = ppForce[s * PP_CAP + si];
          gvx += pf.x; gvy += pf.y; gvz += pf.z;
        }
      }
    }
  };
  int n = (int)((wa - wa0) / rowB);
  if (PLANES > 1) {
    // each list's length in its last row (a slack row, past every write of the walk), then the lists behind
    // list 0 in plane order, unless they overflow together
    cand[w * CAND_ROWS + CAND_ROWS - 1][lane] = n;
    __syncthreads();
    int total = 0, off = 0;
#pragma unroll
    for (int v = 0; v < PLANES; v++) {
      const int nv = cand[v * CAND_ROWS + CAND_ROWS - 1][lane];
      off += v < w ? nv : 0;
      total += nv;
    }
    if (w > 0 && total <= CAND_CAP)
      for (int k = 0; k < n; k++) cand[off + k][lane] = cand[w * CAND_ROWS + k][lane];
    __syncthreads();
    if (w > 0) {   // wavefront 1 takes gravity into list 1's first rows (free now) while wavefront 0 goes on
      if (w == 1) {
        float gx, gy, gz;
        gravity(gx, gy, gz);
        cand[CAND_ROWS][lane] = __float_as_int(gx);
        cand[CAND_ROWS + 1][lane] = __float_as_int(gy);
        cand[CAND_ROWS + 2][lane] = __float_as_int(gz);
      }
      __syncthreads();
      return;
    }
    n = total;
  }
  const bool over = n > CAND_CAP;
  // the particle's own velocity, temperature and mass, needed from the contacts on
  float4 V = vel[me];
  float ri = mat[me];
  float ci = ri * (mod4(P.w) < 0.5f ? u.cpIron : 1.0f);
  float vx = V.x, vy = V.y, vz = V.z, Ti = V.w;
  const PairSelf self = pair_self(u, ri, ci);
  if (!over) {
    // B: linked pairs counted, touching ones kept (in place, in order); positions fetched eight at a time.
    // Branch-free: the eight rows are read at once, rows past the list stand for the particle itself (which
    // the link test drops), and each entry is written at the kept list's end, which only a kept one advances
    // (that row is never above the rows already read).
    int nt = 0;
    for (int t0 = 0; t0 < n; t0 += 8) {
      float4 Pj[8];
      int jj[8];
      uint32_t w[8];
#pragma unroll
      for (int k = 0; k < 8; k++) w[k] = (uint32_t)cand[t0 + k][lane];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        jj[k] = t0 + k < n ? (int)(w[k] & idxMask) : me;
        Pj[k] = bload_f4(posR, (uint32_t)jj[k] << 4);
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        float ddx = Pj[k].x - px, ddy = Pj[k].y - py, ddz = Pj[k].z - pz;
        float r2 = dot3(ddx, ddy, ddz, ddx, ddy, ddz);
        const bool linked = (r2 < u.link2) & (jj[k] != me);
        acc.nb += linked ? 1.0f : 0.0f;
        cand[nt][lane] = jj[k];
        nt += (int)(linked & (r2 < touch2r));
      }
    }
    // C: the contacts, in order; their data fetched four at a time
    for (int t0 = 0; t0 < nt; t0 += 4) {
      float4 Pj[4], Vj[4];
      float rj[4];
      uint32_t w[4];
#pragma unroll
      for (int k = 0; k < 4; k++) w[k] = (uint32_t)cand[t0 + k][lane];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        uint32_t j = t0 + k < nt ? w[k] : (uint32_t)me;
        Pj[k] = bload_f4(posR, j << 4); Vj[k] = bload_f4(velR, j << 4); rj[k] = bload_f1(matR, j << 2);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (t0 + k >= nt) break;
        contact_pair(u, px, py, pz, vx, vy, vz, Ti, ri, ci, self, Pj[k], Vj[k], rj[k], acc);
      }
    }
  } else {
    // the direct path: every cell and seat in order, as the reference walks them
    for (int dz = -1; dz <= 1; dz++) for (int dy = -1; dy <= 1; dy++) for (int dx = -1; dx <= 1; dx++) {
      float lox = c0x + (float)dx * u.cell, loy = c0y + (float)dy * u.cell, loz = c0z + (float)dz * u.cell;
      float qx = gclamp(px, lox, lox + u.cell) - px, qy = gclamp(py, loy, loy + u.cell) - py,
            qz = gclamp(pz, loz, loz + u.cell) - pz;
      if (dot3(qx, qy, qz, qx, qy, qz) > u.link2) continue;
      int h = hash_cell(u.grid, cx + dx, cy + dy, cz + dz);
      const int bx = SEAT_Q * dx + 16 - pqx, by = SEAT_Q * dy + 16 - pqy, bz = SEAT_Q * dz + 16 - pqz;
      for (int k = 0; k < SLOTS; k++) {
        uint32_t sv = seats[(size_t)h * SLOTS + ((k & 1) ? 4 : 0) + (k >> 1)];
        if (sv & 0x80000000u) break;
        int j = (int)(sv & idxMask);
        if (j == me || seat_d2(sv, B, bx, by, bz) > u.reachQ2) continue;
        float4 Pj = pos[j];
        float ddx = Pj.x - px, ddy = Pj.y - py, ddz = Pj.z - pz;
        float r2 = dot3(ddx, ddy, ddz, ddx, ddy, ddz);
        if (r2 < u.link2) {
          if (r2 < touch2r) contact_pair(u, px, py, pz, vx, vy, vz, Ti, ri, ci, self, Pj, vel[j], mat[j], acc);
          acc.nb += 1.0f;
        }
      }
    }
  }
  float pushx = acc.pushx, pushy = acc.pushy, pushz = acc.pushz, nb = acc.nb, lost = acc.lost, flow = acc.flow;
  float body = P.w;
  float gvx, gvy, gvz;
  if (PLANES > 1) {   // wavefront 1's, beside the contacts
    __syncthreads();
    gvx = __int_as_float(cand[CAND_ROWS][lane]);
    gvy = __int_as_float(cand[CAND_ROWS + 1][lane]);
    gvz = __int_as_float(cand[CAND_ROWS + 2][lane]);
  } else {
    gravity(gvx, gvy, gvz);
  }
  float sc = u.invM / ri;
  vx += pushx * sc * u.dt + gvx * u.gdt;
  vy += pushy * sc * u.dt + gvy * u.gdt;
  vz += pushz * sc * u.dt + gvz * u.gdt;
  bool tgt = body < 3.5f;
  float s0 = tgt ? u.spin0[0] : u.spin1[0], s1 = tgt ? u.spin0[1] : u.spin1[1], s2 = tgt ? u.spin0[2] : u.spin1[2];
  float ax = px - (tgt ? u.cen0[0] : u.cen1[0]), ay = py - (tgt ? u.cen0[1] : u.cen1[1]),
        az = pz - (tgt ? u.cen0[2] : u.cen1[2]);
  float vrx = s1 * az - s2 * ay, vry = s2 * ax - s0 * az, vrz = s0 * ay - s1 * ax;
  if (u.settle > 0.5f) {
    float dvx = (vx - vrx) * u.damp, dvy = (vy - vry) * u.damp, dvz = (vz - vrz) * u.damp;
    float spd = sqrtf(dot3(dvx, dvy, dvz, dvx, dvy, dvz));
    float f = gmin(1.0f, u.creep / gmax(spd, 1e-9f));
    vx = vrx + dvx * f; vy = vry + dvy * f; vz = vrz + dvz * f;
  }
  if (u.kick > 0.5f) {
    vx = (tgt ? u.bulk0[0] : u.bulk1[0]) + vrx;
    vy = (tgt ? u.bulk0[1] : u.bulk1[1]) + vry;
    vz = (tgt ? u.bulk0[2] : u.bulk1[2]) + vrz;
  }
  float Q = u.settle > 0.5f ? 0.0f : q[me] + (lost + flow * u.kappa / u.tK) * u.dt;
  if (u.frozenOn && fabsf(body - u.frozen) < 0.5f) { vx = 0.0f; vy = 0.0f; vz = 0.0f; }
  px += vx * u.dt; py += vy * u.dt; pz += vz * u.dt;
  oPos[me] = make_float4(px, py, pz, body);
  oVel[me] = make_float4(vx, vy, vz, u.tK * Q / ci);
  oAux[me] = make_float4(gvx, gvy, gvz, nb);
  oQ[me] = Q;
  if (COUNT) {   // the next rebuild's count, on the position just written (a block is one wavefront, whose
                 // places in the visit order make the occupied list's segment)
    int h;
    uint32_t w = seat_word(px, py, pz, u.invCell, u.grid, me, h);
    count_arrival(at, w, h, cnt);
  }
}

// (at most 128 registers: four wavefronts a SIMD, as the 10 KiB of LDS a wavefront allow).  u.order 1: the
// work-groups in launch order, else spread over the XCDs; PLANES: wavefronts per 64 particles (sim_body).
template <bool COUNT, int PLANES>
__global__ void __launch_bounds__(SIM_BLOCK * PLANES) __attribute__((amdgpu_waves_per_eu(4)))
k_sim(SimU u, const Box* __restrict__ bxp, float touch2r, const float4* __restrict__ pos,
      const float4* __restrict__ vel, const float* __restrict__ mat, const float* __restrict__ q,
      const uint32_t* __restrict__ seats, const float4* __restrict__ force0, const float4* __restrict__ mono0,
      const float4* __restrict__ force1, const float4* __restrict__ mono1, const float* __restrict__ ppSlot,
      const float4* __restrict__ ppForce, float4* oPos, float4* oVel, float4* oAux, float* oQ, SlotsOut cnt) {
  __shared__ int cand[PLANES * CAND_ROWS][SIM_BLOCK];
  const int at = (u.order == 1 ? (int)blockIdx.x : xcd_block()) * SIM_BLOCK + ((int)threadIdx.x & (SIM_BLOCK - 1));
  sim_body<COUNT, PLANES>(cand, u, bxp, touch2r, pos, vel, mat, q, seats, force0, mono0, force1, mono1, ppSlot,
                          ppForce, oPos, oVel, oAux, oQ, cnt, at);
}

// The step passes of several instances in one launch: logical work-group L (spread over the XCDs as
// xcd_block does, so each XCD's L2 holds a contiguous run of some instance's Morton blocks) serves job
// blockJob[L], its local block L - block0.  A batch whose instances step together launches one per step.
template <bool COUNT>
__global__ void __launch_bounds__(SIM_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_sim_jobs(const SimJob* __restrict__ jobs, const uint16_t* __restrict__ blockJob, int blocks) {
  __shared__ int cand[CAND_ROWS][SIM_BLOCK];
  const int L = xcd_block();
  if (L >= blocks) return;
  const SimJob& J = jobs[blockJob[L]];
  sim_body<COUNT>(cand, J.u, J.bx, J.touch2r, J.pos, J.vel, J.mat, J.q, J.seats, J.force0, J.mono0, J.force1,
                  J.mono1, J.ppSlot, J.ppForce, J.oPos, J.oVel, J.oAux, J.oQ, J.cnt, (L - J.block0) * SIM_BLOCK +
                  (int)threadIdx.x);
}

// the smallest r^2 whose correctly rounded square root is not below the
// contact distance: r^2 < touch2r is then the same test as touch - sqrt(r^2) > 0
float touch_r2(float touch) {
  float x = touch * touch;
  while (x > 0.0f && std::sqrt(std::nextafter(x, 0.0f)) >= touch) x = std::nextafter(x, 0.0f);
  while (std::sqrt(x) < touch) x = std::nextafter(x, INFINITY);
  return x;
}

SimJob sim_job(const SimU& u, const Box* bx, const float4* pos, const float4* vel, const float* mat, const float* q,
               const uint32_t* seats, const MeshBuf& m0, const MeshBuf& m1, const float* ppSlot,
               const float4* ppForce, float4* oPos, float4* oVel, float4* oAux, float* oQ, const SlotsOut* count) {
  SimJob j;
  j.u = u; j.bx = bx; j.touch2r = touch_r2(u.touch);
  j.pos = pos; j.vel = vel; j.mat = mat; j.q = q; j.seats = seats;
  j.force0 = m0.force; j.mono0 = m0.mono; j.force1 = m1.force; j.mono1 = m1.mono;
  j.ppSlot = ppSlot; j.ppForce = ppForce;
  j.oPos = oPos; j.oVel = oVel; j.oAux = oAux; j.oQ = oQ;
  j.cnt = count ? *count : SlotsOut{};
  j.block0 = 0;
  return j;
}

void launch_sim_jobs(const SimJob* jobs, const uint16_t* blockJob, int blocks, bool count, hipStream_t s) {
  if (blocks <= 0) return;
  int grid = (blocks + 7) & ~7;   // a multiple of the 8 XCDs, so the block remap applies
  if (count)
    hipLaunchKernelGGL(k_sim_jobs<true>, dim3(grid), dim3(SIM_BLOCK), 0, s, jobs, blockJob, blocks);
  else
    hipLaunchKernelGGL(k_sim_jobs<false>, dim3(grid), dim3(SIM_BLOCK), 0, s, jobs, blockJob, blocks);
}

void launch_sim(const SimU& u, const Box* bx, const float4* pos, const float4* vel, const float* mat, const float* q,
                const uint32_t* seats, const MeshBuf& m0, const MeshBuf& m1, const float* ppSlot,
                const float4* ppForce, float4* oPos, float4* oVel, float4* oAux, float* oQ, const SlotsOut* count,
                hipStream_t s) {
  int blocks = (u.N + SIM_BLOCK - 1) / SIM_BLOCK;
  blocks = (blocks + 7) & ~7;   // a multiple of the 8 XCDs, so the block remap applies
  const SlotsOut o = count ? *count : SlotsOut{};
  const float t2 = touch_r2(u.touch);
  if (u.order == 3) {   // nine wavefronts per 64 particles, one per row of three cells of the neighbourhood
    if (count)
      hipLaunchKernelGGL((k_sim<true, 9>), dim3(blocks), dim3(9 * SIM_BLOCK), 0, s, u, bx, t2, pos, vel, mat, q,
                         seats, m0.force, m0.mono, m1.force, m1.mono, ppSlot, ppForce, oPos, oVel, oAux, oQ, o);
    else
      hipLaunchKernelGGL((k_sim<false, 9>), dim3(blocks), dim3(9 * SIM_BLOCK), 0, s, u, bx, t2, pos, vel, mat, q,
                         seats, m0.force, m0.mono, m1.force, m1.mono, ppSlot, ppForce, oPos, oVel, oAux, oQ, o);
  } else if (u.order == 2) {   // three wavefronts per 64 particles, one per z-plane of the neighbourhood
    if (count)
      hipLaunchKernelGGL((k_sim<true, 3>), dim3(blocks), dim3(3 * SIM_BLOCK), 0, s, u, bx, t2, pos, vel, mat, q,
                         seats, m0.force, m0.mono, m1.force, m1.mono, ppSlot, ppForce, oPos, oVel, oAux, oQ, o);
    else
      hipLaunchKernelGGL((k_sim<false, 3>), dim3(blocks), dim3(3 * SIM_BLOCK), 0, s, u, bx, t2, pos, vel, mat, q,
                         seats, m0.force, m0.mono, m1.force, m1.mono, ppSlot, ppForce, oPos, oVel, oAux, oQ, o);
  } else if (count) {
    hipLaunchKernelGGL((k_sim<true, 1>), dim3(blocks), dim3(SIM_BLOCK), 0, s, u, bx, t2, pos, vel, mat, q, seats,
                       m0.force, m0.mono, m1.force, m1.mono, ppSlot, ppForce, oPos, oVel, oAux, oQ, o);
  } else {
    hipLaunchKernelGGL((k_sim<false, 1>), dim3(blocks), dim3(SIM_BLOCK), 0, s, u, bx, t2, pos, vel, mat, q, seats,
                       m0.force, m0.mono, m1.force, m1.mono, ppSlot, ppForce, oPos, oVel, oAux, oQ, o);
  }
}

// ----------------------------------------------------------- rigid (A14) --
struct Rigid { float bulk0[3], bulk1[3], cen0[3], cen1[3], spin0[3], spin1[3], rot0[9], rot1[9], dt; };

__global__ void __launch_bounds__(256) k_rigid(int N, Rigid r, const float4* __restrict__ pos,
                                               const float4* __restrict__ vel, const float4* __restrict__ aux,
                                               const float* __restrict__ q, float4* oPos, float4* oVel, float4* oAux,
                                               float* oQ) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float4 P = pos[i];
  bool tgt = P.w < 3.5f;
  const float* bulk = tgt ? r.bulk0 : r.bulk1;
  const float* cen = tgt ? r.cen0 : r.cen1;
  const float* sp = tgt ? r.spin0 : r.spin1;
  const float* m = tgt ? r.rot0 : r.rot1;
  float e0 = P.x - cen[0], e1 = P.y - cen[1], e2 = P.z - cen[2];
  float a0 = m[0] * e0 + m[3] * e1 + m[6] * e2;   // column-major mat3 times vec3
  float a1 = m[1] * e0 + m[4] * e1 + m[7] * e2;
  float a2 = m[2] * e0 + m[5] * e1 + m[8] * e2;
  oPos[i] = make_float4(cen[0] + a0 + bulk[0] * r.dt, cen[1] + a1 + bulk[1] * r.dt, cen[2] + a2 + bulk[2] * r.dt, P.w);
  oVel[i] = make_float4(bulk[0] + (sp[1] * a2 - sp[2] * a1), bulk[1] + (sp[2] * a0 - sp[0] * a2),
                        bulk[2] + (sp[0] * a1 - sp[1] * a0), vel[i].w);
  oAux[i] = aux[i];
  oQ[i] = q[i];
}
void launch_rigid(int N, const float4* pos, const float4* vel, const float4* aux, const float* q, const float* bulk0,
                  const float* bulk1, const float* cen0, const float* cen1, const float* spin0, const float* spin1,
                  const float* rot0, const float* rot1, float dt, float4* oPos, float4* oVel, float4* oAux, float* oQ,
                  hipStream_t s) {
  Rigid r;
  for (int a = 0; a < 3; a++) {
    r.bulk0[a] = bulk0[a]; r.bulk1[a] = bulk1[a]; r.cen0[a] = cen0[a]; r.cen1[a] = cen1[a];
    r.spin0[a] = spin0[a]; r.spin1[a] = spin1[a];
  }
  for (int a = 0; a < 9; a++) { r.rot0[a] = rot0[a]; r.rot1[a] = rot1[a]; }
  r.dt = dt;
  hipLaunchKernelGGL(k_rigid, dim3((N + 255) / 256), dim3(256), 0, s, N, r, pos, vel, aux, q, oPos, oVel, oAux, oQ);
}

// ----------------------------------------------- crater coupling (A20) --
// The excavation flow at touchdown (COUPLE_FS, chicxulub.html:255-315): each
// unfrozen grain gets the Housen-Holsapple speed for its range from the point
// source, capped at the impact speed and stalled at the transient radius, in
// the Z-model direction about the downward axis; the impactor becomes melt
// and keeps half its own velocity; the heat is that of the speed.
__global__ void __launch_bounds__(256) k_couple(int N, CoupleU c, const float4* __restrict__ pos,
                                                const float4* __restrict__ vel, const float4* __restrict__ aux,
                                                const float* __restrict__ mat, float4* oPos, float4* oVel,
                                                float4* oAux, float* oQ) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float4 P = pos[i], V = vel[i];
  float w = P.w, ri = mat[i];
  oAux[i] = aux[i];
  if (fabsf(w - c.frozen) < 0.5f) {
    oPos[i] = P; oVel[i] = make_float4(0.0f, 0.0f, 0.0f, c.ambient); oQ[i] = 0.0f;
    return;
  }
  float Rx = P.x - c.centre[0], Ry = P.y - c.centre[1], Rz = P.z - c.centre[2];
  float r = gmax(sqrtf(dot3(Rx, Ry, Rz, Rx, Ry, Rz)), 1e-4f);
  float v = c.c1 * c.vimp * powf(r / c.aimp, -c.invMu);
  v = gmin(v, c.vimp);
  v *= powf(gmax(0.0f, 1.0f - r / c.stall), c.pexp);
  float nx = Rx / r, ny = Ry / r, nz = Rz / r;
  float ct = gclamp(dot3(nx, ny, nz, 0.0f, -1.0f, 0.0f), -1.0f, 1.0f);
  float st = sqrtf(gmax(1.0f - ct * ct, 1e-8f));
  float ex = (ct * nx - 0.0f) / st, ey = (ct * ny - -1.0f) / st, ez = (ct * nz - 0.0f) / st;
  float th = acosf(ct);
  float f = (c.z - 2.0f) * tanf(0.5f * th);
  float dx = nx + f * ex, dy = ny + f * ey, dz = nz + f * ez;
  float inv = 1.0f / sqrtf(dot3(dx, dy, dz, dx, dy, dz));   // (a direction of any length: the general forms)
  dx *= inv; dy *= inv; dz *= inv;
  bool imp = w > 3.5f;
  float vx = dx * v, vy = dy * v, vz = dz * v;
  if (imp) { vx = gmix(vx, V.x, 0.5f); vy = gmix(vy, V.y, 0.5f); vz = gmix(vz, V.z, 0.5f); }
  float speed = sqrtf(dot3(vx, vy, vz, vx, vy, vz));
  float T = c.ambient + c.heat * 0.5f * speed * speed;
  oVel[i] = make_float4(vx, vy, vz, T);
  oQ[i] = T * ri / gmax(c.tK, 1e-9f);
  oPos[i] = make_float4(P.x, P.y, P.z, imp ? c.melt : (T > 1500.0f ? c.melt : w));
}
void launch_couple(int N, const CoupleU& c, const float4* pos, const float4* vel, const float4* aux, const float* mat,
                   float4* oPos, float4* oVel, float4* oAux, float* oQ, hipStream_t s) {
  hipLaunchKernelGGL(k_couple, dim3((N + 255) / 256), dim3(256), 0, s, N, c, pos, vel, aux, mat, oPos, oVel, oAux, oQ);
}

// --------------------------------------------------------- read-out (A16) --
__global__ void __launch_bounds__(256) k_diag(SimU u, const Box* __restrict__ bxp, const float4* __restrict__ pos,
                                              const float* __restrict__ q,
                                              const uint32_t* __restrict__ seats, const float4* __restrict__ force0,
                                              const float4* __restrict__ mono0, const float4* __restrict__ force1,
                                              const float4* __restrict__ mono1, const float* __restrict__ ppSlot,
                                              const float4* __restrict__ ppForce, float4* out) {
  int me = blockIdx.x * blockDim.x + threadIdx.x;
  if (me >= u.N) return;
  float4 P = pos[me];
  int h = hash_cell(u.grid, (int)floorf(P.x * u.invCell), (int)floorf(P.y * u.invCell), (int)floorf(P.z * u.invCell));
  const uint4* s4 = reinterpret_cast<const uint4*>(seats + (size_t)h * SLOTS);
  uint4 ea = s4[0], ob = s4[1];
  uint32_t sv8[8] = { ea.x, ob.x, ea.y, ob.y, ea.z, ob.z, ea.w, ob.w };
  float seated = 0.0f, taken = 0.0f;
  for (int k = 0; k < SLOTS; k++) {
    if (sv8[k] & 0x80000000u) break;
    taken += 1.0f;
    if ((int)(sv8[k] & ((1u << u.grid.IB) - 1u)) == me) seated = 1.0f;
  }
  float phi = 0.0f, off = 0.0f, gx, gy, gz;
  const Box B = *bxp;
  int box = box_of(B, P.x, P.y, P.z, gx, gy, gz);
  float4 m0 = mono0[0], m1 = mono1[0];
  if (box >= 0) {
    const float4* F = box == 1 ? force1 : force0;
    float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float tx = gx - fx, ty = gy - fy, tz = gz - fz;
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    for (int k = 0; k < 8; k++) {
      int o0 = k & 1, o1 = (k >> 1) & 1, o2 = k >> 2;
      float w = gmix(1.0f - tx, tx, (float)o0) * gmix(1.0f - ty, ty, (float)o1) * gmix(1.0f - tz, tz, (float)o2);
      phi += F[fine_at(min(ix + o0, 63), min(iy + o1, 63), min(iz + o2, 63))].w * w;
    }
    if (B.boxB > 0.5f) phi += box == 0 ? pot_at(P.x, P.y, P.z, m1, u.gm, u.eps2) : pot_at(P.x, P.y, P.z, m0, u.gm, u.eps2);
  } else {
    phi = pot_at(P.x, P.y, P.z, m0, u.gm, u.eps2);
    if (B.boxB > 0.5f) phi += pot_at(P.x, P.y, P.z, m1, u.gm, u.eps2);
    off = 2.0f;
  }
  float seat = ppSlot[me];
  if (seat >= 0.0f) {
    int si = (int)seat;
    for (int s = 0; s < 4; s++) phi += ppForce[s * PP_CAP + si].w;
  }
  out[me] = make_float4(phi, seated + off, taken, q[me]);
}
void launch_diag(const SimU& u, const Box* bx, const float4* pos, const float* q, const uint32_t* seats, const MeshBuf& m0,
                 const MeshBuf& m1, const float* ppSlot, const float4* ppForce, float4* out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag, dim3((u.N + 255) / 256), dim3(256), 0, s, u, bx, pos, q, seats, m0.force, m0.mono,
                     m1.force, m1.mono, ppSlot, ppForce, out);
}

// ------------------------------------------------------ Morton order (A15) --
constexpr int KEYS = 1 << 21;

__global__ void __launch_bounds__(256) k_morton_keys(const float4* __restrict__ pos, int N, double cx, double cy,
                                                     double cz, int* keys, int* slot, int* hist) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float4 p = pos[i];
  double c[3] = { cx, cy, cz }, v[3] = { (double)p.x, (double)p.y, (double)p.z };
  int m[3];
  for (int a = 0; a < 3; a++) {
    double t = floor(((v[a] - c[a]) + 6.0) / 12.0 * 128.0);
    t = t < 0.0 ? 0.0 : t;
    t = t > 127.0 ? 127.0 : t;
    m[a] = (int)t;
  }
  int mk = 0;
  for (int k = 0; k < 7; k++)
    mk |= ((m[0] >> k) & 1) << (3 * k) | ((m[1] >> k) & 1) << (3 * k + 1) | ((m[2] >> k) & 1) << (3 * k + 2);
  keys[i] = mk;
  slot[i] = atomicAdd(&hist[mk], 1);
}

// exclusive scan of KEYS+1 counters in place: per-block totals, their scan, then each block
constexpr int SCAN_BLOCK = 1024;
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_blocks(const int* in, int n, int* blockSums) {
  __shared__ int sh[SCAN_BLOCK / 64];
  int i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
  int v = i < n ? in[i] : 0;
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < SCAN_BLOCK / 64; w++) s += sh[w];
    blockSums[blockIdx.x] = s;
  }
}
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_sums(int* blockSums, int nb) {
  __shared__ int sh[SCAN_BLOCK];
  // nb <= SCAN_BLOCK * 4: each thread scans up to four totals
  int loc[4], base = threadIdx.x * 4, s = 0;
  for (int q = 0; q < 4; q++) { loc[q] = base + q < nb ? blockSums[base + q] : 0; s += loc[q]; }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < SCAN_BLOCK; o <<= 1) {
    int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - s;
  for (int q = 0; q < 4; q++) { if (base + q < nb) blockSums[base + q] = run; run += loc[q]; }
}
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_final(int* data, int n, const int* blockSums) {
  __shared__ int sh[SCAN_BLOCK];
  int i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
  int v = i < n ? data[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < SCAN_BLOCK; o <<= 1) {
    int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) data[i] = sh[threadIdx.x] - v + blockSums[blockIdx.x];
}

void launch_scan_excl(int* data, int n, int* blockSums, hipStream_t s) {
  int nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_BLOCK), 0, s, data, n, blockSums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, blockSums, nb);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(SCAN_BLOCK), 0, s, data, n, blockSums);
}

__global__ void __launch_bounds__(256) k_morton_scatter(int N, const int* keys, const int* slot, const int* start,
                                                        int* tmp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) tmp[start[keys[i]] + slot[i]] = i;
}
// within a key, members in index order: the reference's counting sort is stable
__global__ void __launch_bounds__(256) k_morton_rank(int N, const int* keys, const int* start, const int* tmp,
                                                     int* perm, int* into) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int k = keys[i], s = start[k], e = start[k + 1], r = 0;
  for (int t = s; t < e; t++) r += tmp[t] < i;
  perm[s + r] = i;
  into[i] = s + r;
}

__global__ void k_copy_int(int n, const int* __restrict__ a, int* b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

void launch_morton(const float4* pos, int N, double cx, double cy, double cz, int* keys, int* slot, int* hist,
                   int* blockSums, int* perm, int* into, hipStream_t s) {
  int n = KEYS + 1;
  launch_fill_u32(reinterpret_cast<uint32_t*>(hist), (size_t)n, 0u, s);
  int blocks = (N + 255) / 256;
  hipLaunchKernelGGL(k_morton_keys, dim3(blocks), dim3(256), 0, s, pos, N, cx, cy, cz, keys, slot, hist);
  launch_scan_excl(hist, n, blockSums, s);
  // perm is used as the scatter's scratch before the ranked write: into holds the arrival list
  hipLaunchKernelGGL(k_morton_scatter, dim3(blocks), dim3(256), 0, s, N, keys, slot, hist, into);
  hipLaunchKernelGGL(k_morton_rank, dim3(blocks), dim3(256), 0, s, N, keys, hist, into, perm, slot);
  hipLaunchKernelGGL(k_copy_int, dim3(blocks), dim3(256), 0, s, N, slot, into);
}

// The step pass's visit order.  Any permutation gives the same results (a
// particle's sums depend only on the seats, which name particles by index), so
// it is chosen for the caches alone: the particles counting-sorted by the
// Morton code of their contact cell, its coordinates taken modulo 128, so that
// a wavefront's lanes share most of their 27 cells (and a work-group of the
// deposit most of its mesh cells).  Within a code the order is whatever the
// atomics give.
__global__ void __launch_bounds__(256) k_visit_keys(const float4* __restrict__ pos, int N, float invCell, int* keys,
                                                    int* slot, int* hist) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float4 p = pos[i];
  const int m[3] = { (int)floorf(p.x * invCell) & 127, (int)floorf(p.y * invCell) & 127,
                     (int)floorf(p.z * invCell) & 127 };
  int mk = 0;
#pragma unroll
  for (int k = 0; k < 7; k++)
    mk |= ((m[0] >> k) & 1) << (3 * k) | ((m[1] >> k) & 1) << (3 * k + 1) | ((m[2] >> k) & 1) << (3 * k + 2);
  keys[i] = mk;
  slot[i] = atomicAdd(&hist[mk], 1);
}
__global__ void __launch_bounds__(256) k_visit_scatter(int N, const int* __restrict__ keys,
                                                       const int* __restrict__ slot, const int* __restrict__ start,
                                                       int* visit) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) visit[start[keys[i]] + slot[i]] = i;
}
void launch_visit(const float4* pos, int N, float invCell, int* keys, int* slot, int* hist, int* blockSums,
                  int* visit, hipStream_t s) {
  launch_fill_u32(reinterpret_cast<uint32_t*>(hist), (size_t)KEYS + 1, 0u, s);
  int blocks = (N + 255) / 256;
  hipLaunchKernelGGL(k_visit_keys, dim3(blocks), dim3(256), 0, s, pos, N, invCell, keys, slot, hist);
  launch_scan_excl(hist, KEYS + 1, blockSums, s);
  hipLaunchKernelGGL(k_visit_scatter, dim3(blocks), dim3(256), 0, s, N, keys, slot, hist, visit);
}
__global__ void k_iota(int* a, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}
void launch_iota(int* a, int n, hipStream_t s) { hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, a, n); }

__global__ void __launch_bounds__(256) k_permute(int N, const int* __restrict__ perm, const float4* __restrict__ pos,
                                                 const float4* __restrict__ vel, const float4* __restrict__ aux,
                                                 const float* __restrict__ q, const float* __restrict__ mat,
                                                 float4* oPos, float4* oVel, float4* oAux, float* oQ, float* oMat) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int s = perm[i];
  oPos[i] = pos[s]; oVel[i] = vel[s]; oAux[i] = aux[s]; oQ[i] = q[s]; oMat[i] = mat[s];
}
void launch_permute(int N, const int* perm, const float4* pos, const float4* vel, const float4* aux, const float* q,
                    const float* mat, float4* oPos, float4* oVel, float4* oAux, float* oQ, float* oMat,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_permute, dim3((N + 255) / 256), dim3(256), 0, s, N, perm, pos, vel, aux, q, mat, oPos, oVel,
                     oAux, oQ, oMat);
}

// ------------------------------------------------------------ books (A17) --
constexpr int NBOOK = 13;   // M px py pz Jx Jy Jz KE PE Q iron unseated full
__global__ void __launch_bounds__(256) k_books(int N, double mref, const float4* __restrict__ pos,
                                               const float4* __restrict__ vel, const float* __restrict__ mat,
                                               const float4* __restrict__ diag, double* partial) {
  __shared__ double sh[4][NBOOK];
  double acc[NBOOK];
  for (int k = 0; k < NBOOK; k++) acc[k] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double m = mref * (double)mat[i];
    float4 P = pos[i], V = vel[i], D = diag[i];
    double bx = P.x, by = P.y, bz = P.z, bu = V.x, bv = V.y, bw = V.z, w = P.w;
    acc[0] += m; acc[1] += m * bu; acc[2] += m * bv; acc[3] += m * bw;
    acc[4] += m * (by * bw - bz * bv); acc[5] += m * (bz * bu - bx * bw); acc[6] += m * (bx * bv - by * bu);
    acc[7] += 0.5 * m * (bu * bu + bv * bv + bw * bw);
    double flag = D.y;
    bool off = flag >= 2.0;
    acc[8] += (off ? 1.0 : 0.5) * m * (double)D.x;
    acc[9] += 0.5 * (double)D.w;
    if (w - 4.0 * floor(w / 4.0) < 0.5) acc[10] += m;
    if (flag - (off ? 2.0 : 0.0) < 0.5) acc[11] += 1.0;
    if (D.z >= 7.5f) acc[12] += 1.0;
  }
  for (int k = 0; k < NBOOK; k++) acc[k] = wave_sum(acc[k]);
  if ((threadIdx.x & 63) == 0) for (int k = 0; k < NBOOK; k++) sh[threadIdx.x >> 6][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < NBOOK) {
    double s = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
    partial[blockIdx.x * NBOOK + threadIdx.x] = s;
  }
}
__global__ void k_books_final(int nb, const double* partial, double* out) {
  int k = threadIdx.x;
  if (k >= NBOOK) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += partial[b * NBOOK + k];
  out[k] = s;
}
void launch_books(int N, double mref, const float4* pos, const float4* vel, const float* mat, const float4* diag,
                  double* partial, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_books, dim3(256), dim3(256), 0, s, N, mref, pos, vel, mat, diag, partial);
  hipLaunchKernelGGL(k_books_final, dim3(1), dim3(64), 0, s, 256, partial, out);
}

}  // namespace cc
