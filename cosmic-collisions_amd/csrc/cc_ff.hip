// The learned far field (no reference counterpart): a 3D convolutional
// operator on the 16^3 coarse mesh that stands in for the block-to-block far
// field (FAR_FS, src/glsl-sim.js:263-295) on the mesh steps between judges.
// Its input is the coarse mesh now and its change since the last exact pass
// (block masses and centre-of-mass offsets); its output, each block's change
// of far-field acceleration, potential and tidal tensor since that pass, is
// added to the exact far field recorded there (the mesh interface below).
//
// Network (weights FP32 masters, BF16 on the matrix cores, FP32 accumulation):
//   4 levels 16^3 / 8^3 / 4^3 / 2^3, 32 channels, 3x3x3 convolutions, leaky
//   ReLU (slope 1/4), 2^3 average pooling down, nearest upsampling plus the
//   level's skip up, a linear 32 -> 16 output convolution (10 used).
//     c0  16->32  c1 32->32            (16^3)  e0
//     c2, c3                            (8^3)  e1 = from pool(e0)
//     c4, c5                            (4^3)  e2 = from pool(e1)
//     c6, c7                            (2^3)  bottom = from pool(e2)
//     c8  on up(bottom) + e2            (4^3)
//     c9  on up(c8) + e1                (8^3)
//     c10 on up(c9) + e0               (16^3)
//     c11 32->16 linear                (16^3)  out
// Activations are channels-last BF16, [instance][z][y][x][channel].
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16bf16_1k: a wavefront
// owns 16 output voxels x all output channels and walks K = 27 taps x input
// channels in steps of 16.  The data gradient is the same kernel on the
// flipped, transposed weights; the weight gradient reduces over voxels, one
// wavefront per (tap, 256-voxel chunk), into FP32 with atomics.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cc_hip.h"
#include "cc_fpmath.h"
#include "cc_internal.h"

namespace cc {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// a wavefront's sum, in lane 0
template <typename T>
__device__ __forceinline__ T ff_wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// ------------------------------------------------------------ convolution --
// in  [M voxels][CIN] bf16 (M = B*S^3), wt [27][CIN/16][COUT][16] bf16,
// bias [COUT] f32 or null; slope: the leaky ReLU's (1: none).  Writes out (bf16, [M][COUT])
// and/or outf (f32), and pre (f32 pre-activation) when asked.
// Work-groups of 256.  The work-group stages the layer's weights in LDS once; each wavefront issues
// all 27 taps' activation loads before its matrix-core chain, so a launch pays
// about one memory latency instead of one per tap (what bounds a single
// instance's 16^3 mesh, whose widest layer is 256 wavefronts).  The chain adds
// taps and channel steps in the same order either way.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv(const uint16_t* __restrict__ in, int B, int S,
                                              const uint16_t* __restrict__ wt, const float* __restrict__ bias,
                                              float slope, uint16_t* out, float* outf, float* pre,
                                              const int* __restrict__ gate) {
  if (gate && *gate == 0) return;   // the same for every work-group: before the barrier
  constexpr int WN = 27 * (CIN / 16) * COUT * 16, NV = WN / 8, PER = (NV + 255) / 256;
  __shared__ __align__(16) uint16_t wl[WN];
  {   // every 16-byte chunk's load in flight before the stores (a load-store loop waits out each in turn)
    uint4 tmp[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = threadIdx.x + 256 * k;
      tmp[k] = i < NV ? reinterpret_cast<const uint4*>(wt)[i] : uint4{ 0u, 0u, 0u, 0u };
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = threadIdx.x + 256 * k;
      if (i < NV) reinterpret_cast<uint4*>(wl)[i] = tmp[k];
    }
  }
  __syncthreads();
  int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  int lane = threadIdx.x & 63;
  int S3 = S * S * S, M = B * S3;
  int m0 = wave * 16;
  if (m0 >= M) return;
  int row = lane & 15, kq = (lane >> 4) * 4;
  int v = m0 + row;
  bool vin = v < M;
  int b = v / S3, r = v - b * S3;
  int z = r / (S * S), y = (r / S) % S, x = r % S;
  const bf16x4 zero = bf16x4{ 0, 0, 0, 0 };
  bf16x4 av[27][CIN / 16];
#pragma unroll
  for (int t = 0; t < 27; t++) {
    int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
    bool ok = vin && zz >= 0 && zz < S && yy >= 0 && yy < S && xx >= 0 && xx < S;
    size_t src = ok ? ((size_t)b * S3 + ((size_t)zz * S + yy) * S + xx) * CIN : 0;
#pragma unroll
    for (int kc = 0; kc < CIN / 16; kc++)
      av[t][kc] = ok ? *reinterpret_cast<const bf16x4*>(in + src + kc * 16 + kq) : zero;
  }
  f32x4 acc[COUT / 16];
#pragma unroll
  for (int n = 0; n < COUT / 16; n++) acc[n] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
  for (int t = 0; t < 27; t++)
#pragma unroll
    for (int kc = 0; kc < CIN / 16; kc++) {
      const uint16_t* wp = wl + ((t * (CIN / 16) + kc) * COUT + row) * 16 + kq;
#pragma unroll
      for (int n = 0; n < COUT / 16; n++) {
        bf16x4 w = *reinterpret_cast<const bf16x4*>(wp + n * 16 * 16);
        acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[t][kc], w, acc[n], 0, 0, 0);
      }
    }
  for (int n = 0; n < COUT / 16; n++) {
    int co = n * 16 + row;
    float bb = bias ? bias[co] : 0.0f;
    for (int j = 0; j < 4; j++) {
      int vo = m0 + kq + j;
      if (vo >= M) continue;
      float p = acc[n][j] + bb;
      size_t o = (size_t)vo * COUT + co;
      if (pre) pre[o] = p;
      float yv = p > 0.0f ? p : slope * p;
      if (out) out[o] = f2bf(yv);
      if (outf) outf[o] = yv;
    }
  }
}

// dW[t][ci][co] += sum over voxels v of in[v + t][ci] * g[v][co] (g bf16, the
// gradient at the convolution's output); one wavefront per (tap, chunk of 256).
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv_wgrad(const uint16_t* __restrict__ in, int B, int S,
                                                    const uint16_t* __restrict__ g, float* dW) {
  int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  int lane = threadIdx.x & 63;
  int S3 = S * S * S, M = B * S3;
  int chunks = (M + 255) / 256;
  int t = wave / chunks, ch = wave - t * chunks;
  if (t >= 27) return;
  int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
  int row = lane & 15, kq = (lane >> 4) * 4;
  f32x4 acc[CIN / 16][COUT / 16];
  for (int i = 0; i < CIN / 16; i++)
    for (int n = 0; n < COUT / 16; n++) acc[i][n] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
  for (int k0 = ch * 256; k0 < ch * 256 + 256 && k0 < M; k0 += 16) {
    // this lane's four voxels of the K step
    bf16x4 av[CIN / 16], gv[COUT / 16];
    for (int i = 0; i < CIN / 16; i++) av[i] = bf16x4{ 0, 0, 0, 0 };
    for (int n = 0; n < COUT / 16; n++) gv[n] = bf16x4{ 0, 0, 0, 0 };
    for (int j = 0; j < 4; j++) {
      int v = k0 + kq + j;
      if (v >= M) continue;
      int b = v / S3, r = v - b * S3;
      int z = r / (S * S) + dz, y = (r / S) % S + dy, x = r % S + dx;
      for (int n = 0; n < COUT / 16; n++) gv[n][j] = (short)g[(size_t)v * COUT + n * 16 + row];
      if (z < 0 || z >= S || y < 0 || y >= S || x < 0 || x >= S) continue;
      size_t src = ((size_t)b * S3 + ((size_t)z * S + y) * S + x) * CIN;
      for (int i = 0; i < CIN / 16; i++) av[i][j] = (short)in[src + i * 16 + row];
    }
    for (int i = 0; i < CIN / 16; i++)
      for (int n = 0; n < COUT / 16; n++)
        acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[i], gv[n], acc[i][n], 0, 0, 0);
  }
  // acc[i][n] holds D[ci = i*16 + kq + j][co = n*16 + row]
  for (int i = 0; i < CIN / 16; i++)
    for (int n = 0; n < COUT / 16; n++)
      for (int j = 0; j < 4; j++)
        atomicAdd(&dW[((size_t)t * CIN + i * 16 + kq + j) * COUT + n * 16 + row], acc[i][n][j]);
}

// FP32 masters [27][cin][cout] -> the forward layout [27][cin/16][cout][16]
// and the data-gradient layout (taps flipped, in and out swapped)
// [27][cout/16][cin][16], both BF16.
__global__ void k_pack(const float* __restrict__ w, int cin, int cout, uint16_t* fwd, uint16_t* bwd) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 27 * cin * cout) return;
  int co = e % cout, ci = (e / cout) % cin, t = e / (cin * cout);
  uint16_t h = f2bf(w[e]);
  fwd[((size_t)(t * (cin / 16) + ci / 16) * cout + co) * 16 + (ci & 15)] = h;
  bwd[((size_t)((26 - t) * (cout / 16) + co / 16) * cin + ci) * 16 + (co & 15)] = h;
}

// ------------------------------------------------------------ elementwise --
// 2^3 average pooling: [B][S]^3[C] -> [B][S/2]^3[C]
__global__ void k_pool(const uint16_t* __restrict__ in, int B, int S, int C, uint16_t* out,
                       const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  int h = S / 2, e = blockIdx.x * blockDim.x + threadIdx.x;
  int n = B * h * h * h * C;
  if (e >= n) return;
  int c = e % C, v = e / C, b = v / (h * h * h), r = v % (h * h * h);
  int z = r / (h * h), y = (r / h) % h, x = r % h;
  float s = 0.0f;
  for (int k = 0; k < 8; k++) {
    int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
    s += bf2f(in[(((size_t)b * S * S * S) + ((size_t)zz * S + yy) * S + xx) * C + c]);
  }
  out[e] = f2bf(s * 0.125f);
}
// gradient of the pooling: each fine voxel gets an eighth of its parent's
__global__ void k_pool_bwd(const float* __restrict__ gout, int B, int S, int C, float* gin, int accumulate) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  int n = B * S * S * S * C;
  if (e >= n) return;
  int h = S / 2, c = e % C, v = e / C, b = v / (S * S * S), r = v % (S * S * S);
  int z = r / (S * S), y = (r / S) % S, x = r % S;
  float gv = 0.125f * gout[(((size_t)b * h * h * h) + ((size_t)(z / 2) * h + y / 2) * h + x / 2) * C + c];
  gin[e] = accumulate ? gin[e] + gv : gv;
}
// nearest upsampling of [B][S/2]^3[C] to [B][S]^3[C] plus the skip
__global__ void k_up_add(const uint16_t* __restrict__ coarse, const uint16_t* __restrict__ skip, int B, int S, int C,
                         uint16_t* out, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  int n = B * S * S * S * C;
  if (e >= n) return;
  int h = S / 2, c = e % C, v = e / C, b = v / (S * S * S), r = v % (S * S * S);
  int z = r / (S * S), y = (r / S) % S, x = r % S;
  float u = bf2f(coarse[(((size_t)b * h * h * h) + ((size_t)(z / 2) * h + y / 2) * h + x / 2) * C + c]);
  out[e] = f2bf(u + bf2f(skip[e]));
}
// gradient of upsampling: the coarse voxel sums its eight children
__global__ void k_up_bwd(const float* __restrict__ gout, int B, int S, int C, float* gcoarse) {
  int h = S / 2, e = blockIdx.x * blockDim.x + threadIdx.x;
  int n = B * h * h * h * C;
  if (e >= n) return;
  int c = e % C, v = e / C, b = v / (h * h * h), r = v % (h * h * h);
  int z = r / (h * h), y = (r / h) % h, x = r % h;
  float s = 0.0f;
  for (int k = 0; k < 8; k++) {
    int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
    s += gout[(((size_t)b * S * S * S) + ((size_t)zz * S + yy) * S + xx) * C + c];
  }
  gcoarse[e] = s;
}
// gradient through the leaky ReLU (slope 1: none), to BF16 for the matrix cores
__global__ void k_act_bwd(const float* __restrict__ gout, const float* __restrict__ pre, int n, float slope,
                          uint16_t* g) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float d = gout[e];
  if (pre[e] <= 0.0f) d *= slope;
  g[e] = f2bf(d);
}
// bias gradient: sum over voxels of g (bf16) per channel
__global__ void k_bias_grad(const uint16_t* __restrict__ g, int M, int C, float* db) {
  int c = blockIdx.x, lane = threadIdx.x;
  float s = 0.0f;
  for (int v = lane; v < M; v += blockDim.x) s += bf2f(g[(size_t)v * C + c]);
  __shared__ float red[256];
  red[lane] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (lane < w) red[lane] += red[lane + w];
    __syncthreads();
  }
  if (lane == 0) db[c] += red[0];
}
__global__ void k_add_f32(float* a, const float* __restrict__ b, int n) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] += b[e];
}
__global__ void k_zero_f32(float* a, size_t n) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = 0.0f;
}

// ------------------------------------------------------ the mesh interface --
// The network predicts how each block's far field has changed since the last
// exact pass (the base), from how the coarse mesh has changed since then.
// Every judge step (the exact far field is computed there anyway) makes that
// pass the new base; the mesh steps in between take the base plus the
// network's change.  Over the 8 or 16 steps a prediction spans, particles
// move a few hundredths of a block, so the change is a small part of the far
// field and the network has to get only that part right.
//
// Voxel v = (z*16 + y)*16 + x indexes the 16^3 blocks as the network's
// activations do; block (x, y, z) is coarse texel coarse_texel(x, y, z).
//
// Features of a pass, per block: its mass in units of the box's total mass
// times 64, and that times its centre of mass less its centre in block
// lengths (4 FH); meta: the box's corner offset L (xyz) and total mass.
//
// Centre record of a pass, FF_R floats per block: acceleration (3),
// potential, tidal tensor (xx yy zz xy xz yz), the block's occupied fine
// texels, one spare.  An occupied block's record is its far field, which the
// exact pass gives at its centre of mass, carried to the block's centre to
// first order by the tidal tensor -- the inverse of what k_ff_out does, so
// that a record carried back by the same tensor returns the far field it came
// from.  Empty blocks: zero.  A block that fills after the base has no record
// to start from: k_far computes its far field exactly on the mesh steps the
// network stands in (there are few), and the network's output there is unused.
__device__ __forceinline__ int coarse_texel(int x, int y, int z) { return ((z >> 2) * 16 + y) * 64 + (z & 3) * 16 + x; }
constexpr float BLOCK_LEN = (float)(4.0 * FH);
constexpr int FF_R = 12;
constexpr int FF_IN = 10;   // input channels in use (of 16)
// the block's centre of mass less its centre, world units
__device__ __forceinline__ void block_offset(float4 c, int x, int y, int z, const float* L, float& dx, float& dy,
                                             float& dz) {
  dx = c.x / c.w - ((4 * x + 2) * (float)FH - L[0]);
  dy = c.y / c.w - ((4 * y + 2) * (float)FH - L[1]);
  dz = c.z / c.w - ((4 * z + 2) * (float)FH - L[2]);
}
__device__ __forceinline__ float4 ff_feature(const float4* __restrict__ coarse, float mt, const float* L, int v) {
  int x = v & 15, y = (v >> 4) & 15, z = v >> 8;
  float4 c = coarse[coarse_texel(x, y, z)];
  float4 f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (c.w > 0.0f) {
    float dx, dy, dz;
    block_offset(c, x, y, z, L, dx, dy, dz);
    f.x = c.w / mt * 64.0f;
    f.y = f.x * dx / BLOCK_LEN;
    f.z = f.x * dy / BLOCK_LEN;
    f.w = f.x * dz / BLOCK_LEN;
  }
  return f;
}
__global__ void k_ff_feat(const float4* __restrict__ coarse, const float4* __restrict__ mono,
                          const Box* __restrict__ bxp, int which, float4* feat, float* meta) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= 4096) return;
  const Box bx = *bxp;
  const float* L = which ? bx.LB : bx.LA;
  float mt = fmaxf(mono[0].w, 1e-20f);
  feat[v] = ff_feature(coarse, mt, L, v);
  if (v == 0) { meta[0] = L[0]; meta[1] = L[1]; meta[2] = L[2]; meta[3] = mt; }
}

// A pass's centre record (one wavefront per block: its 64 fine texels counted
// by ballot).  With featT given the pass becomes a box's base: its features
// and meta are copied to featB and metaB, and *validB set.
__device__ __forceinline__ void ff_base_body(const float4* __restrict__ fine, const float4* __restrict__ coarse,
                                             const float4* __restrict__ farA, const float4* __restrict__ farT0,
                                             const float4* __restrict__ farT1, const Box* __restrict__ bxp,
                                             int which, float* R, const float4* __restrict__ featT,
                                             const float* __restrict__ metaT, float4* featB, float* metaB,
                                             int* validB) {
  int lane = threadIdx.x & 63;
  int v = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (v >= 4096) return;
  int x = v & 15, y = (v >> 4) & 15, z = v >> 8;
  if (featT && lane == 0) featB[v] = featT[v];
  if (featT && v == 0 && lane == 0) {
    for (int k = 0; k < 4; k++) metaB[k] = metaT[k];
    *validB = 1;
  }
  float* o = R + (size_t)v * FF_R;
  int t = coarse_texel(x, y, z);
  float4 c = coarse[t];
  if (c.w <= 0.0f) {
    if (lane < FF_R) o[lane] = 0.0f;
    return;
  }
  const int cx = 4 * x + (lane & 3), cy = 4 * y + ((lane >> 2) & 3), cz = 4 * z + (lane >> 4);
  const unsigned long long filled = __ballot(fine[((cz >> 3) * 64 + cy) * 512 + (cz & 7) * 64 + cx].w > 0.0f);
  if (lane == 0) {
    const Box bx = *bxp;
    float dx, dy, dz;
    block_offset(c, x, y, z, which ? bx.LB : bx.LA, dx, dy, dz);
    float4 A = farA[t], T0 = farT0[t], T1 = farT1[t];
    float ax = A.x - (T0.x * dx + T0.w * dy + T1.x * dz), ay = A.y - (T0.w * dx + T0.y * dy + T1.y * dz),
          az = A.z - (T1.x * dx + T1.y * dy + T0.z * dz);
    float r[FF_R] = { ax, ay, az, A.w + (ax * dx + ay * dy + az * dz), T0.x, T0.y, T0.z, T0.w, T1.x, T1.y,
                      (float)__popcll(filled), 0.0f };
    for (int k = 0; k < FF_R; k++) o[k] = r[k];
  }
}

__global__ void __launch_bounds__(256) k_ff_base(const float4* __restrict__ fine, const float4* __restrict__ coarse,
                                                 const float4* __restrict__ farA, const float4* __restrict__ farT0,
                                                 const float4* __restrict__ farT1, const Box* __restrict__ bxp,
                                                 int which, float* R, const float4* __restrict__ featT,
                                                 const float* __restrict__ metaT, float4* featB, float* metaB,
                                                 int* validB) {
  ff_base_body(fine, coarse, farA, farT0, farT1, bxp, which, R, featT, metaT, featB, metaB, validB);
}

// The network's input for a pair of passes, the present one (T) and the base
// (B), by voxel: channels 0-3 the present features, 4-7 their change since the
// base over s, 8 and 9 the block's occupancy at T and at B, the rest zero.
// s (to scale[0]) is the root mean square of the change's four channels over
// the blocks occupied at T or B, or 1 when nothing changed.  One work-group.
// With ok given, whether the base can serve -- recorded, and the box's corner
// within a quarter block of where it was -- goes to *ok, and a given gate is
// cleared when it cannot.  MESH: the present features are taken here from box
// `which`'s coarse mesh (k_ff_feat's) and stored to featT and metaT; else read
// from featT and metaT.
constexpr int PAIR_BLOCK = 1024;
template <bool MESH>
__device__ __forceinline__ void ff_pair_in_body(float4* featT, float* metaT, const float4* __restrict__ featB,
                                                const float* __restrict__ metaB, const int* __restrict__ validB,
                                                uint16_t* x0, float* scale, int* ok, int* gate,
                                                const float4* __restrict__ coarse, const float4* __restrict__ mono,
                                                const Box* __restrict__ bxp, int which) {
  __shared__ float redS[PAIR_BLOCK / 64];
  __shared__ int redN[PAIR_BLOCK / 64];
  const int tid = threadIdx.x;
  float4 fT[4096 / PAIR_BLOCK], fB[4096 / PAIR_BLOCK];
  float ss = 0.0f;
  int cnt = 0;
  float mt = 0.0f;
  const float* L = nullptr;
  Box bx;
  if (MESH) {
    bx = *bxp;
    L = which ? bx.LB : bx.LA;
    mt = fmaxf(mono[0].w, 1e-20f);
  }
  for (int j = 0; j < 4096 / PAIR_BLOCK; j++) {
    int v = tid + PAIR_BLOCK * j;
    if (MESH) {
      fT[j] = ff_feature(coarse, mt, L, v);
      featT[v] = fT[j];
    } else {
      fT[j] = featT[v];
    }
    fB[j] = featB[v];
    if (fT[j].x > 0.0f || fB[j].x > 0.0f) {
      float a = fT[j].x - fB[j].x, b = fT[j].y - fB[j].y, c = fT[j].z - fB[j].z, d = fT[j].w - fB[j].w;
      ss += (a * a + b * b) + (c * c + d * d);
      cnt++;
    }
  }
  ss = ff_wave_sum(ss);
  cnt = ff_wave_sum(cnt);
  if ((tid & 63) == 0) { redS[tid >> 6] = ss; redN[tid >> 6] = cnt; }
  __syncthreads();
  ss = 0.0f;
  cnt = 0;
  for (int w = 0; w < PAIR_BLOCK / 64; w++) { ss += redS[w]; cnt += redN[w]; }
  const float s = ss > 0.0f ? sqrtf(ss / (4.0f * (float)cnt)) : 1.0f;
  for (int j = 0; j < 4096 / PAIR_BLOCK; j++) {
    int v = tid + PAIR_BLOCK * j;
    float ch[10] = { fT[j].x, fT[j].y, fT[j].z, fT[j].w, (fT[j].x - fB[j].x) / s, (fT[j].y - fB[j].y) / s,
                     (fT[j].z - fB[j].z) / s, (fT[j].w - fB[j].w) / s, fT[j].x > 0.0f ? 1.0f : 0.0f,
                     fB[j].x > 0.0f ? 1.0f : 0.0f };
    uint16_t* o = x0 + (size_t)v * 16;
    for (int k = 0; k < 16; k++) o[k] = k < 10 ? f2bf(ch[k]) : (uint16_t)0;
  }
  if (tid == 0) {
    scale[0] = s;
    if (MESH) { metaT[0] = L[0]; metaT[1] = L[1]; metaT[2] = L[2]; metaT[3] = mt; }
    if (ok) {
      const float q = 0.25f * BLOCK_LEN;
      int good = (!validB || *validB) && fabsf(metaT[0] - metaB[0]) <= q && fabsf(metaT[1] - metaB[1]) <= q &&
                 fabsf(metaT[2] - metaB[2]) <= q;
      *ok = good;
      if (!good && gate) *gate = 0;
    }
  }
}

template <bool MESH>
__global__ void __launch_bounds__(PAIR_BLOCK) k_ff_pair_in(float4* featT, float* metaT,
                                                           const float4* __restrict__ featB, const float* __restrict__ metaB,
                                                           const int* __restrict__ validB, uint16_t* x0, float* scale,
                                                           int* ok, int* gate, const float4* __restrict__ coarse,
                                                           const float4* __restrict__ mono, const Box* __restrict__ bxp,
                                                           int which) {
  ff_pair_in_body<MESH>(featT, metaT, featB, metaB, validB, x0, scale, ok, gate, coarse, mono, bxp, which);
}

// Output and target units: the change of a block's record per unit of the
// input's change is that of a mass s * (total mass) / 64; accelerations in
// that mass over the block length squared, potentials over the block length,
// tidal tensors over its cube.
constexpr float FF_A = 1.0f / (BLOCK_LEN * BLOCK_LEN), FF_P = 1.0f / BLOCK_LEN,
                FF_T = 1.0f / (BLOCK_LEN * BLOCK_LEN * BLOCK_LEN);
__device__ __forceinline__ float ff_unit(int k) { return k < 3 ? FF_A : (k == 3 ? FF_P : FF_T); }

// The training target of a pair: the change of each block's record since the
// base in the output's units, on the blocks occupied at T and at B (a block
// that fills after the base gets the exact far field); the mask weighs each by
// its occupied fine texels at T over 64, as the judge weighs the texels.
__global__ void k_ff_pair_tgt(const float* __restrict__ RT, const float* __restrict__ RB,
                              const float4* __restrict__ featT, const float4* __restrict__ featB,
                              const float* __restrict__ metaT, const float* __restrict__ scale, float* tgt,
                              float* mask) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= 4096) return;
  bool occ = featT[v].x > 0.0f && featB[v].x > 0.0f;
  float u = scale[0] * metaT[3] / 64.0f;
  float* o = tgt + (size_t)v * 16;
  for (int k = 0; k < 16; k++)
    o[k] = occ && k < 10 ? (RT[(size_t)v * FF_R + k] - RB[(size_t)v * FF_R + k]) / (u * ff_unit(k)) : 0.0f;
  mask[v] = occ ? RT[(size_t)v * FF_R + 10] * (1.0f / 64.0f) : 0.0f;
}

// The network's far field: each occupied block's base record plus the
// network's change, carried from the block's centre to its present centre of
// mass by the tidal tensor; empty blocks zero; blocks empty at the base left
// to k_far.  Skips unless *gate, when given.
__device__ __forceinline__ void ff_out_body(const float* __restrict__ y, const float* __restrict__ RB,
                                            const float4* __restrict__ featB, const float* __restrict__ scale,
                                            const float4* __restrict__ coarse, const float* __restrict__ metaT,
                                            const Box* __restrict__ bxp, int which, float4* farA, float4* farT0,
                                            float4* farT1, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= 4096) return;
  int x = v & 15, yv = (v >> 4) & 15, z = v >> 8;
  int t = coarse_texel(x, yv, z);
  float4 c = coarse[t];
  if (c.w <= 0.0f) {
    farA[t] = make_float4(0, 0, 0, 0); farT0[t] = make_float4(0, 0, 0, 0); farT1[t] = make_float4(0, 0, 0, 0);
    return;
  }
  if (featB[v].x <= 0.0f) return;   // filled after the base: k_far's
  const Box bx = *bxp;
  float dx, dy, dz;
  block_offset(c, x, yv, z, which ? bx.LB : bx.LA, dx, dy, dz);
  const float* o = y + (size_t)v * 16;
  const float* r = RB + (size_t)v * FF_R;
  const float u = scale[0] * metaT[3] / 64.0f, uA = u * FF_A, uP = u * FF_P, uT = u * FF_T;
  float ax = r[0] + o[0] * uA, ay = r[1] + o[1] * uA, az = r[2] + o[2] * uA, p = r[3] + o[3] * uP;
  float txx = r[4] + o[4] * uT, tyy = r[5] + o[5] * uT, tzz = r[6] + o[6] * uT, txy = r[7] + o[7] * uT;
  float txz = r[8] + o[8] * uT, tyz = r[9] + o[9] * uT;
  float sx = ax + txx * dx + txy * dy + txz * dz, sy = ay + txy * dx + tyy * dy + tyz * dz,
        sz = az + txz * dx + tyz * dy + tzz * dz;
  farA[t] = make_float4(sx, sy, sz, p - (ax * dx + ay * dy + az * dz));
  farT0[t] = make_float4(txx, tyy, tzz, txy);
  farT1[t] = make_float4(txz, tyz, 0.0f, 0.0f);
}
__global__ void k_ff_out(const float* __restrict__ y, const float* __restrict__ RB, const float4* __restrict__ featB,
                         const float* __restrict__ scale, const float4* __restrict__ coarse,
                         const float* __restrict__ metaT, const Box* __restrict__ bxp, int which, float4* farA,
                         float4* farT0, float4* farT1, const int* __restrict__ gate) {
  ff_out_body(y, RB, featB, scale, coarse, metaT, bxp, which, farA, farT0, farT1, gate);
}

// The loss's weight of output channel c: accelerations 1, the potential and
// the tidal tensor 1/4 (the judge sees them only through the tidal term's
// short lever), the unused channels 0.
__device__ __forceinline__ float ff_chan_weight(int c) { return c < 3 ? 1.0f : (c < 10 ? 0.25f : 0.0f); }
// Weighted mean squared error (each voxel by its mask, each channel by
// ff_chan_weight) and its gradient; loss accumulated in double.
__global__ void k_ff_loss(const float* __restrict__ y, const float* __restrict__ tgt, const float* __restrict__ mask,
                          int M, float invCount, float* gy, double* loss) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (e < M * 16) {
    int v = e >> 4, c = e & 15;
    float w = mask[v] * ff_chan_weight(c);
    float d = w > 0.0f ? y[e] - tgt[e] : 0.0f;
    gy[e] = 2.0f * w * d * invCount;
    l = w * d * d * invCount;
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(loss, (double)l);
}
// Adam on the FP32 masters
__global__ void k_adam(float* w, const float* __restrict__ g, float* m, float* v, int n, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float gg = g[e];
  float mm = b1 * m[e] + (1.0f - b1) * gg, vv = b2 * v[e] + (1.0f - b2) * gg * gg;
  m[e] = mm; v[e] = vv;
  w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}
// The judge's sums over the force mesh's texels (scratch row left out):
// [0] |learned - exact far-field part|^2, [1] |exact force|^2, in double.  A
// texel's far-field part is its block's far acceleration carried by the
// tidal tensor to the texel's centre of mass, times G m, as k_cell adds it;
// the near part is the same in both, so the difference of the two force
// meshes is the difference of their far parts.  Blocks empty at the base
// count no difference: the mesh steps the network stands in give them k_far's.
constexpr int JUDGE_BLOCKS = 128;   // work-groups of 256 striding over the texels: one atomic pair each
__device__ __forceinline__ void ff_judge_sum_body(const float4* __restrict__ com, const float4* __restrict__ ccom,
                                                  const float4* __restrict__ Al, const float4* __restrict__ T0l,
                                                  const float4* __restrict__ T1l, const float4* __restrict__ Ae,
                                                  const float4* __restrict__ T0e, const float4* __restrict__ T1e,
                                                  const float4* __restrict__ force,
                                                  const float4* __restrict__ featB, float gm, double* sums) {
  __shared__ float redD[4], redR[4];
  float d2 = 0.0f, r2 = 0.0f;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < 512 * 512; e += gridDim.x * blockDim.x) {
    float4 F = force[e];
    r2 += F.x * F.x + F.y * F.y + F.z * F.z;
    float4 mc = com[e];
    int mx = e & 511, my = e >> 9;
    int c0 = mx & 63, c1 = my & 63, c2 = (my >> 6) * 8 + (mx >> 6);
    if (mc.w > 0.0f && featB[((c2 >> 2) * 16 + (c1 >> 2)) * 16 + (c0 >> 2)].x > 0.0f) {
      int ca = coarse_texel(c0 >> 2, c1 >> 2, c2 >> 2);
      float4 mb = ccom[ca];
      float q0 = mc.x - mb.x, q1 = mc.y - mb.y, q2 = mc.z - mb.z;
      float4 a = Al[ca], b = Ae[ca], s0 = T0l[ca], t0 = T0e[ca], s1 = T1l[ca], t1 = T1e[ca];
      float u0 = s0.x - t0.x, u1 = s0.y - t0.y, u2 = s0.z - t0.z, u3 = s0.w - t0.w, u4 = s1.x - t1.x, u5 = s1.y - t1.y;
      float dx = (a.x - b.x) + u0 * q0 + u3 * q1 + u4 * q2;
      float dy = (a.y - b.y) + u3 * q0 + u1 * q1 + u5 * q2;
      float dz = (a.z - b.z) + u4 * q0 + u5 * q1 + u2 * q2;
      d2 += (dx * dx + dy * dy + dz * dz) * (gm * gm);
    }
  }
  d2 = ff_wave_sum(d2);
  r2 = ff_wave_sum(r2);
  if ((threadIdx.x & 63) == 0) { redD[threadIdx.x >> 6] = d2; redR[threadIdx.x >> 6] = r2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&sums[0], (double)redD[0] + (double)redD[1] + (double)redD[2] + (double)redD[3]);
    atomicAdd(&sums[1], (double)redR[0] + (double)redR[1] + (double)redR[2] + (double)redR[3]);
  }
}
__global__ void __launch_bounds__(256) k_ff_judge_sum(const float4* __restrict__ com, const float4* __restrict__ ccom,
                                                      const float4* __restrict__ Al, const float4* __restrict__ T0l,
                                                      const float4* __restrict__ T1l, const float4* __restrict__ Ae,
                                                      const float4* __restrict__ T0e, const float4* __restrict__ T1e,
                                                      const float4* __restrict__ force,
                                                      const float4* __restrict__ featB, float gm, double* sums) {
  ff_judge_sum_body(com, ccom, Al, T0l, T1l, Ae, T0e, T1e, force, featB, gm, sums);
}
// The judge's verdict for the coming mesh steps (the sums put back to zero):
// the network stands in while its relative error is within tol and it was
// measured from a base that could serve.  stats: [0] last error, [1] judged,
// [2] fallbacks, [3] learned mesh steps, [4] largest error measured from a
// base that could serve, [5] their sum, [6] their count.
__device__ __forceinline__ void judge_decide_body(double* sums, float tol, const int* __restrict__ okBase,
                                                  float* stats, int* gate) {
  if (threadIdx.x != 0) return;
  double err = sums[1] > 0.0 ? sqrt(sums[0] / sums[1]) : 0.0;
  stats[0] = (float)err;
  stats[1] += 1.0f;
  int usable = *okBase;
  int ok = usable && err <= tol;
  if (!ok) stats[2] += 1.0f;
  if (usable) {
    stats[4] = fmaxf(stats[4], (float)err);
    stats[5] += (float)err;
    stats[6] += 1.0f;
  }
  *gate = ok;
  sums[0] = 0.0; sums[1] = 0.0;
}
__global__ void k_judge_decide(double* sums, float tol, const int* __restrict__ okBase, float* stats, int* gate) {
  judge_decide_body(sums, tol, okBase, stats, gate);
}
__global__ void k_count_learned(const int* flag, float* stats) {
  if (threadIdx.x == 0 && *flag) stats[3] += 1.0f;
}

// ------------------------------------------- batched: one pass over a batch --
// Every (instance, box) of a batch at its mesh step is a slot of one forward
// pass: the slot's input pair from its own mesh and base (work-group y = slot),
// the forward pass over all slots, and each slot's output to its destination
// (the mesh's far field on a step the network may stand in for, gated; the
// judge's buffer on a judge step).  The judged slots are judged together at
// the end of the batch's cycle, when every instance's exact pass is done.
__global__ void __launch_bounds__(PAIR_BLOCK) k_ff_pair_in_b(const FFSlot* __restrict__ slots, uint16_t* x0,
                                                             float* scale) {
  const FFSlot& S = slots[blockIdx.x];
  ff_pair_in_body<true>(S.featT, S.metaT, S.featB, S.metaB, S.validB, x0 + (size_t)blockIdx.x * 4096 * 16,
                        scale + blockIdx.x, S.okBase, S.gate, S.coarse, S.mono, S.box, S.which);
}
__global__ void k_ff_out_b(const FFSlot* __restrict__ slots, const float* __restrict__ y, const float* __restrict__ scale) {
  const FFSlot& S = slots[blockIdx.y];
  ff_out_body(y + (size_t)blockIdx.y * 4096 * 16, S.RB, S.featB, scale + blockIdx.y, S.coarse, S.metaT, S.box,
              S.which, S.A, S.T0, S.T1, S.gate);
}
__global__ void __launch_bounds__(256) k_ff_judge_sum_b(const FFJudgeSlot* __restrict__ slots) {
  const FFJudgeSlot& J = slots[blockIdx.y];
  ff_judge_sum_body(J.com, J.ccom, J.learned, J.learned + 4096, J.learned + 8192, J.farA, J.farT0, J.farT1, J.force,
                    J.featB, J.gm, J.sums);
}
__global__ void k_judge_decide_b(const FFJudgeSlot* __restrict__ slots) {
  const FFJudgeSlot& J = slots[blockIdx.x];
  judge_decide_body(J.sums, J.tol, J.okBase, J.stats, J.gate);
}
__global__ void __launch_bounds__(256) k_ff_base_b(const FFJudgeSlot* __restrict__ slots) {
  const FFJudgeSlot& J = slots[blockIdx.y];
  ff_base_body(J.fine, J.coarse, J.farA, J.farT0, J.farT1, J.box, J.which, J.RB, J.featT, J.metaT, J.featB, J.metaB,
               J.validB);
}
__global__ void k_zero_d(double* p, int n) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = 0.0;
}


// relative error sums of a forward pass, each voxel weighted by its mask:
// [0] |dA|^2 [1] |A|^2 over the acceleration channels, [2] |dY|^2 [3] |Y|^2
// over all ten
__global__ void k_eval_sum(const float* __restrict__ y, const float* __restrict__ tgt, const float* __restrict__ mask,
                           int M, double* sums) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  if (e < M * 16) {
    int v = e >> 4, c = e & 15;
    if (c < 10 && mask[v] > 0.0f) {
      float d = y[e] - tgt[e], t = tgt[e], w = mask[v];
      s2 = w * d * d; s3 = w * t * t;
      if (c < 3) { s0 = s2; s1 = s3; }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o); s1 += __shfl_xor(s1, o); s2 += __shfl_xor(s2, o); s3 += __shfl_xor(s3, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&sums[0], (double)s0); atomicAdd(&sums[1], (double)s1);
    atomicAdd(&sums[2], (double)s2); atomicAdd(&sums[3], (double)s3);
  }
}

// ------------------------------------------------------------------- host --
struct Layer {
  int cin, cout, act;
  size_t w, b;          // offsets of the layer's weights and bias in the master array
  uint16_t *fwd, *bwd;  // packed BF16 weights
};
// the network's twelve convolutions (see the top of this file)
static const int LAYER_IO[12][3] = { { 16, 32, 1 }, { 32, 32, 1 }, { 32, 32, 1 }, { 32, 32, 1 }, { 32, 32, 1 },
                                     { 32, 32, 1 }, { 32, 32, 1 }, { 32, 32, 1 }, { 32, 32, 1 }, { 32, 32, 1 },
                                     { 32, 32, 1 }, { 32, 16, 0 } };
static const int LAYER_S[12] = { 16, 16, 8, 8, 4, 4, 2, 2, 4, 8, 16, 16 };

size_t ff_weight_count() {
  size_t n = 0;
  for (auto& l : LAYER_IO) n += (size_t)27 * l[0] * l[1] + l[1];
  return n;
}

// Forward activations for B instances (outputs, and pre-activations when
// training) and the gradient buffers.
struct Acts {
  int B = 0;
  uint16_t *x0, *a0, *e0, *p1, *a2, *e1, *p2, *a4, *e2, *p3, *a6, *bot, *u2, *d2, *u1, *d1, *u0, *d0;
  float* y;
  float* pre[12];
  float* scale;
  // gradients (training only)
  float *gy, *gA16, *gB16, *gE16, *gA8, *gB8, *gE8, *gA4, *gB4, *gE4, *gA2, *gB2;
  uint16_t* gbf;
  std::vector<void*> allocs;
};

struct FFNet {
  Layer L[12];
  size_t nw = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  int adamT = 0;
  float slope = 0.25f;          // the leaky ReLU's negative slope (1: the network is linear)
  int cap = 0, count = 0;
  uint16_t* dataIn = nullptr;   // [cap][4096][16]
  float* dataTgt = nullptr;     // [cap][4096][16]
  float* dataMask = nullptr;    // [cap][4096]
  float* dataScale = nullptr;   // [cap] the input's scale s of each sample
  // recorded passes of the exact far field, from which samples are paired
  int passes = 0;
  float4* passFeat = nullptr;   // [cap][4096]
  float* passMeta = nullptr;    // [cap][4]
  float* passR = nullptr;       // [cap][4096][FF_R]
  Acts train;                   // batch activations for training
  int* idx = nullptr;
  int idxCap = 0;
  double* loss = nullptr;
  int lossCap = 0;
  hipStream_t s = nullptr;
  std::vector<void*> allocs;
  std::string err;
};

// A ctx's workspace: the inference activations, and per box the present
// pass's features and the base (features, meta, centre record, whether one is
// recorded, whether it could serve the last prediction).
struct FFWork {
  Acts a;
  float4 *featT[2], *featB[2];
  float *metaT[2], *metaB[2], *RB[2];
  int *validB[2], *okBase[2];
  std::vector<void*> allocs;
};

static bool dev_alloc(std::vector<void*>& list, void** p, size_t bytes, std::string& err) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) { err = std::string("hipMalloc: ") + hipGetErrorString(e); return false; }
  list.push_back(*p);
  return true;
}
#define DA(list, ptr, bytes) if (!dev_alloc(list, (void**)&(ptr), (bytes), err)) return false

static bool acts_alloc(Acts& a, int B, bool training, std::string& err) {
  a.B = B;
  size_t v16 = (size_t)B * 4096, v8 = (size_t)B * 512, v4 = (size_t)B * 64, v2 = (size_t)B * 8;
  DA(a.allocs, a.x0, v16 * 16 * 2);
  DA(a.allocs, a.a0, v16 * 64); DA(a.allocs, a.e0, v16 * 64); DA(a.allocs, a.p1, v8 * 64);
  DA(a.allocs, a.a2, v8 * 64); DA(a.allocs, a.e1, v8 * 64); DA(a.allocs, a.p2, v4 * 64);
  DA(a.allocs, a.a4, v4 * 64); DA(a.allocs, a.e2, v4 * 64); DA(a.allocs, a.p3, v2 * 64);
  DA(a.allocs, a.a6, v2 * 64); DA(a.allocs, a.bot, v2 * 64); DA(a.allocs, a.u2, v4 * 64);
  DA(a.allocs, a.d2, v4 * 64); DA(a.allocs, a.u1, v8 * 64); DA(a.allocs, a.d1, v8 * 64);
  DA(a.allocs, a.u0, v16 * 64); DA(a.allocs, a.d0, v16 * 64);
  DA(a.allocs, a.y, v16 * 16 * 4);
  DA(a.allocs, a.scale, (size_t)B * 4);
  for (int l = 0; l < 12; l++) a.pre[l] = nullptr;
  if (!training) return true;
  for (int l = 0; l < 12; l++) {
    size_t vox = (size_t)B * LAYER_S[l] * LAYER_S[l] * LAYER_S[l];
    DA(a.allocs, a.pre[l], vox * LAYER_IO[l][1] * 4);
  }
  DA(a.allocs, a.gy, v16 * 16 * 4);
  DA(a.allocs, a.gA16, v16 * 32 * 4); DA(a.allocs, a.gB16, v16 * 32 * 4); DA(a.allocs, a.gE16, v16 * 32 * 4);
  DA(a.allocs, a.gA8, v8 * 32 * 4); DA(a.allocs, a.gB8, v8 * 32 * 4); DA(a.allocs, a.gE8, v8 * 32 * 4);
  DA(a.allocs, a.gA4, v4 * 32 * 4); DA(a.allocs, a.gB4, v4 * 32 * 4); DA(a.allocs, a.gE4, v4 * 32 * 4);
  DA(a.allocs, a.gA2, v2 * 32 * 4); DA(a.allocs, a.gB2, v2 * 32 * 4);
  DA(a.allocs, a.gbf, v16 * 32 * 2);
  return true;
}
static void acts_free(Acts& a) {
  for (void* p : a.allocs) hipFree(p);
  a.allocs.clear();
}

static inline unsigned grid_of(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

static void conv(const Layer& L, float slope, const float* bias, const uint16_t* in, int B, int S, uint16_t* out,
                 float* outf, float* pre, const int* gate, hipStream_t s) {
  size_t M = (size_t)B * S * S * S;
  unsigned blocks = grid_of((M + 15) / 16 * 64);
  if (L.cin == 16)
    hipLaunchKernelGGL((k_conv<16, 32>), dim3(blocks), dim3(256), 0, s, in, B, S, L.fwd, bias, slope, out, outf, pre,
                       gate);
  else if (L.cout == 32)
    hipLaunchKernelGGL((k_conv<32, 32>), dim3(blocks), dim3(256), 0, s, in, B, S, L.fwd, bias, slope, out, outf, pre,
                       gate);
  else
    hipLaunchKernelGGL((k_conv<32, 16>), dim3(blocks), dim3(256), 0, s, in, B, S, L.fwd, bias, slope, out, outf, pre,
                       gate);
}
// the data gradient of a layer: g (bf16, at its output) through the flipped,
// transposed weights to f32 at its input
static void conv_dgrad(const Layer& L, const uint16_t* g, int B, int S, float* gin, hipStream_t s) {
  size_t M = (size_t)B * S * S * S;
  unsigned blocks = grid_of((M + 15) / 16 * 64);
  if (L.cout == 16)   // the output layer: 16 gradient channels in, 32 out
    hipLaunchKernelGGL((k_conv<16, 32>), dim3(blocks), dim3(256), 0, s, g, B, S, L.bwd, nullptr, 1.0f, nullptr, gin,
                       nullptr, nullptr);
  else if (L.cin == 32)
    hipLaunchKernelGGL((k_conv<32, 32>), dim3(blocks), dim3(256), 0, s, g, B, S, L.bwd, nullptr, 1.0f, nullptr, gin,
                       nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_conv<32, 16>), dim3(blocks), dim3(256), 0, s, g, B, S, L.bwd, nullptr, 1.0f, nullptr, gin,
                       nullptr, nullptr);
}
static void conv_wgrad(const Layer& L, const uint16_t* in, const uint16_t* g, int B, int S, float* dW, float* db,
                       hipStream_t s) {
  size_t M = (size_t)B * S * S * S;
  unsigned waves = (unsigned)(27 * ((M + 255) / 256));
  unsigned blocks = grid_of((size_t)waves * 64);
  if (L.cin == 16)
    hipLaunchKernelGGL((k_conv_wgrad<16, 32>), dim3(blocks), dim3(256), 0, s, in, B, S, g, dW);
  else if (L.cout == 32)
    hipLaunchKernelGGL((k_conv_wgrad<32, 32>), dim3(blocks), dim3(256), 0, s, in, B, S, g, dW);
  else
    hipLaunchKernelGGL((k_conv_wgrad<32, 16>), dim3(blocks), dim3(256), 0, s, in, B, S, g, dW);
  hipLaunchKernelGGL(k_bias_grad, dim3(L.cout), dim3(256), 0, s, g, (int)M, L.cout, db);
}

static void forward(const FFNet* n, Acts& a, int B, bool training, const int* gate, hipStream_t s) {
  const float* w = n->w;
  auto pre = [&](int l) { return training ? a.pre[l] : nullptr; };
  auto C = [&](int l, const uint16_t* in, uint16_t* out) {
    conv(n->L[l], n->L[l].act ? n->slope : 1.0f, w + n->L[l].b, in, B, LAYER_S[l], out, nullptr, pre(l), gate, s);
  };
  auto pool = [&](const uint16_t* in, int S, uint16_t* out) {
    hipLaunchKernelGGL(k_pool, dim3(grid_of((size_t)B * (S / 2) * (S / 2) * (S / 2) * 32)), dim3(256), 0, s, in, B, S,
                       32, out, gate);
  };
  auto up = [&](const uint16_t* coarse, const uint16_t* skip, int S, uint16_t* out) {
    hipLaunchKernelGGL(k_up_add, dim3(grid_of((size_t)B * S * S * S * 32)), dim3(256), 0, s, coarse, skip, B, S, 32,
                       out, gate);
  };
  C(0, a.x0, a.a0); C(1, a.a0, a.e0);
  pool(a.e0, 16, a.p1); C(2, a.p1, a.a2); C(3, a.a2, a.e1);
  pool(a.e1, 8, a.p2); C(4, a.p2, a.a4); C(5, a.a4, a.e2);
  pool(a.e2, 4, a.p3); C(6, a.p3, a.a6); C(7, a.a6, a.bot);
  up(a.bot, a.e2, 4, a.u2); C(8, a.u2, a.d2);
  up(a.d2, a.e1, 8, a.u1); C(9, a.u1, a.d1);
  up(a.d1, a.e0, 16, a.u0); C(10, a.u0, a.d0);
  conv(n->L[11], 1.0f, w + n->L[11].b, a.d0, B, 16, nullptr, a.y, pre(11), gate, s);
}

// backward from a.gy (f32, at the output) to the weight gradients n->g
static void backward(FFNet* n, Acts& a, int B, hipStream_t s) {
  float* g = n->g;
  auto actb = [&](const float* gout, int l) {   // gradient at layer l's output through its activation, to bf16
    size_t cnt = (size_t)B * LAYER_S[l] * LAYER_S[l] * LAYER_S[l] * n->L[l].cout;
    hipLaunchKernelGGL(k_act_bwd, dim3(grid_of(cnt)), dim3(256), 0, s, gout, a.pre[l], (int)cnt,
                       n->L[l].act ? n->slope : 1.0f, a.gbf);
  };
  auto W = [&](int l, const uint16_t* in) {
    conv_wgrad(n->L[l], in, a.gbf, B, LAYER_S[l], g + n->L[l].w, g + n->L[l].b, s);
  };
  auto D = [&](int l, float* gin) { conv_dgrad(n->L[l], a.gbf, B, LAYER_S[l], gin, s); };
  auto upb = [&](const float* gout, int S, float* gc) {
    hipLaunchKernelGGL(k_up_bwd, dim3(grid_of((size_t)B * (S / 2) * (S / 2) * (S / 2) * 32)), dim3(256), 0, s, gout, B,
                       S, 32, gc);
  };
  auto poolb = [&](const float* gout, int S, float* gin) {   // accumulate into gin
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid_of((size_t)B * S * S * S * 32)), dim3(256), 0, s, gout, B, S, 32, gin, 1);
  };
  auto copy = [&](float* dst, const float* src, size_t cnt) {
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_of(cnt)), dim3(256), 0, s, dst, cnt);
    hipLaunchKernelGGL(k_add_f32, dim3(grid_of(cnt)), dim3(256), 0, s, dst, src, (int)cnt);
  };
  size_t v16 = (size_t)B * 4096 * 32, v8 = (size_t)B * 512 * 32, v4 = (size_t)B * 64 * 32;
  actb(a.gy, 11); W(11, a.d0); D(11, a.gA16);                 // gA16: d0
  actb(a.gA16, 10); W(10, a.u0); D(10, a.gB16);               // gB16: u0 = up(d1) + e0
  copy(a.gE16, a.gB16, v16);                                  // gE16: e0 (skip)
  upb(a.gB16, 16, a.gA8);                                     // gA8: d1
  actb(a.gA8, 9); W(9, a.u1); D(9, a.gB8);                    // gB8: u1 = up(d2) + e1
  copy(a.gE8, a.gB8, v8);
  upb(a.gB8, 8, a.gA4);                                       // gA4: d2
  actb(a.gA4, 8); W(8, a.u2); D(8, a.gB4);                    // gB4: u2 = up(bot) + e2
  copy(a.gE4, a.gB4, v4);
  upb(a.gB4, 4, a.gA2);                                       // gA2: bot
  actb(a.gA2, 7); W(7, a.a6); D(7, a.gB2);                    // gB2: a6
  actb(a.gB2, 6); W(6, a.p3); D(6, a.gA2);                    // gA2: p3
  poolb(a.gA2, 4, a.gE4);                                     // e2 += pool grad
  actb(a.gE4, 5); W(5, a.a4); D(5, a.gB4);                    // gB4: a4
  actb(a.gB4, 4); W(4, a.p2); D(4, a.gA4);                    // gA4: p2
  poolb(a.gA4, 8, a.gE8);                                     // e1
  actb(a.gE8, 3); W(3, a.a2); D(3, a.gB8);                    // a2
  actb(a.gB8, 2); W(2, a.p1); D(2, a.gA8);                    // p1
  poolb(a.gA8, 16, a.gE16);                                   // e0
  actb(a.gE16, 1); W(1, a.a0); D(1, a.gB16);                  // a0
  actb(a.gB16, 0); W(0, a.x0);
}

static void repack(FFNet* n, hipStream_t s) {
  for (int l = 0; l < 12; l++) {
    const Layer& L = n->L[l];
    size_t cnt = (size_t)27 * L.cin * L.cout;
    hipLaunchKernelGGL(k_pack, dim3(grid_of(cnt)), dim3(256), 0, s, n->w + L.w, L.cin, L.cout, L.fwd, L.bwd);
  }
}

static bool ff_init(FFNet* n, int cap, std::string& err) {
  size_t off = 0;
  for (int l = 0; l < 12; l++) {
    Layer& L = n->L[l];
    L.cin = LAYER_IO[l][0]; L.cout = LAYER_IO[l][1]; L.act = LAYER_IO[l][2];
    L.w = off; off += (size_t)27 * L.cin * L.cout;
    L.b = off; off += L.cout;
    DA(n->allocs, L.fwd, (size_t)27 * L.cin * L.cout * 2);
    DA(n->allocs, L.bwd, (size_t)27 * L.cin * L.cout * 2);
  }
  n->nw = off;
  DA(n->allocs, n->w, off * 4); DA(n->allocs, n->g, off * 4); DA(n->allocs, n->m, off * 4); DA(n->allocs, n->v, off * 4);
  n->cap = cap;
  if (cap > 0) {
    DA(n->allocs, n->dataIn, (size_t)cap * 4096 * 16 * 2);
    DA(n->allocs, n->dataTgt, (size_t)cap * 4096 * 16 * 4);
    DA(n->allocs, n->dataMask, (size_t)cap * 4096 * 4);
    DA(n->allocs, n->dataScale, (size_t)cap * 4);
    DA(n->allocs, n->passFeat, (size_t)cap * 4096 * 16);
    DA(n->allocs, n->passMeta, (size_t)cap * 4 * 4);
    DA(n->allocs, n->passR, (size_t)cap * 4096 * FF_R * 4);
  }
  hipError_t e = hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking);
  if (e != hipSuccess) { err = hipGetErrorString(e); return false; }
  // He-style initialisation, seeded: uniform with variance 2 / fan-in (1 / fan-in for the linear output)
  std::vector<float> w(off, 0.0f);
  uint32_t st = 0x5EEDF00Du;
  auto rnd = [&]() { st ^= st << 13; st ^= st >> 17; st ^= st << 5; return (st >> 8) * (1.0f / 16777216.0f); };
  for (int l = 0; l < 12; l++) {
    const Layer& L = n->L[l];
    float fan = 27.0f * (l == 0 ? (float)FF_IN : (float)L.cin);
    float a = sqrtf(3.0f * (L.act ? 2.0f : 1.0f) / fan);
    for (int t = 0; t < 27; t++)
      for (int ci = 0; ci < L.cin; ci++)
        for (int co = 0; co < L.cout; co++)
          w[L.w + ((size_t)t * L.cin + ci) * L.cout + co] = (l == 0 && ci >= FF_IN) ? 0.0f : a * (2.0f * rnd() - 1.0f);
  }
  e = hipMemcpyAsync(n->w, w.data(), off * 4, hipMemcpyHostToDevice, n->s);
  if (e == hipSuccess) e = hipMemsetAsync(n->m, 0, off * 4, n->s);
  if (e == hipSuccess) e = hipMemsetAsync(n->v, 0, off * 4, n->s);
  repack(n, n->s);
  if (e == hipSuccess) e = hipStreamSynchronize(n->s);
  if (e != hipSuccess) { err = hipGetErrorString(e); return false; }
  return true;
}

FFNet* ff_create(int cap, std::string& err) {
  FFNet* n = new FFNet();
  if (!ff_init(n, cap, err)) { ff_destroy(n); return nullptr; }
  return n;
}
void ff_destroy(FFNet* n) {
  if (!n) return;
  if (n->s) hipStreamSynchronize(n->s);
  acts_free(n->train);
  for (void* p : n->allocs) hipFree(p);
  if (n->s) hipStreamDestroy(n->s);
  delete n;
}
static bool work_alloc(FFWork* w, std::string& err) {
  for (int b = 0; b < 2; b++) {
    DA(w->allocs, w->featT[b], 4096 * 16); DA(w->allocs, w->featB[b], 4096 * 16);
    DA(w->allocs, w->metaT[b], 16); DA(w->allocs, w->metaB[b], 16);
    DA(w->allocs, w->RB[b], (size_t)4096 * FF_R * 4);
    DA(w->allocs, w->validB[b], 4); DA(w->allocs, w->okBase[b], 4);
  }
  for (int b = 0; b < 2; b++)   // no base yet: zero features and records, not whatever the allocation held
    if (hipMemset(w->featB[b], 0, 4096 * 16) != hipSuccess || hipMemset(w->RB[b], 0, (size_t)4096 * FF_R * 4) != hipSuccess ||
        hipMemset(w->featT[b], 0, 4096 * 16) != hipSuccess || hipMemset(w->metaB[b], 0, 16) != hipSuccess) {
      err = "hipMemset";
      return false;
    }
  return true;
}
FFWork* ffwork_create(std::string& err) {
  FFWork* w = new FFWork();
  if (!acts_alloc(w->a, 1, false, err) || !work_alloc(w, err)) { ffwork_destroy(w); return nullptr; }
  return w;
}
void ffwork_destroy(FFWork* w) {
  if (!w) return;
  acts_free(w->a);
  for (void* p : w->allocs) hipFree(p);
  delete w;
}
__global__ void k_fill_i32(int* p, int n, int v) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}
// both boxes' bases forgotten: the next judge of each records one, and no
// mesh step stands in before it (a fill kernel, so a caller may capture it)
const float4* ff_base_features(const FFWork* w, int which) { return w->featB[which]; }
void ff_forget_bases(FFWork* w, hipStream_t s) {
  for (int b = 0; b < 2; b++) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, w->validB[b], 1, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, w->okBase[b], 1, 0);
  }
}

// The network's far field for box `which` at a mesh step: the present
// pass's features, the input pair against the box's base, the forward pass and
// the output into farA/T0/T1.  gate (the mesh steps between judges): cleared
// when the base cannot serve, and the forward pass and the output skip unless
// it is set.  Without a gate (a judge step) everything runs.
void ff_predict(const FFNet* n, FFWork* wk, int which, const float4* coarse, const float4* mono, const Box* bx,
                float4* farA, float4* farT0, float4* farT1, int* gate, hipStream_t s) {
  Acts& a = wk->a;
  hipLaunchKernelGGL(k_ff_pair_in<true>, dim3(1), dim3(PAIR_BLOCK), 0, s, wk->featT[which], wk->metaT[which],
                     wk->featB[which], wk->metaB[which], wk->validB[which], a.x0, a.scale, wk->okBase[which], gate,
                     coarse, mono, bx, which);
  forward(n, a, 1, false, gate, s);
  hipLaunchKernelGGL(k_ff_out, dim3(16), dim3(256), 0, s, a.y, wk->RB[which], wk->featB[which], a.scale, coarse,
                     wk->metaT[which], bx, which, farA, farT0, farT1, gate);
}

// A judge step, after ff_predict (into learned: farA, farT0, farT1 of 4096
// each) and the exact far field and force mesh: the judge's sums and verdict
// (gate, stats), then the exact pass becomes the box's base.
void ff_judge_rebase(FFWork* wk, int which, const float4* fine, const float4* com, const float4* ccom,
                     const float4* coarse, const float4* learned, const float4* farA, const float4* farT0,
                     const float4* farT1, const float4* force, const Box* bx, float gm, double* sums, float tol,
                     float* stats, int* gate, hipStream_t s) {
  hipLaunchKernelGGL(k_ff_judge_sum, dim3(JUDGE_BLOCKS), dim3(256), 0, s, com, ccom, learned, learned + 4096,
                     learned + 8192, farA, farT0, farT1, force, wk->featB[which], gm, sums);
  hipLaunchKernelGGL(k_judge_decide, dim3(1), dim3(64), 0, s, sums, tol, wk->okBase[which], stats, gate);
  hipLaunchKernelGGL(k_ff_base, dim3(1024), dim3(256), 0, s, fine, coarse, farA, farT0, farT1, bx, which,
                     wk->RB[which], wk->featT[which], wk->metaT[which], wk->featB[which], wk->metaB[which],
                     wk->validB[which]);
}
void launch_count_learned(const int* gate, float* stats, hipStream_t s) {
  hipLaunchKernelGGL(k_count_learned, dim3(1), dim3(64), 0, s, gate, stats);
}

FFSlot ff_slot(FFWork* wk, int which, const float4* coarse, const float4* mono, const Box* box, int* gate, float4* A,
               float4* T0, float4* T1) {
  FFSlot o;
  o.coarse = coarse; o.mono = mono; o.box = box; o.which = which;
  o.featT = wk->featT[which]; o.metaT = wk->metaT[which]; o.featB = wk->featB[which]; o.metaB = wk->metaB[which];
  o.validB = wk->validB[which]; o.okBase = wk->okBase[which]; o.gate = gate; o.RB = wk->RB[which];
  o.A = A; o.T0 = T0; o.T1 = T1;
  return o;
}
FFJudgeSlot ff_judge_slot(FFWork* wk, int which, const float4* fine, const float4* com, const float4* ccom,
                          const float4* coarse, const float4* learned, const float4* farA, const float4* farT0,
                          const float4* farT1, const float4* force, const Box* box, float gm, double* sums, float tol,
                          float* stats, int* gate) {
  FFJudgeSlot o;
  o.fine = fine; o.com = com; o.ccom = ccom; o.coarse = coarse; o.learned = learned;
  o.farA = farA; o.farT0 = farT0; o.farT1 = farT1; o.force = force; o.box = box; o.which = which;
  o.gm = gm; o.tol = tol; o.sums = sums; o.stats = stats; o.gate = gate; o.okBase = wk->okBase[which];
  o.RB = wk->RB[which]; o.featT = wk->featT[which]; o.metaT = wk->metaT[which]; o.featB = wk->featB[which];
  o.metaB = wk->metaB[which]; o.validB = wk->validB[which];
  return o;
}

// A batch's workspace: activations for up to cap slots and, per captured cycle
// of the batch's graph cache, the slot tables that cycle's launches read.
struct FFBatchWork {
  int cap = 0;
  Acts a;
  FFSlot* slots[FF_BATCH_TABLES + 1] = {};
  FFJudgeSlot* judge[FF_BATCH_TABLES + 1] = {};
  std::vector<void*> allocs;
};
FFBatchWork* ffbatch_create(int cap, std::string& err) {
  FFBatchWork* w = new FFBatchWork();
  w->cap = cap;
  bool ok = acts_alloc(w->a, cap, false, err);
  for (int t = 0; ok && t <= FF_BATCH_TABLES; t++)
    ok = dev_alloc(w->allocs, (void**)&w->slots[t], sizeof(FFSlot) * cap, err) &&
         dev_alloc(w->allocs, (void**)&w->judge[t], sizeof(FFJudgeSlot) * cap, err);
  if (!ok) { ffbatch_destroy(w); return nullptr; }
  return w;
}
void ffbatch_destroy(FFBatchWork* w) {
  if (!w) return;
  acts_free(w->a);
  for (void* p : w->allocs) hipFree(p);
  delete w;
}
int ffbatch_cap(const FFBatchWork* w) { return w->cap; }
int ffbatch_set(FFBatchWork* w, int t, const std::vector<FFSlot>& slots, const std::vector<FFJudgeSlot>& judge,
                std::string& err) {
  if (t < 0 || t > FF_EAGER_TABLE || (int)slots.size() > w->cap || (int)judge.size() > w->cap) {
    err = "batched far field: slot table out of range";
    return -1;
  }
  hipError_t e = hipSuccess;
  if (!slots.empty()) e = hipMemcpy(w->slots[t], slots.data(), sizeof(FFSlot) * slots.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !judge.empty())
    e = hipMemcpy(w->judge[t], judge.data(), sizeof(FFJudgeSlot) * judge.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) { err = hipGetErrorString(e); return -1; }
  return 0;
}
void ff_predict_batch(const FFNet* n, FFBatchWork* w, int t, int nSlots, hipStream_t s) {
  if (nSlots <= 0) return;
  Acts& a = w->a;
  hipLaunchKernelGGL(k_ff_pair_in_b, dim3(nSlots), dim3(PAIR_BLOCK), 0, s, w->slots[t], a.x0, a.scale);
  forward(n, a, nSlots, false, nullptr, s);
  hipLaunchKernelGGL(k_ff_out_b, dim3(16, nSlots), dim3(256), 0, s, w->slots[t], a.y, a.scale);
}
void ff_judge_batch(FFBatchWork* w, int t, int nJudge, hipStream_t s) {
  if (nJudge <= 0) return;
  hipLaunchKernelGGL(k_ff_judge_sum_b, dim3(JUDGE_BLOCKS, nJudge), dim3(256), 0, s, w->judge[t]);
  hipLaunchKernelGGL(k_judge_decide_b, dim3(nJudge), dim3(64), 0, s, w->judge[t]);
  hipLaunchKernelGGL(k_ff_base_b, dim3(1024, nJudge), dim3(256), 0, s, w->judge[t]);
}

// a recorded pass: box `which`'s features and the centre record of its exact
// far field; -1 when the record is full
int ff_record_pass(FFNet* n, const float4* fine, const float4* coarse, const float4* mono, const float4* farA, const float4* farT0, const float4* farT1, const Box* bx, int which,
                   hipStream_t s) {
  if (n->passes >= n->cap) return -1;
  int k = n->passes++;
  float4* feat = n->passFeat + (size_t)k * 4096;
  float* meta = n->passMeta + (size_t)k * 4;
  hipLaunchKernelGGL(k_ff_feat, dim3(16), dim3(256), 0, s, coarse, mono, bx, which, feat, meta);
  hipLaunchKernelGGL(k_ff_base, dim3(1024), dim3(256), 0, s, fine, coarse, farA, farT0, farT1, bx, which,
                     n->passR + (size_t)k * 4096 * FF_R, nullptr, nullptr, nullptr, nullptr, nullptr);
  return k;
}

// a training sample from recorded passes `base` and `cur`: the network's input
// and target as a mesh step would form them; -1 when the record is full
static int ff_pair(FFNet* n, int base, int cur) {
  if (n->count >= n->cap) return -1;
  int k = n->count++;
  hipStream_t s = n->s;
  const float4 *fT = n->passFeat + (size_t)cur * 4096, *fB = n->passFeat + (size_t)base * 4096;
  const float *mT = n->passMeta + (size_t)cur * 4, *mB = n->passMeta + (size_t)base * 4;
  hipLaunchKernelGGL(k_ff_pair_in<false>, dim3(1), dim3(PAIR_BLOCK), 0, s, const_cast<float4*>(fT),
                     const_cast<float*>(mT), fB, mB, nullptr, n->dataIn + (size_t)k * 4096 * 16, n->dataScale + k,
                     nullptr, nullptr, nullptr, nullptr, nullptr, 0);
  hipLaunchKernelGGL(k_ff_pair_tgt, dim3(16), dim3(256), 0, s, n->passR + (size_t)cur * 4096 * FF_R,
                     n->passR + (size_t)base * 4096 * FF_R, fT, fB, mT, n->dataScale + k, n->dataTgt + (size_t)k * 4096 * 16,
                     n->dataMask + (size_t)k * 4096);
  return k;
}

__global__ void k_gather(const uint16_t* __restrict__ dataIn, const float* __restrict__ dataTgt,
                         const float* __restrict__ dataMask, const int* __restrict__ idx, int B, uint16_t* x0,
                         float* tgt, float* mask) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)B * 4096 * 16) return;
  int b = (int)(e / (4096 * 16));
  size_t off = e - (size_t)b * 4096 * 16, src = (size_t)idx[b] * 4096 * 16 + off;
  x0[e] = dataIn[src];
  tgt[e] = dataTgt[src];
  if (off % 16 == 0) mask[(size_t)b * 4096 + off / 16] = dataMask[(size_t)idx[b] * 4096 + off / 16];
}

// iters steps of Adam on batches of B samples drawn (seeded) from the record;
// losses[iters] the mean squared error of each step's batch
int ff_train(FFNet* n, int iters, int B, float lr, uint32_t seed, double* losses, std::string& err) {
  if (n->count <= 0) { err = "no training samples recorded"; return -1; }
  if (n->train.B != B) {
    acts_free(n->train);
    if (!acts_alloc(n->train, B, true, err)) return -2;
  }
  Acts& a = n->train;
  float *tgt = nullptr, *mask = nullptr;
  std::vector<void*> tmp;
  if (!dev_alloc(tmp, (void**)&tgt, (size_t)B * 4096 * 16 * 4, err) ||
      !dev_alloc(tmp, (void**)&mask, (size_t)B * 4096 * 4, err)) {
    for (void* p : tmp) hipFree(p);
    return -2;
  }
  if (n->idxCap < iters * B) {
    if (n->idx) hipFree(n->idx);
    if (hipMalloc((void**)&n->idx, (size_t)iters * B * 4) != hipSuccess) { err = "hipMalloc idx"; return -2; }
    n->idxCap = iters * B;
  }
  if (n->lossCap < iters) {
    if (n->loss) hipFree(n->loss);
    if (hipMalloc((void**)&n->loss, (size_t)iters * 8) != hipSuccess) { err = "hipMalloc loss"; return -2; }
    n->lossCap = iters;
  }
  std::vector<int> idx((size_t)iters * B);
  uint32_t st = seed ? seed : 1u;
  for (auto& i : idx) {
    st ^= st << 13; st ^= st >> 17; st ^= st << 5;
    i = (int)(st % (uint32_t)n->count);
  }
  hipStream_t s = n->s;
  hipMemcpyAsync(n->idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, s);
  hipLaunchKernelGGL(k_zero_d, dim3(grid_of(iters)), dim3(256), 0, s, n->loss, iters);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int it = 0; it < iters; it++) {
    hipLaunchKernelGGL(k_gather, dim3(grid_of((size_t)B * 4096 * 16)), dim3(256), 0, s, n->dataIn, n->dataTgt,
                       n->dataMask, n->idx + (size_t)it * B, B, a.x0, tgt, mask);
    forward(n, a, B, true, nullptr, s);
    // the loss over the used channels of occupied blocks, per sample mean
    hipLaunchKernelGGL(k_ff_loss, dim3(grid_of((size_t)B * 4096 * 16)), dim3(256), 0, s, a.y, tgt, mask, B * 4096,
                       1.0f / (B * 4096.0f), a.gy, n->loss + it);
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_of(n->nw)), dim3(256), 0, s, n->g, n->nw);
    backward(n, a, B, s);
    n->adamT++;
    float c1 = 1.0f - powf(b1, (float)n->adamT), c2 = 1.0f - powf(b2, (float)n->adamT);
    hipLaunchKernelGGL(k_adam, dim3(grid_of(n->nw)), dim3(256), 0, s, n->w, n->g, n->m, n->v, (int)n->nw, lr, b1, b2,
                       c1, c2, 1e-8f);
    repack(n, s);
  }
  hipError_t e = hipMemcpyAsync(losses, n->loss, (size_t)iters * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (void* p : tmp) hipFree(p);
  if (e != hipSuccess) { err = hipGetErrorString(e); return -3; }
  return 0;
}



// forward over recorded samples [first, first+count) in batches of up to 16;
// optional outputs (count*4096*16 f32) and the error sums of k_eval_sum
static int ff_forward_samples(FFNet* n, int first, int count, float* out, double sums[4], std::string& err) {
  if (first < 0 || count <= 0 || first + count > n->count) { err = "sample range outside the record"; return -1; }
  const int B = count < 16 ? count : 16;
  if (n->train.B < B) {
    acts_free(n->train);
    n->train.B = 0;
    if (!acts_alloc(n->train, B, true, err)) return -2;
  }
  Acts& a = n->train;
  hipStream_t s = n->s;
  double* ds = nullptr;
  int* idx = nullptr;
  float *tgt = nullptr, *mask = nullptr;
  std::vector<void*> tmp;
  if (!dev_alloc(tmp, (void**)&ds, 32, err) || !dev_alloc(tmp, (void**)&idx, (size_t)count * 4, err) ||
      !dev_alloc(tmp, (void**)&tgt, (size_t)B * 4096 * 16 * 4, err) || !dev_alloc(tmp, (void**)&mask, (size_t)B * 4096 * 4, err)) {
    for (void* p : tmp) hipFree(p);
    return -2;
  }
  std::vector<int> ix(count);
  for (int k = 0; k < count; k++) ix[k] = first + k;
  hipMemcpyAsync(idx, ix.data(), (size_t)count * 4, hipMemcpyHostToDevice, s);
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, s, ds, 4);
  for (int k = 0; k < count; k += B) {
    int b = count - k < B ? count - k : B;
    hipLaunchKernelGGL(k_gather, dim3(grid_of((size_t)b * 4096 * 16)), dim3(256), 0, s, n->dataIn, n->dataTgt,
                       n->dataMask, idx + k, b, a.x0, tgt, mask);
    forward(n, a, b, false, nullptr, s);
    hipLaunchKernelGGL(k_eval_sum, dim3(grid_of((size_t)b * 4096 * 16)), dim3(256), 0, s, a.y, tgt, mask, b * 4096, ds);
    if (out) hipMemcpyAsync(out + (size_t)k * 4096 * 16, a.y, (size_t)b * 4096 * 16 * 4, hipMemcpyDeviceToHost, s);
  }
  hipError_t e = hipMemcpyAsync(sums, ds, 32, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (void* p : tmp) hipFree(p);
  if (e != hipSuccess) { err = hipGetErrorString(e); return -3; }
  return 0;
}

}  // namespace cc

using namespace cc;

struct cc_ff {
  FFNet* n = nullptr;
  std::string err;
};
static std::string g_fferr;

#define FFFAIL(f, code, msg)                \
  do {                                      \
    if (f) (f)->err = (msg); else g_fferr = (msg); \
    return code;                            \
  } while (0)

FFNet* ff_net(cc_ff* f) { return f ? f->n : nullptr; }

extern "C" {

const char* cc_ff_last_error(cc_ff* f) { return f ? f->err.c_str() : g_fferr.c_str(); }

int cc_ff_create(int capacity, cc_ff** out) {
  if (!out || capacity < 0) return CC_EINVAL;
  *out = nullptr;
  std::string err;
  FFNet* n = ff_create(capacity, err);
  if (!n) FFFAIL((cc_ff*)nullptr, CC_EDEVICE, "cc_ff_create: " + err);
  cc_ff* f = new cc_ff();
  f->n = n;
  *out = f;
  return 0;
}

int cc_ff_destroy(cc_ff* f) {
  if (!f) return 0;
  ff_destroy(f->n);
  delete f;
  return 0;
}

int cc_ff_weight_count(int64_t* n) {
  if (!n) return CC_EINVAL;
  *n = (int64_t)ff_weight_count();
  return 0;
}

int cc_ff_get_weights(cc_ff* f, float* w, int64_t n) {
  if (!f || !w || n != (int64_t)f->n->nw) FFFAIL(f, CC_EINVAL, "weight count mismatch");
  hipError_t e = hipMemcpyAsync(w, f->n->w, (size_t)n * 4, hipMemcpyDeviceToHost, f->n->s);
  if (e == hipSuccess) e = hipStreamSynchronize(f->n->s);
  if (e != hipSuccess) FFFAIL(f, CC_EDEVICE, hipGetErrorString(e));
  return 0;
}

int cc_ff_set_weights(cc_ff* f, const float* w, int64_t n) {
  if (!f || !w || n != (int64_t)f->n->nw) FFFAIL(f, CC_EINVAL, "weight count mismatch");
  hipError_t e = hipMemcpyAsync(f->n->w, w, (size_t)n * 4, hipMemcpyHostToDevice, f->n->s);
  if (e == hipSuccess) e = hipMemsetAsync(f->n->m, 0, (size_t)n * 4, f->n->s);
  if (e == hipSuccess) e = hipMemsetAsync(f->n->v, 0, (size_t)n * 4, f->n->s);
  repack(f->n, f->n->s);
  if (e == hipSuccess) e = hipStreamSynchronize(f->n->s);
  if (e != hipSuccess) FFFAIL(f, CC_EDEVICE, hipGetErrorString(e));
  f->n->adamT = 0;
  return 0;
}

int cc_ff_set_slope(cc_ff* f, double slope) {
  if (!f || !(slope >= 0.0 && slope <= 1.0)) FFFAIL(f, CC_EINVAL, "the slope must be in [0, 1]");
  f->n->slope = (float)slope;
  return 0;
}

int cc_ff_samples(cc_ff* f, int* count) {
  if (!f || !count) return CC_EINVAL;
  *count = f->n->count;
  return 0;
}

int cc_ff_clear(cc_ff* f) {
  if (!f) return CC_EINVAL;
  f->n->count = 0;
  return 0;
}

int cc_ff_passes(cc_ff* f, int* count) {
  if (!f || !count) return CC_EINVAL;
  *count = f->n->passes;
  return 0;
}

int cc_ff_clear_passes(cc_ff* f) {
  if (!f) return CC_EINVAL;
  f->n->passes = 0;
  return 0;
}

int cc_ff_pair(cc_ff* f, int base, int cur, int* sample) {
  if (!f || base < 0 || cur < 0 || base >= f->n->passes || cur >= f->n->passes)
    FFFAIL(f, CC_EINVAL, "cc_ff_pair: pass index outside the record");
  int k = ff_pair(f->n, base, cur);
  if (k < 0) FFFAIL(f, CC_ESTATE, "cc_ff_pair: the sample record is full");
  hipError_t e = hipStreamSynchronize(f->n->s);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) FFFAIL(f, CC_EDEVICE, hipGetErrorString(e));
  if (sample) *sample = k;
  return 0;
}

int cc_ff_read_pass(cc_ff* f, int k, float* feat, float* meta, float* record) {
  if (!f || k < 0 || k >= f->n->passes) return CC_EINVAL;
  FFNet* n = f->n;
  hipError_t e = hipSuccess;
  if (feat) e = hipMemcpyAsync(feat, n->passFeat + (size_t)k * 4096, 4096 * 16, hipMemcpyDeviceToHost, n->s);
  if (e == hipSuccess && meta) e = hipMemcpyAsync(meta, n->passMeta + (size_t)k * 4, 16, hipMemcpyDeviceToHost, n->s);
  if (e == hipSuccess && record)
    e = hipMemcpyAsync(record, n->passR + (size_t)k * 4096 * FF_R, (size_t)4096 * FF_R * 4, hipMemcpyDeviceToHost, n->s);
  if (e == hipSuccess) e = hipStreamSynchronize(n->s);
  if (e != hipSuccess) FFFAIL(f, CC_EDEVICE, hipGetErrorString(e));
  return 0;
}

int cc_ff_push_sample(cc_ff* f, const float* in, const float* target, const float* mask) {
  if (!f || !in || !target || !mask) return CC_EINVAL;
  FFNet* n = f->n;
  if (n->count >= n->cap) FFFAIL(f, CC_EINVAL, "the sample record is full");
  std::vector<uint16_t> h((size_t)4096 * 16);
  for (size_t k = 0; k < h.size(); k++) {   // round to nearest even, as the device does
    uint32_t u;
    std::memcpy(&u, &in[k], 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    h[k] = (uint16_t)(u >> 16);
  }
  int k = n->count;
  hipError_t e = hipMemcpyAsync(n->dataIn + (size_t)k * 4096 * 16, h.data(), h.size() * 2, hipMemcpyHostToDevice, n->s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(n->dataTgt + (size_t)k * 4096 * 16, target, (size_t)4096 * 16 * 4, hipMemcpyHostToDevice, n->s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(n->dataMask + (size_t)k * 4096, mask, (size_t)4096 * 4, hipMemcpyHostToDevice, n->s);
  if (e == hipSuccess) e = hipStreamSynchronize(n->s);
  if (e != hipSuccess) FFFAIL(f, CC_EDEVICE, hipGetErrorString(e));
  n->count++;
  return 0;
}

int cc_ff_read_sample(cc_ff* f, int k, float* in, float* target, float* mask) {
  if (!f || k < 0 || k >= f->n->count) return CC_EINVAL;
  FFNet* n = f->n;
  std::vector<uint16_t> h((size_t)4096 * 16);
  hipError_t e = hipMemcpyAsync(h.data(), n->dataIn + (size_t)k * 4096 * 16, h.size() * 2, hipMemcpyDeviceToHost, n->s);
  if (e == hipSuccess && target)
    e = hipMemcpyAsync(target, n->dataTgt + (size_t)k * 4096 * 16, (size_t)4096 * 16 * 4, hipMemcpyDeviceToHost, n->s);
  if (e == hipSuccess && mask)
    e = hipMemcpyAsync(mask, n->dataMask + (size_t)k * 4096, (size_t)4096 * 4, hipMemcpyDeviceToHost, n->s);
  if (e == hipSuccess) e = hipStreamSynchronize(n->s);
  if (e != hipSuccess) FFFAIL(f, CC_EDEVICE, hipGetErrorString(e));
  if (in)
    for (size_t q = 0; q < h.size(); q++) {
      uint32_t u = (uint32_t)h[q] << 16;
      std::memcpy(&in[q], &u, 4);
    }
  return 0;
}

int cc_ff_train(cc_ff* f, int iters, int batch, double lr, uint32_t seed, double* losses) {
  if (!f || iters <= 0 || batch <= 0 || batch > 64 || !losses) FFFAIL(f, CC_EINVAL, "bad training arguments");
  std::string err;
  int r = ff_train(f->n, iters, batch, (float)lr, seed, losses, err);
  if (r) FFFAIL(f, r == -1 ? CC_ESTATE : CC_EDEVICE, "cc_ff_train: " + err);
  return 0;
}

int cc_ff_eval(cc_ff* f, int first, int count, float* out, double* relAcc, double* relAll) {
  if (!f) return CC_EINVAL;
  std::string err;
  double sums[4] = { 0, 0, 0, 0 };
  int r = ff_forward_samples(f->n, first, count, out, sums, err);
  if (r) FFFAIL(f, r == -1 ? CC_EINVAL : CC_EDEVICE, "cc_ff_eval: " + err);
  if (relAcc) *relAcc = sums[1] > 0 ? sqrt(sums[0] / sums[1]) : 0.0;
  if (relAll) *relAll = sums[3] > 0 ? sqrt(sums[2] / sums[3]) : 0.0;
  return 0;
}

// the loss and the full weight gradient on samples [first, first+count) at
// the present weights (no update)
int cc_ff_gradient(cc_ff* f, int first, int count, float* grad, double* loss) {
  if (!f || !grad || !loss || first < 0 || count <= 0 || count > 64 || first + count > f->n->count)
    FFFAIL(f, CC_EINVAL, "bad gradient arguments");
  FFNet* n = f->n;
  std::string err;
  if (n->train.B != count) {
    acts_free(n->train);
    n->train.B = 0;
    if (!acts_alloc(n->train, count, true, err)) FFFAIL(f, CC_EDEVICE, err);
  }
  Acts& a = n->train;
  hipStream_t s = n->s;
  std::vector<void*> tmp;
  int* idx = nullptr;
  float *tgt = nullptr, *mask = nullptr;
  double* l = nullptr;
  if (!dev_alloc(tmp, (void**)&idx, (size_t)count * 4, err) || !dev_alloc(tmp, (void**)&tgt, (size_t)count * 4096 * 64, err) ||
      !dev_alloc(tmp, (void**)&mask, (size_t)count * 4096 * 4, err) || !dev_alloc(tmp, (void**)&l, 8, err)) {
    for (void* p : tmp) hipFree(p);
    FFFAIL(f, CC_EDEVICE, err);
  }
  std::vector<int> ix(count);
  for (int k = 0; k < count; k++) ix[k] = first + k;
  hipMemcpyAsync(idx, ix.data(), (size_t)count * 4, hipMemcpyHostToDevice, s);
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, s, l, 1);
  hipLaunchKernelGGL(k_gather, dim3(grid_of((size_t)count * 4096 * 16)), dim3(256), 0, s, n->dataIn, n->dataTgt,
                     n->dataMask, idx, count, a.x0, tgt, mask);
  forward(n, a, count, true, nullptr, s);
  hipLaunchKernelGGL(k_ff_loss, dim3(grid_of((size_t)count * 4096 * 16)), dim3(256), 0, s, a.y, tgt, mask,
                     count * 4096, 1.0f / (count * 4096.0f), a.gy, l);
  hipLaunchKernelGGL(k_zero_f32, dim3(grid_of(n->nw)), dim3(256), 0, s, n->g, n->nw);
  backward(n, a, count, s);
  hipError_t e = hipMemcpyAsync(grad, n->g, n->nw * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(loss, l, 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (void* p : tmp) hipFree(p);
  if (e != hipSuccess) FFFAIL(f, CC_EDEVICE, hipGetErrorString(e));
  return 0;
}

}  // extern "C"
