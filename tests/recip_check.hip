// Exhaustive and random checks, on the GPU, of the correctly rounded FP32
// helpers the step kernels use (cosmic-collisions_amd/csrc/cc_fpmath.h,
// included here so the check runs the library's own code).
//
// A result y of an operation with exact value v is the correctly rounded
// (round to nearest) FP32 result iff v lies strictly between the midpoints
// from y to its two neighbours (no tie can occur for a quotient or root of
// FP32 operands).  The comparisons are made in double, where the midpoints
// (25 significant bits) times a divisor, or squared, are exact (at most 50
// bits), so the check does not rest on the GPU's own double division or
// root.
//   qrcp   every significand at every exponent in [-100, 100]
//   qdiv   2^33 random signed pairs, exponents in [-30, 30], then [-60, 60]
//   qsqrt  every x in [2^-100, FLT_MAX]
// As controls, the raw hardware reciprocal (one ulp) must fail the same check
// for a share of the significands, and qsqrt is reported below 2^-100, where
// its residual underflows and it is not meant to be used.
// Exit status 0 iff every helper passes and the control fails.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../cosmic-collisions_amd/csrc/cc_fpmath.h"

// the midpoints between |y| and its neighbours below and above, in double
__device__ __forceinline__ void mids(float y, double& lo, double& hi) {
  const uint32_t u = __float_as_uint(fabsf(y));
  const double a = (double)__uint_as_float(u), below = (double)__uint_as_float(u - 1u),
               above = (double)__uint_as_float(u + 1u);
  lo = 0.5 * (a + below);
  hi = 0.5 * (a + above);
}
// is y the correctly rounded n / d (signed operands, d != 0)?
__device__ __forceinline__ bool rounds_quotient(float y, float n, float d) {
  if ((__float_as_uint(y) ^ __float_as_uint(n) ^ __float_as_uint(d)) & 0x80000000u) return false;   // sign
  double lo, hi;
  mids(y, lo, hi);
  const double an = fabs((double)n), ad = fabs((double)d);
  return ad * lo < an && an < ad * hi;
}
__device__ __forceinline__ bool rounds_root(float y, float x) {
  if (y < 0.0f) return false;
  double lo, hi;
  mids(y, lo, hi);
  return lo * lo < (double)x && (double)x < hi * hi;
}

// one thread per significand of 2^e (the grid covers exactly 2^23 threads)
__global__ void k_recip(int e, int raw, unsigned long long* bad, uint32_t* first) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= (1u << 23)) return;
  const float b = __uint_as_float(((uint32_t)(e + 127) << 23) | m);
  const float y = raw ? __builtin_amdgcn_rcpf(b) : cc::qrcp(b);
  if (!rounds_quotient(y, 1.0f, b)) {
    const unsigned long long k = atomicAdd(bad, 1ull);
    if (k < 4) first[k] = __float_as_uint(b);
  }
}

__device__ __forceinline__ uint32_t mix(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
  return (uint32_t)x;
}
__device__ __forceinline__ float draw(uint32_t r, int span) {
  const int e = (int)((r >> 23) % (uint32_t)(2 * span + 1)) - span;
  return __uint_as_float((r & 0x80000000u) | ((uint32_t)(e + 127) << 23) | (r & 0x7fffffu));
}
// 64 pairs a thread
__global__ void k_div(uint64_t seed, int span, unsigned long long* bad, uint32_t* first) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int it = 0; it < 64; it++) {
    const uint64_t id = t * 64 + it;
    const float a = draw(mix(seed ^ (2 * id + 1)), span), b = draw(mix(seed ^ (2 * id + 2) ^ 0x9e3779b97f4a7c15ull), span);
    if (!rounds_quotient(cc::qdiv(a, b), a, b)) {
      const unsigned long long k = atomicAdd(bad, 1ull);
      if (k < 2) { first[2 * k] = __float_as_uint(a); first[2 * k + 1] = __float_as_uint(b); }
    }
  }
}

// x over the normal floats above 0: bit patterns base + i (the grid covers 2^23 threads x 254 exponents)
__global__ void k_sqrt(uint32_t base, unsigned long long* bad, uint32_t* first) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << 23)) return;
  const float x = __uint_as_float(base + i);
  if (!rounds_root(cc::qsqrt(x), x)) {
    const unsigned long long k = atomicAdd(bad, 1ull);
    if (k < 4) first[k] = __float_as_uint(x);
  }
}

#define CK(x)                                                                          \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);           \
      return 2;                                                                        \
    }                                                                                  \
  } while (0)

int main() {
  unsigned long long* bad;
  uint32_t* first;
  CK(hipMalloc(&bad, 8));
  CK(hipMalloc(&first, 16 * 4));
  unsigned long long h;
  uint32_t f[16];
  bool ok = true;
  for (int raw = 0; raw < 2; raw++) {
    unsigned long long total = 0;
    for (int e = -100; e <= 100; e++) {
      CK(hipMemset(bad, 0, 8));
      hipLaunchKernelGGL(k_recip, dim3((1u << 23) / 256), dim3(256), 0, 0, e, raw, bad, first);
      CK(hipGetLastError());
      CK(hipMemcpy(&h, bad, 8, hipMemcpyDeviceToHost));
      CK(hipMemcpy(f, first, sizeof f, hipMemcpyDeviceToHost));
      if (h && !raw) printf("qrcp exponent %d: %llu differ (first b = %08x)\n", e, h, f[0]);
      total += h;
    }
    printf("%s: %llu of %llu reciprocals not correctly rounded (every significand, exponents -100..100)\n",
           raw ? "hardware reciprocal alone (control)" : "qrcp", total, 201ull << 23);
    ok = ok && (raw ? total > 0 : total == 0);
  }
  for (int span : {30, 60}) {
    unsigned long long total = 0;
    for (int rep = 0; rep < 8; rep++) {
      CK(hipMemset(bad, 0, 8));
      hipLaunchKernelGGL(k_div, dim3(1 << 16), dim3(256), 0, 0, (uint64_t)(0x5eedull * (rep + 1) + span), span, bad,
                         first);
      CK(hipGetLastError());
      CK(hipMemcpy(&h, bad, 8, hipMemcpyDeviceToHost));
      CK(hipMemcpy(f, first, sizeof f, hipMemcpyDeviceToHost));
      if (h) printf("qdiv: %llu differ (first a, b = %08x %08x)\n", h, f[0], f[1]);
      total += h;
    }
    printf("qdiv: %llu of %llu random quotients not correctly rounded (exponents -%d..%d)\n", total,
           8ull * (1ull << 24) * 64, span, span);
    ok = ok && total == 0;
  }
  {
    unsigned long long total = 0, tiny = 0;
    for (uint32_t ex = 1; ex <= 254; ex++) {
      CK(hipMemset(bad, 0, 8));
      hipLaunchKernelGGL(k_sqrt, dim3((1u << 23) / 256), dim3(256), 0, 0, ex << 23, bad, first);
      CK(hipGetLastError());
      CK(hipMemcpy(&h, bad, 8, hipMemcpyDeviceToHost));
      CK(hipMemcpy(f, first, sizeof f, hipMemcpyDeviceToHost));
      if (ex < 27) { tiny += h; continue; }   // x below 2^-100
      if (h) printf("qsqrt exponent field %u: %llu differ (first x = %08x)\n", ex, h, f[0]);
      total += h;
    }
    printf("qsqrt: %llu of %llu roots not correctly rounded (every x in [2^-100, FLT_MAX])\n", total, 228ull << 23);
    printf("qsqrt below 2^-100 (outside its domain, shown only): %llu of %llu not correctly rounded\n", tiny,
           26ull << 23);
    ok = ok && total == 0;
  }
  CK(hipFree(bad));
  CK(hipFree(first));
  printf(ok ? "all correctly rounded\n" : "FAILED\n");
  return ok ? 0 : 1;
}
